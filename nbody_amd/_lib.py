"""ctypes binding of libnbody_hip.so (the C ABI declared in include/nbody_hip.h).

The library is built in-tree (nbody_amd/libnbody_hip.so); importing this
module never falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbody_hip.so")

NBH_OK = 0
NBH_E_INVALID = -1
NBH_E_HIP = -2
NBH_E_NOMEM = -3
NBH_E_NONFINITE = -4
NBH_E_INTERNAL = -5
NBH_E_OUTSIDE = -6
NOALLOC_MAX_TARGETS = 65536

SOLVER_FMM, SOLVER_DIRECT = 0, 1
M2L_PLAIN, M2L_GROUPED, M2L_BF16X3, M2L_MFMA_F32 = 0, 1, 2, 3
M2L_TERMS_ALL, M2L_TERMS_LOW = 0, 1
P2P_LEAF, P2P_LANE = 0, 1
INT_KDK, INT_EULER, INT_KICK_DRIFT = 0, 1, 2
LAW_SOFTENED, LAW_COMPAT_NAIVE = 0, 1


class Config(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("solver", ctypes.c_uint32),
        ("G", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("leaf_capacity", ctypes.c_uint32),
        ("order", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32),
        ("m2l", ctypes.c_uint32),
        ("integrator", ctypes.c_uint32),
        ("force_law", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("fixed_root", ctypes.c_uint32),
        ("root_lo", ctypes.c_double * 3),
        ("root_size", ctypes.c_double),
        ("p2p", ctypes.c_uint32),
        ("m2l_order", ctypes.c_uint32),
        ("m2l_terms", ctypes.c_uint32),
        ("p2p_chunk", ctypes.c_uint32),
        ("graph", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_uint64),
        ("num_leaves", ctypes.c_uint64),
        ("depth", ctypes.c_uint64),
        ("m2l_pairs", ctypes.c_uint64),
        ("p2p_pairs", ctypes.c_uint64),
        ("p2p_interactions", ctypes.c_uint64),
        ("split_pairs", ctypes.c_uint64),
        ("ms_tree", ctypes.c_double),
        ("ms_moments", ctypes.c_double),
        ("ms_walk", ctypes.c_double),
        ("ms_m2l", ctypes.c_double),
        ("ms_l2l", ctypes.c_double),
        ("ms_p2p", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("m2l_same_order", ctypes.c_uint64),
        ("m2l_same_pairs", ctypes.c_uint64),
        ("m2l_same_launches", ctypes.c_uint64),
        ("ms_m2l_same", ctypes.c_double),
        ("graph_captures", ctypes.c_uint64),
        ("graph_replays", ctypes.c_uint64),
        ("graph_fallbacks", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("leaf_capacity", ctypes.c_uint32),
        ("p2p_chunk", ctypes.c_uint32),
        ("ms_per_eval", ctypes.c_double),
        ("tried", ctypes.c_uint32),
    ]


_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t

# name -> (restype, argtypes); every symbol declared in include/nbody_hip.h
SIGNATURES = {
    "nbh_abi_version": (ctypes.c_int, []),
    "nbh_config_size": (ctypes.c_size_t, []),
    "nbh_config_default": (ctypes.c_int, [ctypes.POINTER(Config)]),
    "nbh_device_count": (ctypes.c_int, []),
    "nbh_create": (ctypes.c_int, [ctypes.POINTER(Config), _sz, _dp, _dp, _dp, _dp, ctypes.POINTER(_vp)]),
    "nbh_create_ensemble": (ctypes.c_int, [ctypes.POINTER(Config), _sz, _u32p, _dp, _dp, _dp, _dp,
                                           ctypes.POINTER(_vp)]),
    "nbh_destroy": (None, [_vp]),
    "nbh_last_error": (ctypes.c_char_p, [_vp]),
    "nbh_step": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "nbh_get_state": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "nbh_set_state": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_double]),
    "nbh_compute_forces": (ctypes.c_int, [_vp, _fp, _fp]),
    "nbh_autotune": (ctypes.c_int, [_vp, _u32p, _sz, _u32p, _sz, ctypes.c_int, ctypes.POINTER(TuneResult), _dp]),
    "nbh_direct_fp64": (ctypes.c_int, [_vp, _sz, _i64p, _dp, _dp]),
    "nbh_direct_fp64_noalloc": (ctypes.c_int, [_vp, _sz, _i64p, _dp, _dp]),
    "nbh_diagnostics": (ctypes.c_int, [_vp, _dp]),
    "nbh_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "nbh_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nbh_force_vjp": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "nbh_adjoint_forward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "nbh_adjoint_backward": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "nbh_debug_sizes": (ctypes.c_int, [_vp, _i64p]),
    "nbh_debug_keys": (ctypes.c_int, [_vp, _u64p, _u32p, _dp]),
    "nbh_debug_tree": (ctypes.c_int, [_vp] + [_i32p] * 10),
    "nbh_debug_lists": (ctypes.c_int, [_vp, _i64p, _i32p, _i64p, _i32p]),
    "nbh_debug_expansions": (ctypes.c_int, [_vp, _fp, _fp]),
    "nbh_ic_uniform": (ctypes.c_int, [ctypes.c_uint64, _sz, _dp, _dp, _dp]),
    "nbh_ic_plummer": (ctypes.c_int, [ctypes.c_uint64, _sz, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _dp, _dp, _dp]),
}

_LIB = None


def lib():
    """Load libnbody_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                               " or `make -C nbody_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def default_config(**overrides) -> Config:
    c = Config()
    rc = lib().nbh_config_default(ctypes.byref(c))
    if rc != NBH_OK:
        raise RuntimeError("nbh_config_default failed")
    for k, v in overrides.items():
        if k == "root_lo":
            for i in range(3):
                c.root_lo[i] = v[i]
        else:
            if not hasattr(c, k):
                raise AttributeError(f"unknown config field {k}")
            setattr(c, k, v)
    return c
