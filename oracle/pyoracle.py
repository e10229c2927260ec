"""ctypes wrapper of oracle/liboracle.so -- the CPU checker (TEST ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product path never does.  Function-level citations of the
reference live in oracle.h / oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_num_nodes.restype = ctypes.c_size_t
        L.orc_tree_num_nodes.argtypes = [ctypes.c_void_p]
        L.orc_tree_depth.argtypes = [ctypes.c_void_p]
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_nodes.argtypes = [ctypes.c_void_p] + [_i32p] * 10
        L.orc_walk.restype = ctypes.c_void_p
        L.orc_walk.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_lists_free.argtypes = [ctypes.c_void_p]
        L.orc_lists_sizes.argtypes = [ctypes.c_void_p, _i64p]
        L.orc_lists_get.argtypes = [ctypes.c_void_p, _i64p, _i32p, _i64p, _i32p]
        L.orc_fmm.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                              _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        _LIB = L
    return _LIB


_REF_O0 = None


def ref_lib(as_shipped=False):
    """The reference's own NaiveSimulation (oracle/_ref), or None if not built:
    built -O2 -ffp-contract=off, or as_shipped (-O0, the reference's build)."""
    global _REF, _REF_O0
    lib_ = _REF_O0 if as_shipped else _REF
    if lib_ is None:
        path = os.path.join(_HERE, "_ref", "libnaive_ref_O0.so" if as_shipped else "libnaive_ref.so")
        if not os.path.exists(path):
            return None
        lib_ = ctypes.CDLL(path)
        if as_shipped:
            _REF_O0 = lib_
        else:
            _REF = lib_
    return lib_


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def direct(xs, qs, xt=None, self_idx=None, G=1.0, eps=0.0):
    """FP64 softened direct sum: returns (acc[nt,3] without q_i/m_i, pot[nt])."""
    xs = c64(xs).reshape(-1, 3)
    qs = c64(qs)
    if xt is None:
        xt = xs
        if self_idx is None:
            self_idx = np.arange(len(xs), dtype=np.int64)
    xt = c64(xt).reshape(-1, 3)
    acc = np.zeros((len(xt), 3))
    pot = np.zeros(len(xt))
    si = None if self_idx is None else np.ascontiguousarray(self_idx, dtype=np.int64)
    lib().orc_direct(ctypes.c_size_t(len(xs)), _ptr(xs, _dp), _ptr(qs, _dp), ctypes.c_size_t(len(xt)),
                     _ptr(xt, _dp), _ptr(si, _i64p), ctypes.c_double(G), ctypes.c_double(eps),
                     _ptr(acc, _dp), _ptr(pot, _dp))
    return acc, pot


def direct_vjp(x, m, w, G=1.0, eps=0.0):
    x = c64(x).reshape(-1, 3)
    m = c64(m)
    w = c64(w).reshape(-1, 3)
    gx = np.zeros_like(x)
    gm = np.zeros_like(m)
    lib().orc_direct_vjp(ctypes.c_size_t(len(x)), _ptr(x, _dp), _ptr(m, _dp), _ptr(w, _dp),
                         ctypes.c_double(G), ctypes.c_double(eps), _ptr(gx, _dp), _ptr(gm, _dp))
    return gx, gm


def morton(x, sys_off=None, bits=21):
    x = c64(x).reshape(-1, 3)
    n = len(x)
    if sys_off is None:
        sys_off = np.array([0, n], dtype=np.uint32)
    sys_off = np.ascontiguousarray(sys_off, dtype=np.uint32)
    nsys = len(sys_off) - 1
    keys = np.zeros(n, dtype=np.uint64)
    box = np.zeros(4 * nsys)
    lib().orc_morton(ctypes.c_size_t(n), _ptr(x, _dp), ctypes.c_size_t(nsys), _ptr(sys_off, _u32p),
                     ctypes.c_int(bits), _ptr(keys, _u64p), _ptr(box, _dp))
    return keys, box.reshape(nsys, 4)


@dataclass
class Tree:
    handle: int
    level: np.ndarray
    sys: np.ndarray
    ix: np.ndarray
    iy: np.ndarray
    iz: np.ndarray
    begin: np.ndarray
    count: np.ndarray
    first_child: np.ndarray
    nchild: np.ndarray
    parent: np.ndarray
    depth: int

    def __del__(self):
        if self.handle and _LIB is not None:
            _LIB.orc_tree_free(ctypes.c_void_p(self.handle))
            self.handle = 0


def build_tree(sorted_keys, sys_off=None, bits=21, cap=8, max_depth=21):
    keys = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    n = len(keys)
    if sys_off is None:
        sys_off = np.array([0, n], dtype=np.uint32)
    sys_off = np.ascontiguousarray(sys_off, dtype=np.uint32)
    L = lib()
    h = L.orc_tree_build(ctypes.c_size_t(n), _ptr(keys, _u64p), ctypes.c_size_t(len(sys_off) - 1),
                         _ptr(sys_off, _u32p), ctypes.c_int(bits), ctypes.c_int(cap),
                         ctypes.c_int(max_depth))
    nn = L.orc_tree_num_nodes(ctypes.c_void_p(h))
    arrs = [np.zeros(nn, dtype=np.int32) for _ in range(10)]
    L.orc_tree_nodes(ctypes.c_void_p(h), *[_ptr(a, _i32p) for a in arrs])
    return Tree(h, *arrs, depth=L.orc_tree_depth(ctypes.c_void_p(h)))


@dataclass
class Lists:
    handle: int
    m2l_off: np.ndarray
    m2l_src: np.ndarray
    p2p_off: np.ndarray
    p2p_src: np.ndarray
    unordered_m2l: int
    unordered_p2p: int
    visited: int

    def __del__(self):
        if self.handle and _LIB is not None:
            _LIB.orc_lists_free(ctypes.c_void_p(self.handle))
            self.handle = 0


def walk(tree: Tree, theta_num: int, theta_den: int):
    L = lib()
    h = L.orc_walk(ctypes.c_void_p(tree.handle), theta_num, theta_den)
    sizes = np.zeros(5, dtype=np.int64)
    L.orc_lists_sizes(ctypes.c_void_p(h), _ptr(sizes, _i64p))
    nn = len(tree.level)
    m2l_off = np.zeros(nn + 1, dtype=np.int64)
    p2p_off = np.zeros(nn + 1, dtype=np.int64)
    m2l_src = np.zeros(sizes[0], dtype=np.int32)
    p2p_src = np.zeros(sizes[1], dtype=np.int32)
    L.orc_lists_get(ctypes.c_void_p(h), _ptr(m2l_off, _i64p), _ptr(m2l_src, _i32p),
                    _ptr(p2p_off, _i64p), _ptr(p2p_src, _i32p))
    return Lists(h, m2l_off, m2l_src, p2p_off, p2p_src, int(sizes[2]), int(sizes[3]), int(sizes[4]))


def ncoef(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def multi_indices(p):
    mi = np.zeros((ncoef(p), 3), dtype=np.int32)
    lib().orc_multi_indices(ctypes.c_int(p), _ptr(mi, _i32p))
    return mi


def derivs(p, R, eps=0.0):
    R = c64(R)
    D = np.zeros(ncoef(p))
    lib().orc_derivs(ctypes.c_int(p), _ptr(R, _dp), ctypes.c_double(eps), _ptr(D, _dp))
    return D


def p2m(p, z, x, q):
    z = c64(z); x = c64(x).reshape(-1, 3); q = c64(q)
    M = np.zeros(ncoef(p))
    lib().orc_p2m(ctypes.c_int(p), _ptr(z, _dp), ctypes.c_size_t(len(x)), _ptr(x, _dp), _ptr(q, _dp), _ptr(M, _dp))
    return M


def m2m(p, zc, zp, Mc):
    zc = c64(zc); zp = c64(zp); Mc = c64(Mc)
    Mp = np.zeros(ncoef(p))
    lib().orc_m2m(ctypes.c_int(p), _ptr(zc, _dp), _ptr(zp, _dp), _ptr(Mc, _dp), _ptr(Mp, _dp))
    return Mp


def _m2l_code(q, m2l_terms):
    """orc_m2l's truncation code: q | 16 for every term of degree <= q
    ("all"), q for the low-degree set ("low")."""
    if m2l_terms not in ("all", "low"):
        raise ValueError(f"m2l_terms must be 'all' or 'low', not {m2l_terms!r}")
    return q | (16 if m2l_terms == "all" else 0)


def m2l(p, zA, zB, M, eps=0.0, m2l_order=None, m2l_terms="all"):
    """m2l_order q (default p): keep every term with |k| + |n| <= q
    (m2l_terms "all"), or |k| + |n| <= p and up to |k| + |n| <= q only the
    terms with min(|k|, |n|) <= q - p ("low")."""
    zA = c64(zA); zB = c64(zB); M = c64(M)
    L = np.zeros(ncoef(p))
    q = _m2l_code(p if m2l_order is None else int(m2l_order), m2l_terms)
    lib().orc_m2l(ctypes.c_int(p), ctypes.c_int(q), _ptr(zA, _dp), _ptr(zB, _dp), ctypes.c_double(eps), _ptr(M, _dp), _ptr(L, _dp))
    return L


def l2l(p, zp, zc, Lp):
    zp = c64(zp); zc = c64(zc); Lp = c64(Lp)
    Lc = np.zeros(ncoef(p))
    lib().orc_l2l(ctypes.c_int(p), _ptr(zp, _dp), _ptr(zc, _dp), _ptr(Lp, _dp), _ptr(Lc, _dp))
    return Lc


def l2p(p, z, L, x):
    z = c64(z); L = c64(L); x = c64(x)
    phi = ctypes.c_double(0)
    g = np.zeros(3)
    lib().orc_l2p(ctypes.c_int(p), _ptr(z, _dp), _ptr(L, _dp), _ptr(x, _dp), ctypes.byref(phi), _ptr(g, _dp))
    return phi.value, g


def _qcode(p, m2l_order, same_level_order, m2l_terms="all"):
    q = _m2l_code(p if m2l_order is None else int(m2l_order), m2l_terms)
    qs = _m2l_code(int(same_level_order), m2l_terms) if same_level_order else 0
    return q | (qs << 8)


def fmm(tree: Tree, lists: Lists, p, x_sorted, q_sorted, box, G=1.0, eps=0.0, m2l_order=None,
        same_level_order=None, m2l_terms="all"):
    """m2l_order, m2l_terms: the M2L truncation (m2l); same_level_order: that
    of the pairs of one level (None: m2l_order); 2p keeps every term |k|,
    |n| <= p (the dense operator)."""
    x = c64(x_sorted).reshape(-1, 3)
    q = c64(q_sorted)
    box = c64(box).reshape(-1)
    acc = np.zeros_like(x)
    pot = np.zeros(len(x))
    lib().orc_fmm(ctypes.c_void_p(tree.handle), ctypes.c_void_p(lists.handle), p,
                  _qcode(p, m2l_order, same_level_order, m2l_terms), len(x),
                  _ptr(x, _dp), _ptr(q, _dp), _ptr(box, _dp), G, eps, _ptr(acc, _dp), _ptr(pot, _dp))
    return acc, pot


def fmm_vjp(tree: Tree, lists: Lists, p, x_sorted, m_sorted, w_sorted, box, G=1.0, eps=0.0, m2l_order=None,
            same_level_order=None, m2l_terms="all"):
    x = c64(x_sorted).reshape(-1, 3)
    m = c64(m_sorted)
    w = c64(w_sorted).reshape(-1, 3)
    box = c64(box).reshape(-1)
    gx = np.zeros_like(x)
    gm = np.zeros(len(x))
    L = lib()
    L.orc_fmm_vjp.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                              _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.orc_fmm_vjp(ctypes.c_void_p(tree.handle), ctypes.c_void_p(lists.handle), p,
                  _qcode(p, m2l_order, same_level_order, m2l_terms), len(x), _ptr(x, _dp),
                  _ptr(m, _dp), _ptr(w, _dp), _ptr(box, _dp), G, eps, _ptr(gx, _dp), _ptr(gm, _dp))
    return gx, gm


def integrate_direct(x, v, m, q, G, eps, dt, nsteps, integrator=0):
    x = c64(x).reshape(-1, 3).copy(); v = c64(v).reshape(-1, 3).copy()
    m = c64(m); q = c64(q)
    lib().orc_integrate_direct(ctypes.c_size_t(len(x)), _ptr(x, _dp), _ptr(v, _dp), _ptr(m, _dp), _ptr(q, _dp),
                               ctypes.c_double(G), ctypes.c_double(eps), ctypes.c_double(dt),
                               ctypes.c_int(nsteps), ctypes.c_int(integrator))
    return x, v


def particles_to_p12(x, v, m, q):
    n = len(m)
    P = np.zeros((n, 12), dtype=np.float32)
    P[:, 0:3] = x
    P[:, 4:7] = v
    P[:, 8] = m
    P[:, 9] = q
    return P


def naive_compat(P12, force_constant, dt, nsteps):
    """Oracle FP32 restatement of NaiveSimulation::step (bug for bug)."""
    P = np.ascontiguousarray(P12, dtype=np.float32).copy()
    t = ctypes.c_float(0)
    lib().orc_naive_compat_steps(ctypes.c_size_t(len(P)), _ptr(P, _fp), ctypes.c_float(force_constant),
                                 ctypes.c_float(dt), ctypes.c_int(nsteps), ctypes.byref(t))
    return P, t.value


def naive_reference(P12, force_constant, dt, nsteps, as_shipped=False):
    """The reference's own NaiveSimulation (compiled from /root/reference;
    as_shipped: the -O0 build)."""
    R = ref_lib(as_shipped)
    if R is None:
        return None
    P = np.ascontiguousarray(P12, dtype=np.float32).copy()
    t = ctypes.c_float(0)
    R.ref_naive_run(ctypes.c_size_t(len(P)), _ptr(P, _fp), ctypes.c_float(force_constant),
                    ctypes.c_float(dt), ctypes.c_int(nsteps), ctypes.byref(t))
    return P, t.value
