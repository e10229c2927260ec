"""Python mirror of the reference's simulation interface over libnbody_hip.

The reference interface is C++ (include/nbody/simulation.h:6-37):
``Simulation<Scalar, Vector>`` with ``struct Particle {position, velocity,
mass, charge}``, ``Scalar step()`` returning the new time and
``std::vector<Particle> particles() const`` returning a copy.  Backends are
NaiveSimulation(particles, forceConstant, timeStep) (naive_simulation.h:25-33)
and OpenClSimulation(bounds, particles, timeStep, log)
(open_cl_simulation.h:194-198).  This module keeps those names and argument
meanings; failures raise RuntimeError like the reference's std::runtime_error.
Array-level accessors (numpy, FP64, input order) sit beside them for tests and
benchmarks.
"""
from __future__ import annotations

import abc
import ctypes
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import Config, Stats, default_config, lib, ptr

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


@dataclass
class Particle:
    """simulation.h:16-32: position, velocity, mass, charge."""
    position: Sequence[float]
    velocity: Sequence[float]
    mass: float
    charge: float


class Simulation(abc.ABC):
    """simulation.h:6-37."""

    @abc.abstractmethod
    def step(self) -> float: ...

    @abc.abstractmethod
    def particles(self) -> List[Particle]: ...


def _check(h, rc, what):
    if rc != _lib.NBH_OK:
        msg = lib().nbh_last_error(h)
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _c64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class HipSimulation(Simulation):
    """Drop-in for OpenClSimulation (open_cl_simulation.h:194-198) backed by
    the HIP FMM pipeline.  ``bounds`` sets a fixed root cube [0, max(bounds))
    when ``fixed_root`` is requested in ``config``; by default the root cube is
    the particles' bounding cube, recomputed every step."""

    def __init__(self, bounds=None, particles: Optional[Iterable[Particle]] = None, time_step: float = 1e-3,
                 log=None, config: Optional[Config] = None, *, arrays=None, sys_off=None):
        cfg = config if config is not None else default_config()
        cfg.dt = float(time_step) if config is None else cfg.dt
        if arrays is not None:
            pos, vel, mass, charge = arrays
        else:
            plist = list(particles or [])
            pos = np.array([[p.position[0], p.position[1], p.position[2]] for p in plist], dtype=np.float64)
            vel = np.array([[p.velocity[0], p.velocity[1], p.velocity[2]] for p in plist], dtype=np.float64)
            mass = np.array([p.mass for p in plist], dtype=np.float64)
            charge = np.array([p.charge for p in plist], dtype=np.float64)
        if bounds is not None and cfg.fixed_root:
            cfg.root_size = float(max(bounds[0], bounds[1], bounds[2]))
        self._log = log
        self.n = int(len(mass))
        self._pos = _c64(pos, (self.n, 3))
        vel = _c64(vel, (self.n, 3))
        mass = _c64(mass)
        charge = None if charge is None else _c64(charge)
        self.config = cfg
        # M2L truncation q = m2l_order (0 selects order + 1) and term set:
        # "all" keeps every term of degree <= q, "low" beyond degree p only
        # those with min(|k|, |n|) <= q - p (the oracle's m2l_terms)
        self.m2l_order = int(cfg.m2l_order) or int(cfg.order) + 1
        self.m2l_terms = "low" if cfg.m2l_terms == _lib.M2L_TERMS_LOW else "all"
        h = ctypes.c_void_p()
        L = lib()
        if sys_off is None:
            rc = L.nbh_create(ctypes.byref(cfg), self.n, ptr(self._pos, _dp), ptr(vel, _dp), ptr(mass, _dp),
                              ptr(charge, _dp), ctypes.byref(h))
            self.nsys = 1
        else:
            so = np.ascontiguousarray(sys_off, dtype=np.uint32)
            rc = L.nbh_create_ensemble(ctypes.byref(cfg), len(so) - 1, ptr(so, _u32p), ptr(self._pos, _dp),
                                       ptr(vel, _dp), ptr(mass, _dp), ptr(charge, _dp), ctypes.byref(h))
            self.nsys = len(so) - 1
        if rc != _lib.NBH_OK:
            raise RuntimeError(f"nbh_create failed ({rc}): {L.nbh_last_error(None).decode()}")
        self._h = h
        self._del = L.nbh_destroy
        # truncation of the pairs of one level (2p on the matrix cores: every term |k|, |n| <= p)
        self.same_level_order = int(self.stats()["m2l_same_order"]) or None
        if self._log is not None:
            self._log.write(f"HipSimulation: {self.n} particles, solver {cfg.solver}, p={cfg.order}\n")

    @classmethod
    def from_arrays(cls, pos, vel, mass, charge=None, config: Optional[Config] = None, sys_off=None):
        return cls(None, None, config.dt if config else 1e-3, None, config, arrays=(pos, vel, mass, charge),
                   sys_off=sys_off)

    def close(self):
        if getattr(self, "_h", None):
            self._del(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- the reference interface ------------------------------------------------
    def step(self, nsteps: int = 1) -> float:
        """Advances nsteps steps; returns the new time (simulation.h:34)."""
        t = ctypes.c_double(0)
        _check(self._h, lib().nbh_step(self._h, int(nsteps), ctypes.byref(t)), "nbh_step")
        if self._log is not None:
            self._log.write(f"Step finished (t={t.value}).\n")
        return t.value

    def particles(self) -> List[Particle]:
        """Copy of the particles in input order (simulation.h:35)."""
        x, v, m, q, _ = self.state()
        return [Particle(tuple(x[i]), tuple(v[i]), float(m[i]), float(q[i])) for i in range(self.n)]

    # -- arrays ---------------------------------------------------------------------
    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        m = np.zeros(self.n)
        q = np.zeros(self.n)
        t = ctypes.c_double(0)
        _check(self._h, lib().nbh_get_state(self._h, ptr(x, _dp), ptr(v, _dp), ptr(m, _dp), ptr(q, _dp),
                                            ctypes.byref(t)), "nbh_get_state")
        return x, v, m, q, t.value

    def positions(self):
        """Positions only (FP64, input order): one device read instead of four."""
        x = np.zeros((self.n, 3))
        _check(self._h, lib().nbh_get_state(self._h, ptr(x, _dp), None, None, None, None), "nbh_get_state")
        return x

    def set_state(self, pos, vel, t=0.0):
        pos = _c64(pos, (self.n, 3))
        vel = _c64(vel, (self.n, 3))
        _check(self._h, lib().nbh_set_state(self._h, ptr(pos, _dp), ptr(vel, _dp), float(t)), "nbh_set_state")

    def forces(self):
        """(acc[n,3], pot[n]) at the current state, FP32, input order."""
        a = np.zeros((self.n, 3), dtype=np.float32)
        p = np.zeros(self.n, dtype=np.float32)
        _check(self._h, lib().nbh_compute_forces(self._h, ptr(a, _fp), ptr(p, _fp)), "nbh_compute_forces")
        return a, p

    def autotune(self, caps=None, chunks=None, evals=2):
        """nbh_autotune: times `evals` force evaluations for every leaf
        capacity in caps and P2P chunk (64 or 128) in chunks (None: keep the
        current value), keeps the fastest pair.  Returns (leaf_capacity,
        p2p_chunk, ms_per_eval, {(cap, chunk): ms})."""
        cs = np.ascontiguousarray(caps if caps is not None else [], dtype=np.uint32)
        ks = np.ascontiguousarray(chunks if chunks is not None else [], dtype=np.uint32)
        tab = np.zeros(max(len(cs), 1) * max(len(ks), 1))
        r = _lib.TuneResult()
        _check(self._h, lib().nbh_autotune(self._h, ptr(cs, _u32p) if len(cs) else None, len(cs),
                                           ptr(ks, _u32p) if len(ks) else None, len(ks), int(evals),
                                           ctypes.byref(r), ptr(tab, _dp)), "nbh_autotune")
        self.config.leaf_capacity = r.leaf_capacity
        self.config.p2p_chunk = r.p2p_chunk
        cl = list(cs) if len(cs) else [None]
        kl = list(ks) if len(ks) else [None]
        table = {(None if c is None else int(c), None if k is None else int(k)): float(tab[i * len(kl) + j])
                 for i, c in enumerate(cl) for j, k in enumerate(kl)}
        return int(r.leaf_capacity), int(r.p2p_chunk), float(r.ms_per_eval), table

    def direct_fp64(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        a = np.zeros((len(t), 3))
        p = np.zeros(len(t))
        _check(self._h, lib().nbh_direct_fp64(self._h, len(t), ptr(t, _i64p), ptr(a, _dp), ptr(p, _dp)),
               "nbh_direct_fp64")
        return a, p

    def direct_fp64_noalloc(self, targets):
        """nbh_direct_fp64_noalloc: the same FP64 sums with no device
        allocation (at most NOALLOC_MAX_TARGETS targets per call)."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        a = np.zeros((len(t), 3))
        p = np.zeros(len(t))
        _check(self._h, lib().nbh_direct_fp64_noalloc(self._h, len(t), ptr(t, _i64p), ptr(a, _dp), ptr(p, _dp)),
               "nbh_direct_fp64_noalloc")
        return a, p

    def diagnostics(self):
        out = np.zeros(8)
        _check(self._h, lib().nbh_diagnostics(self._h, ptr(out, _dp)), "nbh_diagnostics")
        return {"kinetic": out[0], "potential": out[1], "momentum": out[2:5].copy(),
                "angular_momentum": out[5:8].copy()}

    def stats(self):
        s = Stats()
        _check(self._h, lib().nbh_get_stats(self._h, ctypes.byref(s)), "nbh_get_stats")
        return s.as_dict()

    def set_timing(self, on=True, serial=False):
        """Per-phase timing; serial=True also runs each evaluation on one stream."""
        mode = (2 if serial else 1) if on else 0
        _check(self._h, lib().nbh_set_timing(self._h, mode), "nbh_set_timing")

    def force_vjp(self, w):
        w = _c64(w, (self.n, 3))
        gx = np.zeros((self.n, 3))
        gm = np.zeros(self.n)
        _check(self._h, lib().nbh_force_vjp(self._h, ptr(w, _dp), ptr(gx, _dp), ptr(gm, _dp)), "nbh_force_vjp")
        return gx, gm

    # -- parity exports ---------------------------------------------------------------
    def debug_sizes(self):
        s = np.zeros(6, dtype=np.int64)
        _check(self._h, lib().nbh_debug_sizes(self._h, ptr(s, _i64p)), "nbh_debug_sizes")
        return dict(nodes=int(s[0]), leaves=int(s[1]), depth=int(s[2]), m2l=int(s[3]), p2p=int(s[4]),
                    nsys=int(s[5]))

    def debug_keys(self):
        keys = np.zeros(self.n, dtype=np.uint64)
        ids = np.zeros(self.n, dtype=np.uint32)
        box = np.zeros((self.nsys, 4))
        _check(self._h, lib().nbh_debug_keys(self._h, ptr(keys, _u64p), ptr(ids, _u32p), ptr(box, _dp)),
               "nbh_debug_keys")
        return keys, ids, box

    def debug_tree(self):
        nn = self.debug_sizes()["nodes"]
        names = ["level", "sys", "ix", "iy", "iz", "begin", "count", "first_child", "nchild", "parent"]
        arrs = [np.zeros(nn, dtype=np.int32) for _ in names]
        _check(self._h, lib().nbh_debug_tree(self._h, *[ptr(a, _i32p) for a in arrs]), "nbh_debug_tree")
        return dict(zip(names, arrs))

    def debug_lists(self):
        s = self.debug_sizes()
        nn = s["nodes"]
        mo = np.zeros(nn + 1, dtype=np.int64)
        po = np.zeros(nn + 1, dtype=np.int64)
        ms = np.zeros(s["m2l"], dtype=np.int32)
        ps = np.zeros(s["p2p"], dtype=np.int32)
        _check(self._h, lib().nbh_debug_lists(self._h, ptr(mo, _i64p), ptr(ms, _i32p), ptr(po, _i64p),
                                              ptr(ps, _i32p)), "nbh_debug_lists")
        return mo, ms, po, ps

    def debug_expansions(self):
        nn = self.debug_sizes()["nodes"]
        p = int(self.config.order)
        nc = (p + 1) * (p + 2) * (p + 3) // 6
        M = np.zeros((nn, nc), dtype=np.float32)
        Lc = np.zeros((nn, nc), dtype=np.float32)
        _check(self._h, lib().nbh_debug_expansions(self._h, ptr(M, _fp), ptr(Lc, _fp)), "nbh_debug_expansions")
        return M, Lc


def device_count() -> int:
    return int(lib().nbh_device_count())
