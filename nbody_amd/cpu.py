"""CpuSimulation: the multi-core FP64 CPU backend (libnbody_cpu.so,
include/nbody_cpu.h) with the interface of HipSimulation -- step() returning
the new time and particles() in input order (simulation.h:6-37).  Forces come
from the direct sum (method="direct") or from the FP64 fast multipole method
of include/nbody/cpu_fmm.h (method="fmm": the HIP library's tree, lists and
expansions on host cores)."""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, List, Optional

import numpy as np

from .simulation import Particle, Simulation

_dp = ctypes.POINTER(ctypes.c_double)
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnbody_cpu.so")
_LIB = None

INT_KDK, INT_EULER, INT_KICK_DRIFT = 0, 1, 2


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C nbody_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        L.ncpu_create.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.ncpu_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.ncpu_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ncpu_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.ncpu_forces.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.ncpu_set_fmm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_int]
        L.ncpu_fmm_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.ncpu_fmm_times.argtypes = [ctypes.c_void_p, _dp]
        L.ncpu_destroy.argtypes = [ctypes.c_void_p]
        L.ncpu_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class CpuSimulation(Simulation):
    """Same constructor prefix as HipSimulation / OpenClSimulation; options:
    G, eps (softening), integrator (INT_*), threads (0 = OpenMP default),
    method ("direct" or "fmm") and, for the FMM, order, m2l_order (0: order +
    1), m2l_terms ("all": every term of degree <= m2l_order; "low": beyond
    degree order only those with min(|k|, |n|) <= m2l_order - order), theta
    and leaf_capacity."""

    def __init__(self, bounds=None, particles: Optional[Iterable[Particle]] = None, time_step: float = 1e-3,
                 log=None, *, G=1.0, eps=0.01, integrator=INT_KDK, threads=0, arrays=None, method="direct",
                 order=4, m2l_order=0, m2l_terms="all", theta=0.5, leaf_capacity=32):
        if arrays is not None:
            pos, vel, mass, charge = arrays
        else:
            plist = list(particles or [])
            pos = [p.position[:3] for p in plist]
            vel = [p.velocity[:3] for p in plist]
            mass = [p.mass for p in plist]
            charge = [p.charge for p in plist]
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.n = len(pos)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        charge = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        self._mass = mass
        self._charge = mass if charge is None else charge
        h = ctypes.c_void_p()
        rc = lib().ncpu_create(self.n, _p(pos), _p(vel), _p(mass), _p(charge), float(time_step), float(G),
                               float(eps), int(integrator), int(threads), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"ncpu_create failed: {lib().ncpu_last_error().decode()}")
        self._h = h
        if method == "fmm":
            if m2l_terms not in ("all", "low"):
                self.close()
                raise ValueError(f"m2l_terms must be 'all' or 'low', not {m2l_terms!r}")
            self._check(lib().ncpu_set_fmm(self._h, int(order), int(m2l_order), int(m2l_terms == "low"),
                                           float(theta), int(leaf_capacity)), "ncpu_set_fmm")
        elif method != "direct":
            self.close()
            raise ValueError(f"method must be 'direct' or 'fmm', not {method!r}")
        if log is not None:
            log.write(f"CpuSimulation: {self.n} particles ({method})\n")

    @classmethod
    def from_arrays(cls, pos, vel, mass, charge=None, time_step=1e-3, **opts):
        return cls(None, None, time_step, None, arrays=(pos, vel, mass, charge), **opts)

    def close(self):
        if getattr(self, "_h", None):
            lib().ncpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed: {lib().ncpu_last_error().decode()}")

    def step(self, nsteps: int = 1) -> float:
        t = ctypes.c_double(0)
        self._check(lib().ncpu_step(self._h, int(nsteps), ctypes.byref(t)), "ncpu_step")
        return t.value

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        t = ctypes.c_double(0)
        self._check(lib().ncpu_get_state(self._h, _p(x), _p(v), ctypes.byref(t)), "ncpu_get_state")
        return x, v, self._mass.copy(), self._charge.copy(), t.value

    def set_state(self, pos, vel, t=0.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        self._check(lib().ncpu_set_state(self._h, _p(pos), _p(vel), float(t)), "ncpu_set_state")

    def particles(self) -> List[Particle]:
        x, v, m, q, _ = self.state()
        return [Particle(tuple(x[i]), tuple(v[i]), float(m[i]), float(q[i])) for i in range(self.n)]

    def forces(self):
        a = np.zeros((self.n, 3))
        p = np.zeros(self.n)
        self._check(lib().ncpu_forces(self._h, _p(a), _p(p)), "ncpu_forces")
        return a, p

    def fmm_stats(self):
        """Statistics of the last FMM evaluation: nodes, leaves, depth, unordered
        M2L and P2P cell pairs, visited cell pairs."""
        out = (ctypes.c_longlong * 6)()
        self._check(lib().ncpu_fmm_stats(self._h, out), "ncpu_fmm_stats")
        return dict(zip(("nodes", "leaves", "depth", "m2l_unordered", "p2p_unordered", "visited"), list(out)))

    def fmm_times(self):
        """Seconds per phase of the last FMM evaluation."""
        out = np.zeros(6)
        self._check(lib().ncpu_fmm_times(self._h, _p(out)), "ncpu_fmm_times")
        return dict(zip(("tree", "walk", "upward", "m2l", "downward", "leaves"), out.tolist()))
