"""Seeded initial conditions for the benchmark and parity configurations.

The reference's driver (main.cpp:16-66) draws unseeded rand() positions in a
cube; these generators are seeded and reproducible (SplitMix64 in the C ABI).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib, ptr

_dp = ctypes.POINTER(ctypes.c_double)


def uniform(n, seed=0):
    """Positions U[0,1)^3, zero velocities, masses = charges = 1/n."""
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    m = np.zeros(n)
    rc = lib().nbh_ic_uniform(seed, n, ptr(x, _dp), ptr(v, _dp), ptr(m, _dp))
    if rc:
        raise RuntimeError("nbh_ic_uniform failed")
    return x, v, m


def plummer(n, seed=1, a=1.0, M=1.0, G=1.0, rmax=20.0):
    """Plummer sphere truncated at rmax; masses M/n (charges = masses)."""
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    m = np.zeros(n)
    rc = lib().nbh_ic_plummer(seed, n, a, M, G, rmax, ptr(x, _dp), ptr(v, _dp), ptr(m, _dp))
    if rc:
        raise RuntimeError("nbh_ic_plummer failed")
    return x, v, m


def clustered_pair(n_each, seeds=(2, 3), a=1.0, M_each=1.0, separation=10.0, speed=0.3, rmax=20.0):
    """Two Plummer spheres centred at (+-separation/2, 0, 0) moving towards
    each other with bulk velocities (-+speed, 0, 0)."""
    xa, va, ma = plummer(n_each, seeds[0], a, M_each, 1.0, rmax)
    xb, vb, mb = plummer(n_each, seeds[1], a, M_each, 1.0, rmax)
    xa[:, 0] -= separation / 2
    va[:, 0] += speed
    xb[:, 0] += separation / 2
    vb[:, 0] -= speed
    return np.concatenate([xa, xb]), np.concatenate([va, vb]), np.concatenate([ma, mb])


def ensemble(nsys=256, seed0=100, nmin=512, nmax=65536, size_seed=99):
    """nsys independent Plummer systems with N log-uniform in [nmin, nmax].
    Returns (x, v, m, sys_off)."""
    rng = np.random.default_rng(size_seed)
    ns = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(np.int64)
    xs, vs, ms = [], [], []
    for s, n in enumerate(ns):
        x, v, m = plummer(int(n), seed0 + s)
        xs.append(x); vs.append(v); ms.append(m)
    off = np.zeros(nsys + 1, dtype=np.uint32)
    off[1:] = np.cumsum(ns)
    return np.concatenate(xs), np.concatenate(vs), np.concatenate(ms), off
