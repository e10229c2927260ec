"""Shared helpers for the parity tests: run the HIP path and the oracle on the
same seeded inputs."""
import numpy as np

import nbody_amd as nb
import pyoracle as po


def make_sim(x, v, m, q=None, **cfg):
    c = nb.default_config(**cfg)
    return nb.HipSimulation.from_arrays(x, v, m, q if q is not None else m, config=c)


def rms_rel(a, ref):
    """sqrt(mean |a - ref|^2 / mean |ref|^2) over rows."""
    a = np.asarray(a, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    return float(np.sqrt(np.mean(np.sum((a - ref) ** 2, 1)) / np.mean(np.sum(ref ** 2, 1))))


def max_rel(a, ref):
    a = np.asarray(a, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    return float(np.max(np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)))


def max_rel_floor(a, ref, floor=0.1):
    """Largest per-particle error |a_i - ref_i| / max(|ref_i|, floor * rms|ref|):
    the relative error of each particle, except that particles whose reference
    is far below the typical magnitude (near-cancelling sums) are measured
    against floor times the RMS magnitude instead of their own."""
    a = np.asarray(a, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    nr = np.linalg.norm(ref, axis=1)
    den = np.maximum(nr, floor * np.sqrt(np.mean(nr ** 2)))
    return float(np.max(np.linalg.norm(a - ref, axis=1) / den))


def oracle_pipeline(x_sorted, q_sorted, box, cap, max_depth, theta_num, theta_den, p, eps, G=1.0,
                    sys_off=None, bits=21, m2l_order=None):
    """Oracle tree + lists + FP64 FMM on particles already in the GPU's order."""
    keys, box_o = po.morton(x_sorted, sys_off=sys_off, bits=bits)
    assert np.all(np.diff(keys.astype(np.float64)) >= 0) or np.all(keys[1:] >= keys[:-1])
    tree = po.build_tree(keys, sys_off=sys_off, bits=bits, cap=cap, max_depth=max_depth)
    lists = po.walk(tree, theta_num, theta_den)
    acc, pot = po.fmm(tree, lists, p, x_sorted, q_sorted, box_o, G=G, eps=eps, m2l_order=m2l_order)
    return keys, box_o, tree, lists, acc, pot
