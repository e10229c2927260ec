"""Parity of the HIP FMM pipeline with the CPU oracle, through the C ABI.

Bit-exact: Morton keys, root boxes, particle order, tree topology, M2L/P2P
lists.  Tolerances (FP32 kernels vs FP64 oracle on the same tree and lists):
forces rtol 2e-5 RMS; against the FP64 direct sum the FMM truncation error
dominates and is bounded per configuration below.
"""
import numpy as np
import pytest

import nbody_amd as nb
import pyoracle as po
from helpers import make_sim, max_rel, max_rel_floor, oracle_pipeline, rms_rel

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


def _gpu_sorted(sim, x_in, q_in):
    keys, ids, box = sim.debug_keys()
    return keys, ids, box, x_in[ids], q_in[ids]


CASES = [
    # name, generator, n, cap, p, theta (num, den), eps
    ("uniform4096_cap16", "uniform", 4096, 16, 4, (1, 2), 1e-2),
    ("plummer20k_cap32", "plummer", 20000, 32, 4, (1, 2), 1e-3),
    ("plummer8k_cap8_p6_theta04", "plummer", 8192, 8, 6, (2, 5), 5e-4),
    # p = 4 at theta = 0.4: the same-level table (483 offsets, 116 KiB)
    # fits one workgroup per CU, the 512-thread form of the kernel
    ("plummer8k_cap16_p4_theta04", "plummer", 8192, 16, 4, (2, 5), 5e-4),
    ("uniform3000_cap1_p2", "uniform", 3000, 1, 2, (1, 2), 1e-2),
    ("uniform2048_cap16_unsoftened", "uniform", 2048, 16, 4, (1, 2), 0.0),
    # P2P lane layouts at the ends: leaves of 33..48 particles (one source
    # stride per target) and of 3 (21 strides, the most staged slots)
    ("plummer6000_cap48", "plummer", 6000, 48, 4, (1, 2), 1e-3),
    ("uniform2500_cap3", "uniform", 2500, 3, 4, (1, 2), 1e-2),
]


def _ics(kind, n, seed=0):
    if kind == "uniform":
        return nb.ics.uniform(n, seed)
    return nb.ics.plummer(n, seed + 1)


# name, M2L kernel, P2P kernel, M2L truncation (0: q = p+1 with every degree-(p+1)
# term, the default; 1: q = p+1 with the low-degree term set; -1: q = p),
# RMS tolerances (acc, pot) against the FP64 oracle FMM.  The MFMA kernels take
# the same-level pairs by translation offset with the dense operator (every
# term |k|, |n| <= p; the oracle is run with that truncation for those pairs)
# and are held to the FP32 tolerances: bf16x3 drops the lo*lo term and rounds
# both halves to bf16 (about 2^-16 relative per product), which the far
# field's share of each force keeps below them.
KERNELS = [("tuned", nb.M2L_GROUPED, nb.P2P_LANE, 0, 2e-5, 2e-6),
           ("tuned_qp", nb.M2L_GROUPED, nb.P2P_LANE, -1, 2e-5, 2e-6),
           ("tuned_low", nb.M2L_GROUPED, nb.P2P_LANE, 1, 2e-5, 2e-6),
           ("plain_low", nb.M2L_PLAIN, nb.P2P_LEAF, 1, 2e-5, 2e-6),
           ("plain", nb.M2L_PLAIN, nb.P2P_LEAF, 0, 2e-5, 2e-6),
           ("mfma_f32", nb.M2L_MFMA_F32, nb.P2P_LANE, 0, 2e-5, 2e-6),
           ("mfma_bf16x3", nb.M2L_BF16X3, nb.P2P_LANE, 0, 2e-5, 2e-6),
           ("mfma_bf16x3_qp", nb.M2L_BF16X3, nb.P2P_LANE, -1, 2e-5, 2e-6),
           # P2P tile of 128 staged sources per round (the autotuner's other shape)
           ("tuned_chunk128", nb.M2L_GROUPED, nb.P2P_LANE, 0, 2e-5, 2e-6)]


@pytest.mark.parametrize("kern,m2l,p2p,qsel,tol_acc,tol_pot", KERNELS)
@pytest.mark.parametrize("name,kind,n,cap,p,th,eps", CASES)
def test_tree_lists_forces_match_oracle(name, kind, n, cap, p, th, eps, kern, m2l, p2p, qsel, tol_acc, tol_pot):
    _case(f"{name}/{kern}")
    x, v, m = _ics(kind, n)
    sim = make_sim(x, v, m, leaf_capacity=cap, order=p, theta=th[0] / th[1], eps=eps, m2l=m2l, p2p=p2p,
                   m2l_order=p if qsel < 0 else 0, m2l_terms=nb.M2L_TERMS_LOW if qsel == 1 else nb.M2L_TERMS_ALL,
                   p2p_chunk=128 if kern.endswith("_chunk128") else 0)
    acc, pot = sim.forces()
    keys, ids, box, xs, qs = _gpu_sorted(sim, x, m)
    # keys, box and the stable particle order
    keys_o, box_o = po.morton(x)
    order = np.argsort(keys_o, kind="stable")
    assert np.array_equal(box, box_o), (box, box_o)
    assert np.array_equal(ids, order.astype(np.uint32))
    assert np.array_equal(keys, keys_o[order])
    # tree
    tree = po.build_tree(keys_o[order], cap=cap, max_depth=21)
    t = sim.debug_tree()
    for f in ["level", "sys", "ix", "iy", "iz", "begin", "count", "first_child", "nchild", "parent"]:
        assert np.array_equal(t[f], getattr(tree, f)), f
    # lists
    lists = po.walk(tree, th[0], th[1])
    mo, ms, pofs, ps = sim.debug_lists()
    assert np.array_equal(mo, lists.m2l_off)
    assert np.array_equal(ms, lists.m2l_src)
    assert np.array_equal(pofs, lists.p2p_off)
    assert np.array_equal(ps, lists.p2p_src)
    # forces: FP32 kernels vs FP64 oracle FMM on the same tree and lists
    acc_o, pot_o = po.fmm(tree, lists, p, xs, qs, box_o, G=1.0, eps=eps, m2l_order=sim.m2l_order, m2l_terms=sim.m2l_terms,
                          same_level_order=sim.same_level_order)
    inv = np.empty_like(ids)
    inv[ids] = np.arange(n, dtype=np.uint32)
    acc_s = acc[ids]
    pot_s = pot[ids]
    e_acc = rms_rel(acc_s, acc_o)
    e_pot = rms_rel(pot_s, pot_o)
    # per particle: an indexing slip that damages a few particles (a chunk or
    # leaf edge) would pass an RMS bound at these sizes
    m_acc = max_rel_floor(acc_s, acc_o)
    m_pot = max_rel_floor(pot_s, pot_o)
    print(f"  fp32 vs fp64-fmm: acc rms {e_acc:.2e} max {m_acc:.2e}; pot rms {e_pot:.2e} max {m_pot:.2e}",
          flush=True)
    assert e_acc < tol_acc
    assert e_pot < tol_pot
    assert m_acc < 1e-3 and m_pot < 1e-4
    st = sim.stats()
    assert st["m2l_pairs"] == len(lists.m2l_src)
    assert st["p2p_pairs"] == len(lists.p2p_src)
    sim.close()


def test_expansions_match_oracle():
    _case("expansions_uniform2000_p4")
    n, p = 2000, 4
    x, v, m = nb.ics.uniform(n, 3)
    sim = make_sim(x, v, m, leaf_capacity=16, order=p, eps=1e-2)
    sim.forces()
    keys, ids, box, xs, qs = _gpu_sorted(sim, x, m)
    tree = po.build_tree(keys, cap=16)
    M, L = sim.debug_expansions()
    nn = len(tree.level)
    cen = np.stack([box[0, a] + (getattr(tree, f) + 0.5) * box[0, 3] / 2.0 ** tree.level
                    for a, f in enumerate(["ix", "iy", "iz"])], 1)
    # P2M of every node from its particles equals the GPU's M (leaves: P2M, internal: M2M)
    for a in range(0, nn, max(1, nn // 60)):
        b, c = tree.begin[a], tree.count[a]
        Mo = po.p2m(p, cen[a], xs[b:b + c], qs[b:b + c])
        scale = np.max(np.abs(Mo)) + 1e-30
        assert np.max(np.abs(M[a] - Mo)) / scale < 1e-5, a
    sim.close()


def test_direct_fp64_matches_oracle():
    _case("direct_fp64_plummer5000")
    n = 5000
    x, v, m = nb.ics.plummer(n, 7)
    sim = make_sim(x, v, m, eps=1e-3)
    tg = np.arange(0, n, 7)
    a, pth = sim.direct_fp64(tg)
    ao, po_ = po.direct(x, m, xt=x[tg], self_idx=tg, eps=1e-3)
    assert rms_rel(a, ao) < 1e-12
    assert rms_rel(pth, po_) < 1e-12
    sim.close()


@pytest.mark.parametrize("p,err", [(2, 5e-3), (4, 8e-5), (6, 1e-5)])
def test_fmm_converges_to_direct(p, err):
    """Uniform cube N=4096 (seed 0), eps=1e-2, theta=0.5, cap 16, the default
    M2L (q = p+1, every degree-(p+1) term): RMS error against the FP64 direct
    sum, 6.6e-5 at p=4 (the FP64 CPU FMM gives 6.65e-5; the low-degree term
    set 9.96e-5, at the 1e-4 gate) and 3.9e-6 at p=6."""
    _case(f"convergence_uniform4096_p{p}")
    n = 4096
    x, v, m = nb.ics.uniform(n, 0)
    sim = make_sim(x, v, m, leaf_capacity=16, order=p, eps=1e-2)
    acc, pot = sim.forces()
    ao, _ = po.direct(x, m, eps=1e-2)
    e = rms_rel(acc, ao)
    print(f"  p={p}: rms err vs fp64 direct {e:.3e}", flush=True)
    assert e < err
    sim.close()


def test_kdk_trajectory_uniform4096():
    """Uniform cube N=4096 (seed 0), eps=1e-2, theta=0.5, cap 16, 10 KDK steps."""
    _case("uniform4096_kdk_10_steps")
    n = 4096
    x, v, m = nb.ics.uniform(n, 0)
    sim = make_sim(x, v, m, leaf_capacity=16, order=4, eps=1e-2, dt=1e-3)
    t = sim.step(10)
    assert abs(t - 0.01) < 1e-12
    xg, vg, _, _, _ = sim.state()
    xo, vo = po.integrate_direct(x, v, m, m, 1.0, 1e-2, 1e-3, 10, 0)
    dx = np.max(np.abs(xg - xo))
    dv = rms_rel(vg, vo)
    print(f"  max |dx| {dx:.2e}, rms dv {dv:.2e}", flush=True)
    # the FMM's force error (~1e-4 rms at p=4) bounds the trajectory error:
    # displacements are ~5e-5 after 10 steps
    assert dx < 1e-7
    assert dv < 3e-4
    sim.close()


def test_direct_solver_kdk_matches_oracle():
    _case("direct_solver_kdk_2048")
    n = 2048
    x, v, m = nb.ics.plummer(n, 11)
    sim = make_sim(x, v, m, solver=nb.SOLVER_DIRECT, eps=1e-2, dt=1e-3)
    sim.step(5)
    xg, vg, _, _, _ = sim.state()
    xo, vo = po.integrate_direct(x, v, m, m, 1.0, 1e-2, 1e-3, 5, 0)
    assert np.max(np.abs(xg - xo)) < 1e-9
    assert rms_rel(vg - v, vo - v) < 1e-4
    sim.close()


@pytest.mark.gpu
def test_direct_solver_unsoftened_forces():
    """eps = 0: the self pair must drop out exactly (no 0 * inf)."""
    _case("direct_solver_unsoftened_1024")
    n = 1024
    x, v, m = nb.ics.uniform(n, 5)
    sim = make_sim(x, v, m, solver=nb.SOLVER_DIRECT, eps=0.0, dt=1e-3)
    acc, pot = sim.forces()
    assert np.all(np.isfinite(acc)) and np.all(np.isfinite(pot))
    ao, po_ = po.direct(x, m, eps=0.0)
    assert rms_rel(acc, ao) < 1e-5
    assert rms_rel(pot, po_) < 1e-6
    sim.close()


@pytest.mark.parametrize("integ", [nb.INT_EULER, nb.INT_KICK_DRIFT])
def test_compat_integrators(integ):
    _case(f"integrator_{integ}")
    n = 1024
    x, v, m = nb.ics.plummer(n, 12)
    sim = make_sim(x, v, m, solver=nb.SOLVER_DIRECT, eps=1e-2, dt=1e-3, integrator=integ)
    sim.step(3)
    xg, vg, _, _, _ = sim.state()
    xo, vo = po.integrate_direct(x, v, m, m, 1.0, 1e-2, 1e-3, 3, integ)
    assert np.max(np.abs(xg - xo)) < 1e-9
    assert rms_rel(vg - v, vo - v) < 1e-4
    sim.close()
