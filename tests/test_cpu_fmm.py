"""The FP64 CPU fast multipole method (include/nbody/cpu_fmm.h, CpuSimulation
method="fmm", through libnbody_cpu.so) against the oracle: the same tree and
interaction lists as the oracle's restatement of the reference traversal
(interaction.cl:42-92), forces equal to the oracle FMM on those lists to FP64
rounding, and convergence to the direct sum.  No GPU needed."""
import numpy as np
import pytest

import nbody_amd as nb
import pyoracle as po
from nbody_amd.cpu import INT_KDK, CpuSimulation


def _oracle_fmm(x, q, cap, num, den, p, m2l_order, eps, m2l_terms="all"):
    """Oracle tree, lists and FMM field (G = 1) in input order."""
    keys, box = po.morton(x)
    order = np.argsort(keys, kind="stable")
    tree = po.build_tree(keys[order], cap=cap, max_depth=21)
    lists = po.walk(tree, num, den)
    acc, pot = po.fmm(tree, lists, p, x[order], q[order], box, G=1.0, eps=eps, m2l_order=m2l_order,
                      m2l_terms=m2l_terms)
    a = np.zeros_like(acc)
    ph = np.zeros_like(pot)
    a[order] = acc
    ph[order] = pot
    return tree, lists, a, ph


def _cases():
    rng = np.random.default_rng(11)
    uni = rng.random((2048, 3))
    pl, _, _ = nb.ics.plummer(4096, 3)
    cl, _, _ = nb.ics.clustered_pair(1500, seeds=(5, 6))
    return {"uniform2048": uni, "plummer4096": pl, "clustered3000": cl}


CASES = _cases()


@pytest.mark.parametrize("name,cap,theta", [("uniform2048", 8, (1, 2)), ("plummer4096", 16, (1, 2)),
                                            ("plummer4096", 32, (1, 3)), ("clustered3000", 8, (1, 2))])
def test_tree_and_lists_match_oracle(name, cap, theta):
    x = CASES[name]
    n = len(x)
    m = np.full(n, 1.0 / n)
    tree, lists, _, _ = _oracle_fmm(x, m, cap, theta[0], theta[1], 2, 3, 1e-2)
    sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m, eps=1e-2, method="fmm", order=2,
                                    theta=theta[0] / theta[1], leaf_capacity=cap)
    sim.forces()
    st = sim.fmm_stats()
    assert st["nodes"] == len(tree.level) and st["depth"] == tree.depth
    assert st["leaves"] == int(np.sum(tree.nchild == 0))
    assert st["m2l_unordered"] == lists.unordered_m2l
    assert st["p2p_unordered"] == lists.unordered_p2p
    assert st["visited"] == lists.visited
    sim.close()


@pytest.mark.parametrize("p,q,terms", [(4, 5, "all"), (4, 5, "low"), (4, 4, "all"), (3, 4, "all"), (6, 6, "all"),
                                       (6, 7, "low")])
def test_forces_match_oracle_fmm(p, q, terms):
    x = CASES["plummer4096"]
    n = len(x)
    rng = np.random.default_rng(2)
    m = rng.random(n) / n + 1e-5
    eps = 5e-3
    _, _, ao, po_ = _oracle_fmm(x, m, 16, 1, 2, p, q, eps, terms)
    sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m, G=1.0, eps=eps, method="fmm", order=p, m2l_order=q,
                                    m2l_terms=terms, theta=0.5, leaf_capacity=16)
    a, ph = sim.forces()
    # both give G q_i/m_i * field and -G * phi; here q = m and G = 1
    err = np.linalg.norm(a - ao, axis=1) / np.linalg.norm(ao, axis=1)
    assert err.max() < 1e-11
    assert np.max(np.abs(ph - po_) / np.abs(po_)) < 1e-12
    sim.close()


def test_converges_to_direct_sum():
    """Uniform N=2048, eps=1e-2, theta=1/2, cap 8, RMS of the per-particle
    relative error: the default truncation (q = p+1, every term of degree
    p+1) at p = 2, 4, 6 decreases and ends below 1e-5 (9.5e-6 at p = 6); the
    low-degree set at q = p+1 lies between it and q = p (at p = 6 it gains
    about 1% over q = p on this uniform input)."""
    x = CASES["uniform2048"]
    n = len(x)
    m = np.full(n, 1.0 / n)
    ad, pd = po.direct(x, m, eps=1e-2)

    def err(p, **kw):
        sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m, eps=1e-2, method="fmm", order=p, theta=0.5,
                                        leaf_capacity=8, **kw)
        a, ph = sim.forces()
        sim.close()
        return float(np.sqrt(np.mean((np.linalg.norm(a - ad, axis=1) / np.linalg.norm(ad, axis=1)) ** 2)))

    errs = [err(p) for p in (2, 4, 6)]
    print("all terms, p = 2, 4, 6:", errs)
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-5, errs
    for i, p in enumerate((4, 6)):
        e_low, e_qp = err(p, m2l_terms="low"), err(p, m2l_order=p)
        print(f"p = {p}: q = p {e_qp:.3e}, low set {e_low:.3e}, all terms {errs[i + 1]:.3e}")
        assert errs[i + 1] < e_low < e_qp


@pytest.mark.parametrize("n", [1, 2, 20])
def test_small_systems_are_one_leaf(n):
    """n <= leaf capacity: the root is a leaf and the pair sum is the direct sum
    (the root's self pair, which the reference drops, defect A7)."""
    rng = np.random.default_rng(n)
    x = rng.random((n, 3))
    m = rng.random(n) + 0.1
    sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m, eps=1e-2, method="fmm", order=4, leaf_capacity=32)
    a, ph = sim.forces()
    ad, pd = po.direct(x, m, eps=1e-2)
    assert np.allclose(a, ad, rtol=1e-13, atol=1e-300) and np.allclose(ph, pd, rtol=1e-13)
    st = sim.fmm_stats()
    assert st["nodes"] == 1 and st["p2p_unordered"] == 1 and st["m2l_unordered"] == 0
    sim.close()


def test_coincident_particles_stop_at_max_depth():
    """More coincident particles than the capacity: cells split to the deepest
    level and stop there with an over-full leaf; softened forces stay finite."""
    x = np.vstack([np.full((40, 3), 0.25), np.random.default_rng(0).random((200, 3))])
    n = len(x)
    m = np.full(n, 1.0 / n)
    sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m, eps=1e-2, method="fmm", order=4, leaf_capacity=8)
    a, ph = sim.forces()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(ph))
    assert sim.fmm_stats()["depth"] == 21
    tree, _, ao, _ = _oracle_fmm(x, m, 8, 1, 2, 4, 5, 1e-2)
    assert tree.depth == 21
    err = np.linalg.norm(a - ao, axis=1) / np.linalg.norm(ao, axis=1)
    assert err.max() < 1e-11
    sim.close()


def test_kdk_steps_follow_direct_sum():
    x, v, m = nb.ics.plummer(1024, 4)
    d = CpuSimulation.from_arrays(x, v, m, eps=1e-2, time_step=1e-3, integrator=INT_KDK)
    f = CpuSimulation.from_arrays(x, v, m, eps=1e-2, time_step=1e-3, integrator=INT_KDK, method="fmm", order=6,
                                  theta=1.0 / 3.0, leaf_capacity=16)
    assert d.step(5) == pytest.approx(f.step(5), abs=1e-15)
    xd, vd, _, _, _ = d.state()
    xf, vf, _, _, _ = f.state()
    assert np.max(np.abs(xf - xd)) < 1e-9 and np.max(np.abs(vf - vd)) / np.max(np.abs(vd)) < 1e-5
    d.close()
    f.close()


def test_invalid_options_rejected():
    x = np.random.default_rng(0).random((10, 3))
    m = np.ones(10)
    with pytest.raises(RuntimeError, match="theta"):
        CpuSimulation.from_arrays(x, np.zeros_like(x), m, method="fmm", theta=0.3141592)
    with pytest.raises(RuntimeError, match="order"):
        CpuSimulation.from_arrays(x, np.zeros_like(x), m, method="fmm", order=9)
    with pytest.raises(RuntimeError, match="m2l_order"):
        CpuSimulation.from_arrays(x, np.zeros_like(x), m, method="fmm", order=4, m2l_order=3)
    with pytest.raises(ValueError):
        CpuSimulation.from_arrays(x, np.zeros_like(x), m, method="tree")
    sim = CpuSimulation.from_arrays(x, np.zeros_like(x), m)
    with pytest.raises(RuntimeError, match="direct sum"):
        sim.fmm_stats()
    sim.close()
