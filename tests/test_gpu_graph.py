"""Captured steps (nbh_config.graph = 1): each KDK step recorded once per
state buffer in a HIP graph and replayed, with no host read-back inside the
step; the same trajectory as the uncaptured steps, and a step that outgrows
the captured launch shapes rolled back and run uncaptured."""
import numpy as np
import pytest

import nbody_amd as nb
from helpers import make_sim, rms_rel

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


def _pair(x, v, m, steps, sys_off=None, **cfg):
    out = []
    for graph in (0, 1):
        c = nb.default_config(graph=graph, **cfg)
        sim = nb.HipSimulation.from_arrays(x, v, m, m, config=c, sys_off=sys_off)
        t = sim.step(steps)
        xs, vs, _, _, _ = sim.state()
        out.append((xs, vs, t, sim.stats()))
        sim.close()
    return out


@pytest.mark.parametrize("kind,n,cap", [("uniform", 4096, 16), ("plummer", 20000, 32)])
def test_captured_steps_follow_the_uncaptured_trajectory(kind, n, cap):
    _case(f"graph_{kind}{n}")
    x, v, m = nb.ics.uniform(n, 0) if kind == "uniform" else nb.ics.plummer(n, 1)
    steps = 12
    (x0, v0, t0, s0), (x1, v1, t1, s1) = _pair(x, v, m, steps, eps=1e-2, theta=0.5, order=4, leaf_capacity=cap,
                                               dt=1e-3)
    ex, ev = rms_rel(x1, x0), rms_rel(v1, v0)
    print(f"  captures {s1['graph_captures']} replays {s1['graph_replays']} fallbacks {s1['graph_fallbacks']}; "
          f"x {ex:.2e} v {ev:.2e}", flush=True)
    assert t0 == t1
    # the first step's opening evaluation runs uncaptured; one graph per
    # state buffer; every step replayed
    assert s0["graph_replays"] == 0 and s1["graph_replays"] == steps and s1["graph_captures"] == 2
    assert s1["graph_fallbacks"] == 0
    # the same kernels; only the float summation order of atomics differs
    assert ex < 1e-9 and ev < 1e-5
    assert s1["num_nodes"] > 0 and s1["m2l_pairs"] > 0 and s1["p2p_interactions"] > 0


def test_captured_steps_on_an_ensemble():
    _case("graph_ensemble_16")
    x, v, m, off = nb.ics.ensemble(nsys=16, seed0=100, nmin=512, nmax=6000, size_seed=7)
    (x0, v0, _, _), (x1, v1, _, s1) = _pair(x, v, m, 6, sys_off=off, eps=1e-2, theta=0.5, order=4,
                                            leaf_capacity=16, dt=1e-3)
    print(f"  replays {s1['graph_replays']}; x {rms_rel(x1, x0):.2e} v {rms_rel(v1, v0):.2e}", flush=True)
    assert s1["graph_replays"] == 6
    assert rms_rel(x1, x0) < 1e-9 and rms_rel(v1, v0) < 1e-5


def test_a_step_that_outgrows_the_capture_is_rolled_back():
    """Half of a uniform cloud in the fixed box [0, 1)^3 falls through the
    centre (each step takes 30% of its initial offset from it): the tree gets
    deeper step by step, beyond the level range a step was captured with, so
    such a step is rolled back and run uncaptured (and the next captures
    anew); every step is either replayed or rolled back, and the trajectory
    equals the uncaptured one.  The particles stay inside the box for the 6
    steps (offsets shrink to 10% and come back to 80%)."""
    _case("graph_fallback_collapse")
    n, dt, steps = 4096, 1e-3, 6
    x, _, m = nb.ics.uniform(n, 3)
    x = 0.05 + 0.9 * x
    v = np.zeros_like(x)
    c = np.array([0.5, 0.5, 0.5])
    v[n // 2:] = 0.3 * (c - x[n // 2:]) / dt
    kw = dict(eps=1e-2, theta=0.5, order=4, leaf_capacity=16, dt=dt, fixed_root=1, root_lo=(0.0, 0.0, 0.0),
              root_size=1.0)
    (x0, v0, _, s0), (x1, v1, _, s1) = _pair(x, v, m, steps, **kw)
    print(f"  depth {s1['depth']} captures {s1['graph_captures']} replays {s1['graph_replays']} "
          f"fallbacks {s1['graph_fallbacks']}; x {rms_rel(x1, x0):.2e} v {rms_rel(v1, v0):.2e}", flush=True)
    assert s1["graph_fallbacks"] >= 1 and s1["graph_replays"] >= 2
    assert s1["graph_replays"] + s1["graph_fallbacks"] == steps
    assert rms_rel(x1, x0) < 1e-9 and rms_rel(v1, v0) < 1e-5


def test_graph_is_off_where_it_does_not_apply():
    """Timing on, the plain kernels or the Euler integrator: steps run
    uncaptured and give the same results as without the flag."""
    _case("graph_not_applicable")
    x, v, m = nb.ics.plummer(3000, 2)
    sim = make_sim(x, v, m, graph=1, leaf_capacity=16, order=4, eps=1e-2)
    sim.set_timing(True)
    sim.step(3)
    assert sim.stats()["graph_replays"] == 0
    sim.set_timing(False)
    sim.step(2)
    assert sim.stats()["graph_replays"] == 2
    sim.close()
    for extra in (dict(m2l=nb.M2L_PLAIN), dict(integrator=nb.INT_EULER)):
        sim = make_sim(x, v, m, graph=1, leaf_capacity=16, order=4, eps=1e-2, **extra)
        sim.step(3)
        assert sim.stats()["graph_replays"] == 0
        sim.close()
