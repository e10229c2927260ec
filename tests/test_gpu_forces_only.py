"""Steps evaluate accelerations only, as the reference's step does
(open_cl_simulation.cpp:81-98 computes the field, no potential): the tuned
P2P and M2L kernels leave out the potential's terms (the M2L products into
L[0]), and nbh_diagnostics evaluates the potentials of the current positions
when the last evaluation skipped them, keeping the accelerations the
integrator holds."""
import numpy as np
import pytest

import nbody_amd as nb
from helpers import make_sim, max_rel, rms_rel

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


@pytest.mark.parametrize("graph", [0, 1])
def test_step_accelerations_match_a_full_evaluation(graph):
    """A KDK step from x3 kicks with a3 and, after the drift, with the
    accelerations of its own forces-only evaluation at x4: v4 - v3 =
    dt/2 (a3 + a4), where a3 and a4 come from nbh_compute_forces (which
    computes the potential too) at x3 and x4.  Equal to the FP32 summation
    order of the two evaluations."""
    _case(f"forces_only_acc_graph{graph}")
    x, v, m = nb.ics.plummer(20000, 3)
    dt = 1e-3
    sim = make_sim(x, v, m, eps=1e-2, theta=0.5, order=4, leaf_capacity=32, dt=dt, graph=graph)
    sim.step(3)
    _, v3, _, _, _ = sim.state()
    a3, pot3 = sim.forces()
    sim.step(1)
    _, v4, _, _, _ = sim.state()
    a4, _ = sim.forces()
    dv = v4 - v3
    want = 0.5 * dt * (a3.astype(np.float64) + a4.astype(np.float64))
    e = rms_rel(dv, want)
    print(f"  dv vs dt/2 (a3 + a4): rms {e:.2e} max {max_rel(dv, want):.2e}", flush=True)
    assert e < 1e-5
    assert np.all(np.isfinite(pot3)) and np.all(pot3 < 0)
    sim.close()


def test_diagnostics_potential_after_steps_and_trajectory_unchanged():
    """nbh_diagnostics after forces-only steps: its potential energy equals
    0.5 sum q_i phi_i of a full evaluation at the same positions, and calling
    it between steps leaves the trajectory as without the call (up to the
    float summation order of the deferred pairs' atomics, as two runs of
    the same steps)."""
    _case("forces_only_diagnostics")
    x, v, m = nb.ics.plummer(12000, 8)
    kw = dict(eps=1e-2, theta=0.5, order=4, leaf_capacity=16, dt=1e-3)
    a = make_sim(x, v, m, **kw)
    b = make_sim(x, v, m, **kw)
    a.step(4)
    b.step(4)
    dg = a.diagnostics()
    a.step(4)
    b.step(4)
    xa, va, _, _, _ = a.state()
    xb, vb, _, _, _ = b.state()
    print(f"  x {rms_rel(xa, xb):.2e} v {rms_rel(va, vb):.2e}", flush=True)
    assert rms_rel(xa, xb) < 1e-9 and rms_rel(va, vb) < 1e-5
    # the potential of diagnostics at step 4 against a full evaluation there
    c = make_sim(x, v, m, **kw)
    c.step(4)
    _, pot = c.forces()
    pe = 0.5 * float(np.sum(m * pot))
    print(f"  PE diagnostics {dg['potential']:.8e} full evaluation {pe:.8e}", flush=True)
    assert abs(dg["potential"] - pe) <= 1e-6 * abs(pe)
    for s in (a, b, c):
        s.close()
