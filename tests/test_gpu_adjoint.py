"""Adjoint of the FMM force (nbh_force_vjp) and of KDK trajectories on the GPU.

Tolerances: the GPU VJP is the FP32 transpose of the FMM on the same tree and
lists, so it matches the oracle's FP64 transpose (orc_fmm_vjp, itself pinned by
central differences with frozen lists in test_oracle.py) to FP32 rounding
(rtol 1e-4 RMS); against the direct-sum adjoint it differs by the FMM's
truncation (rtol 3e-3 RMS at p=4, theta=0.5).
"""
import time

import numpy as np
import pytest

import nbody_amd as nb
import pyoracle as po
from helpers import make_sim, rms_rel
from nbody_amd.adjoint import HipForceProvider, TrajectoryAdjoint

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


@pytest.mark.parametrize("kind,n,cap,p,m2l", [("plummer", 6000, 16, 4, nb.M2L_GROUPED),
                                              ("uniform", 3000, 8, 3, nb.M2L_GROUPED),
                                              ("plummer", 6000, 16, 4, nb.M2L_BF16X3)])
def test_force_vjp_matches_oracle_transpose(kind, n, cap, p, m2l):
    """The transposed far-field chain runs the handle's M2L kernels too (with
    M2L_BF16X3 its same-level pairs on the matrix cores)."""
    _case(f"force_vjp_{kind}{n}_p{p}_m2l{m2l}")
    x, v, m = nb.ics.plummer(n, 5) if kind == "plummer" else nb.ics.uniform(n, 5)
    eps = 1e-2
    sim = make_sim(x, v, m, leaf_capacity=cap, order=p, eps=eps, m2l=m2l)
    w = np.random.default_rng(6).standard_normal((n, 3))
    gx, gm = sim.force_vjp(w)
    keys, ids, box = sim.debug_keys()
    tree = po.build_tree(keys, cap=cap)
    lists = po.walk(tree, 1, 2)
    gxo, gmo = po.fmm_vjp(tree, lists, p, x[ids], m[ids], w[ids], box, eps=eps, m2l_order=sim.m2l_order, m2l_terms=sim.m2l_terms,
                          same_level_order=sim.same_level_order)
    inv = np.empty(n, dtype=np.int64)
    inv[ids] = np.arange(n)
    e_x = rms_rel(gx, gxo[inv])
    e_m = rms_rel(gm, gmo[inv])
    print(f"  vs fp64 transpose: gx {e_x:.2e} gm {e_m:.2e}", flush=True)
    assert e_x < 1e-4 and e_m < 1e-4
    gxd, gmd = po.direct_vjp(x, m, w, eps=eps)
    assert rms_rel(gx, gxd) < 3e-3 and rms_rel(gm, gmd) < 3e-3
    sim.close()


def test_trajectory_gradient_against_direct_adjoint_and_fd():
    """Trajectory gradient, Plummer N=2048, eps=1e-2, theta=0.5, p=4, 40 KDK steps,
    checkpoints every 10.  Loss = mean |x_N - x*|^2 with x* from initial
    velocities perturbed by sigma=1e-3 noise.  GPU FMM adjoint vs the FP64
    direct-sum adjoint; FP64 central differences (h=1e-5) of the direct-sum
    trajectory on 64 random components."""
    _case("trajectory_adjoint_2048x40")
    n, steps, dt, eps = 2048, 40, 1e-3, 1e-2
    x0, v0, m = nb.ics.plummer(n, 4)
    rng = np.random.default_rng(5)
    xt, _ = po.integrate_direct(x0, v0 + 1e-3 * rng.standard_normal((n, 3)), m, m, 1.0, eps, dt, steps, 0)
    cfg = nb.default_config(eps=eps, theta=0.5, order=4, leaf_capacity=16, dt=dt)
    prov = HipForceProvider(x0, v0, m, cfg)
    adj = TrajectoryAdjoint(prov, dt, steps, ckpt_every=10)
    t0 = time.perf_counter()
    xN, vN = adj.forward(x0, v0)
    t_fwd = time.perf_counter() - t0
    xbar = 2.0 * (xN - xt) / (3 * n)
    t0 = time.perf_counter()
    gx, gv, gm = adj.backward(xbar, np.zeros_like(vN))
    t_bwd = time.perf_counter() - t0
    prov.close()
    print(f"  forward {t_fwd:.3f}s backward {t_bwd:.3f}s ratio {t_bwd / t_fwd:.2f}", flush=True)

    class Direct:
        def run(self, x, v, k, record=False):
            xs = [np.array(x)] if record else None
            for _ in range(k):
                x, v = po.integrate_direct(x, v, m, m, 1.0, eps, dt, 1, 0)
                if record:
                    xs.append(x.copy())
            return x, v, xs

        def vjp(self, x, w):
            return po.direct_vjp(x, m, w, eps=eps)

    ad = TrajectoryAdjoint(Direct(), dt, steps, ckpt_every=10)
    xNd, vNd = ad.forward(x0, v0)
    gxd, gvd, gmd = ad.backward(2.0 * (xNd - xt) / (3 * n), np.zeros_like(vNd))
    ex, ev, em = rms_rel(gx, gxd), rms_rel(gv, gvd), rms_rel(gm, gmd)
    print(f"  vs fp64 direct adjoint: x {ex:.2e} v {ev:.2e} m {em:.2e}", flush=True)
    assert ex < 5e-3 and ev < 5e-3 and em < 5e-3

    def loss(x, v):
        xN_, _ = po.integrate_direct(x, v, m, m, 1.0, eps, dt, steps, 0)
        return float(np.sum((xN_ - xt) ** 2) / (3 * n))

    sel = rng.choice(3 * n, 64, replace=False)
    h = 1e-5
    worst = 0.0
    for s in sel:
        i, k = divmod(int(s), 3)
        e = np.zeros((n, 3)); e[i, k] = h
        fd = (loss(x0 + e, v0) - loss(x0 - e, v0)) / (2 * h)
        worst = max(worst, abs(fd - gxd[i, k]) / (np.abs(gxd).max() + 1e-30))
    print(f"  direct adjoint vs central differences: worst {worst:.2e}", flush=True)
    assert worst < 1e-5


def _traj_setup(n, steps, seed=4):
    x0, v0, m = nb.ics.plummer(n, seed)
    rng = np.random.default_rng(5)
    return x0, v0, m, rng


def test_device_trajectory_matches_oracle_fmm_adjoint():
    """Reverse leapfrog on the GPU (nbh_adjoint_forward/backward, checkpoints
    every 10) against the FP64 oracle of the same FMM trajectory gradient
    (tests/fmm_trajectory.py, itself pinned to central differences with frozen
    lists in test_adjoint_cpu.py).  Plummer N=2048 (seed 4), eps=1e-2,
    theta=0.5, p=4, cap 16, 40 KDK steps at dt=1e-3; loss = mean |x_N - x*|^2
    with x* from velocities perturbed by sigma=1e-3 (seed 5).  The two differ
    by FP32 arithmetic (and a tree that can differ where a particle sits within
    rounding of a cell face): rtol 3e-4 RMS."""
    from fmm_trajectory import OracleFMMTrajectory
    from nbody_amd.adjoint import DeviceTrajectory
    _case("device_trajectory_vs_oracle_fmm_2048x40")
    n, steps, dt, eps = 2048, 40, 1e-3, 1e-2
    x0, v0, m, rng = _traj_setup(n, steps)
    cfg = nb.default_config(eps=eps, theta=0.5, order=4, leaf_capacity=16, dt=dt)
    orc = OracleFMMTrajectory(m, p=4, cap=16, eps=eps, dt=dt, m2l_order=5)
    xt, _, _, _ = orc.forward(x0, v0 + 1e-3 * rng.standard_normal((n, 3)), steps)
    adj = DeviceTrajectory(x0, v0, m, steps, ckpt_every=10, config=cfg)
    xN, vN = adj.forward()
    xbar = 2.0 * (xN - xt) / (3 * n)
    gx, gv, gm = adj.backward(xbar, np.zeros_like(vN))
    same = adj.sim.same_level_order
    adj.close()
    orc.qs = same
    xo, vo, xs, sts = orc.forward(x0, v0, steps)
    print(f"  final state vs oracle: x {np.max(np.abs(xN - xo)):.2e} v {rms_rel(vN, vo):.2e}", flush=True)
    gxo, gvo, gmo = orc.gradient(xs, sts, 2.0 * (xo - xt) / (3 * n), np.zeros_like(vo))
    ex, ev, em = rms_rel(gx, gxo), rms_rel(gv, gvo), rms_rel(gm, gmo)
    print(f"  gradient vs fp64 oracle fmm adjoint: x {ex:.2e} v {ev:.2e} m {em:.2e}", flush=True)
    assert ex < 3e-4 and ev < 3e-4 and em < 3e-4


def test_trajectory_function_autograd_equals_device_adjoint():
    """torch.autograd.Function over the C ABI (trajectory_function): the
    gradients .backward() leaves on x0, v0 and m equal DeviceTrajectory's."""
    import torch
    from nbody_amd.adjoint import DeviceTrajectory, trajectory_function
    _case("trajectory_function_autograd_1024x12")
    n, steps, dt, eps = 1024, 12, 1e-3, 1e-2
    x0, v0, m, rng = _traj_setup(n, steps, seed=7)
    target = x0 + 1e-3 * rng.standard_normal((n, 3))
    cfg = nb.default_config(eps=eps, theta=0.5, order=4, leaf_capacity=16, dt=dt)
    xt = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    vt = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    Traj = trajectory_function()
    xN, vN = Traj.apply(xt, vt, mt, steps, 4, cfg)
    loss = ((xN - torch.from_numpy(target)) ** 2).mean() + 0.5 * (vN ** 2).sum()
    loss.backward()
    adj = DeviceTrajectory(x0, v0, m, steps, ckpt_every=4, config=cfg)
    xN2, vN2 = adj.forward()
    gx, gv, gm = adj.backward(2.0 * (xN2 - target) / (3 * n), vN2)
    adj.close()
    assert np.max(np.abs(xN.detach().numpy() - xN2)) < 1e-12
    ex, ev, em = (rms_rel(xt.grad.numpy(), gx), rms_rel(vt.grad.numpy(), gv), rms_rel(mt.grad.numpy(), gm))
    print(f"  autograd vs device adjoint: x {ex:.2e} v {ev:.2e} m {em:.2e}", flush=True)
    assert ex < 1e-6 and ev < 1e-6 and em < 1e-6


def test_adjoint_costs_at_most_three_forward_passes():
    """N=16384 Plummer (seed 4), eps=1e-2, theta=0.5, p=4, 200 KDK steps at
    dt=1e-3, checkpoints every 20: the reverse pass (segment recomputation plus
    one force VJP per state, all on the device) takes at most 3x the forward."""
    from nbody_amd.adjoint import DeviceTrajectory
    _case("adjoint_cost_16384x200")
    n, steps = 16384, 200
    x0, v0, m, rng = _traj_setup(n, steps)
    cfg = nb.default_config(eps=1e-2, theta=0.5, order=4, leaf_capacity=16, dt=1e-3)
    adj = DeviceTrajectory(x0, v0, m, steps, ckpt_every=20, config=cfg)
    adj.forward()  # warm: allocations and first-call setup
    adj.sim.set_state(x0, v0)
    xN, vN = adj.forward()
    gx, gv, gm = adj.backward(np.ones_like(xN) / (3 * n), np.zeros_like(vN))
    ratio = adj.seconds["backward"] / adj.seconds["forward"]
    print(f"  forward {adj.seconds['forward']:.3f}s backward {adj.seconds['backward']:.3f}s ratio {ratio:.2f}",
          flush=True)
    adj.close()
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gm))
    assert ratio <= 3.0
