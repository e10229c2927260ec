"""Reverse-mode gradients of KDK leapfrog trajectories through the HIP FMM.

Forward: x_{n+1}, v_{n+1} = KDK(x_n, v_n) with a(x) from the FMM
(nbh_step).  Backward: the transposed leapfrog.  For one step,
    v_h = v + dt/2 a(x);  x' = x + dt v_h;  v' = v_h + dt/2 a(x'),
the cotangents propagate as
    vbar_h = vbar' + dt xbar';  xbar = xbar' + VJP_x(x, dt/2 vbar_h);  vbar = vbar_h
with VJP_x(x', dt/2 vbar') already added to xbar' when x' was visited.  The two
half kicks that read a(x_n) (end of step n-1, start of step n) get the same
cotangent, so each state needs exactly one force VJP (nbh_force_vjp, the
exact transpose of the FMM on the tree and lists of that state): with
cotangent dt * vbar_h for interior states and dt/2 * vbar_h at both ends.
Masses enter only through a(x; m), so mbar sums the mass parts of the VJPs.

Memory: states are checkpointed every ``ckpt_every`` steps; the backward pass
recomputes each segment from its checkpoint before reversing it.
"""
from __future__ import annotations

import time
from typing import Callable, Optional

import numpy as np

from ._lib import default_config, lib, ptr
from .simulation import HipSimulation, _check, _dp


class HipForceProvider:
    """Forward steps and force VJPs on one GPU through the C ABI."""

    def __init__(self, x0, v0, m, config=None):
        cfg = config if config is not None else default_config()
        self.sim = HipSimulation.from_arrays(x0, v0, m, m, config=cfg)
        self.dt = float(cfg.dt)
        self._zeros = np.zeros((len(m), 3))

    def run(self, x, v, nsteps, record=False):
        self.sim.set_state(x, v)
        xs = [np.array(x, dtype=np.float64)] if record else None
        for _ in range(nsteps):
            self.sim.step(1)
            if record:
                xs.append(self.sim.positions())
        xf, vf, _, _, _ = self.sim.state()
        return xf, vf, xs

    def vjp(self, x, w):
        self.sim.set_state(x, self._zeros if self._zeros.shape == np.shape(x) else np.zeros_like(x))
        return self.sim.force_vjp(w)

    def close(self):
        self.sim.close()


class TrajectoryAdjoint:
    """Gradient of a scalar loss of (x_N, v_N) with respect to (x_0, v_0, m)."""

    def __init__(self, provider, dt: float, nsteps: int, ckpt_every: int = 20):
        self.p = provider
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.k = max(1, int(ckpt_every))
        self.ckpts = None
        self.vjp_calls = 0
        self.seconds = {"recompute": 0.0, "vjp": 0.0}

    def forward(self, x0, v0):
        x = np.array(x0, dtype=np.float64)
        v = np.array(v0, dtype=np.float64)
        self.ckpts = [(0, x.copy(), v.copy())]
        done = 0
        while done < self.nsteps:
            s = min(self.k, self.nsteps - done)
            x, v, _ = self.p.run(x, v, s)
            done += s
            if done < self.nsteps:
                self.ckpts.append((done, x.copy(), v.copy()))
        self.xN, self.vN = x, v
        return x, v

    def backward(self, xbar, vbar):
        dt = self.dt
        xbar = np.array(xbar, dtype=np.float64)
        vbar = np.array(vbar, dtype=np.float64)
        mbar = None
        # state N: second half kick of the last step
        gx, gm = self.p.vjp(self.xN, 0.5 * dt * vbar)
        self.vjp_calls += 1
        xbar += gx
        mbar = gm.copy()
        for start, xc, vc in reversed(self.ckpts):
            seg = min(self.k, self.nsteps - start)
            t0 = time.perf_counter()
            _, _, xs = self.p.run(xc, vc, seg, record=True)  # xs[0..seg]
            self.seconds["recompute"] += time.perf_counter() - t0
            for j in range(seg - 1, -1, -1):
                vh = vbar + dt * xbar
                n = start + j
                c = (0.5 if n == 0 else 1.0) * dt
                t0 = time.perf_counter()
                gx, gm = self.p.vjp(xs[j], c * vh)
                self.seconds["vjp"] += time.perf_counter() - t0
                self.vjp_calls += 1
                xbar = xbar + gx
                mbar += gm
                vbar = vh
        return xbar, vbar, mbar


class DeviceTrajectory:
    """The same gradient with the whole reverse pass on the GPU
    (nbh_adjoint_forward / nbh_adjoint_backward): checkpoints, recomputed
    states and cotangents stay in device memory; the host sees only the
    initial state, the cotangents of the final state and the gradient."""

    def __init__(self, x0, v0, m, nsteps: int, ckpt_every: int = 20, config=None):
        cfg = config if config is not None else default_config()
        self.sim = HipSimulation.from_arrays(x0, v0, m, m, config=cfg)
        self.n = len(m)
        self.nsteps = int(nsteps)
        self.k = max(1, int(ckpt_every))
        self.seconds = {"forward": 0.0, "backward": 0.0}

    def forward(self):
        t0 = time.perf_counter()
        rc = lib().nbh_adjoint_forward(self.sim._h, self.nsteps, self.k)
        self.seconds["forward"] = time.perf_counter() - t0
        _check(self.sim._h, rc, "nbh_adjoint_forward")
        x, v, _, _, _ = self.sim.state()
        return x, v

    def backward(self, xbar, vbar):
        xb = np.ascontiguousarray(xbar, dtype=np.float64).reshape(self.n, 3)
        vb = np.ascontiguousarray(vbar, dtype=np.float64).reshape(self.n, 3)
        gx = np.zeros((self.n, 3))
        gv = np.zeros((self.n, 3))
        gm = np.zeros(self.n)
        t0 = time.perf_counter()
        rc = lib().nbh_adjoint_backward(self.sim._h, ptr(xb, _dp), ptr(vb, _dp), ptr(gx, _dp), ptr(gv, _dp),
                                        ptr(gm, _dp))
        self.seconds["backward"] = time.perf_counter() - t0
        _check(self.sim._h, rc, "nbh_adjoint_backward")
        return gx, gv, gm

    def close(self):
        self.sim.close()


def trajectory_function():
    """torch.autograd.Function computing (x_N, v_N) from (x0, v0, m) on the GPU;
    the backward pass is DeviceTrajectory's reverse leapfrog.  Import torch
    lazily (only the harness needs it)."""
    import torch

    class Trajectory(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0, m, nsteps: int, ckpt_every: int, config):
            adj = DeviceTrajectory(x0.detach().cpu().double().numpy(), v0.detach().cpu().double().numpy(),
                                   m.detach().cpu().double().numpy(), nsteps, ckpt_every, config)
            xN, vN = adj.forward()
            ctx.adj = adj
            return (torch.from_numpy(xN).to(x0.device, x0.dtype), torch.from_numpy(vN).to(v0.device, v0.dtype))

        @staticmethod
        def backward(ctx, gxN, gvN):
            adj = ctx.adj
            gx0, gv0, gm = adj.backward(gxN.detach().cpu().double().numpy(), gvN.detach().cpu().double().numpy())
            adj.close()
            dev, dt_ = gxN.device, gxN.dtype
            return (torch.from_numpy(gx0).to(dev, dt_), torch.from_numpy(gv0).to(dev, dt_),
                    torch.from_numpy(gm).to(dev, dt_), None, None, None)

    return Trajectory
