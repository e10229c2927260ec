"""FP64 oracle of a KDK trajectory through the FMM and of its gradient (test
infrastructure over pyoracle).

Each force evaluation builds the oracle's own Morton keys, root cube, octree
and lists at the current positions (as the GPU does every evaluation) and
evaluates the FP64 FMM (orc_fmm); the gradient transposes the leapfrog step by
step, with two force VJPs (orc_fmm_vjp) per step on the structures of the
state they belong to:
    v_h = v + dt/2 a(x);  x' = x + dt v_h;  v' = v_h + dt/2 a(x')
    xbar' += dt/2 J(x')^T vbar';  vbar_h = vbar' + dt xbar';  xbar = xbar' + dt/2 J(x)^T vbar_h
(textbook reverse mode, written independently of the product's merged form).
With `frozen`, the structures of a previous run are reused at every state,
which makes the trajectory a smooth function of the initial positions, so
central differences pin the gradient.
"""
import numpy as np

import pyoracle as po


class OracleFMMTrajectory:
    def __init__(self, m, p=4, theta=(1, 2), cap=16, eps=1e-2, dt=1e-3, G=1.0, m2l_order=None,
                 same_level_order=None, m2l_terms="all"):
        self.m = np.asarray(m, dtype=np.float64)
        self.p, self.theta, self.cap, self.eps, self.dt, self.G = p, theta, cap, eps, dt, G
        self.q = m2l_order if m2l_order is not None else p + 1
        self.qs = same_level_order
        self.terms = m2l_terms

    def structures(self, x):
        keys, box = po.morton(x)
        order = np.argsort(keys, kind="stable")
        tree = po.build_tree(keys[order], cap=self.cap)
        lists = po.walk(tree, *self.theta)
        return order, tree, lists, box

    def acc(self, x, st, m=None):
        order, tree, lists, box = st
        mm = self.m if m is None else m
        a_s, _ = po.fmm(tree, lists, self.p, x[order], mm[order], box, G=self.G, eps=self.eps,
                        m2l_order=self.q, same_level_order=self.qs, m2l_terms=self.terms)
        a = np.empty_like(a_s)
        a[order] = a_s
        return a

    def vjp(self, x, w, st):
        order, tree, lists, box = st
        gx_s, gm_s = po.fmm_vjp(tree, lists, self.p, x[order], self.m[order], w[order], box, G=self.G,
                                eps=self.eps, m2l_order=self.q, same_level_order=self.qs, m2l_terms=self.terms)
        gx = np.empty_like(gx_s)
        gm = np.empty_like(gm_s)
        gx[order] = gx_s
        gm[order] = gm_s
        return gx, gm

    def forward(self, x0, v0, nsteps, frozen=None, m=None):
        """(x_N, v_N, states, structures); frozen: structures to reuse per state."""
        dt = self.dt
        x = np.array(x0, dtype=np.float64)
        v = np.array(v0, dtype=np.float64)
        xs = [x.copy()]
        sts = [frozen[0] if frozen else self.structures(x)]
        a = self.acc(x, sts[0], m)
        for s in range(nsteps):
            vh = v + 0.5 * dt * a
            x = x + dt * vh
            st = frozen[s + 1] if frozen else self.structures(x)
            a = self.acc(x, st, m)
            v = vh + 0.5 * dt * a
            xs.append(x.copy())
            sts.append(st)
        return x, v, xs, sts

    def gradient(self, xs, sts, xbar, vbar):
        """Gradient of <xbar, x_N> + <vbar, v_N> w.r.t. (x_0, v_0, m)."""
        dt = self.dt
        xbar = np.array(xbar, dtype=np.float64)
        vbar = np.array(vbar, dtype=np.float64)
        mbar = np.zeros(len(self.m))
        for s in range(len(xs) - 2, -1, -1):
            gx, gm = self.vjp(xs[s + 1], 0.5 * dt * vbar, sts[s + 1])
            xbar += gx
            mbar += gm
            vh = vbar + dt * xbar
            gx, gm = self.vjp(xs[s], 0.5 * dt * vh, sts[s])
            xbar += gx
            mbar += gm
            vbar = vh
        return xbar, vbar, mbar
