"""CPU tests that pin the oracle (oracle/oracle.cpp) before it is trusted as
the parity checker.  Every oracle function is checked against something it
does not compute itself: golden vectors of the compiled reference, closed
forms, sympy, brute force, or invariants chosen so that a dropped term, a
wrong sign/index or a transposed operand fails.
"""
import glob
import itertools
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# NaiveSimulation compat restatement vs the reference's own output
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "naive_*.npz"))))
def test_naive_compat_bit_exact_vs_reference_golden(path):
    d = np.load(path)
    out, t = po.naive_compat(d["inp"], float(d["force_constant"]), float(d["dt"]), int(d["steps"]))
    assert np.array_equal(out.view(np.uint32), d["out"].view(np.uint32))
    assert np.float32(t) == d["time"]


def test_naive_compat_matches_live_reference_build():
    if po.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    n = 97
    x = rng.random((n, 3)).astype(np.float32)
    P = po.particles_to_p12(x, np.zeros((n, 3)), np.full(n, 2.0), np.full(n, 0.5))
    a, ta = po.naive_compat(P, 1.0, 1e-3, 4)
    b, tb = po.naive_reference(P, 1.0, 1e-3, 4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and ta == tb


def test_naive_reference_defect_is_reproduced():
    """naive_simulation.cpp:10-14 uses (dx, dx, dx): a pair separated along y
    only gives 0/0 (NaN) in the reference and in the restatement."""
    P = po.particles_to_p12(np.array([[0.5, 0.1, 0.5], [0.5, 0.9, 0.5]]), np.zeros((2, 3)), np.ones(2), np.ones(2))
    out, _ = po.naive_compat(P, 1.0, 1e-3, 1)
    assert np.isnan(out[:, 4:7]).all()


# ---------------------------------------------------------------------------
# direct sum
def test_direct_two_body_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    q = np.array([2.0, 5.0])
    eps = 0.5
    a, p = po.direct(x, q, eps=eps, G=1.5)
    s3 = (25.0 + eps ** 2) ** 1.5
    assert np.allclose(a[0], 1.5 * 5.0 * np.array([3.0, 4.0, 0.0]) / s3, rtol=1e-14)
    assert np.allclose(a[1], -1.5 * 2.0 * np.array([3.0, 4.0, 0.0]) / s3, rtol=1e-14)
    assert np.allclose(p, [-1.5 * 5.0 / np.sqrt(25.25), -1.5 * 2.0 / np.sqrt(25.25)], rtol=1e-14)


def test_direct_newton_third_law_and_brute_force():
    rng = np.random.default_rng(1)
    n = 200
    x = rng.random((n, 3))
    q = rng.random(n) + 0.1
    a, p = po.direct(x, q, eps=1e-2)
    assert np.abs((q[:, None] * a).sum(0)).max() < 1e-10 * np.abs(q[:, None] * a).max() * n
    r = x[None, :, :] - x[:, None, :]
    s = np.sqrt((r ** 2).sum(-1) + 1e-4)
    w = q[None, :] / s ** 3
    np.fill_diagonal(w, 0)
    a_bf = (w[:, :, None] * r).sum(1)
    assert np.allclose(a, a_bf, rtol=1e-12, atol=1e-12)


def test_direct_vjp_against_finite_differences():
    rng = np.random.default_rng(2)
    n = 12
    x = rng.random((n, 3))
    m = rng.random(n) + 0.5
    w = rng.standard_normal((n, 3))
    gx, gm = po.direct_vjp(x, m, w, eps=0.1)

    def f(xx, mm):
        a, _ = po.direct(xx, mm, eps=0.1)
        return float((a * w).sum())

    h = 1e-6
    for i, k in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        xp = x.copy(); xp[i, k] += h
        xm = x.copy(); xm[i, k] -= h
        assert abs((f(xp, m) - f(xm, m)) / (2 * h) - gx[i, k]) < 1e-6 * max(1, abs(gx[i, k]))
    for i in [0, 5, 11]:
        mp = m.copy(); mp[i] += h
        mm = m.copy(); mm[i] -= h
        assert abs((f(x, mp) - f(x, mm)) / (2 * h) - gm[i]) < 1e-6 * max(1, abs(gm[i]))


# ---------------------------------------------------------------------------
# Morton keys: floor quantisation in the bounding cube and bit interleave
def test_morton_brute_force():
    rng = np.random.default_rng(3)
    x = rng.random((500, 3)) * [2.0, 1.0, 0.5] - 1.0
    keys, box = po.morton(x, bits=21)
    lo = x.min(0)
    size = (x.max(0) - lo).max()
    assert np.array_equal(box[0, :3], lo) and box[0, 3] == size
    q = np.minimum(np.floor((x - lo) * (2.0 ** 21 / size)), 2 ** 21 - 1).astype(np.int64)
    for i in range(0, 500, 7):
        k = 0
        for b in range(21):
            for ax, sh in ((0, 2), (1, 1), (2, 0)):
                k |= ((int(q[i, ax]) >> b) & 1) << (3 * b + sh)
        assert int(keys[i]) == k


def test_morton_ensemble_prefix():
    rng = np.random.default_rng(4)
    x = rng.random((300, 3))
    off = np.array([0, 100, 250, 300], dtype=np.uint32)
    keys, box = po.morton(x, sys_off=off, bits=18)
    for s in range(3):
        assert np.all((keys[off[s]:off[s + 1]] >> np.uint64(54)) == s)
        assert np.array_equal(box[s, :3], x[off[s]:off[s + 1]].min(0))


# ---------------------------------------------------------------------------
# tree
def _tree(x, cap, max_depth=21, sys_off=None, bits=21):
    keys, box = po.morton(x, sys_off=sys_off, bits=bits)
    o = np.argsort(keys, kind="stable")
    return keys[o], o, po.build_tree(keys[o], sys_off=sys_off, bits=bits, cap=cap, max_depth=max_depth), box


@pytest.mark.parametrize("cap,max_depth", [(1, 21), (8, 21), (16, 3), (40, 21)])
def test_tree_invariants(cap, max_depth):
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random((600, 3)), 0.5 + 1e-4 * rng.random((200, 3))])
    keys, _, t, _ = _tree(x, cap, max_depth)
    n = len(x)
    leaves = np.where(t.nchild == 0)[0]
    # leaves partition [0, n)
    cover = np.zeros(n, dtype=int)
    for a in leaves:
        cover[t.begin[a]:t.begin[a] + t.count[a]] += 1
    assert np.all(cover == 1)
    for a in range(len(t.level)):
        l = t.level[a]
        sh = np.uint64(3 * (21 - l))
        pref = keys[t.begin[a]:t.begin[a] + t.count[a]] >> sh
        assert np.all(pref == pref[0])  # all keys of the cell share its prefix
        # cell coordinates are the deinterleaved prefix
        if t.nchild[a] == 0:
            assert t.count[a] <= cap or l == max_depth
        else:
            assert t.count[a] > cap and l < max_depth
            ch = np.arange(t.first_child[a], t.first_child[a] + t.nchild[a])
            assert np.all(t.parent[ch] == a) and np.all(t.level[ch] == l + 1)
            assert t.count[ch].sum() == t.count[a]
            assert np.all(np.diff(t.begin[ch]) > 0)
    # BFS order: levels non-decreasing, begins increasing within a level
    assert np.all(np.diff(t.level) >= 0)
    for l in range(t.depth + 1):
        b = t.begin[t.level == l]
        assert np.all(np.diff(b) > 0)


def test_tree_brute_force_levels():
    """Per level, the cells are exactly the distinct key prefixes whose parent
    cell holds more than cap particles."""
    rng = np.random.default_rng(6)
    x = rng.random((400, 3)) ** 2
    cap = 5
    keys, _, t, _ = _tree(x, cap)
    k = keys.astype(np.uint64)
    for l in range(1, t.depth + 1):
        par = k >> np.uint64(3 * (21 - l + 1))
        pu, pc = np.unique(par, return_counts=True)
        split = set(pu[pc > cap].tolist())
        pref = k >> np.uint64(3 * (21 - l))
        expect = sorted({int(p) for p, pp in zip(pref, par) if int(pp) in split})
        got = []
        for a in np.where(t.level == l)[0]:
            got.append(int(k[t.begin[a]] >> np.uint64(3 * (21 - l))))
        assert got == expect


# ---------------------------------------------------------------------------
# walk: coverage, MAC restated independently, lattice counts
def _mac_exact(t, a, b, num, den):
    """interaction.cl:63-79 restated in exact rational arithmetic from the cell
    geometry (centres and sides in units of the root cube), independent of the
    oracle's integer form.  Ties (e.g. offset (2,2,2) at theta=1/2, where the
    ratio is exactly theta^2) are not approximated, as in the reference's FP32
    form where 3 d^2 / 12 d^2 rounds to exactly 0.25."""
    from fractions import Fraction as Fr

    def geo(i):
        h = Fr(1, 2 ** int(t.level[i]))
        return [(int(v) * 2 + 1) * h / 2 for v in (t.ix[i], t.iy[i], t.iz[i])], h
    ca, da = geo(a)
    cb, db = geo(b)
    d2 = sum((p - q) ** 2 for p, q in zip(cb, ca))
    return a != b and Fr(3, 4) * (da + db) ** 2 / d2 < Fr(num, den) ** 2


@pytest.mark.parametrize("cap,th", [(1, (1, 2)), (4, (1, 2)), (3, (2, 5))])
def test_walk_covers_every_pair_once(cap, th):
    rng = np.random.default_rng(7)
    n = 300
    x = np.concatenate([rng.random((250, 3)), 0.2 + 0.01 * rng.random((50, 3))])
    keys, _, t, box = _tree(x, cap)
    L = po.walk(t, *th)
    nn = len(t.level)
    leaf_of = np.empty(n, dtype=int)
    anc = [[] for _ in range(n)]
    for a in range(nn):
        for i in range(t.begin[a], t.begin[a] + t.count[a]):
            anc[i].append(a)
            if t.nchild[a] == 0:
                leaf_of[i] = a
    cover = np.zeros((n, n), dtype=int)
    for A in range(nn):
        for e in range(L.p2p_off[A], L.p2p_off[A + 1]):
            B = L.p2p_src[e]
            assert t.nchild[A] == 0 and t.nchild[B] == 0
            ia = np.arange(t.begin[A], t.begin[A] + t.count[A])
            ib = np.arange(t.begin[B], t.begin[B] + t.count[B])
            cover[np.ix_(ia, ib)] += 1
        for e in range(L.m2l_off[A], L.m2l_off[A + 1]):
            B = L.m2l_src[e]
            ia = np.arange(t.begin[A], t.begin[A] + t.count[A])
            ib = np.arange(t.begin[B], t.begin[B] + t.count[B])
            cover[np.ix_(ia, ib)] += 1
    # (i, i) comes once from the leaf's own P2P self pair; every i != j exactly once
    assert np.all(cover.diagonal() == 1)
    off = ~np.eye(n, dtype=bool)
    assert np.all(cover[off] == 1)
    # lists ascending per target, MAC holds for M2L pairs and fails for P2P pairs
    num, den = th
    for A in range(nn):
        s = L.m2l_src[L.m2l_off[A]:L.m2l_off[A + 1]]
        assert np.all(np.diff(s) > 0)
        for B in s[:20]:
            assert _mac_exact(t, A, B, num, den)
        p = L.p2p_src[L.p2p_off[A]:L.p2p_off[A + 1]]
        assert np.all(np.diff(p) > 0)
        for B in p:
            assert not _mac_exact(t, A, B, num, den)


def test_walk_lattice_counts():
    """Uniform lattice: an interior level-5 cell (complete neighbourhood at
    levels 4 and 5) sees 179 near cells (P2P) and 1253 M2L partners at
    theta=0.5 (sum-of-half-diagonals MAC, interaction.cl:71-79), and 341 /
    2387 at theta=0.4."""
    g = 32
    ijk = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)
    ijk = ijk[np.all(np.abs(ijk - 16) <= 10, axis=1)]
    x = np.concatenate([(ijk + 0.5) / g, [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]])  # corners pin the cube to [0, 1]
    keys, o, t, box = _tree(x, cap=1, max_depth=5)
    target = [a for a in np.where(t.level == 5)[0] if (t.ix[a], t.iy[a], t.iz[a]) == (16, 16, 16)][0]
    for (num, den), near, far in [((1, 2), 179, 1253), ((2, 5), 341, 2387)]:
        L = po.walk(t, num, den)
        assert L.p2p_off[target + 1] - L.p2p_off[target] == near
        assert L.m2l_off[target + 1] - L.m2l_off[target] == far


def _same_level_offsets(num, den):
    """(|o_x|, |o_y|, |o_z|) of the same-level M2L pairs that can occur: the
    criterion accepts o and rejects the parent offset floor((c + o) / 2) for
    some target position c in its parent (the set the library tables)."""
    K = 2 * int(np.floor(np.sqrt(3.0) * den / num)) + 1

    def acc(x, y, z):
        return num * num * (x * x + y * y + z * z) > 3 * den * den

    out = set()
    for o in itertools.product(range(-K, K + 1), repeat=3):
        if acc(*o) and any(not acc(*[(ci + oi) // 2 for ci, oi in zip(c, o)])
                           for c in itertools.product((0, 1), repeat=3)):
            out.add(tuple(abs(v) for v in o))
    return out


def test_same_level_offsets_that_occur_are_the_tabled_set():
    """Uniform lattice with complete neighbourhoods around the 8 level-5
    cells (16|17)^3 (one per position in the parent): the folded offsets of
    their same-level M2L partners in the oracle's walk are exactly the set;
    a clustered cloud only meets offsets of the set."""
    g = 32
    ijk = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)
    ijk = ijk[np.all(np.abs(ijk - 16) <= 10, axis=1)]
    x = np.concatenate([(ijk + 0.5) / g, [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]])
    _, _, t, _ = _tree(x, cap=1, max_depth=5)
    rng = np.random.default_rng(11)
    xc = np.concatenate([rng.random((1500, 3)), 0.3 + 0.02 * rng.standard_normal((1500, 3))]).clip(0, 1)
    _, _, tc, _ = _tree(xc, cap=4)
    for (num, den), rows in [((1, 2), 273), ((2, 5), 483)]:
        want = _same_level_offsets(num, den)
        assert len(want) == rows
        L = po.walk(t, num, den)
        seen = set()
        for A in np.where(t.level == 5)[0]:
            if not all(v in (16, 17) for v in (t.ix[A], t.iy[A], t.iz[A])):
                continue
            for B in L.m2l_src[L.m2l_off[A]:L.m2l_off[A + 1]]:
                if t.level[B] == 5:
                    seen.add((abs(int(t.ix[B]) - int(t.ix[A])), abs(int(t.iy[B]) - int(t.iy[A])),
                              abs(int(t.iz[B]) - int(t.iz[A]))))
        assert seen == want
        Lc = po.walk(tc, num, den)
        for A in range(len(tc.level)):
            for B in Lc.m2l_src[Lc.m2l_off[A]:Lc.m2l_off[A + 1]]:
                if tc.level[B] == tc.level[A]:
                    o = (abs(int(tc.ix[B]) - int(tc.ix[A])), abs(int(tc.iy[B]) - int(tc.iy[A])),
                         abs(int(tc.iz[B]) - int(tc.iz[A])))
                    assert o in want


# ---------------------------------------------------------------------------
# expansions
def test_derivative_tensor_vs_sympy():
    sp = pytest.importorskip("sympy")
    X, Y, Z = sp.symbols("x y z", real=True)
    eps = sp.Rational(3, 10)
    f = 1 / sp.sqrt(X ** 2 + Y ** 2 + Z ** 2 + eps ** 2)
    R = np.array([0.7, -1.3, 0.4])
    p = 4
    D = po.derivs(p, R, eps=0.3)
    mi = po.multi_indices(p)
    for k, (a, b, c) in enumerate(mi):
        expr = sp.diff(f, X, int(a), Y, int(b), Z, int(c)) if a + b + c else f
        val = float(expr.subs({X: R[0], Y: R[1], Z: R[2]}))
        assert abs(D[k] - val) <= 1e-12 * max(1.0, abs(val)), (a, b, c)


def test_p2m_matches_reference_moment_definitions():
    """moment.cl:43-54: charge, dipole, 3 q (yz, xz, xy), q (2x^2-y^2-z^2, ...)."""
    rng = np.random.default_rng(8)
    x = rng.random((30, 3))
    q = rng.random(30)
    z = np.array([0.4, 0.6, 0.5])
    M = po.p2m(2, z, x, q)
    mi = [tuple(m) for m in po.multi_indices(2)]
    Mi = dict(zip(mi, M))
    r = x - z
    assert np.isclose(Mi[(0, 0, 0)], q.sum())
    assert np.allclose([Mi[(1, 0, 0)], Mi[(0, 1, 0)], Mi[(0, 0, 1)]], (q[:, None] * r).sum(0))
    cross = 3 * np.array([(q * r[:, 1] * r[:, 2]).sum(), (q * r[:, 0] * r[:, 2]).sum(), (q * r[:, 0] * r[:, 1]).sum()])
    assert np.allclose(cross, 3 * np.array([Mi[(0, 1, 1)], Mi[(1, 0, 1)], Mi[(1, 1, 0)]]))
    tr = (q[:, None] * (2 * r ** 2 - (r ** 2).sum(1, keepdims=True) + r ** 2)).sum(0)
    xx, yy, zz = 2 * Mi[(2, 0, 0)], 2 * Mi[(0, 2, 0)], 2 * Mi[(0, 0, 2)]
    assert np.allclose(tr, [2 * xx - yy - zz, 2 * yy - xx - zz, 2 * zz - xx - yy])


def test_m2m_translation_identity():
    rng = np.random.default_rng(9)
    p = 5
    zc = [np.array([0.25, 0.25, 0.75]), np.array([0.75, 0.25, 0.25])]
    zp = np.array([0.5, 0.5, 0.5])
    Mp = np.zeros(po.ncoef(p))
    xs, qs = [], []
    for c in zc:
        x = c + 0.25 * (rng.random((10, 3)) - 0.5)
        q = rng.random(10)
        Mp += po.m2m(p, c, zp, po.p2m(p, c, x, q))
        xs.append(x); qs.append(q)
    direct = po.p2m(p, zp, np.concatenate(xs), np.concatenate(qs))
    assert np.allclose(Mp, direct, rtol=1e-12, atol=1e-15)


def test_l2l_is_exact_for_polynomials():
    rng = np.random.default_rng(10)
    p = 4
    Lp = rng.standard_normal(po.ncoef(p))
    zp = np.array([0.0, 0.0, 0.0])
    zc = np.array([0.1, -0.2, 0.05])
    Lc = po.l2l(p, zp, zc, Lp)
    for _ in range(5):
        x = zc + 0.1 * rng.standard_normal(3)
        a = po.l2p(p, zp, Lp, x)
        b = po.l2p(p, zc, Lc, x)
        assert np.isclose(a[0], b[0], rtol=1e-12) and np.allclose(a[1], b[1], rtol=1e-11)


def test_m2l_converges_to_direct_with_sign_and_gradient():
    rng = np.random.default_rng(11)
    zB = np.array([0.0, 0.0, 0.0])
    zA = np.array([2.0, 1.0, -0.5])
    xs = zB + 0.3 * (rng.random((40, 3)) - 0.5)
    qs = rng.random(40)
    xt = zA + 0.3 * (rng.random(3) - 0.5)
    a_d, p_d = po.direct(xs, qs, xt=xt[None], eps=0.05)
    errs = []
    for p in range(1, 8):
        L = po.m2l(p, zA, zB, po.p2m(p, zB, xs, qs), eps=0.05)
        phi, g = po.l2p(p, zA, L, xt)
        assert np.sign(-phi) == np.sign(p_d[0])
        errs.append(np.linalg.norm(g - a_d[0]) / np.linalg.norm(a_d[0]))
    assert errs[-1] < 1e-5 and all(e2 < e1 for e1, e2 in zip(errs[1:], errs[2:]))


def test_fmm_converges_to_direct():
    rng = np.random.default_rng(12)
    x = rng.random((1500, 3))
    q = np.full(1500, 1 / 1500)
    keys, o, t, box = _tree(x, cap=12)
    L = po.walk(t, 1, 2)
    ad, pd = po.direct(x[o], q, eps=1e-2)
    prev = None
    for p in [2, 4, 6]:
        a, pot = po.fmm(t, L, p, x[o], q, box, eps=1e-2)
        e = np.sqrt(np.mean(np.sum((a - ad) ** 2, 1)) / np.mean(np.sum(ad ** 2, 1)))
        assert prev is None or e < prev / 4
        prev = e
    assert prev < 3e-5


# ---------------------------------------------------------------------------
# integrators
def test_kdk_kepler_energy_and_reversibility():
    x = np.array([[0.5, 0.0, 0.0], [-0.5, 0.0, 0.0]])
    v = np.array([[0.0, 0.5, 0.0], [0.0, -0.5, 0.0]])
    m = np.array([0.5, 0.5])

    def energy(x, v):
        r = np.linalg.norm(x[0] - x[1])
        return 0.5 * (m * (v ** 2).sum(1)).sum() - m[0] * m[1] / r

    e0 = energy(x, v)
    x1, v1 = po.integrate_direct(x, v, m, m, 1.0, 0.0, 1e-3, 6283, 0)
    assert abs(energy(x1, v1) - e0) < 1e-6 * abs(e0)
    x2, v2 = po.integrate_direct(x1, -v1, m, m, 1.0, 0.0, 1e-3, 6283, 0)
    assert np.allclose(x2, x, atol=1e-9) and np.allclose(-v2, v, atol=1e-9)


def test_euler_and_kick_drift_one_step_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v = np.array([[0.0, 0.1, 0.0], [0.0, 0.0, 0.2]])
    m = np.array([1.0, 2.0])
    dt = 0.01
    a0 = np.array([[2.0, 0, 0], [-1.0, 0, 0]])
    xe, ve = po.integrate_direct(x, v, m, m, 1.0, 0.0, dt, 1, 1)
    assert np.allclose(xe, x + v * dt) and np.allclose(ve, v + a0 * dt)
    xk, vk = po.integrate_direct(x, v, m, m, 1.0, 0.0, dt, 1, 2)
    assert np.allclose(vk, v + a0 * dt) and np.allclose(xk, x + (v + a0 * dt) * dt)


def test_m2l_extended_truncation_is_the_higher_order_operator_restricted():
    """orc_m2l(p, q=p+1) with every term keeps |k| + |n| <= p+1, which is the
    order-(p+1) M2L applied to M padded with zeros and read at |k| <= p.  The
    low-degree set keeps the terms of total degree <= p and those of degree
    p+1 with |n| = 1 or |k| = 1; against the same operator: local orders
    |k| <= 1 receive every moment and equal it; on a monopole + dipole every
    local order equals it; and on moments of order >= 2 the orders |k| >= 2
    equal the total-degree truncation q = p.  (The order-(p+1) operator is
    pinned by the sympy derivative test and the convergence tests above.)"""
    rng = np.random.default_rng(21)
    zA, zB = np.array([0.3, -0.2, 0.1]), np.array([2.1, 0.7, -1.4])
    for p in [1, 2, 3, 4, 5]:
        nc = po.ncoef(p)
        M = rng.standard_normal(nc)
        low = np.zeros(nc)
        low[:4] = M[:4]            # monopole and dipole
        high = M - low             # orders >= 2

        def full(Mx):
            pad = np.zeros(po.ncoef(p + 1))
            pad[:nc] = Mx
            return po.m2l(p + 1, zA, zB, pad, eps=0.03)[:nc]

        ref = full(M)
        tol = 1e-13 * np.abs(ref).max()
        every = po.m2l(p, zA, zB, M, eps=0.03, m2l_order=p + 1, m2l_terms="all")
        assert np.allclose(every, ref, rtol=1e-13, atol=tol)
        ext = po.m2l(p, zA, zB, M, eps=0.03, m2l_order=p + 1, m2l_terms="low")
        assert np.allclose(ext[:4], ref[:4], rtol=1e-13, atol=tol)
        ext_low = po.m2l(p, zA, zB, low, eps=0.03, m2l_order=p + 1, m2l_terms="low")
        assert np.allclose(ext_low, full(low), rtol=1e-13, atol=tol)
        ext_high = po.m2l(p, zA, zB, high, eps=0.03, m2l_order=p + 1, m2l_terms="low")
        tot_high = po.m2l(p, zA, zB, high, eps=0.03)
        assert np.allclose(ext_high[4:], tot_high[4:], rtol=1e-13, atol=tol)
        assert np.allclose(ext, ext_low + ext_high, rtol=1e-12, atol=tol)
        # and it differs from the total-degree truncation (p >= 1), and from
        # the full degree-(p+1) set where that has more terms (p >= 3)
        assert not np.allclose(ext, po.m2l(p, zA, zB, M, eps=0.03))
        if p >= 3:
            assert not np.allclose(ext, every)


def test_fmm_extended_m2l_is_more_accurate():
    rng = np.random.default_rng(12)
    x = rng.random((1500, 3))
    q = np.full(1500, 1 / 1500)
    keys, o, t, box = _tree(x, cap=12)
    L = po.walk(t, 1, 2)
    ad, _ = po.direct(x[o], q, eps=1e-2)

    def err(p, qq, terms="all"):
        a, _ = po.fmm(t, L, p, x[o], q, box, eps=1e-2, m2l_order=qq, m2l_terms=terms)
        return np.sqrt(np.mean(np.sum((a - ad) ** 2, 1)) / np.mean(np.sum(ad ** 2, 1)))

    for p in [2, 3, 4]:
        e0, e_low, e_all, e2 = err(p, p), err(p, p + 1, "low"), err(p, p + 1), err(p + 1, p + 1)
        assert e2 < e_all < e0 and e_all <= e_low < e0, (p, e0, e_low, e_all, e2)


@pytest.mark.parametrize("ext,terms", [(0, "all"), (1, "all"), (1, "low")])
def test_fmm_vjp_is_the_exact_transpose_of_the_fmm(ext, terms):
    """orc_fmm_vjp against FP64 central differences of orc_fmm with the tree
    and lists held fixed, and against the direct-sum adjoint (to the FMM's
    truncation accuracy); M2L truncated at q = p + ext with either term set."""
    rng = np.random.default_rng(13)
    n = 600
    x = np.concatenate([rng.random((500, 3)), 0.3 + 0.05 * rng.random((100, 3))])
    m = rng.random(n) / n + 1e-4
    keys, o, t, box = _tree(x, cap=8)
    L = po.walk(t, 1, 2)
    xs, ms = x[o], m[o]
    w = rng.standard_normal((n, 3))
    p, eps = 4, 1e-2
    gx, gm = po.fmm_vjp(t, L, p, xs, ms, w, box, eps=eps, m2l_order=p + ext, m2l_terms=terms)

    def f(xx, mm):
        a, _ = po.fmm(t, L, p, xx, mm, box, eps=eps, m2l_order=p + ext, m2l_terms=terms)
        return float((a * w).sum())

    h = 1e-6
    for i in rng.choice(n, 12, replace=False):
        for k in range(3):
            xp = xs.copy(); xp[i, k] += h
            xm = xs.copy(); xm[i, k] -= h
            fd = (f(xp, ms) - f(xm, ms)) / (2 * h)
            assert abs(fd - gx[i, k]) < 1e-6 * np.abs(gx).max(), (i, k, fd, gx[i, k])
        mp = ms.copy(); mp[i] += h * 1e-3
        mm_ = ms.copy(); mm_[i] -= h * 1e-3
        fd = (f(xs, mp) - f(xs, mm_)) / (2 * h * 1e-3)
        assert abs(fd - gm[i]) < 1e-6 * np.abs(gm).max(), (i, fd, gm[i])
    gxd, gmd = po.direct_vjp(xs, ms, w, eps=eps)
    assert np.sqrt(np.sum((gx - gxd) ** 2) / np.sum(gxd ** 2)) < 2e-3
    assert np.sqrt(np.sum((gm - gmd) ** 2) / np.sum(gmd ** 2)) < 2e-3


def test_m2l_term_set_is_symmetric_with_the_counted_size():
    """The kept (k, n) terms of orc_m2l, probed column by column with unit
    moments: 210 at p = q = 4, 420 at p = 4, q = 5 with every term and 300 with
    the low-degree set (the counts bench.py turns into FLOP per pair), and the
    pattern is symmetric in (k, n), which is what lets the adjoint reuse the
    same operator."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "bench_mod", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    zA, zB = np.array([0.3, -0.2, 0.1]), np.array([2.1, 0.7, -1.4])
    p = 4
    nc = po.ncoef(p)
    for q, terms, count in [(4, "all", 210), (5, "all", 420), (5, "low", 300)]:
        T = np.zeros((nc, nc))
        for n in range(nc):
            e = np.zeros(nc)
            e[n] = 1.0
            T[:, n] = po.m2l(p, zA, zB, e, eps=0.03, m2l_order=q, m2l_terms=terms)
        nz = np.abs(T) > 0
        assert int(nz.sum()) == count
        assert np.array_equal(nz, nz.T)
        assert bench.m2l_flop(p, q, terms) == 2 * count
