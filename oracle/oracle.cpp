// oracle.cpp -- CPU restatement of the nbody FMM force pipeline (TEST ONLY).
//
// This file is test infrastructure: it is the checker that tests/, smoke()
// and the cpu_baseline leg of bench.py compare the HIP library against.  The
// product never links it.  See oracle.h for the reference file:line map.
//
// Parity status of each function, and how it is pinned (tests/test_oracle.py):
//   orc_direct ........ closed forms (two-body, Newton's third law, symmetric
//                       shells), brute force in numpy on tiny inputs.
//   orc_morton ........ closed form (floor quantisation + bit interleave),
//                       brute force.
//   orc_tree_* ........ invariants (ranges partition [0,n), every particle's key
//                       inside its cell, capacity rule, child ordering), and a
//                       brute-force per-level enumeration in numpy.
//   orc_walk .......... brute-force coverage: every ordered particle pair (i,j)
//                       is covered exactly once by a P2P pair of its leaves or
//                       an M2L pair of ancestors; MAC restated in numpy floats
//                       as in interaction.cl:63-79; lattice counts (179 near
//                       cells, 1253 M2L partners at theta=0.5).
//   expansions ........ derivative tensor vs sympy; P2M moments vs moment.cl
//                       monopole/dipole/quadrupole closed forms; M2M/L2L
//                       translation identities; convergence in p to the direct
//                       sum.
//   orc_fmm ........... converges to orc_direct as p grows / theta shrinks.
//   integrators ....... Kepler two-body energy/period, time reversibility.
//   naive compat ...... bit-exact against golden vectors made by compiling
//                       the reference's src/naive_simulation.cpp.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Multi-index bookkeeping.  Coefficients are ordered by total degree, then by
// descending nx, then descending ny (nz = d - nx - ny).
struct MultiIndex {
    int p = 0;
    std::vector<std::array<int, 3>> mi;
    std::vector<int> lut;  // (p+1)^3 -> index or -1
    explicit MultiIndex(int p_) : p(p_), lut((p_ + 1) * (p_ + 1) * (p_ + 1), -1) {
        for (int d = 0; d <= p; ++d)
            for (int nx = d; nx >= 0; --nx)
                for (int ny = d - nx; ny >= 0; --ny) {
                    int nz = d - nx - ny;
                    lut[(nx * (p + 1) + ny) * (p + 1) + nz] = (int)mi.size();
                    mi.push_back({nx, ny, nz});
                }
    }
    int size() const { return (int)mi.size(); }
    int at(int nx, int ny, int nz) const {
        if (nx < 0 || ny < 0 || nz < 0 || nx + ny + nz > p) return -1;
        return lut[(nx * (p + 1) + ny) * (p + 1) + nz];
    }
};

double factorial(int n) {
    double f = 1;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

// x^n / n! for all multi-indices
void monomials(const MultiIndex& T, const double* a, std::vector<double>& out) {
    out.resize(T.size());
    for (int k = 0; k < T.size(); ++k) {
        const auto& n = T.mi[k];
        double v = 1;
        for (int ax = 0; ax < 3; ++ax) v *= std::pow(a[ax], n[ax]) / factorial(n[ax]);
        out[k] = v;
    }
}

// Derivatives of f(R) = s^(-1/2), s = |R|^2 + eps^2.  From s d_i f + R_i f = 0
// differentiated by Leibniz (only s, 2R_j and 2 survive as derivatives of s):
//   s D_n = -[ sum_{j!=i} (2 n_j R_j D_{n-e_j} + n_j(n_j-1) D_{n-2e_j})
//              + (2n_i-1) R_i D_{n-e_i} + (n_i-1)^2 D_{n-2e_i} ]
// with i any axis where n_i > 0.
void derivs(const MultiIndex& T, const double* R, double eps, std::vector<double>& D) {
    D.assign(T.size(), 0.0);
    double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps * eps;
    D[0] = 1.0 / std::sqrt(s);
    for (int k = 1; k < T.size(); ++k) {
        auto n = T.mi[k];
        int i = n[0] > 0 ? 0 : (n[1] > 0 ? 1 : 2);
        double acc = 0;
        for (int j = 0; j < 3; ++j) {
            auto m1 = n;
            m1[j] -= 1;
            auto m2 = n;
            m2[j] -= 2;
            int i1 = T.at(m1[0], m1[1], m1[2]);
            int i2 = T.at(m2[0], m2[1], m2[2]);
            double d1 = i1 >= 0 ? D[i1] : 0.0, d2 = i2 >= 0 ? D[i2] : 0.0;
            if (j != i) {
                acc += 2.0 * n[j] * R[j] * d1 + double(n[j]) * (n[j] - 1) * d2;
            } else {
                acc += (2.0 * n[i] - 1.0) * R[i] * d1 + double(n[i] - 1) * (n[i] - 1) * d2;
            }
        }
        D[k] = -acc / s;
    }
}

// ---------------------------------------------------------------------------
struct Tree {
    int bits = 21, cap = 8, max_depth = 21, depth = 0;
    size_t n = 0, nsys = 1;
    std::vector<int32_t> level, sys, ix, iy, iz, begin, count, first_child, nchild, parent;
};

struct Lists {
    std::vector<int64_t> m2l_off, p2p_off;
    std::vector<int32_t> m2l_src, p2p_src;
    int64_t unordered_m2l = 0, unordered_p2p = 0, visited = 0;
};

uint64_t spread3(uint32_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= (uint64_t)((v >> b) & 1u) << (3 * b);
    return r;
}
uint32_t compact3(uint64_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 21; ++b) r |= (uint32_t)((v >> (3 * b)) & 1u) << b;
    return r;
}

}  // namespace

struct orc_tree { Tree t; };
struct orc_lists { Lists l; };

extern "C" {

// ---------------------------------------------------------------------------
// Direct sum.  Softened field as field.cl:17-32: r = x_b - x_a,
// |r|_eps = sqrt(r.r + eps^2), field on a = -q_b * F * r / |r|^3 with F = -1,
// i.e. +G q_b r / |r|^3 (attractive).
void orc_direct(size_t ns, const double* xs, const double* qs, size_t nt,
                const double* xt, const int64_t* self, double G, double eps,
                double* acc, double* pot) {
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < (long long)nt; ++i) {
        double ax = 0, ay = 0, az = 0, ph = 0;
        const double tx = xt[3 * i], ty = xt[3 * i + 1], tz = xt[3 * i + 2];
        const int64_t skip = self ? self[i] : -1;
        for (size_t j = 0; j < ns; ++j) {
            if ((int64_t)j == skip) continue;
            double rx = xs[3 * j] - tx, ry = xs[3 * j + 1] - ty, rz = xs[3 * j + 2] - tz;
            double s2 = rx * rx + ry * ry + rz * rz + e2;
            double inv = 1.0 / std::sqrt(s2);
            double w = qs[j] * inv * inv * inv;
            ax += w * rx;
            ay += w * ry;
            az += w * rz;
            ph += qs[j] * inv;
        }
        acc[3 * i] = G * ax;
        acc[3 * i + 1] = G * ay;
        acc[3 * i + 2] = G * az;
        if (pot) pot[i] = -G * ph;
    }
}

// gx_k = G sum_{j!=k} T(x_j - x_k) (m_k w_j - m_j w_k),  T(r) = I/s^3 - 3 r r^T/s^5
// gm_j = G sum_{i!=j} w_i . (x_j - x_i)/s^3
void orc_direct_vjp(size_t n, const double* x, const double* m, const double* w,
                    double G, double eps, double* gx, double* gm) {
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long k = 0; k < (long long)n; ++k) {
        double g[3] = {0, 0, 0}, gmk = 0;
        for (size_t j = 0; j < n; ++j) {
            if ((long long)j == k) continue;
            double r[3] = {x[3 * j] - x[3 * k], x[3 * j + 1] - x[3 * k + 1], x[3 * j + 2] - x[3 * k + 2]};
            double s2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + e2;
            double inv = 1.0 / std::sqrt(s2);
            double i3 = inv * inv * inv, i5 = i3 * inv * inv;
            double u[3];
            for (int a = 0; a < 3; ++a) u[a] = m[k] * w[3 * j + a] - m[j] * w[3 * k + a];
            double ru = r[0] * u[0] + r[1] * u[1] + r[2] * u[2];
            for (int a = 0; a < 3; ++a) g[a] += u[a] * i3 - 3.0 * r[a] * ru * i5;
            // mass gradient: d a_j / d m_k contributes w_j . (x_k - x_j)/s^3 = -(w_j . r) i3
            gmk += -(w[3 * j] * r[0] + w[3 * j + 1] * r[1] + w[3 * j + 2] * r[2]) * i3;
        }
        for (int a = 0; a < 3; ++a) gx[3 * k + a] = G * g[a];
        gm[k] = G * gmk;
    }
}

// ---------------------------------------------------------------------------
void orc_morton(size_t n, const double* x, size_t nsys, const uint32_t* sys_off,
                int bits, uint64_t* keys, double* box_out) {
    for (size_t s = 0; s < nsys; ++s) {
        size_t b = sys_off ? sys_off[s] : 0, e = sys_off ? sys_off[s + 1] : n;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], x[3 * i + a]);
                hi[a] = std::max(hi[a], x[3 * i + a]);
            }
        double size = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        if (!(size > 0)) size = 1.0;
        double inv = std::ldexp(1.0, bits) / size;
        const double qmax = std::ldexp(1.0, bits) - 1.0;
        for (size_t i = b; i < e; ++i) {
            uint32_t q[3];
            for (int a = 0; a < 3; ++a) {
                double f = std::floor((x[3 * i + a] - lo[a]) * inv);
                q[a] = (uint32_t)std::min(f, qmax);
            }
            uint64_t m = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
            keys[i] = ((uint64_t)s << (3 * bits)) | m;
        }
        box_out[4 * s + 0] = lo[0];
        box_out[4 * s + 1] = lo[1];
        box_out[4 * s + 2] = lo[2];
        box_out[4 * s + 3] = size;
    }
}

// ---------------------------------------------------------------------------
// Top-down recursive octree build over sorted keys (independent of the GPU's
// particle-parallel emission).  Children are the non-empty octants in digit
// order; a cell splits iff count > cap and level < max_depth.
orc_tree* orc_tree_build(size_t n, const uint64_t* keys, size_t nsys,
                         const uint32_t* sys_off, int bits, int cap, int max_depth) {
    auto* out = new orc_tree;
    Tree& t = out->t;
    t.bits = bits; t.cap = cap; t.max_depth = max_depth; t.n = n; t.nsys = nsys;
    struct Cell { int32_t sys, level, begin, count; uint64_t prefix; int32_t parent_slot; };
    // per-level lists, filled in DFS (Morton) order
    std::vector<std::vector<Cell>> levels(max_depth + 1);
    std::vector<std::vector<std::vector<int>>> children_of(max_depth + 1);  // per level, per cell -> child slots
    struct Frame { int sys, level, begin, end; uint64_t prefix; int parent_slot; };
    for (size_t s = 0; s < nsys; ++s) {
        int b = sys_off ? (int)sys_off[s] : 0, e = sys_off ? (int)sys_off[s + 1] : (int)n;
        if (e <= b) continue;
        std::vector<Frame> stack;
        stack.push_back({(int)s, 0, b, e, 0, -1});
        while (!stack.empty()) {
            Frame f = stack.back();
            stack.pop_back();
            int slot = (int)levels[f.level].size();
            levels[f.level].push_back({f.sys, f.level, f.begin, f.end - f.begin, f.prefix, f.parent_slot});
            children_of[f.level].emplace_back();
            if (f.level > 0) children_of[f.level - 1][f.parent_slot].push_back(slot);
            if (f.end - f.begin > cap && f.level < max_depth) {
                int shift = 3 * (bits - f.level - 1);
                std::vector<Frame> kids;
                int i = f.begin;
                while (i < f.end) {
                    uint64_t digit = (keys[i] >> shift) & 7u;
                    int j = i;
                    while (j < f.end && ((keys[j] >> shift) & 7u) == digit) ++j;
                    kids.push_back({f.sys, f.level + 1, i, j, (f.prefix << 3) | digit, slot});
                    i = j;
                }
                // push in reverse so that children are visited in digit order
                for (auto it = kids.rbegin(); it != kids.rend(); ++it) stack.push_back(*it);
            }
        }
    }
    std::vector<int> level_base(max_depth + 2, 0);
    for (int l = 0; l <= max_depth; ++l) level_base[l + 1] = level_base[l] + (int)levels[l].size();
    size_t nn = level_base[max_depth + 1];
    t.level.resize(nn); t.sys.resize(nn); t.ix.resize(nn); t.iy.resize(nn); t.iz.resize(nn);
    t.begin.resize(nn); t.count.resize(nn); t.first_child.resize(nn); t.nchild.resize(nn); t.parent.resize(nn);
    t.depth = 0;
    for (int l = 0; l <= max_depth; ++l) {
        if (!levels[l].empty()) t.depth = l;
        for (size_t k = 0; k < levels[l].size(); ++k) {
            const Cell& c = levels[l][k];
            size_t id = level_base[l] + k;
            t.level[id] = l;
            t.sys[id] = c.sys;
            // prefix holds l octal digits, x bit highest in each digit
            uint64_t m = c.prefix;
            t.ix[id] = (int32_t)compact3(m >> 2);
            t.iy[id] = (int32_t)compact3(m >> 1);
            t.iz[id] = (int32_t)compact3(m);
            t.begin[id] = c.begin;
            t.count[id] = c.count;
            t.parent[id] = l > 0 ? level_base[l - 1] + c.parent_slot : -1;
            const auto& ch = children_of[l][k];
            t.nchild[id] = (int32_t)ch.size();
            t.first_child[id] = ch.empty() ? -1 : level_base[l + 1] + ch.front();
            for (size_t q = 1; q < ch.size(); ++q)
                if (ch[q] != ch[q - 1] + 1) throw std::logic_error("children not contiguous");
        }
    }
    return out;
}

void orc_tree_free(orc_tree* t) { delete t; }
size_t orc_tree_num_nodes(const orc_tree* t) { return t->t.level.size(); }
int orc_tree_depth(const orc_tree* t) { return t->t.depth; }
void orc_tree_nodes(const orc_tree* tr, int32_t* level, int32_t* sys, int32_t* ix,
                    int32_t* iy, int32_t* iz, int32_t* begin, int32_t* count,
                    int32_t* first_child, int32_t* nchild, int32_t* parent) {
    const Tree& t = tr->t;
    size_t nn = t.level.size();
    std::copy(t.level.begin(), t.level.end(), level);
    std::copy(t.sys.begin(), t.sys.end(), sys);
    std::copy(t.ix.begin(), t.ix.end(), ix);
    std::copy(t.iy.begin(), t.iy.end(), iy);
    std::copy(t.iz.begin(), t.iz.end(), iz);
    std::copy(t.begin.begin(), t.begin.end(), begin);
    std::copy(t.count.begin(), t.count.end(), count);
    std::copy(t.first_child.begin(), t.first_child.end(), first_child);
    std::copy(t.nchild.begin(), t.nchild.end(), nchild);
    std::copy(t.parent.begin(), t.parent.end(), parent);
    (void)nn;
}

// ---------------------------------------------------------------------------
// Reference opening criterion (interaction.cl:63-79):
//   approx = (a != b) && 0.75 (d_a + d_b)^2 / |c_b - c_a|^2 < theta^2
// restated exactly in integers.  Centres and sides are integers in units of
// half the finest cell: c = (2 i + 1) 2^(bits-l), d = 2^(bits-l+1), and
// theta = num/den gives  3 den^2 (d_a+d_b)^2 < 4 num^2 |dc|^2.
static bool mac_approx(const Tree& t, int a, int b, int num, int den) {
    if (a == b) return false;
    typedef unsigned __int128 u128;
    int64_t ca[3], cb[3];
    int la = t.level[a], lb = t.level[b];
    const int32_t* ia[3] = {&t.ix[a], &t.iy[a], &t.iz[a]};
    const int32_t* ib[3] = {&t.ix[b], &t.iy[b], &t.iz[b]};
    for (int k = 0; k < 3; ++k) {
        ca[k] = (2 * (int64_t)*ia[k] + 1) << (t.bits - la);
        cb[k] = (2 * (int64_t)*ib[k] + 1) << (t.bits - lb);
    }
    u128 dsq = 0;
    for (int k = 0; k < 3; ++k) {
        int64_t d = cb[k] - ca[k];
        dsq += (u128)((__int128)d * d);
    }
    u128 side = ((u128)1 << (t.bits - la + 1)) + ((u128)1 << (t.bits - lb + 1));
    u128 lhs = (u128)3 * den * den * side * side;
    u128 rhs = (u128)4 * num * num * dsq;
    return lhs < rhs;
}

// Restates find_interactions (interaction.cl:42-92) and the host classification
// (open_cl_simulation.cpp:252-265) as a recursion over unordered pairs seeded
// with (root, root) (open_cl_simulation.h:127-133).  A childless root gives a
// P2P self pair here (the reference drops the (0,0) pair, defect A7).
orc_lists* orc_walk(const orc_tree* tr, int num, int den) {
    const Tree& t = tr->t;
    auto* out = new orc_lists;
    Lists& L = out->l;
    size_t nn = t.level.size();
    std::vector<std::vector<int32_t>> m2l(nn), p2p(nn);
    std::vector<std::pair<int, int>> stack;
    for (size_t r = 0; r < nn && t.level[r] == 0; ++r) stack.push_back({(int)r, (int)r});
    while (!stack.empty()) {
        auto [a, b] = stack.back();
        stack.pop_back();
        ++L.visited;
        std::vector<int> ca, cb;
        if (t.nchild[a]) for (int k = 0; k < t.nchild[a]; ++k) ca.push_back(t.first_child[a] + k);
        else ca.push_back(a);
        if (t.nchild[b]) for (int k = 0; k < t.nchild[b]; ++k) cb.push_back(t.first_child[b] + k);
        else cb.push_back(b);
        for (size_t ib = 0; ib < cb.size(); ++ib)
            for (size_t ia = 0; ia < ca.size(); ++ia) {
                if (a == b && ib > ia) continue;  // interaction.cl:57
                int A = ca[ia], B = cb[ib];
                bool approx = mac_approx(t, A, B, num, den);
                bool reduce = !approx && (t.nchild[A] || t.nchild[B]);
                if (reduce) {
                    stack.push_back({A, B});
                } else if (!approx) {
                    ++L.unordered_p2p;
                    p2p[A].push_back(B);
                    if (A != B) p2p[B].push_back(A);
                } else {
                    ++L.unordered_m2l;
                    m2l[A].push_back(B);
                    m2l[B].push_back(A);
                }
            }
    }
    L.m2l_off.assign(nn + 1, 0);
    L.p2p_off.assign(nn + 1, 0);
    for (size_t i = 0; i < nn; ++i) {
        std::sort(m2l[i].begin(), m2l[i].end());
        std::sort(p2p[i].begin(), p2p[i].end());
        if (std::adjacent_find(m2l[i].begin(), m2l[i].end()) != m2l[i].end() ||
            std::adjacent_find(p2p[i].begin(), p2p[i].end()) != p2p[i].end())
            throw std::logic_error("duplicate directed pair");
        L.m2l_off[i + 1] = L.m2l_off[i] + (int64_t)m2l[i].size();
        L.p2p_off[i + 1] = L.p2p_off[i] + (int64_t)p2p[i].size();
    }
    L.m2l_src.reserve(L.m2l_off[nn]);
    L.p2p_src.reserve(L.p2p_off[nn]);
    for (size_t i = 0; i < nn; ++i) {
        L.m2l_src.insert(L.m2l_src.end(), m2l[i].begin(), m2l[i].end());
        L.p2p_src.insert(L.p2p_src.end(), p2p[i].begin(), p2p[i].end());
    }
    return out;
}

void orc_lists_free(orc_lists* l) { delete l; }
void orc_lists_sizes(const orc_lists* l, int64_t* s) {
    s[0] = (int64_t)l->l.m2l_src.size();
    s[1] = (int64_t)l->l.p2p_src.size();
    s[2] = l->l.unordered_m2l;
    s[3] = l->l.unordered_p2p;
    s[4] = l->l.visited;
}
void orc_lists_get(const orc_lists* l, int64_t* m2l_off, int32_t* m2l_src,
                   int64_t* p2p_off, int32_t* p2p_src) {
    std::copy(l->l.m2l_off.begin(), l->l.m2l_off.end(), m2l_off);
    std::copy(l->l.m2l_src.begin(), l->l.m2l_src.end(), m2l_src);
    std::copy(l->l.p2p_off.begin(), l->l.p2p_off.end(), p2p_off);
    std::copy(l->l.p2p_src.begin(), l->l.p2p_src.end(), p2p_src);
}

// ---------------------------------------------------------------------------
int orc_ncoef(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

void orc_multi_indices(int p, int32_t* mi) {
    MultiIndex T(p);
    for (int k = 0; k < T.size(); ++k)
        for (int a = 0; a < 3; ++a) mi[3 * k + a] = T.mi[k][a];
}

void orc_derivs(int p, const double* R, double eps, double* D) {
    MultiIndex T(p);
    std::vector<double> d;
    derivs(T, R, eps, d);
    std::copy(d.begin(), d.end(), D);
}

void orc_p2m(int p, const double* z, size_t n, const double* x, const double* q, double* M) {
    MultiIndex T(p);
    std::vector<double> mono;
    std::fill(M, M + T.size(), 0.0);
    for (size_t j = 0; j < n; ++j) {
        double b[3] = {x[3 * j] - z[0], x[3 * j + 1] - z[1], x[3 * j + 2] - z[2]};
        monomials(T, b, mono);
        for (int k = 0; k < T.size(); ++k) M[k] += q[j] * mono[k];
    }
}

void orc_m2m(int p, const double* zc, const double* zp, const double* Mc, double* Mp) {
    MultiIndex T(p);
    double d[3] = {zc[0] - zp[0], zc[1] - zp[1], zc[2] - zp[2]};
    std::vector<double> mono;
    monomials(T, d, mono);
    for (int k = 0; k < T.size(); ++k) {
        const auto& n = T.mi[k];
        double acc = 0;
        for (int mx = 0; mx <= n[0]; ++mx)
            for (int my = 0; my <= n[1]; ++my)
                for (int mz = 0; mz <= n[2]; ++mz) {
                    int im = T.at(mx, my, mz), ir = T.at(n[0] - mx, n[1] - my, n[2] - mz);
                    acc += Mc[im] * mono[ir];
                }
        Mp[k] += acc;
    }
}

void orc_m2l(int p, int qc, const double* zA, const double* zB, double eps, const double* M, double* L) {
    // qc = q | 16: every term of total degree <= q; qc = q: the low-degree set
    const int q = qc & 15;
    const bool all = (qc & 16) != 0;
    MultiIndex T(p), TQ(q);
    double R[3] = {zA[0] - zB[0], zA[1] - zB[1], zA[2] - zB[2]};
    std::vector<double> D;
    derivs(TQ, R, eps, D);
    for (int k = 0; k < T.size(); ++k) {
        const auto& kk = T.mi[k];
        double acc = 0;
        for (int n = 0; n < T.size(); ++n) {
            const auto& nn = T.mi[n];
            const int dk = kk[0] + kk[1] + kk[2], dn = nn[0] + nn[1] + nn[2];
            // low-degree set: beyond total degree p only the terms with
            // min(|k|, |n|) <= q - p (both sets stop at total degree q)
            if (!all && dk + dn > p && std::min(dk, dn) > q - p) continue;
            int s = TQ.at(kk[0] + nn[0], kk[1] + nn[1], kk[2] + nn[2]);
            if (s < 0) continue;
            double sign = (dn & 1) ? -1.0 : 1.0;
            acc += sign * D[s] * M[n];
        }
        L[k] += acc;
    }
}

void orc_l2l(int p, const double* zp, const double* zc, const double* Lp, double* Lc) {
    MultiIndex T(p);
    double d[3] = {zc[0] - zp[0], zc[1] - zp[1], zc[2] - zp[2]};
    std::vector<double> mono;
    monomials(T, d, mono);
    for (int j = 0; j < T.size(); ++j) {
        const auto& jj = T.mi[j];
        double acc = 0;
        for (int m = 0; m < T.size(); ++m) {
            const auto& mm = T.mi[m];
            int s = T.at(jj[0] + mm[0], jj[1] + mm[1], jj[2] + mm[2]);
            if (s < 0) continue;
            acc += Lp[s] * mono[m];
        }
        Lc[j] += acc;
    }
}

void orc_l2p(int p, const double* z, const double* L, const double* x, double* phi, double* g) {
    MultiIndex T(p);
    double a[3] = {x[0] - z[0], x[1] - z[1], x[2] - z[2]};
    std::vector<double> mono;
    monomials(T, a, mono);
    double ph = 0, gg[3] = {0, 0, 0};
    for (int k = 0; k < T.size(); ++k) {
        ph += L[k] * mono[k];
        const auto& kk = T.mi[k];
        for (int ax = 0; ax < 3; ++ax) {
            int e[3] = {kk[0], kk[1], kk[2]};
            e[ax] += 1;
            int s = T.at(e[0], e[1], e[2]);
            if (s >= 0) gg[ax] += L[s] * mono[k];
        }
    }
    *phi = ph;
    for (int ax = 0; ax < 3; ++ax) g[ax] = gg[ax];
}

// ---------------------------------------------------------------------------
static void cell_centre(const Tree& t, const double* box, int node, double* c) {
    int s = t.sys[node], l = t.level[node];
    double h = box[4 * s + 3] / std::ldexp(1.0, l);
    c[0] = box[4 * s + 0] + (t.ix[node] + 0.5) * h;
    c[1] = box[4 * s + 1] + (t.iy[node] + 0.5) * h;
    c[2] = box[4 * s + 2] + (t.iz[node] + 0.5) * h;
}

// qm2l = q | q_same << 8: the M2L truncation of pairs of different levels (q)
// and of pairs of one level (q_same; 0 = q).  q_same = 2p keeps every term
// |k|, |n| <= p, the dense operator of the matrix-core same-level kernel.
static int pair_q(const Tree& t, int qm2l, int a, int b) {
    const int q = qm2l & 0xff, qs = (qm2l >> 8) & 0xff;
    return (qs && t.level[a] == t.level[b]) ? qs : q;
}

void orc_fmm(const orc_tree* tr, const orc_lists* li, int p, int qm2l, size_t n, const double* x,
             const double* q, const double* box, double G, double eps,
             double* acc, double* pot) {
    const Tree& t = tr->t;
    const Lists& L = li->l;
    MultiIndex T(p);
    const int nc = T.size();
    size_t nn = t.level.size();
    std::vector<double> M(nn * nc, 0.0), Lc(nn * nc, 0.0), cen(nn * 3);
    for (size_t i = 0; i < nn; ++i) cell_centre(t, box, (int)i, &cen[3 * i]);
    // P2M at leaves, M2M upwards (deepest level first: BFS order reversed)
    for (size_t i = nn; i-- > 0;) {
        if (t.nchild[i] == 0) {
            orc_p2m(p, &cen[3 * i], (size_t)t.count[i], x + 3 * (size_t)t.begin[i],
                    q + t.begin[i], &M[i * nc]);
        } else {
            for (int c = 0; c < t.nchild[i]; ++c) {
                int ch = t.first_child[i] + c;
                orc_m2m(p, &cen[3 * ch], &cen[3 * i], &M[(size_t)ch * nc], &M[i * nc]);
            }
        }
    }
    // M2L
#pragma omp parallel for schedule(dynamic, 8)
    for (long long a = 0; a < (long long)nn; ++a)
        for (int64_t e = L.m2l_off[a]; e < L.m2l_off[a + 1]; ++e) {
            int b = L.m2l_src[e];
            orc_m2l(p, pair_q(t, qm2l, (int)a, b), &cen[3 * a], &cen[3 * b], eps, &M[(size_t)b * nc],
                    &Lc[(size_t)a * nc]);
        }
    // L2L downwards (BFS order: parents before children)
    for (size_t i = 0; i < nn; ++i)
        for (int c = 0; c < t.nchild[i]; ++c) {
            int ch = t.first_child[i] + c;
            orc_l2l(p, &cen[3 * i], &cen[3 * ch], &Lc[i * nc], &Lc[(size_t)ch * nc]);
        }
    // L2P + P2P per leaf
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long a = 0; a < (long long)nn; ++a) {
        if (t.nchild[a]) continue;
        for (int i = t.begin[a]; i < t.begin[a] + t.count[a]; ++i) {
            double ph, g[3];
            orc_l2p(p, &cen[3 * a], &Lc[(size_t)a * nc], x + 3 * (size_t)i, &ph, g);
            double ax = 0, ay = 0, az = 0, pp = 0;
            for (int64_t e = L.p2p_off[a]; e < L.p2p_off[a + 1]; ++e) {
                int b = L.p2p_src[e];
                for (int j = t.begin[b]; j < t.begin[b] + t.count[b]; ++j) {
                    if (j == i) continue;
                    double rx = x[3 * j] - x[3 * i], ry = x[3 * j + 1] - x[3 * i + 1],
                           rz = x[3 * j + 2] - x[3 * i + 2];
                    double inv = 1.0 / std::sqrt(rx * rx + ry * ry + rz * rz + e2);
                    double w = q[j] * inv * inv * inv;
                    ax += w * rx; ay += w * ry; az += w * rz;
                    pp += q[j] * inv;
                }
            }
            acc[3 * i] = G * (g[0] + ax);
            acc[3 * i + 1] = G * (g[1] + ay);
            acc[3 * i + 2] = G * (g[2] + az);
            pot[i] = -G * (ph + pp);
        }
    }
}

// ---------------------------------------------------------------------------
// VJP of orc_fmm (gravity, q == m) with cotangent w, on the same tree and
// lists: the exact transpose of the discrete FMM.  Near field: pair terms
// T(r)(m_i w_j - m_j w_i) and -(w_j . r)/s^3.  Far field: Hessian of the mass
// local expansion applied to w_i, plus m_i grad psi and psi from the dipole
// moments M^w_n = sum_i sum_e w_ie b^(n-e)/(n-e)! run through M2M, M2L, L2L.
void orc_fmm_vjp(const orc_tree* tr, const orc_lists* li, int p, int qm2l, size_t n, const double* x, const double* m,
                 const double* w, const double* box, double G, double eps, double* gx, double* gm) {
    const Tree& t = tr->t;
    const Lists& L = li->l;
    MultiIndex T(p);
    const int nc = T.size();
    size_t nn = t.level.size();
    std::vector<double> M(nn * nc, 0.0), Lc(nn * nc, 0.0), Mw(nn * nc, 0.0), Lw(nn * nc, 0.0), cen(nn * 3);
    for (size_t i = 0; i < nn; ++i) cell_centre(t, box, (int)i, &cen[3 * i]);
    std::vector<double> mono;
    for (size_t a = nn; a-- > 0;) {
        if (t.nchild[a] == 0) {
            orc_p2m(p, &cen[3 * a], (size_t)t.count[a], x + 3 * (size_t)t.begin[a], m + t.begin[a], &M[a * nc]);
            for (int i = t.begin[a]; i < t.begin[a] + t.count[a]; ++i) {
                double b[3] = {x[3 * i] - cen[3 * a], x[3 * i + 1] - cen[3 * a + 1], x[3 * i + 2] - cen[3 * a + 2]};
                monomials(T, b, mono);
                for (int k = 0; k < nc; ++k) {
                    const auto& nk = T.mi[k];
                    for (int e = 0; e < 3; ++e) {
                        int d[3] = {nk[0], nk[1], nk[2]};
                        if (d[e] == 0) continue;
                        d[e] -= 1;
                        Mw[a * nc + k] += w[3 * i + e] * mono[T.at(d[0], d[1], d[2])];
                    }
                }
            }
        } else {
            for (int c = 0; c < t.nchild[a]; ++c) {
                int ch = t.first_child[a] + c;
                orc_m2m(p, &cen[3 * ch], &cen[3 * a], &M[(size_t)ch * nc], &M[a * nc]);
                orc_m2m(p, &cen[3 * ch], &cen[3 * a], &Mw[(size_t)ch * nc], &Mw[a * nc]);
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (long long a = 0; a < (long long)nn; ++a)
        for (int64_t e = L.m2l_off[a]; e < L.m2l_off[a + 1]; ++e) {
            int b = L.m2l_src[e];
            const int qp = pair_q(t, qm2l, (int)a, b);
            orc_m2l(p, qp, &cen[3 * a], &cen[3 * b], eps, &M[(size_t)b * nc], &Lc[(size_t)a * nc]);
            orc_m2l(p, qp, &cen[3 * a], &cen[3 * b], eps, &Mw[(size_t)b * nc], &Lw[(size_t)a * nc]);
        }
    for (size_t a = 0; a < nn; ++a)
        for (int c = 0; c < t.nchild[a]; ++c) {
            int ch = t.first_child[a] + c;
            orc_l2l(p, &cen[3 * a], &cen[3 * ch], &Lc[a * nc], &Lc[(size_t)ch * nc]);
            orc_l2l(p, &cen[3 * a], &cen[3 * ch], &Lw[a * nc], &Lw[(size_t)ch * nc]);
        }
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long a = 0; a < (long long)nn; ++a) {
        if (t.nchild[a]) continue;
        std::vector<double> mo;
        for (int i = t.begin[a]; i < t.begin[a] + t.count[a]; ++i) {
            double g[3] = {0, 0, 0}, gmi = 0;
            for (int64_t e = L.p2p_off[a]; e < L.p2p_off[a + 1]; ++e) {
                int b = L.p2p_src[e];
                for (int j = t.begin[b]; j < t.begin[b] + t.count[b]; ++j) {
                    if (j == i) continue;
                    double r[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
                    double inv = 1.0 / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + e2);
                    double i3 = inv * inv * inv, i5 = i3 * inv * inv;
                    double u[3], ru = 0, wr = 0;
                    for (int k = 0; k < 3; ++k) {
                        u[k] = m[i] * w[3 * j + k] - m[j] * w[3 * i + k];
                        ru += r[k] * u[k];
                        wr += w[3 * j + k] * r[k];
                    }
                    for (int k = 0; k < 3; ++k) g[k] += u[k] * i3 - 3.0 * r[k] * ru * i5;
                    gmi -= wr * i3;
                }
            }
            double b[3] = {x[3 * i] - cen[3 * a], x[3 * i + 1] - cen[3 * a + 1], x[3 * i + 2] - cen[3 * a + 2]};
            monomials(T, b, mo);
            double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, psi = 0, dpsi[3] = {0, 0, 0};
            for (int k = 0; k < nc; ++k) {
                const auto& kk = T.mi[k];
                psi += Lw[(size_t)a * nc + k] * mo[k];
                for (int u = 0; u < 3; ++u) {
                    int d1[3] = {kk[0], kk[1], kk[2]};
                    d1[u] += 1;
                    int s1 = T.at(d1[0], d1[1], d1[2]);
                    if (s1 >= 0) dpsi[u] += Lw[(size_t)a * nc + s1] * mo[k];
                    for (int v = 0; v < 3; ++v) {
                        int d2[3] = {d1[0], d1[1], d1[2]};
                        d2[v] += 1;
                        int s2 = T.at(d2[0], d2[1], d2[2]);
                        if (s2 >= 0) H[u][v] += Lc[(size_t)a * nc + s2] * mo[k];
                    }
                }
            }
            for (int u = 0; u < 3; ++u) {
                double hw = H[u][0] * w[3 * i] + H[u][1] * w[3 * i + 1] + H[u][2] * w[3 * i + 2];
                gx[3 * i + u] = G * (g[u] + m[i] * dpsi[u] + hw);
            }
            gm[i] = G * (gmi + psi);
        }
    }
}

// ---------------------------------------------------------------------------
static void accel_direct(size_t n, const double* x, const double* m, const double* q,
                         double G, double eps, std::vector<double>& a) {
    a.assign(3 * n, 0.0);
    std::vector<int64_t> self(n);
    for (size_t i = 0; i < n; ++i) self[i] = (int64_t)i;
    orc_direct(n, x, q, n, x, self.data(), G, eps, a.data(), nullptr);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) a[3 * i + k] *= q[i] / m[i];  // force.cl:4-19, F = q E; a = F/m
}

void orc_integrate_direct(size_t n, double* x, double* v, const double* m,
                          const double* q, double G, double eps, double dt,
                          int nsteps, int integrator) {
    std::vector<double> a;
    if (integrator == 0) {
        accel_direct(n, x, m, q, G, eps, a);
        for (int s = 0; s < nsteps; ++s) {
            for (size_t i = 0; i < 3 * n; ++i) v[i] += 0.5 * dt * a[i];
            for (size_t i = 0; i < 3 * n; ++i) x[i] += dt * v[i];
            accel_direct(n, x, m, q, G, eps, a);
            for (size_t i = 0; i < 3 * n; ++i) v[i] += 0.5 * dt * a[i];
        }
    } else if (integrator == 1) {
        // open_cl_simulation.cpp:600-607: v' = v + F/m dt ; x' = x + v_old dt
        for (int s = 0; s < nsteps; ++s) {
            accel_direct(n, x, m, q, G, eps, a);
            for (size_t i = 0; i < 3 * n; ++i) {
                double vold = v[i];
                v[i] += a[i] * dt;
                x[i] += vold * dt;
            }
        }
    } else {
        // naive_simulation.cpp:28-41: kick with all pair forces, then drift with new v
        for (int s = 0; s < nsteps; ++s) {
            accel_direct(n, x, m, q, G, eps, a);
            for (size_t i = 0; i < 3 * n; ++i) {
                v[i] += a[i] * dt;
                x[i] += v[i] * dt;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// NaiveSimulation::step (naive_simulation.cpp:7-46), FP32, same operation
// order, including its separation (dx, dx, dx) and the missing softening.
// Build with -ffp-contract=off so that no FMA is formed.
void orc_naive_compat_steps(size_t n, float* P, float fc, float dt, int nsteps, float* time) {
    auto px = [&](size_t i) -> float& { return P[12 * i + 0]; };
    auto vel = [&](size_t i, int k) -> float& { return P[12 * i + 4 + k]; };
    auto pos = [&](size_t i, int k) -> float& { return P[12 * i + k]; };
    auto mass = [&](size_t i) -> float { return P[12 * i + 8]; };
    auto charge = [&](size_t i) -> float { return P[12 * i + 9]; };
    for (int s = 0; s < nsteps; ++s) {
        for (unsigned i = 1; i < n; ++i) {
            for (unsigned j = 0; j < i; ++j) {
                float d0 = px(j) - px(i);
                float d[3] = {d0, d0, d0};
                float r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                float cf = charge(i) * charge(j);
                float mA = mass(i), mB = mass(j);
                float f[3];
                for (int k = 0; k < 3; ++k) f[k] = fc * cf * d[k] / (r * r * r);
                for (int k = 0; k < 3; ++k) vel(i, k) += f[k] / mA * dt;
                for (int k = 0; k < 3; ++k) vel(j, k) -= f[k] / mB * dt;
            }
        }
        for (unsigned i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) pos(i, k) += vel(i, k) * dt;
        *time += dt;
    }
}

}  // extern "C"
