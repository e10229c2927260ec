// expansion.h -- Cartesian Taylor expansions of order P for the softened
// kernel g(R) = (|R|^2 + eps^2)^(-1/2), as device templates.
//
// The reference evaluates the far field at order 0 only: the monopole of the
// source cell at the centre of the target cell (field.cl:35-47, moments about
// the geometric centre moment.cl:27).  Here the same cell-centred scheme is
// carried to order P:
//
//   P2M  M_n   = sum_j q_j b_j^n / n!                      (b = x_j - z_B)
//   M2M  M'_n  = sum_{m<=n} M_m d^(n-m)/(n-m)!             (d = z_child - z_parent)
//   M2L  L_k  += sum_{(k,n) kept} (-1)^|n| D_{k+n}(R) M_n   (R = z_A - z_B)
//        kept: |k|, |n| <= P and |k|+|n| <= q (every term of total degree
//        <= q), or, with the low-degree term set, either |k|+|n| <= P or
//        (|k|+|n| <= q and min(|k|, |n|) <= q-P) (q = P: the total-degree
//        truncation; the low set at q = P+1: plus the degree-(P+1) terms of
//        the dipole and of the field gradient only)
//   L2L  L'_j += sum_m L_{j+m} d^m / m!                    (d = z_child - z_parent)
//   L2P  phi   = sum_k L_k a^k/k!,  g_e = sum_k L_{k+e} a^k/k!   (a = x - z_A)
//
// with potential -G q_i phi and acceleration G (q_i/m_i) g (field.cl:22-30 sign
// convention: attractive for like charges).  D_n are the derivatives of g,
// built by the recurrence from s d_i g + R_i g = 0 (s = |R|^2 + eps^2):
//   s D_n = -[ sum_{j!=i} (2 n_j R_j D_{n-e_j} + n_j(n_j-1) D_{n-2e_j})
//              + (2n_i-1) R_i D_{n-e_i} + (n_i-1)^2 D_{n-2e_i} ]
//
// Coefficient order: total degree, then descending nx, then descending ny.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace nbh {

constexpr int ncoef(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
constexpr int npairs(int p) {  // #{(k,n): |k|+|n| <= p} = C(p+6, 6)
    return (p + 1) * (p + 2) * (p + 3) * (p + 4) * (p + 5) * (p + 6) / 720;
}

template <int P>
struct Tables {
    static constexpr int NC = ncoef(P);
    static constexpr int NP = npairs(P);
    int mx[NC], my[NC], mz[NC], deg[NC];
    int lut[(P + 1) * (P + 1) * (P + 1)];
    // monomial recurrence: mono[k] = mono[mpar[k]] * b[max_axis] / n[max_axis]
    int mpar[NC], maxis[NC];
    // derivative recurrence: axis i, and for each axis j: index of n-e_j, n-2e_j
    int dax[NC], d1[NC][3], d2[NC][3];
    float c1[NC][3], c2[NC][3];
    // M2L / L2L pair table: (k, n, k+n)
    int pk[NP], pn[NP], pkn[NP];
    // M2M pair table: (n, m, n-m) for m <= n componentwise
    static constexpr int NM = [] {
        int c = 0;
        for (int d = 0; d <= P; ++d)
            for (int nx = d; nx >= 0; --nx)
                for (int ny = d - nx; ny >= 0; --ny) c += (nx + 1) * (ny + 1) * (d - nx - ny + 1);
        return c;
    }();
    int qn[NM], qm[NM], qr[NM];

    constexpr int at(int x, int y, int z) const {
        return (x < 0 || y < 0 || z < 0 || x + y + z > P) ? -1 : lut[(x * (P + 1) + y) * (P + 1) + z];
    }

    constexpr Tables() : mx(), my(), mz(), deg(), lut(), mpar(), maxis(), dax(), d1(), d2(), c1(), c2(),
                         pk(), pn(), pkn(), qn(), qm(), qr() {
        for (int i = 0; i < (P + 1) * (P + 1) * (P + 1); ++i) lut[i] = -1;
        int k = 0;
        for (int d = 0; d <= P; ++d)
            for (int nx = d; nx >= 0; --nx)
                for (int ny = d - nx; ny >= 0; --ny) {
                    int nz = d - nx - ny;
                    mx[k] = nx; my[k] = ny; mz[k] = nz; deg[k] = d;
                    lut[(nx * (P + 1) + ny) * (P + 1) + nz] = k;
                    ++k;
                }
        for (int q = 0; q < NC; ++q) {
            int n[3] = {mx[q], my[q], mz[q]};
            // monomial parent along the axis with the largest exponent
            int ax = 0;
            if (n[1] > n[ax]) ax = 1;
            if (n[2] > n[ax]) ax = 2;
            maxis[q] = ax;
            int pn_[3] = {n[0], n[1], n[2]};
            pn_[ax] -= 1;
            mpar[q] = q == 0 ? 0 : at(pn_[0], pn_[1], pn_[2]);
            // derivative recurrence pivot: first axis with n_i > 0
            int i = n[0] > 0 ? 0 : (n[1] > 0 ? 1 : 2);
            dax[q] = i;
            for (int j = 0; j < 3; ++j) {
                int a[3] = {n[0], n[1], n[2]}, b[3] = {n[0], n[1], n[2]};
                a[j] -= 1;
                b[j] -= 2;
                d1[q][j] = at(a[0], a[1], a[2]);
                d2[q][j] = at(b[0], b[1], b[2]);
                if (j != i) {
                    c1[q][j] = 2.0f * n[j];
                    c2[q][j] = float(n[j] * (n[j] - 1));
                } else {
                    c1[q][j] = 2.0f * n[i] - 1.0f;
                    c2[q][j] = float((n[i] - 1) * (n[i] - 1));
                }
            }
        }
        // ordered by source coefficient n (outer) so that M2L consumes each
        // M_n in one run and L2L each mono_n likewise
        int t = 0;
        for (int b = 0; b < NC; ++b)
            for (int a = 0; a < NC; ++a) {
                int s = at(mx[a] + mx[b], my[a] + my[b], mz[a] + mz[b]);
                if (s >= 0) { pk[t] = a; pn[t] = b; pkn[t] = s; ++t; }
            }
        int u = 0;
        for (int m = 0; m < NC; ++m)
            for (int a = 0; a < NC; ++a) {
                if (mx[m] > mx[a] || my[m] > my[a] || mz[m] > mz[a]) continue;
                qn[u] = a;
                qm[u] = m;
                qr[u] = at(mx[a] - mx[m], my[a] - my[m], mz[a] - mz[m]);
                ++u;
            }
    }
};

template <int P>
struct Tab {
    static constexpr Tables<P> t{};
};

// M2L truncation code qc = q | M2L_ALL or q: the term (k, n) of orders
// dk = |k|, dn = |n| (<= p) is kept when dk + dn <= q (qc = q | M2L_ALL:
// every term of total degree <= q), or, with the low-degree set (qc = q),
// when dk + dn <= p, or dk + dn <= q and min(dk, dn) <= q - p.  q = p is the
// usual total-degree truncation either way.  The low set at q = p + 1 adds
// only the degree-(p + 1) terms with |n| = 1 (the dipole into the order-p
// local coefficients; about the geometric cell centre the dipole is not zero)
// and their mirror images |k| = 1: 90 of the 210 terms of degree 5 at p = 4.
// Both sets are symmetric in (k, n), so the same operator serves the adjoint.
constexpr int M2L_ALL = 16;
constexpr int m2l_deg(int qc) { return qc & (M2L_ALL - 1); }
constexpr bool m2l_keeps(int p, int qc, int dk, int dn) {
    const int q = m2l_deg(qc);
    const int r = (qc & M2L_ALL) ? q : q - p;
    return dk + dn <= p || (dk + dn <= q && (dk < dn ? dk : dn) <= r);
}
// M2L pair table: (k, n, k+n) for the kept terms, with k+n indexed in the
// order-q coefficient order (orders <= P are a prefix of it); ordered by
// source coefficient n.
constexpr int m2l_npairs(int p, int q) {
    int c = 0;
    for (int dk = 0; dk <= p; ++dk)
        for (int dn = 0; dn <= p; ++dn)
            if (m2l_keeps(p, q, dk, dn)) c += (dk + 1) * (dk + 2) / 2 * ((dn + 1) * (dn + 2) / 2);
    return c;
}
template <int P, int Q>
struct M2LTables {
    static constexpr int NP = m2l_npairs(P, Q);
    int pk[NP], pn[NP], pkn[NP];
    constexpr M2LTables() : pk(), pn(), pkn() {
        constexpr auto& TP = Tab<P>::t;
        constexpr auto& TQ = Tab<m2l_deg(Q)>::t;
        int t = 0;
        for (int b = 0; b < ncoef(P); ++b)
            for (int a = 0; a < ncoef(P); ++a) {
                if (!m2l_keeps(P, Q, TP.deg[a], TP.deg[b])) continue;
                pk[t] = a;
                pn[t] = b;
                pkn[t] = TQ.at(TP.mx[a] + TP.mx[b], TP.my[a] + TP.my[b], TP.mz[a] + TP.mz[b]);
                ++t;
            }
    }
};
template <int P, int Q>
struct M2LTab {
    static constexpr M2LTables<P, Q> t{};
};
// The same pairs ordered by tensor coefficient k+n (a stable counting sort of
// the table above): a tensor coefficient is dead after its group of products,
// so a kernel that streams the tensor in holds the moments, the local
// expansion and a few tensor coefficients at a time, not the whole tensor.
template <int P, int Q>
struct M2LTablesByD {
    static constexpr int NP = m2l_npairs(P, Q);
    int pk[NP], pn[NP], pkn[NP];
    constexpr M2LTablesByD() : pk(), pn(), pkn() {
        constexpr auto& T = M2LTab<P, Q>::t;
        constexpr int NCQ = ncoef(m2l_deg(Q));
        int start[NCQ + 1] = {};
        for (int i = 0; i < NP; ++i) ++start[T.pkn[i] + 1];
        for (int c = 0; c < NCQ; ++c) start[c + 1] += start[c];
        for (int i = 0; i < NP; ++i) {
            const int t = start[T.pkn[i]]++;
            pk[t] = T.pk[i];
            pn[t] = T.pn[i];
            pkn[t] = T.pkn[i];
        }
    }
};
template <int P, int Q>
struct M2LTabByD {
    static constexpr M2LTablesByD<P, Q> t{};
};

// ---------------------------------------------------------------------------
// Straight-line code generation: each table entry becomes a constexpr
// template argument via std::index_sequence fold expressions, so every
// coefficient array is indexed by constants and lives in registers.

template <int P, int K>
__device__ __forceinline__ void mono_step(float bx, float by, float bz, float* mono) {
    constexpr auto& T = Tab<P>::t;
    constexpr int ax = T.maxis[K];
    constexpr int e = ax == 0 ? T.mx[K] : (ax == 1 ? T.my[K] : T.mz[K]);
    constexpr int par = T.mpar[K];
    constexpr float inv = 1.0f / float(e);
    const float b = ax == 0 ? bx : (ax == 1 ? by : bz);
    mono[K] = mono[par] * (b * inv);
}
template <int P, size_t... I>
__device__ __forceinline__ void mono_all(float bx, float by, float bz, float* mono, std::index_sequence<I...>) {
    (mono_step<P, (int)I + 1>(bx, by, bz, mono), ...);
}
// b^n / n! for all n (NC-1 multiplies)
template <int P>
__device__ __forceinline__ void monomials(float bx, float by, float bz, float* mono) {
    mono[0] = 1.0f;
    mono_all<P>(bx, by, bz, mono, std::make_index_sequence<ncoef(P) - 1>{});
}

template <int P, int K, typename T_>
__device__ __forceinline__ void deriv_step(const T_* R, T_ inv_s, T_* D) {
    constexpr auto& T = Tab<P>::t;
    T_ acc = T_(0);
    if constexpr (T.d1[K][0] >= 0) acc += (T_(T.c1[K][0]) * R[0]) * D[T.d1[K][0] < 0 ? 0 : T.d1[K][0]];
    if constexpr (T.d1[K][1] >= 0) acc += (T_(T.c1[K][1]) * R[1]) * D[T.d1[K][1] < 0 ? 0 : T.d1[K][1]];
    if constexpr (T.d1[K][2] >= 0) acc += (T_(T.c1[K][2]) * R[2]) * D[T.d1[K][2] < 0 ? 0 : T.d1[K][2]];
    if constexpr (T.d2[K][0] >= 0 && T.c2[K][0] != 0.0f) acc += T_(T.c2[K][0]) * D[T.d2[K][0] < 0 ? 0 : T.d2[K][0]];
    if constexpr (T.d2[K][1] >= 0 && T.c2[K][1] != 0.0f) acc += T_(T.c2[K][1]) * D[T.d2[K][1] < 0 ? 0 : T.d2[K][1]];
    if constexpr (T.d2[K][2] >= 0 && T.c2[K][2] != 0.0f) acc += T_(T.c2[K][2]) * D[T.d2[K][2] < 0 ? 0 : T.d2[K][2]];
    D[K] = -acc * inv_s;
}
template <int P, typename T_, size_t... I>
__device__ __forceinline__ void deriv_all(const T_* R, T_ inv_s, T_* D, std::index_sequence<I...>) {
    (deriv_step<P, (int)I + 1, T_>(R, inv_s, D), ...);
}
// D_n(R), |n| <= P, of (|R|^2 + eps^2)^(-1/2); T_ = float or double
template <int P, typename T_ = float>
__device__ __forceinline__ void derivs(T_ Rx, T_ Ry, T_ Rz, T_ eps2, T_* D) {
    const T_ s = Rx * Rx + Ry * Ry + Rz * Rz + eps2;
    const T_ inv_s = T_(1) / s;
    if constexpr (sizeof(T_) == 4) D[0] = rsqrtf(s);
    else D[0] = T_(1) / sqrt(s);
    const T_ R[3] = {Rx, Ry, Rz};
    deriv_all<P, T_>(R, inv_s, D, std::make_index_sequence<ncoef(P) - 1>{});
}

// POT = false leaves out the products into L[0], the coefficient that only
// the potential uses (the acceleration is the gradient: L2L never moves L[0]
// into another coefficient, and L2P reads it for the potential alone), but
// for the one with D[0]: with D[0] dead the compiler narrows the 16-byte
// tensor loads to pairs of 4-byte LDS reads, twice the LDS instructions
// (measured: 35% slower same-level kernel)
template <int P, int Q, int I, bool POT>
__device__ __forceinline__ void m2l_step(const float* D, const float* M, float* L) {
    constexpr auto& T = M2LTab<P, Q>::t;
    constexpr int k = T.pk[I], n = T.pn[I], kn = T.pkn[I];
    if constexpr (!POT && k == 0 && kn != 0) return;
    else if constexpr (Tab<P>::t.deg[n] & 1) L[k] = fmaf(-D[kn], M[n], L[k]);
    else L[k] = fmaf(D[kn], M[n], L[k]);
}
template <int P, int Q, bool POT, size_t... I>
__device__ __forceinline__ void m2l_all(const float* D, const float* M, float* L, std::index_sequence<I...>) {
    (m2l_step<P, Q, (int)I, POT>(D, M, L), ...);
}
// D: derivatives to order Q; M, L: order P
template <int P, int Q, bool POT = true>
__device__ __forceinline__ void m2l_apply(const float* D, const float* M, float* L) {
    m2l_all<P, Q, POT>(D, M, L, std::make_index_sequence<m2l_npairs(P, Q)>{});
}
// The same products in tensor-coefficient order (M2LTabByD)
template <int P, int Q, int I, bool POT>
__device__ __forceinline__ void m2l_step_d(const float* D, const float* M, float* L) {
    constexpr auto& T = M2LTabByD<P, Q>::t;
    constexpr int k = T.pk[I], n = T.pn[I], kn = T.pkn[I];
    if constexpr (!POT && k == 0 && kn != 0) return;
    else if constexpr (Tab<P>::t.deg[n] & 1) L[k] = fmaf(-D[kn], M[n], L[k]);
    else L[k] = fmaf(D[kn], M[n], L[k]);
}
template <int P, int Q, bool POT, size_t... I>
__device__ __forceinline__ void m2l_all_d(const float* D, const float* M, float* L, std::index_sequence<I...>) {
    (m2l_step_d<P, Q, (int)I, POT>(D, M, L), ...);
}
template <int P, int Q, bool POT = true>
__device__ __forceinline__ void m2l_apply_by_d(const float* D, const float* M, float* L) {
    m2l_all_d<P, Q, POT>(D, M, L, std::make_index_sequence<m2l_npairs(P, Q)>{});
}

// Octant symmetry of the derivative tensor: with S = diag(s) a sign flip of
// the axes, D_m(S R) = s^m D_m(R), s^m = prod_i s_i^{m_i}.  neg holds bit i
// when s_i = -1; coefficients whose odd-exponent axes meet neg an odd number
// of times change sign.
template <int Q, int K>
__device__ __forceinline__ void octant_step(const uint32_t* fl, float* D) {
    constexpr auto& T = Tab<Q>::t;
    constexpr int par = (T.mx[K] & 1) | ((T.my[K] & 1) << 1) | ((T.mz[K] & 1) << 2);
    if constexpr (par != 0) D[K] = __uint_as_float(__float_as_uint(D[K]) ^ fl[par]);
}
template <int Q, size_t... I>
__device__ __forceinline__ void octant_all(const uint32_t* fl, float* D, std::index_sequence<I...>) {
    (octant_step<Q, (int)I>(fl, D), ...);
}
template <int Q>
__device__ __forceinline__ void octant_signs(uint32_t neg, float* D) {
    uint32_t fl[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) fl[c] = (uint32_t)(__popc(neg & (uint32_t)c) & 1) << 31;
    octant_all<Q>(fl, D, std::make_index_sequence<ncoef(Q)>{});
}

template <int P, int I>
__device__ __forceinline__ void m2m_step(const float* mono, const float* Mc, float* Mp) {
    constexpr auto& T = Tab<P>::t;
    constexpr int n = T.qn[I], m = T.qm[I], r = T.qr[I];
    Mp[n] = fmaf(Mc[m], mono[r], Mp[n]);
}
template <int P, size_t... I>
__device__ __forceinline__ void m2m_all(const float* mono, const float* Mc, float* Mp, std::index_sequence<I...>) {
    (m2m_step<P, (int)I>(mono, Mc, Mp), ...);
}
template <int P>
__device__ __forceinline__ void m2m_apply(float dx, float dy, float dz, const float* Mc, float* Mp) {
    float mono[ncoef(P)];
    monomials<P>(dx, dy, dz, mono);
    m2m_all<P>(mono, Mc, Mp, std::make_index_sequence<Tables<P>::NM>{});
}

template <int P, int I>
__device__ __forceinline__ void l2l_step(const float* mono, const float* Lp, float* Lc) {
    constexpr auto& T = Tab<P>::t;
    constexpr int k = T.pk[I], n = T.pn[I], kn = T.pkn[I];
    Lc[k] = fmaf(Lp[kn], mono[n], Lc[k]);
}
template <int P, size_t... I>
__device__ __forceinline__ void l2l_all(const float* mono, const float* Lp, float* Lc, std::index_sequence<I...>) {
    (l2l_step<P, (int)I>(mono, Lp, Lc), ...);
}
template <int P>
__device__ __forceinline__ void l2l_apply(float dx, float dy, float dz, const float* Lp, float* Lc) {
    float mono[ncoef(P)];
    monomials<P>(dx, dy, dz, mono);
    l2l_all<P>(mono, Lp, Lc, std::make_index_sequence<npairs(P)>{});
}

template <int P, int K>
__device__ __forceinline__ void l2p_step(const float* mono, const float* L, float& p, float& x, float& y, float& z) {
    constexpr auto& T = Tab<P>::t;
    p = fmaf(L[K], mono[K], p);
    if constexpr (T.deg[K] < P) {
        constexpr int ix = T.at(T.mx[K] + 1, T.my[K], T.mz[K]);
        constexpr int iy = T.at(T.mx[K], T.my[K] + 1, T.mz[K]);
        constexpr int iz = T.at(T.mx[K], T.my[K], T.mz[K] + 1);
        x = fmaf(L[ix], mono[K], x);
        y = fmaf(L[iy], mono[K], y);
        z = fmaf(L[iz], mono[K], z);
    }
}
template <int P, size_t... I>
__device__ __forceinline__ void l2p_all(const float* mono, const float* L, float& p, float& x, float& y, float& z,
                                        std::index_sequence<I...>) {
    (l2p_step<P, (int)I>(mono, L, p, x, y, z), ...);
}
// phi = sum_k L_k a^k/k!, g = grad (sum over |k| <= P-1 of L_{k+e} a^k/k!)
template <int P>
__device__ __forceinline__ void l2p_eval(float ax, float ay, float az, const float* L, float& phi,
                                         float& gx, float& gy, float& gz) {
    float mono[ncoef(P)];
    monomials<P>(ax, ay, az, mono);
    float p = 0.f, x = 0.f, y = 0.f, z = 0.f;
    l2p_all<P>(mono, L, p, x, y, z, std::make_index_sequence<ncoef(P)>{});
    phi = p; gx = x; gy = y; gz = z;
}

// ---------------------------------------------------------------------------
// Adjoint helpers.
//
// Dipole P2M: moments of point dipoles w at offsets b,
//   M_n = sum_i sum_e w_{i,e} b_i^(n - e_e) / (n - e_e)!     (n_e >= 1)
// i.e. the derivative of the monopole moments with respect to the source
// position; through the unchanged M2M/M2L/L2L it yields the local expansion of
// psi(x) = sum_i w_i . grad g(x_i - x).
template <int P, int K>
__device__ __forceinline__ void dip_step(const float* mono, float wx, float wy, float wz, float* M) {
    constexpr auto& T = Tab<P>::t;
    if constexpr (T.mx[K] > 0) M[K] = fmaf(wx, mono[T.at(T.mx[K] - 1, T.my[K], T.mz[K])], M[K]);
    if constexpr (T.my[K] > 0) M[K] = fmaf(wy, mono[T.at(T.mx[K], T.my[K] - 1, T.mz[K])], M[K]);
    if constexpr (T.mz[K] > 0) M[K] = fmaf(wz, mono[T.at(T.mx[K], T.my[K], T.mz[K] - 1)], M[K]);
}
template <int P, size_t... I>
__device__ __forceinline__ void dip_all(const float* mono, float wx, float wy, float wz, float* M,
                                        std::index_sequence<I...>) {
    (dip_step<P, (int)I + 1>(mono, wx, wy, wz, M), ...);
}
template <int P>
__device__ __forceinline__ void dipole_p2m_accum(float bx, float by, float bz, float wx, float wy, float wz,
                                                 float* M) {
    float mono[ncoef(P)];
    monomials<P>(bx, by, bz, mono);
    dip_all<P>(mono, wx, wy, wz, M, std::make_index_sequence<ncoef(P) - 1>{});
}

// Hessian of phi_L(a) = sum_k L_k a^k/k!:  H_ab = sum_{|k| <= P-2} L_{k+e_a+e_b} a^k/k!
template <int P, int K>
__device__ __forceinline__ void hess_step(const float* mono, const float* L, float* H) {
    constexpr auto& T = Tab<P>::t;
    if constexpr (T.deg[K] + 2 <= P) {
        constexpr int x = T.mx[K], y = T.my[K], z = T.mz[K];
        H[0] = fmaf(L[T.at(x + 2, y, z)], mono[K], H[0]);
        H[1] = fmaf(L[T.at(x, y + 2, z)], mono[K], H[1]);
        H[2] = fmaf(L[T.at(x, y, z + 2)], mono[K], H[2]);
        H[3] = fmaf(L[T.at(x + 1, y + 1, z)], mono[K], H[3]);
        H[4] = fmaf(L[T.at(x + 1, y, z + 1)], mono[K], H[4]);
        H[5] = fmaf(L[T.at(x, y + 1, z + 1)], mono[K], H[5]);
    }
}
template <int P, size_t... I>
__device__ __forceinline__ void hess_all(const float* mono, const float* L, float* H, std::index_sequence<I...>) {
    (hess_step<P, (int)I>(mono, L, H), ...);
}
// H = {xx, yy, zz, xy, xz, yz}
template <int P>
__device__ __forceinline__ void l2p_hessian(float ax, float ay, float az, const float* L, float* H) {
    float mono[ncoef(P)];
    monomials<P>(ax, ay, az, mono);
#pragma unroll
    for (int k = 0; k < 6; ++k) H[k] = 0.f;
    if constexpr (P >= 2) hess_all<P>(mono, L, H, std::make_index_sequence<ncoef(P)>{});
}

}  // namespace nbh
