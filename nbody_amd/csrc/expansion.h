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
//   M2L  L_k  += sum_{|k|+|n|<=P} (-1)^|n| D_{k+n}(R) M_n  (R = z_A - z_B)
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
        int t = 0;
        for (int a = 0; a < NC; ++a)
            for (int b = 0; b < NC; ++b) {
                int s = at(mx[a] + mx[b], my[a] + my[b], mz[a] + mz[b]);
                if (s >= 0) { pk[t] = a; pn[t] = b; pkn[t] = s; ++t; }
            }
        int u = 0;
        for (int a = 0; a < NC; ++a)
            for (int x = 0; x <= mx[a]; ++x)
                for (int y = 0; y <= my[a]; ++y)
                    for (int z = 0; z <= mz[a]; ++z) {
                        qn[u] = a;
                        qm[u] = at(x, y, z);
                        qr[u] = at(mx[a] - x, my[a] - y, mz[a] - z);
                        ++u;
                    }
    }
};

template <int P>
struct Tab {
    static constexpr Tables<P> t{};
};

// b^n / n! for all n (unrolled; NC multiplies)
template <int P>
__device__ __forceinline__ void monomials(float bx, float by, float bz, float* mono) {
    constexpr auto& T = Tab<P>::t;
    mono[0] = 1.0f;
#pragma unroll
    for (int k = 1; k < T.NC; ++k) {
        const int ax = T.maxis[k];
        const float b = ax == 0 ? bx : (ax == 1 ? by : bz);
        const int e = ax == 0 ? T.mx[k] : (ax == 1 ? T.my[k] : T.mz[k]);
        mono[k] = mono[T.mpar[k]] * (b * (1.0f / float(e)));
    }
}

template <int P>
__device__ __forceinline__ void derivs(float Rx, float Ry, float Rz, float eps2, float* D) {
    constexpr auto& T = Tab<P>::t;
    const float s = Rx * Rx + Ry * Ry + Rz * Rz + eps2;
    const float inv_s = 1.0f / s;
    D[0] = rsqrtf(s);
    const float R[3] = {Rx, Ry, Rz};
#pragma unroll
    for (int k = 1; k < T.NC; ++k) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (T.d1[k][j] >= 0) acc += T.c1[k][j] * R[j] * D[T.d1[k][j]];
            if (T.d2[k][j] >= 0 && T.c2[k][j] != 0.0f) acc += T.c2[k][j] * D[T.d2[k][j]];
        }
        D[k] = -acc * inv_s;
    }
}

template <int P>
__device__ __forceinline__ void m2l_apply(const float* D, const float* M, float* L) {
    constexpr auto& T = Tab<P>::t;
#pragma unroll
    for (int t = 0; t < T.NP; ++t) {
        const float v = D[T.pkn[t]] * M[T.pn[t]];
        if (T.deg[T.pn[t]] & 1) L[T.pk[t]] -= v;
        else L[T.pk[t]] += v;
    }
}

template <int P>
__device__ __forceinline__ void m2m_apply(float dx, float dy, float dz, const float* Mc, float* Mp) {
    constexpr auto& T = Tab<P>::t;
    float mono[T.NC];
    monomials<P>(dx, dy, dz, mono);
#pragma unroll
    for (int u = 0; u < T.NM; ++u) Mp[T.qn[u]] += Mc[T.qm[u]] * mono[T.qr[u]];
}

template <int P>
__device__ __forceinline__ void l2l_apply(float dx, float dy, float dz, const float* Lp, float* Lc) {
    constexpr auto& T = Tab<P>::t;
    float mono[T.NC];
    monomials<P>(dx, dy, dz, mono);
    // Lc[j] += sum_m Lp[j+m] mono[m]  (pair table with k=j, n=m)
#pragma unroll
    for (int t = 0; t < T.NP; ++t) Lc[T.pk[t]] += Lp[T.pkn[t]] * mono[T.pn[t]];
}

// phi = sum_k L_k a^k/k!, g = grad (sum over |k| <= P-1 of L_{k+e} a^k/k!)
template <int P>
__device__ __forceinline__ void l2p_eval(float ax, float ay, float az, const float* L, float& phi,
                                         float& gx, float& gy, float& gz) {
    constexpr auto& T = Tab<P>::t;
    float mono[T.NC];
    monomials<P>(ax, ay, az, mono);
    float p = 0.f, x = 0.f, y = 0.f, z = 0.f;
#pragma unroll
    for (int k = 0; k < T.NC; ++k) {
        p += L[k] * mono[k];
        if (T.deg[k] < P) {
            x += L[T.at(T.mx[k] + 1, T.my[k], T.mz[k])] * mono[k];
            y += L[T.at(T.mx[k], T.my[k] + 1, T.mz[k])] * mono[k];
            z += L[T.at(T.mx[k], T.my[k], T.mz[k] + 1)] * mono[k];
        }
    }
    phi = p; gx = x; gy = y; gz = z;
}

}  // namespace nbh
