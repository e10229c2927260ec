// vjp.hip -- adjoint of the FMM force evaluation (no counterpart in the
// reference, whose pipeline is forward only).
//
// For gravity (charge == mass), a_i = G sum_j m_j f(x_j - x_i) with
// f(r) = r / (|r|^2+eps^2)^(3/2).  Its vector-Jacobian product with a
// cotangent w is
//   gx_k = G sum_j T(x_j - x_k) (m_k w_j - m_j w_k),   T(r) = df/dr (tidal tensor)
//   gm_k = G psi(x_k),   psi(x) = sum_j w_j . grad g(x_j - x),  g = (|r|^2+eps^2)^(-1/2)
// and gx_k = G [ m_k grad psi(x_k) + Hess(phi)(x_k) w_k ] with phi the mass
// potential of the forward pass.  The far field is the exact transpose of
// the forward FMM on the same tree and lists: L2L^T has the form of M2M,
// M2M^T of L2L, and M2L^T is M2L between the same (symmetric) cell pairs,
// so the cotangents enter as point dipoles (dipole P2M) and run through the
// unchanged M2M -> M2L -> L2L kernels; the near field is the exact pair VJP
// over the P2P lists, and Hess(phi) comes from the forward local expansions.
#include <algorithm>

#include "common.h"
#include "expansion.h"
#include "launch.h"

namespace nbh {

// cotangent in sorted order: pw[i] = (w[id[i]], 0), FP32
__global__ void vjp_gather_kernel(const double* w_in, const uint32_t* id, float4* pw, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = id[i];
    pw[i] = make_float4((float)w_in[3 * (size_t)j], (float)w_in[3 * (size_t)j + 1], (float)w_in[3 * (size_t)j + 2],
                        0.f);
}

// dipole P2M per leaf (one wavefront), offsets from the forward pass (pf)
template <int P>
__global__ __launch_bounds__(256) void dipole_p2m_kernel(FmmParams F, const float4* pw, float* Mw,
                                                         uint32_t num_leaves) {
    constexpr int NC = ncoef(P);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (wid >= num_leaves) return;
    const int A = (int)F.leaf_ids[wid];
    const int4 lk = F.nodes.link[A];
    float M[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) M[k] = 0.f;
    for (int j = lane_id(); j < lk.y; j += WAVE) {
        const int i = lk.x + j;
        const float4 b = F.pf[i];
        const float4 w = pw[i];
        dipole_p2m_accum<P>(b.x, b.y, b.z, w.x, w.y, w.z, M);
    }
    group_reduce_scatter<NC, 6>(M);
    int rs0, rsn;
    rs_slot(NC, 6, lane_id(), rs0, rsn);
    float* out = Mw + (size_t)A * F.ncp + rs0;
#pragma unroll
    for (int j = 0; j < rs_len(NC, 6); ++j)
        if (j < rsn) out[j] = M[j];
}

// near field (pair VJP over the P2P lists) + far field (Hessian of the mass
// local expansion, gradient and value of the dipole local expansion).  One
// wavefront per target leaf, organised like the forward P2P kernel: the
// leaf's P2P sources are flattened over its row segments, 64 at a time
// staged in LDS (offset in the target leaf's frame, mass, cotangent), lanes
// = (target particle) x (source stride), partial sums combined by shuffles.
// As there, every lane runs the chunk's ceil(nchunk / splits) iterations and
// the slots past the chunk hold zero mass and cotangent (index -1, skipped).
constexpr int VJP_WAVES = 4;
constexpr int VJP_STAGE = 96;  // splits * ceil(nchunk / splits) <= 84

template <int P>
__global__ __launch_bounds__(WAVE * VJP_WAVES) void vjp_staged_kernel(FmmParams F, const float4* pw, uint32_t num_leaves,
                                                                      double* gx, double* gy, double* gz, double* gm) {
    __shared__ float4 stage_p[VJP_WAVES][VJP_STAGE];
    __shared__ float4 stage_w[VJP_WAVES][VJP_STAGE];
    __shared__ int stage_i[VJP_WAVES][VJP_STAGE];
    __shared__ int2 ent_s[VJP_WAVES][WAVE];      // list entry: first particle, flattened start
    __shared__ float4 entd_s[VJP_WAVES][WAVE];   // list entry: source centre - target centre
    __shared__ unsigned long long start_s[VJP_WAVES];  // entries starting in the chunk
    const int wslot = threadIdx.x / WAVE;
    const int lane = lane_id();
    float4* sp4 = stage_p[wslot];
    float4* sw4 = stage_w[wslot];
    int* sidx = stage_i[wslot];
    int2* ent = ent_s[wslot];
    float4* entd = entd_s[wslot];
    if (lane < VJP_STAGE - WAVE) {  // never staged: empty for good
        sp4[WAVE + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        sw4[WAVE + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        sidx[WAVE + lane] = -1;
    }
    const uint32_t nw = gridDim.x * VJP_WAVES;
    const float eps2 = F.eps2;
    for (uint32_t w = blockIdx.x * VJP_WAVES + wslot; w < num_leaves; w += nw) {
        const int A = (int)F.leaf_ids[w];
        const int4 ga = F.nodes.geo[A];
        const int4 la = F.nodes.link[A];
        const float unit = F.unit[F.nodes.sys[A]];
        for (int t0 = 0; t0 < la.y; t0 += WAVE) {
            const int nt = min(WAVE, la.y - t0);
            const int splits = WAVE / nt;
            // ceil(x / splits) = umulhi(x + splits - 1, magic) for x <= 64
            const uint32_t magic = splits == 1 ? 0u : 0xffffffffu / (uint32_t)splits + 1u;
            const int ti = lane % nt;
            const int sp = lane / nt;
            const int i = la.x + t0 + ti;
            const float4 me = F.pf[i];
            const float4 wi = pw[i];
            const float mi = me.w;
            float g0 = 0.f, g1 = 0.f, g2 = 0.f, gmm = 0.f;
            for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r])
            for (uint32_t eb = F.lists.p_start[r], e_end = eb + F.lists.p_cnt[r]; eb < e_end; eb += WAVE) {
                const int ne = (int)min((uint32_t)WAVE, e_end - eb);
                int cnt = 0, beg = 0;
                float dx = 0.f, dy = 0.f, dz = 0.f;
                if (lane < ne) {
                    const int B = F.lists.p_pool[eb + lane];
                    const int4 lb = F.nodes.link[B];
                    const float3 d = centre_diff(ga, F.nodes.geo[B], F.bits, unit);
                    cnt = lb.y;
                    beg = lb.x;
                    dx = d.x; dy = d.y; dz = d.z;
                }
                int incl = cnt;
#pragma unroll
                for (int o = 1; o < WAVE; o <<= 1) {
                    const int y = __shfl_up(incl, o, WAVE);
                    if (lane >= o) incl += y;
                }
                const int total = __shfl(incl, WAVE - 1, WAVE);
                const int excl = incl - cnt;
                if (lane < ne) {
                    ent[lane] = make_int2(beg, excl);
                    entd[lane] = make_float4(dx, dy, dz, 0.f);
                }
                for (int f0 = 0; f0 < total; f0 += WAVE) {
                    const int fidx = f0 + lane;
                    // entry of flattened source fidx (as in p2p_staged_kernel):
                    // the entries before the chunk plus the starts inside it
                    const int kb0 = __popcll(__ballot(lane < ne && excl < f0));
                    if (lane == 0) start_s[wslot] = 0ull;
                    wave_lds_sync();
                    if (lane < ne && excl >= f0 && excl < f0 + WAVE)
                        atomicOr(&start_s[wslot], 1ull << (excl - f0));
                    wave_lds_sync();
                    const unsigned long long starts = start_s[wslot];
                    {
                        float4 sv = make_float4(0.f, 0.f, 0.f, 0.f), wv = sv;  // past the sources: empty
                        int src = -1;
                        if (fidx < total) {
                            const int k = kb0 + __popcll(starts & ((2ull << lane) - 1ull)) - 1;
                            const int2 e = ent[k];
                            const float4 dk = entd[k];
                            src = e.x + (fidx - e.y);
                            const float4 s = F.pf[src];
                            sv = make_float4(dk.x + s.x, dk.y + s.y, dk.z + s.z, s.w);
                            wv = pw[src];
                        }
                        sp4[lane] = sv;
                        sw4[lane] = wv;
                        sidx[lane] = src;
                    }
                    wave_lds_sync();
                    const int nchunk = min(WAVE, total - f0);
                    const int Tc = __builtin_amdgcn_readfirstlane(
                        splits == 1 ? nchunk : (int)__umulhi((uint32_t)(nchunk + splits - 1), magic));
                    if (sp < splits) {
                        for (int kk = 0; kk < Tc; ++kk) {
                            const int j = sp + kk * splits;
                            const float4 s = sp4[j];
                            const float4 wj = sw4[j];
                            const float rx = s.x - me.x, ry = s.y - me.y, rz = s.z - me.z;
                            const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                            const int sj = sidx[j];
                            const float inv = (sj == i || sj < 0) ? 0.f : rsqrtf(r2);
                            const float i2 = inv * inv;
                            const float i3 = i2 * inv;
                            const float i5 = i3 * i2;
                            const float ux = mi * wj.x - s.w * wi.x, uy = mi * wj.y - s.w * wi.y,
                                        uz = mi * wj.z - s.w * wi.z;
                            const float ru = rx * ux + ry * uy + rz * uz;
                            const float c = 3.f * ru * i5;
                            g0 += ux * i3 - rx * c;
                            g1 += uy * i3 - ry * c;
                            g2 += uz * i3 - rz * c;
                            gmm -= (wj.x * rx + wj.y * ry + wj.z * rz) * i3;
                        }
                    }
                    wave_lds_sync();
                }
            }
            {
                const float a0 = g0, a1 = g1, a2 = g2, a3 = gmm;
                for (int q = 1; q < splits; ++q) {
                    const int srcl = ti + q * nt;
                    g0 += __shfl(a0, srcl, WAVE);
                    g1 += __shfl(a1, srcl, WAVE);
                    g2 += __shfl(a2, srcl, WAVE);
                    gmm += __shfl(a3, srcl, WAVE);
                }
            }
            if (sp == 0) {
                // raw near-field sums (FP32 values); vjp_far_kernel adds the far field
                gx[i] = g0;
                gy[i] = g1;
                gz[i] = g2;
                gm[i] = gmm;
            }
        }
    }
}

// Far-field part of the VJP, one lane per particle, after the near-field
// kernel (raw sums in g) and the transposed far-field chain (Lw): the tidal
// term from the forward local expansion's Hessian, the field of the
// cotangent dipoles from Lw, then G.
template <int P>
__global__ __launch_bounds__(256) void vjp_far_kernel(FmmParams F, const float4* pw, const float* Lw, uint32_t n,
                                                      double* gx, double* gy, double* gz, double* gm) {
    constexpr int NC = ncoef(P);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 me = F.pf[i];
    const float4 wi = pw[i];
    const float mi = me.w;
    const uint32_t A = F.pleaf[i];
    float L[NC];
    float H[6];
    {
        const float* Ls = F.L + (size_t)A * F.ncp;
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = Ls[k];
        l2p_hessian<P>(me.x, me.y, me.z, L, H);
    }
    float psi, px, py, pz;
    {
        const float* Ls = Lw + (size_t)A * F.ncp;
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = Ls[k];
        l2p_eval<P>(me.x, me.y, me.z, L, psi, px, py, pz);
    }
    float g0 = (float)gx[i], g1 = (float)gy[i], g2 = (float)gz[i], gmm = (float)gm[i];
    g0 += mi * px + H[0] * wi.x + H[3] * wi.y + H[4] * wi.z;
    g1 += mi * py + H[3] * wi.x + H[1] * wi.y + H[5] * wi.z;
    g2 += mi * pz + H[4] * wi.x + H[5] * wi.y + H[2] * wi.z;
    gmm += psi;
    const double G = (double)F.G;
    gx[i] = G * g0;
    gy[i] = G * g1;
    gz[i] = G * g2;
    gm[i] = G * gmm;
}

template <int P>
static void launch_vjp_near_t(const FmmParams& F, const float4* pw, uint32_t n, uint32_t num_leaves, double* g,
                              hipStream_t st) {
    const unsigned grid = std::min<unsigned>(ceil_div(num_leaves, VJP_WAVES), 65536u);
    hipLaunchKernelGGL((vjp_staged_kernel<P>), dim3(grid), dim3(WAVE * VJP_WAVES), 0, st, F, pw, num_leaves, g, g + n,
                       g + 2 * (size_t)n, g + 3 * (size_t)n);
}

template <int P>
static void launch_vjp_chain_t(int q, const M2LPlan& plan, const FmmParams& F, const FmmParams& Fw, const float4* pw,
                               uint32_t num_leaves, const uint32_t* level_count, const uint32_t* level_off,
                               int depth, hipStream_t st) {
    hipLaunchKernelGGL((dipole_p2m_kernel<P>), dim3(ceil_div(num_leaves, 4)), dim3(256), 0, st, F, pw, Fw.M,
                       num_leaves);
    for (int l = depth - 1; l >= 0; --l) launch_m2m_level(P, Fw, l, level_count[l], level_off[l], st);
    // the matrix-core same-level kernels read the dipole moments split (into plan.off.mb)
    if (plan.mfma)
        launch_off_moments(P, plan.mfma == 1, Fw.M, Fw.ncp, plan.num_nodes, Fw.nodes.geo, Fw.box, (void*)plan.off.mb, st);
    launch_m2l(P, q, Fw, plan, st);
    for (int l = 1; l <= depth; ++l) launch_l2l_level(P, Fw, l, level_count[l], level_off[l], st);
}

#define NBH_VJP_DISPATCH(p, call)                      \
    switch (p) {                                       \
        case 1: { constexpr int PP = 1; call; } break; \
        case 2: { constexpr int PP = 2; call; } break; \
        case 3: { constexpr int PP = 3; call; } break; \
        case 4: { constexpr int PP = 4; call; } break; \
        case 5: { constexpr int PP = 5; call; } break; \
        case 6: { constexpr int PP = 6; call; } break; \
        default: break;                                \
    }

// ---- reverse leapfrog on the device (trajectory adjoint) --------------------
// cotangents in input order: update: vbar += dt xbar; then w = c vbar
__global__ void adj_kick_kernel(double* __restrict__ xb, double* __restrict__ vb, double* __restrict__ w, size_t m,
                                double dt, double c, int update) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double v = vb[i];
    if (update) {
        v += dt * xb[i];
        vb[i] = v;
    }
    w[i] = c * v;
}
// xbar[id] += g (sorted order, gx | gy | gz | gm arrays of n), mbar[id] += gm
__global__ void adj_accum_kernel(const double* __restrict__ g, const uint32_t* __restrict__ id, uint32_t n,
                                 double* __restrict__ xb, double* __restrict__ mb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t j = id[i];
    xb[3 * j] += g[i];
    xb[3 * j + 1] += g[(size_t)n + i];
    xb[3 * j + 2] += g[2 * (size_t)n + i];
    mb[j] += g[3 * (size_t)n + i];
}
__global__ void fill_f64_kernel(double* p, size_t m, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}
void launch_adj_kick(double* xb, double* vb, double* w, size_t m, double dt, double c, bool update, hipStream_t st) {
    hipLaunchKernelGGL(adj_kick_kernel, dim3((unsigned)ceil_div(m, (size_t)256)), dim3(256), 0, st, xb, vb, w, m, dt, c,
                       update ? 1 : 0);
}
void launch_adj_accum(const double* g, const uint32_t* id, uint32_t n, double* xb, double* mb, hipStream_t st) {
    hipLaunchKernelGGL(adj_accum_kernel, dim3(ceil_div(n, 256u)), dim3(256), 0, st, g, id, n, xb, mb);
}
void launch_fill_f64(double* p, size_t m, double v, hipStream_t st) {
    if (!m) return;
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)ceil_div(m, (size_t)256)), dim3(256), 0, st, p, m, v);
}

void launch_vjp_gather(const double* w_in, const uint32_t* id, float4* pw, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(vjp_gather_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, w_in, id, pw, n);
}
void launch_vjp_near(int p, const FmmParams& F, const float4* pw, uint32_t n, uint32_t num_leaves, double* g,
                     hipStream_t st) {
    NBH_VJP_DISPATCH(p, launch_vjp_near_t<PP>(F, pw, n, num_leaves, g, st));
}
void launch_vjp_chain(int p, int q, const M2LPlan& plan, const FmmParams& F, const FmmParams& Fw, const float4* pw,
                      uint32_t num_leaves, const uint32_t* level_count, const uint32_t* level_off, int depth,
                      hipStream_t st) {
    NBH_VJP_DISPATCH(p, launch_vjp_chain_t<PP>(q, plan, F, Fw, pw, num_leaves, level_count, level_off, depth, st));
}
void launch_vjp_far(int p, const FmmParams& F, const FmmParams& Fw, const float4* pw, uint32_t n, double* g,
                    hipStream_t st) {
    NBH_VJP_DISPATCH(p, hipLaunchKernelGGL((vjp_far_kernel<PP>), dim3(ceil_div(n, 256)), dim3(256), 0, st, F, pw,
                                           Fw.L, n, g, g + n, g + 2 * (size_t)n, g + 3 * (size_t)n));
}

}  // namespace nbh
