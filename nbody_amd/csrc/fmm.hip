// fmm.hip -- moments, far field, near field and integration kernels.
//
// Replaces moment.cl (P2M/M2M, :6-160), field.cl (P2P and cell fields,
// :17-211), force.cl (per-particle reductions, :21-81) and the host
// integration loop (open_cl_simulation.cpp:572-616).  Fields are accumulated
// in registers and reduced inside the producing kernel; there are no
// per-interaction slot buffers.
//
// Precision: particle positions and velocities are FP64.  Each particle is
// stored for the force kernels as an FP32 offset from the centre of its leaf
// cell; centre differences between cells are exact integers (units of half
// the finest cell) scaled once to FP32, so FP32 separations keep their
// relative precision however large the root box is.
#include <algorithm>

#include "common.h"
#include "expansion.h"
#include "launch.h"

namespace nbh {

// ---------------------------------------------------------------------------
__global__ void gather_state_kernel(State a, State b, const uint32_t* __restrict__ perm, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    b.x[i] = a.x[j]; b.y[i] = a.y[j]; b.z[i] = a.z[j];
    b.vx[i] = a.vx[j]; b.vy[i] = a.vy[j]; b.vz[i] = a.vz[j];
    b.mass[i] = a.mass[j]; b.charge[i] = a.charge[j];
    b.ax[i] = a.ax[j]; b.ay[i] = a.ay[j]; b.az[i] = a.az[j];
    b.pot[i] = a.pot[j];
    b.id[i] = a.id[j];
}

void launch_gather_state(const State& src, const State& dst, const uint32_t* perm, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(gather_state_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, src, dst, perm, n);
}

// ---------------------------------------------------------------------------
// P2M (moment.cl:6-67 at order P): one wavefront per leaf cell.  Also writes
// the particles' FP32 offsets from the leaf centre (used by P2P and L2P).
template <int P>
__global__ __launch_bounds__(256) void p2m_kernel(FmmParams F, State S, uint32_t num_leaves) {
    constexpr int NC = ncoef(P);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (wid >= num_leaves) return;
    const int A = (int)F.leaf_ids[wid];
    const int4 g = F.nodes.geo[A];
    const int4 lk = F.nodes.link[A];
    const int s = F.nodes.sys[A];
    double c[3];
    centre64(g, F.box[s], F.bits, c);
    float M[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) M[k] = 0.f;
    for (int j = lane_id(); j < lk.y; j += WAVE) {
        const int i = lk.x + j;
        const float bx = (float)(S.x[i] - c[0]), by = (float)(S.y[i] - c[1]), bz = (float)(S.z[i] - c[2]);
        const float q = (float)S.charge[i];
        F.pf[i] = make_float4(bx, by, bz, q);
        F.pleaf[i] = A;
        float mono[NC];
        monomials<P>(bx, by, bz, mono);
#pragma unroll
        for (int k = 0; k < NC; ++k) M[k] += q * mono[k];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) M[k] = wave_sum(M[k]);
    if (lane_id() == 0) {
        float* out = F.M + (size_t)A * F.ncp;
#pragma unroll
        for (int k = 0; k < NC; ++k) out[k] = M[k];
    }
}

// M2M (replaces moment.cl:72-160, which sums untranslated moments): one
// thread per cell of the level; children's moments shifted to the parent
// centre.  Levels run bottom-up.
template <int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2m_kernel(FmmParams F, uint32_t count,
                                                                                           uint32_t off) {
    constexpr int NC = ncoef(P);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int A = (int)(off + t);
    const int4 g = F.nodes.geo[A];
    const int nch = g.w >> 8;
    if (nch == 0) return;
    const int4 lk = F.nodes.link[A];
    const float unit = F.unit[F.nodes.sys[A]];
    float Mp[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) Mp[k] = 0.f;
    for (int c = 0; c < nch; ++c) {
        const int ch = lk.z + c;
        const int4 gc = F.nodes.geo[ch];
        const float3 d = centre_diff(g, gc, F.bits, unit);  // c_child - c_parent
        m2m_apply<P>(d.x, d.y, d.z, F.M + (size_t)ch * F.ncp, Mp);
    }
    float* out = F.M + (size_t)A * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k] = Mp[k];
}

// M2L, plain: one thread per target cell, the derivative tensor rebuilt per
// pair (replaces field.cl:153-211's order-0 cell field).
template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2l_plain_kernel(FmmParams F,
                                                                                                  uint32_t num_nodes) {
    constexpr int NC = ncoef(P);
    const uint32_t A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= num_nodes) return;
    const int4 ga = F.nodes.geo[A];
    const float unit = F.unit[F.nodes.sys[A]];
    float L[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) L[k] = 0.f;
    for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r]) {
        const uint32_t s0 = F.lists.m_start[r], n = F.lists.m_cnt[r];
        for (uint32_t e = 0; e < n; ++e) {
            const int B = F.lists.m_pool[s0 + e];
            const int4 gb = F.nodes.geo[B];
            const float3 d = centre_diff(gb, ga, F.bits, unit);  // R = c_A - c_B
            float D[NC];
            derivs<P>(d.x, d.y, d.z, F.eps2, D);
            // source moments read straight from memory; the pair table is ordered
            // by source coefficient so each M_n is loaded once and used at once
            m2l_apply<P>(D, F.M + (size_t)B * F.ncp, L);
        }
    }
    float* out = F.L + (size_t)A * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k] = L[k];
}

// L2L: one thread per cell of the level (top-down), adds the parent's local
// expansion shifted to this cell's centre.
template <int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void l2l_kernel(FmmParams F, uint32_t count,
                                                                                           uint32_t off) {
    constexpr int NC = ncoef(P);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int A = (int)(off + t);
    const int4 g = F.nodes.geo[A];
    const int par = F.nodes.link[A].w;
    if (par < 0) return;
    const int4 gp = F.nodes.geo[par];
    const float3 d = centre_diff(gp, g, F.bits, F.unit[F.nodes.sys[A]]);  // c_child - c_parent
    float Lp[NC], Lc[NC];
    const float* sp = F.L + (size_t)par * F.ncp;
    float* sc = F.L + (size_t)A * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) { Lp[k] = sp[k]; Lc[k] = sc[k]; }
    l2l_apply<P>(d.x, d.y, d.z, Lp, Lc);
#pragma unroll
    for (int k = 0; k < NC; ++k) sc[k] = Lc[k];
}

// ---------------------------------------------------------------------------
// P2P + L2P + force + (optional) closing half kick: one wavefront per target
// leaf.  Replaces field.cl:49-148 (pair fields into slots), force.cl:21-81
// (slot reductions) and the host integration loop.  Lanes are split as
// (target particle) x (source stride): with nt targets in a chunk, lane l
// handles target l % nt and every (64/nt)-th particle of each source leaf, so
// all 64 lanes work for any leaf occupancy; partial sums meet in LDS.
template <int P>
__global__ __launch_bounds__(256) void p2p_kernel(FmmParams F, State S, uint32_t num_leaves, double kick_dt) {
    constexpr int NC = ncoef(P);
    const int lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (wid >= num_leaves) return;  // whole wave exits together
    const int A = (int)F.leaf_ids[wid];
    const int4 ga = F.nodes.geo[A];
    const int4 la = F.nodes.link[A];
    const float unit = F.unit[F.nodes.sys[A]];
    const float eps2 = F.eps2;
    unsigned long long inter = 0;
    for (int t0 = 0; t0 < la.y; t0 += WAVE) {
        const int nt = min(WAVE, la.y - t0);
        const int splits = WAVE / nt;
        const int ti = lane % nt;
        const int sp = lane / nt;
        const bool active = sp < splits;
        const int i = la.x + t0 + ti;
        const float4 me = F.pf[i];
        float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
        for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r])
        for (uint32_t e = F.lists.p_start[r], e1 = e + F.lists.p_cnt[r]; e < e1; ++e) {
            const int B = F.lists.p_pool[e];
            const int4 gb = F.nodes.geo[B];
            const int4 lb = F.nodes.link[B];
            const float3 d = centre_diff(ga, gb, F.bits, unit);
            const float dx = d.x - me.x, dy = d.y - me.y, dz = d.z - me.z;
            const float4* src = F.pf + lb.x;
            const int jn = active ? lb.y : 0;
            if (t0 == 0 && lane == 0) inter += (unsigned long long)la.y * (unsigned long long)(B == A ? lb.y - 1 : lb.y);
            if (B != A) {
                for (int j = sp; j < jn; j += splits) {
                    const float4 s = src[j];
                    const float rx = dx + s.x, ry = dy + s.y, rz = dz + s.z;
                    const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                    const float inv = rsqrtf(r2);
                    const float qi = s.w * inv;
                    const float qi3 = qi * inv * inv;
                    ax = fmaf(rx, qi3, ax);
                    ay = fmaf(ry, qi3, ay);
                    az = fmaf(rz, qi3, az);
                    ph += qi;
                }
            } else {
                const int self = t0 + ti;
                for (int j = sp; j < jn; j += splits) {
                    const float4 s = src[j];
                    const float rx = dx + s.x, ry = dy + s.y, rz = dz + s.z;
                    const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                    const float inv = rsqrtf(r2);
                    const float q = j == self ? 0.f : s.w;
                    const float qi = q * inv;
                    const float qi3 = qi * inv * inv;
                    ax = fmaf(rx, qi3, ax);
                    ay = fmaf(ry, qi3, ay);
                    az = fmaf(rz, qi3, az);
                    ph += qi;
                }
            }
        }
        // combine the source strides of each target (lane ti + q*nt holds
        // stride q of target ti; every lane takes part in the shuffles)
        {
            const float ax0 = ax, ay0 = ay, az0 = az, ph0 = ph;
            for (int q = 1; q < splits; ++q) {
                const int srcl = ti + q * nt;
                const float ox = __shfl(ax0, srcl, WAVE), oy = __shfl(ay0, srcl, WAVE);
                const float oz = __shfl(az0, srcl, WAVE), ow = __shfl(ph0, srcl, WAVE);
                ax += ox; ay += oy; az += oz; ph += ow;
            }
        }
        if (sp == 0) {
            float lphi, gx, gy, gz;
            const float* Lsrc = F.L + (size_t)A * F.ncp;
            float Lr[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) Lr[k] = Lsrc[k];
            l2p_eval<P>(me.x, me.y, me.z, Lr, lphi, gx, gy, gz);
            const double qm = S.charge[i] / S.mass[i];
            const float fac = (float)qm * F.G;
            const float axo = fac * (ax + gx), ayo = fac * (ay + gy), azo = fac * (az + gz);
            S.ax[i] = axo;
            S.ay[i] = ayo;
            S.az[i] = azo;
            S.pot[i] = -F.G * (ph + lphi);
            if (kick_dt != 0.0) {
                S.vx[i] += (double)axo * kick_dt;
                S.vy[i] += (double)ayo * kick_dt;
                S.vz[i] += (double)azo * kick_dt;
            }
        }
    }
    if (lane == 0 && inter) atomicAdd(F.p2p_inter, inter);
}

// ---------------------------------------------------------------------------
// P2P, staged: one wavefront per target leaf, sources flattened across the
// leaf's P2P list.  Up to 64 list entries at a time are read by the lanes
// (one entry each: centre offset, first particle, count) and prefix-summed;
// then 64 source particles at a time are gathered by a shuffle binary search
// over that prefix, shifted into the target leaf's frame and staged in LDS.
// Lanes are (target particle) x (source stride): with nt targets, lane l
// takes target l % nt and staged sources j == l / nt (mod 64 / nt), so the
// wave stays full for any leaf size and every lane runs the same trip count.
constexpr int P2P_WAVES = 4;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int P>
__global__ __launch_bounds__(WAVE * P2P_WAVES) void p2p_staged_kernel(FmmParams F, State S, uint32_t num_leaves,
                                                                     double kick_dt) {
    constexpr int NC = ncoef(P);
    __shared__ float4 stage_s[P2P_WAVES][WAVE];
    __shared__ int stage_i[P2P_WAVES][WAVE];
    const int wslot = threadIdx.x / WAVE;
    const int lane = lane_id();
    float4* stage = stage_s[wslot];
    int* sidx = stage_i[wslot];
    const uint32_t nw = gridDim.x * P2P_WAVES;
    unsigned long long inter = 0;  // particle pairs, summed over the lanes at the end
    for (uint32_t w = blockIdx.x * P2P_WAVES + wslot; w < num_leaves; w += nw) {
        const int A = (int)F.leaf_ids[w];
        const int4 ga = F.nodes.geo[A];
        const int4 la = F.nodes.link[A];
        const float unit = F.unit[F.nodes.sys[A]];
        const float eps2 = F.eps2;
        for (int t0 = 0; t0 < la.y; t0 += WAVE) {
            const int nt = min(WAVE, la.y - t0);
            const int splits = WAVE / nt;
            const int ti = lane % nt;
            const int sp = lane / nt;
            const int i = la.x + t0 + ti;
            const float4 me = F.pf[i];
            float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
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
                    if (t0 == 0) inter += (unsigned long long)la.y * (unsigned long long)(B == A ? cnt - 1 : cnt);
                }
                // inclusive prefix of the counts across the wave
                int incl = cnt;
#pragma unroll
                for (int o = 1; o < WAVE; o <<= 1) {
                    const int y = __shfl_up(incl, o, WAVE);
                    if (lane >= o) incl += y;
                }
                const int total = __shfl(incl, WAVE - 1, WAVE);
                const int excl = incl - cnt;
                for (int f0 = 0; f0 < total; f0 += WAVE) {
                    const int fidx = f0 + lane;
                    // entry k: the last lane with excl <= fidx (binary search over lanes)
                    int lo = 0, hi = ne - 1;
#pragma unroll
                    for (int step = 0; step < 6; ++step) {
                        const int mid = (lo + hi + 1) >> 1;
                        const int em = __shfl(excl, mid, WAVE);
                        if (em <= fidx) lo = mid; else hi = mid - 1;
                    }
                    const int kb = __shfl(beg, lo, WAVE), ke = __shfl(excl, lo, WAVE);
                    const float kx = __shfl(dx, lo, WAVE), ky = __shfl(dy, lo, WAVE), kz = __shfl(dz, lo, WAVE);
                    if (fidx < total) {
                        const int src = kb + (fidx - ke);
                        const float4 s = F.pf[src];
                        stage[lane] = make_float4(kx + s.x, ky + s.y, kz + s.z, s.w);
                        sidx[lane] = src;
                    }
                    wave_lds_sync();
                    const int nchunk = min(WAVE, total - f0);
                    if (sp < splits) {
                        if (eps2 > 0.f) {
                            // the self pair has r = 0 exactly: no force, potential
                            // q_i/eps (removed after the loop)
#pragma unroll 4
                            for (int j = sp; j < nchunk; j += splits) {
                                const float4 s = stage[j];
                                const float rx = s.x - me.x, ry = s.y - me.y, rz = s.z - me.z;
                                const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                                const float inv = rsqrtf(r2);
                                const float qi = s.w * inv;
                                const float qi3 = qi * inv * inv;
                                ax = fmaf(rx, qi3, ax);
                                ay = fmaf(ry, qi3, ay);
                                az = fmaf(rz, qi3, az);
                                ph += qi;
                            }
                        } else {
                            for (int j = sp; j < nchunk; j += splits) {
                                const float4 s = stage[j];
                                const float rx = s.x - me.x, ry = s.y - me.y, rz = s.z - me.z;
                                const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
                                const float inv = sidx[j] == i ? 0.f : rsqrtf(r2);
                                const float qi = s.w * inv;
                                const float qi3 = qi * inv * inv;
                                ax = fmaf(rx, qi3, ax);
                                ay = fmaf(ry, qi3, ay);
                                az = fmaf(rz, qi3, az);
                                ph += qi;
                            }
                        }
                    }
                    wave_lds_sync();
                }
            }
            // the self term q_i/eps sits in exactly one stride's partial sum
            if (eps2 > 0.f && sp == 0) ph -= me.w * rsqrtf(eps2);
            // combine the source strides of each target
            {
                const float ax0 = ax, ay0 = ay, az0 = az, ph0 = ph;
                for (int q = 1; q < splits; ++q) {
                    const int srcl = ti + q * nt;
                    const float ox = __shfl(ax0, srcl, WAVE), oy = __shfl(ay0, srcl, WAVE);
                    const float oz = __shfl(az0, srcl, WAVE), ow = __shfl(ph0, srcl, WAVE);
                    ax += ox; ay += oy; az += oz; ph += ow;
                }
            }
            if (sp == 0) {
                float lphi, gx, gy, gz;
                const float* Lsrc = F.L + (size_t)A * F.ncp;
                float Lr[NC];
#pragma unroll
                for (int k = 0; k < NC; ++k) Lr[k] = Lsrc[k];
                l2p_eval<P>(me.x, me.y, me.z, Lr, lphi, gx, gy, gz);
                const float fac = (float)(S.charge[i] / S.mass[i]) * F.G;
                const float axo = fac * (ax + gx), ayo = fac * (ay + gy), azo = fac * (az + gz);
                S.ax[i] = axo;
                S.ay[i] = ayo;
                S.az[i] = azo;
                S.pot[i] = -F.G * (ph + lphi);
                if (kick_dt != 0.0) {
                    S.vx[i] += (double)axo * kick_dt;
                    S.vy[i] += (double)ayo * kick_dt;
                    S.vz[i] += (double)azo * kick_dt;
                }
            }
        }
    }
    inter = wave_sum(inter);
    if (lane == 0 && inter) atomicAdd(F.p2p_inter, inter);
}

// ---------------------------------------------------------------------------
// M2L grouped by translation offset.  Same-level pairs of one system share
// their derivative tensor whenever they share the offset o = i_B - i_A
// (cells at level l, R = -o h_l): the tensor of every offset in [-K, K]^3 is
// built once per level and step (FP64, stored FP32) and the M2L kernel only
// contracts it.  Cross-level pairs (a leaf kept against a finer or coarser
// cell) build their tensor per pair.  One wavefront per target cell; lanes
// stride its source list and meet in a wave reduction.
template <int P>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 2))) void dtable_kernel(FmmParams F, int level0,
                                                                                                int nlev, int K) {
    constexpr int NC = ncoef(P);
    const int side = 2 * K + 1;
    const int per = side * side * side;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * nlev) return;
    const int lv = t / per, o = t % per;
    const int ox = o / (side * side) - K, oy = (o / side) % side - K, oz = o % side - K;
    const double h = ldexp(F.box[0].w, -(level0 + lv));
    double D[NC];
    derivs<P, double>(-ox * h, -oy * h, -oz * h, (double)F.eps2_64, D);
    float* out = F.dtab + (size_t)t * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k] = (float)D[k];
}

template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2l_grouped_kernel(
    FmmParams F, uint32_t num_nodes) {
    constexpr int NC = ncoef(P);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const int lane = lane_id();
    const int K = F.dK, side = 2 * K + 1;
    for (uint32_t A = wid; A < num_nodes; A += nw) {
        const int4 ga = F.nodes.geo[A];
        const int la = ga.w & 0xff;
        const float unit = F.unit[F.nodes.sys[A]];
        float L[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = 0.f;
        for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r])
        for (uint32_t e = F.lists.m_start[r] + lane, e1 = F.lists.m_start[r] + F.lists.m_cnt[r]; e < e1; e += WAVE) {
            const int B = F.lists.m_pool[e];
            const int4 gb = F.nodes.geo[B];
            const int ox = gb.x - ga.x, oy = gb.y - ga.y, oz = gb.z - ga.z;
            float D[NC];
            const bool tab = F.dtab && (gb.w & 0xff) == la && la >= F.dlevel0 && la < F.dlevel0 + F.dnlev &&
                             abs(ox) <= K && abs(oy) <= K && abs(oz) <= K;
            if (tab) {
                const int o = ((ox + K) * side + (oy + K)) * side + (oz + K);
                const float4* src = reinterpret_cast<const float4*>(
                    F.dtab + ((size_t)(la - F.dlevel0) * side * side * side + o) * F.ncp);
#pragma unroll
                for (int q = 0; q < (NC + 3) / 4; ++q) {
                    const float4 v = src[q];
                    if (4 * q + 0 < NC) D[4 * q + 0] = v.x;
                    if (4 * q + 1 < NC) D[4 * q + 1] = v.y;
                    if (4 * q + 2 < NC) D[4 * q + 2] = v.z;
                    if (4 * q + 3 < NC) D[4 * q + 3] = v.w;
                }
            } else {
                const float3 d = centre_diff(gb, ga, F.bits, unit);  // R = c_A - c_B
                derivs<P>(d.x, d.y, d.z, F.eps2, D);
            }
            m2l_apply<P>(D, F.M + (size_t)B * F.ncp, L);
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = wave_sum(L[k]);
        if (lane == 0) {
            float* out = F.L + (size_t)A * F.ncp;
#pragma unroll
            for (int k = 0; k < NC; ++k) out[k] = L[k];
        }
    }
}

template <int P>
static void launch_p2m_t(const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st) {
    hipLaunchKernelGGL((p2m_kernel<P>), dim3(ceil_div(num_leaves, 4)), dim3(256), 0, st, F, S, num_leaves);
}
template <int P>
static void launch_m2m_t(const FmmParams& F, uint32_t count, uint32_t off, hipStream_t st) {
    hipLaunchKernelGGL((m2m_kernel<P>), dim3(ceil_div(count, 64)), dim3(64), 0, st, F, count, off);
}
template <int P>
static void launch_m2l_t(const FmmParams& F, uint32_t num_nodes, hipStream_t st) {
    hipLaunchKernelGGL((m2l_plain_kernel<P>), dim3(ceil_div(num_nodes, 256)), dim3(256), 0, st, F, num_nodes);
}
template <int P>
static void launch_l2l_t(const FmmParams& F, uint32_t count, uint32_t off, hipStream_t st) {
    hipLaunchKernelGGL((l2l_kernel<P>), dim3(ceil_div(count, 64)), dim3(64), 0, st, F, count, off);
}
template <int P>
static void launch_p2p_t(const FmmParams& F, const State& S, uint32_t num_leaves, double kick_dt, hipStream_t st) {
    hipLaunchKernelGGL((p2p_kernel<P>), dim3(ceil_div(num_leaves, 4)), dim3(256), 0, st, F, S, num_leaves, kick_dt);
}

template <int P>
static void launch_p2p_staged_t(const FmmParams& F, const State& S, uint32_t num_leaves, double kick_dt,
                                hipStream_t st) {
    const unsigned grid = std::min<unsigned>(ceil_div(num_leaves, P2P_WAVES), 65536u);
    hipLaunchKernelGGL((p2p_staged_kernel<P>), dim3(grid), dim3(WAVE * P2P_WAVES), 0, st, F, S, num_leaves, kick_dt);
}
template <int P>
static void launch_dtable_t(const FmmParams& F, int level0, int nlev, int K, hipStream_t st) {
    const int side = 2 * K + 1;
    const unsigned total = (unsigned)(side * side * side * nlev);
    hipLaunchKernelGGL((dtable_kernel<P>), dim3(ceil_div(total, 128)), dim3(128), 0, st, F, level0, nlev, K);
}
template <int P>
static void launch_m2l_grouped_t(const FmmParams& F, uint32_t num_nodes, hipStream_t st) {
    const unsigned grid = std::min<unsigned>(ceil_div(num_nodes, 4), 65536u);
    hipLaunchKernelGGL((m2l_grouped_kernel<P>), dim3(grid), dim3(256), 0, st, F, num_nodes);
}

#define NBH_DISPATCH_P(p, call)          \
    switch (p) {                         \
        case 1: { constexpr int PP = 1; call; } break; \
        case 2: { constexpr int PP = 2; call; } break; \
        case 3: { constexpr int PP = 3; call; } break; \
        case 4: { constexpr int PP = 4; call; } break; \
        case 5: { constexpr int PP = 5; call; } break; \
        case 6: { constexpr int PP = 6; call; } break; \
        default: break;                  \
    }

void launch_p2m(int p, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st) {
    if (!num_leaves) return;
    NBH_DISPATCH_P(p, launch_p2m_t<PP>(F, S, num_leaves, st));
}
void launch_m2m_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st) {
    (void)level;
    if (!count) return;
    NBH_DISPATCH_P(p, launch_m2m_t<PP>(F, count, off, st));
}
void launch_m2l_plain(int p, const FmmParams& F, uint32_t num_nodes, hipStream_t st) {
    if (!num_nodes) return;
    NBH_DISPATCH_P(p, launch_m2l_t<PP>(F, num_nodes, st));
}
void launch_l2l_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st) {
    (void)level;
    if (!count) return;
    NBH_DISPATCH_P(p, launch_l2l_t<PP>(F, count, off, st));
}
void launch_p2p_staged(int p, const FmmParams& F, const State& S, uint32_t num_leaves, double kick_dt,
                       hipStream_t st) {
    if (!num_leaves) return;
    NBH_DISPATCH_P(p, launch_p2p_staged_t<PP>(F, S, num_leaves, kick_dt, st));
}
void launch_dtable(int p, const FmmParams& F, int level0, int nlev, int K, hipStream_t st) {
    if (nlev <= 0) return;
    NBH_DISPATCH_P(p, launch_dtable_t<PP>(F, level0, nlev, K, st));
}
void launch_m2l_grouped(int p, const FmmParams& F, uint32_t num_nodes, hipStream_t st) {
    if (!num_nodes) return;
    NBH_DISPATCH_P(p, launch_m2l_grouped_t<PP>(F, num_nodes, st));
}
void launch_p2p(int p, const FmmParams& F, const State& S, uint32_t num_leaves, double kick_dt, hipStream_t st) {
    if (!num_leaves) return;
    NBH_DISPATCH_P(p, launch_p2p_t<PP>(F, S, num_leaves, kick_dt, st));
}

// ---------------------------------------------------------------------------
__global__ void integrate_kernel(State S, uint32_t n, int mode, double dtk, double dtd) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3] = {S.vx[i], S.vy[i], S.vz[i]};
    const double a[3] = {S.ax[i], S.ay[i], S.az[i]};
    double x[3] = {S.x[i], S.y[i], S.z[i]};
    if (mode == 0) {          // KDK opening half kick + drift
        for (int k = 0; k < 3; ++k) { v[k] += a[k] * dtk; x[k] += v[k] * dtd; }
    } else if (mode == 1) {   // explicit Euler, open_cl_simulation.cpp:602-607
        for (int k = 0; k < 3; ++k) { x[k] += v[k] * dtd; v[k] += a[k] * dtk; }
    } else {                  // kick-drift, naive_simulation.cpp:28-41
        for (int k = 0; k < 3; ++k) { v[k] += a[k] * dtk; x[k] += v[k] * dtd; }
    }
    S.vx[i] = v[0]; S.vy[i] = v[1]; S.vz[i] = v[2];
    S.x[i] = x[0]; S.y[i] = x[1]; S.z[i] = x[2];
}

void launch_integrate(const State& S, uint32_t n, int mode, double dt_k, double dt_d, hipStream_t st) {
    hipLaunchKernelGGL(integrate_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, S, n, mode, dt_k, dt_d);
}

// ---------------------------------------------------------------------------
// Direct sum on the GPU (replaces NaiveSimulation::step's O(N^2) loop,
// naive_simulation.cpp:8-35).  Source tiles staged in LDS; FP32 positions
// relative to the system's box centre.  compat_law reproduces the reference
// naive force law: separation (dx, dx, dx), no softening (:10-17).
constexpr int DS_THREADS = 256;

__global__ __launch_bounds__(DS_THREADS) void direct_f32_kernel(State S, uint32_t n, const uint32_t* sys_off,
                                                                int nsys, const double4* box, float G, float eps2,
                                                                int compat, double kick_dt) {
    __shared__ float4 tile[DS_THREADS];
    const int s = blockIdx.y;
    const uint32_t b = sys_off[s], e = sys_off[s + 1];
    const uint32_t i = b + blockIdx.x * DS_THREADS + threadIdx.x;
    if (b + blockIdx.x * DS_THREADS >= e) return;
    const double4 bx = box[s];
    const double h = 0.5 * bx.w;
    const double cx = bx.x + h, cy = bx.y + h, cz = bx.z + h;
    const bool valid = i < e;
    float px = 0, py = 0, pz = 0;
    if (valid) { px = (float)(S.x[i] - cx); py = (float)(S.y[i] - cy); pz = (float)(S.z[i] - cz); }
    float ax = 0, ay = 0, az = 0, ph = 0;
    for (uint32_t j0 = b; j0 < e; j0 += DS_THREADS) {
        const uint32_t j = j0 + threadIdx.x;
        __syncthreads();
        tile[threadIdx.x] = j < e ? make_float4((float)(S.x[j] - cx), (float)(S.y[j] - cy), (float)(S.z[j] - cz),
                                                (float)S.charge[j])
                                  : make_float4(0, 0, 0, 0);
        __syncthreads();
        const uint32_t jn = min((uint32_t)DS_THREADS, e - j0);
        if (!valid) continue;
        if (!compat) {
            for (uint32_t k = 0; k < jn; ++k) {
                const float4 q = tile[k];
                const float rx = q.x - px, ry = q.y - py, rz = q.z - pz;
                const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                const float inv = rsqrtf(r2);
                const float w = (j0 + k == i) ? 0.f : q.w * inv;
                const float w3 = w * inv * inv;
                ax = fmaf(rx, w3, ax); ay = fmaf(ry, w3, ay); az = fmaf(rz, w3, az);
                ph += w;
            }
        } else {
            for (uint32_t k = 0; k < jn; ++k) {
                if (j0 + k == i) continue;
                const float4 q = tile[k];
                const float d = q.x - px;
                const float r = sqrtf(3.0f * d * d);
                const float w = q.w * d / (r * r * r);
                ax += w; ay += w; az += w;
                ph += q.w / r;
            }
        }
    }
    if (!valid) return;
    const float fac = (float)(S.charge[i] / S.mass[i]) * G;
    S.ax[i] = fac * ax; S.ay[i] = fac * ay; S.az[i] = fac * az;
    S.pot[i] = -G * ph;
    if (kick_dt != 0.0) {
        S.vx[i] += (double)(fac * ax) * kick_dt;
        S.vy[i] += (double)(fac * ay) * kick_dt;
        S.vz[i] += (double)(fac * az) * kick_dt;
    }
}

void launch_direct_f32(const State& S, uint32_t n, const uint32_t* sys_off, int nsys, const double4* box, float G,
                       float eps2, int compat_law, double kick_dt, hipStream_t st) {
    // grid.x sized for the largest system: callers pass n = max system size
    hipLaunchKernelGGL(direct_f32_kernel, dim3(ceil_div(n, DS_THREADS), nsys), dim3(DS_THREADS), 0, st, S, n, sys_off,
                       nsys, box, G, eps2, compat_law, kick_dt);
}

// FP64 direct sum for selected targets (sorted slots): one wavefront per target.
__global__ __launch_bounds__(256) void direct_fp64_kernel(State S, const uint32_t* sys_off, int nsys,
                                                          const uint32_t* targets, uint32_t nt, double G,
                                                          double eps2, double* acc, double* pot) {
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (w >= nt) return;
    const uint32_t i = targets[w];
    int s = 0;
    {
        int lo = 0, hi = nsys - 1;
        while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (sys_off[mid] <= i) lo = mid; else hi = mid - 1; }
        s = lo;
    }
    const uint32_t b = sys_off[s], e = sys_off[s + 1];
    const double px = S.x[i], py = S.y[i], pz = S.z[i];
    double ax = 0, ay = 0, az = 0, ph = 0;
    for (uint32_t j = b + lane_id(); j < e; j += WAVE) {
        if (j == i) continue;
        const double rx = S.x[j] - px, ry = S.y[j] - py, rz = S.z[j] - pz;
        const double r2 = rx * rx + ry * ry + rz * rz + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double q = S.charge[j] * inv;
        const double q3 = q * inv * inv;
        ax += rx * q3; ay += ry * q3; az += rz * q3;
        ph += q;
    }
    ax = wave_sum(ax); ay = wave_sum(ay); az = wave_sum(az); ph = wave_sum(ph);
    if (lane_id() == 0) {
        const double fac = S.charge[i] / S.mass[i] * G;
        acc[3 * w] = fac * ax; acc[3 * w + 1] = fac * ay; acc[3 * w + 2] = fac * az;
        pot[w] = -G * ph;
    }
}

void launch_direct_fp64(const State& S, const uint32_t* sys_off, int nsys, const uint32_t* targets, uint32_t nt,
                        double G, double eps2, double* acc, double* pot, hipStream_t st) {
    hipLaunchKernelGGL(direct_fp64_kernel, dim3(ceil_div(nt, 4)), dim3(256), 0, st, S, sys_off, nsys, targets, nt, G,
                       eps2, acc, pot);
}

// ---------------------------------------------------------------------------
// Diagnostics partials per block: KE, PE, P[3], Lang[3]
constexpr int DG_THREADS = 256;
uint32_t diagnostics_blocks(uint32_t n) { return min(1024u, ceil_div(n, DG_THREADS)); }

__global__ __launch_bounds__(DG_THREADS) void diagnostics_kernel(State S, uint32_t n, double* partial) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * DG_THREADS + threadIdx.x; i < n; i += gridDim.x * DG_THREADS) {
        const double m = S.mass[i];
        const double vx = S.vx[i], vy = S.vy[i], vz = S.vz[i];
        const double x = S.x[i], y = S.y[i], z = S.z[i];
        v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[1] += 0.5 * S.charge[i] * (double)S.pot[i];
        v[2] += m * vx; v[3] += m * vy; v[4] += m * vz;
        v[5] += m * (y * vz - z * vy); v[6] += m * (z * vx - x * vz); v[7] += m * (x * vy - y * vx);
    }
    __shared__ double sh[8][DG_THREADS / WAVE];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double r = wave_sum(v[k]);
        if (lane_id() == 0) sh[k][threadIdx.x / WAVE] = r;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double r = 0;
        for (int w = 0; w < DG_THREADS / WAVE; ++w) r += sh[threadIdx.x][w];
        partial[blockIdx.x * 8 + threadIdx.x] = r;
    }
}

void launch_diagnostics(const State& S, uint32_t n, double* partial, uint32_t nblk, hipStream_t st) {
    hipLaunchKernelGGL(diagnostics_kernel, dim3(nblk), dim3(DG_THREADS), 0, st, S, n, partial);
}

__global__ void inverse_perm_kernel(const uint32_t* id, uint32_t* slot_of, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slot_of[id[i]] = i;
}
void launch_inverse_perm(const uint32_t* id, uint32_t* slot_of, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(inverse_perm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, id, slot_of, n);
}

}  // namespace nbh
