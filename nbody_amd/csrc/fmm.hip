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
// P2M (moment.cl:6-67 at order P): one lane per leaf cell, streaming its
// particles (leaves hold ~10 particles on average, so a wavefront per leaf
// would leave most lanes idle).  Also writes the particles' FP32 offsets from
// the leaf centre (used by P2P and L2P) and their leaf.
template <int P>
__global__ __launch_bounds__(256) void p2m_kernel(FmmParams F, State S, uint32_t num_leaves) {
    constexpr int NC = ncoef(P);
    if (num_leaves == DEV_COUNT) {
        if (F.C->overflow & OVF_NODES) return;
        num_leaves = F.C->num_leaves;
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < num_leaves; t += gridDim.x * blockDim.x) {
    const int A = (int)F.leaf_ids[t];
    const int4 g = F.nodes.geo[A];
    const int4 lk = F.nodes.link[A];
    const int s = F.nodes.sys[A];
    double c[3];
    centre64(g, F.box[s], F.bits, c);
    float M[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) M[k] = 0.f;
    for (int i = lk.x; i < lk.x + lk.y; ++i) {
        const float bx = (float)(S.x[i] - c[0]), by = (float)(S.y[i] - c[1]), bz = (float)(S.z[i] - c[2]);
        const float q = (float)S.charge[i];
        F.pf[i] = make_float4(bx, by, bz, q);
        F.pleaf[i] = A;
        float mono[NC];
        monomials<P>(bx, by, bz, mono);
#pragma unroll
        for (int k = 0; k < NC; ++k) M[k] = fmaf(q, mono[k], M[k]);
    }
    float* out = F.M + (size_t)A * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k] = M[k];
    }
}

// L2P of the far field, one lane per particle: phi and grad phi of its
// leaf's local expansion, written to (ax, ay, az, pot) for the P2P kernel to
// complete.
template <int P>
__global__ __launch_bounds__(256) void l2p_kernel(FmmParams F, State S, uint32_t n, double kick_dt) {
    constexpr int NC = ncoef(P);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 me = F.pf[i];
    const float* Ls = F.L + (size_t)F.pleaf[i] * F.ncp;
    float L[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) L[k] = Ls[k];
    float phi, gx, gy, gz;
    l2p_eval<P>(me.x, me.y, me.z, L, phi, gx, gy, gz);
    // the P2P kernel left the raw near-field sums in (ax, ay, az, pot)
    const float nx = S.ax[i], ny = S.ay[i], nz = S.az[i], nph = S.pot[i];
    const float fac = (float)(S.charge[i] / S.mass[i]) * F.G;
    const float axo = fac * (nx + gx), ayo = fac * (ny + gy), azo = fac * (nz + gz);
    S.ax[i] = axo;
    S.ay[i] = ayo;
    S.az[i] = azo;
    if (F.pot) S.pot[i] = -F.G * (nph + phi);
    if (kick_dt != 0.0) {
        S.vx[i] += (double)axo * kick_dt;
        S.vy[i] += (double)ayo * kick_dt;
        S.vz[i] += (double)azo * kick_dt;
    }
}

// M2M (replaces moment.cl:72-160, which sums untranslated moments): one
// thread per cell of the level; children's moments shifted to the parent
// centre.  Levels run bottom-up.
template <int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2m_kernel(FmmParams F, uint32_t count,
                                                                                           uint32_t off) {
    constexpr int NC = ncoef(P);
    if (count == DEV_COUNT) {  // off is the level
        if (F.C->overflow & OVF_NODES) return;
        count = F.C->level_count[off];
        off = F.C->level_off[off];
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
    const int A = (int)(off + t);
    const int4 g = F.nodes.geo[A];
    const int nch = g.w >> 8;
    if (nch == 0) continue;
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
}

// M2L, plain: one thread per target cell, the derivative tensor rebuilt per
// pair (replaces field.cl:153-211's order-0 cell field).
template <int P, int Q, bool POT>
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
            float D[ncoef(m2l_deg(Q))];
            derivs<m2l_deg(Q)>(d.x, d.y, d.z, F.eps2, D);
            // source moments read straight from memory; the pair table is ordered
            // by source coefficient so each M_n is loaded once and used at once
            m2l_apply<P, Q, POT>(D, F.M + (size_t)B * F.ncp, L);
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
    if (count == DEV_COUNT) {  // off is the level
        if (F.C->overflow & OVF_WALK) return;
        count = F.C->level_count[off];
        off = F.C->level_off[off];
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
    const int A = (int)(off + t);
    const int4 g = F.nodes.geo[A];
    const int par = F.nodes.link[A].w;
    if (par < 0) continue;
    const int4 gp = F.nodes.geo[par];
    const float3 d = centre_diff(gp, g, F.bits, F.unit[F.nodes.sys[A]]);  // c_child - c_parent
    float Lp[NC], Lc[NC];
    const float* sp = F.L + (size_t)par * F.ncp;
    float* sc = F.L + (size_t)A * F.ncp;
#pragma unroll
    for (int k = 0; k < NC; ++k) { Lp[k] = sp[k]; Lc[k] = sc[k]; }
    if (F.Lx) {  // the other-level and deferred pairs' part
        const float* sx = F.Lx + (size_t)A * F.ncp;
#pragma unroll
        for (int k = 0; k < NC; ++k) Lc[k] += sx[k];
    }
    l2l_apply<P>(d.x, d.y, d.z, Lp, Lc);
#pragma unroll
    for (int k = 0; k < NC; ++k) sc[k] = Lc[k];
    }
}

// ---------------------------------------------------------------------------
// P2P + L2P + force + (optional) closing half kick: one wavefront per target
// leaf.  Replaces field.cl:49-148 (pair fields into slots), force.cl:21-81
// (slot reductions) and the host integration loop.  Lanes are split as
// (target particle) x (source stride): with nt targets in a chunk, lane l
// handles target l % nt and every (64/nt)-th particle of each source leaf, so
// all 64 lanes work for any leaf occupancy; partial sums meet in LDS.
template <int P, bool POT>
__global__ __launch_bounds__(256) void p2p_kernel(FmmParams F, State S, uint32_t num_leaves) {
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
                    if (POT) ph += qi;
                }
            } else {
                const int self = t0 + ti;
                for (int j = sp; j < jn; j += splits) {
                    const float4 s = src[j];
                    const float rx = dx + s.x, ry = dy + s.y, rz = dz + s.z;
                    const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                    // the self pair contributes nothing (r2 = 0 when unsoftened)
                    const float inv = j == self ? 0.f : rsqrtf(r2);
                    const float qi = s.w * inv;
                    const float qi3 = qi * inv * inv;
                    ax = fmaf(rx, qi3, ax);
                    ay = fmaf(ry, qi3, ay);
                    az = fmaf(rz, qi3, az);
                    if (POT) ph += qi;
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
            // raw near-field sums; l2p_kernel adds the far field
            S.ax[i] = ax;
            S.ay[i] = ay;
            S.az[i] = az;
            if (POT) S.pot[i] = ph;
        }
    }
    if (lane == 0 && inter) atomicAdd(F.p2p_inter, inter);
}

// ---------------------------------------------------------------------------
// P2P, staged: one wavefront per target leaf, sources flattened across the
// leaf's P2P list.  Up to 64 list entries at a time are read by the lanes
// (one entry each: centre offset, first particle, count) and prefix-summed;
// then 64 source particles at a time are gathered (each lane finds its
// source's entry from a bitmask of the entry starts), shifted into the target
// leaf's frame and staged in LDS.
// Lanes are (target particle) x (source stride): with nt targets, lane l
// takes target l % nt and the staged sources j = sp + k splits, k < Tc, of
// stride sp = l / nt (splits = 64 / nt), so the wave stays full for any leaf
// size.  Tc = ceil(nchunk / splits) is the same for every lane: slots past the
// chunk's nchunk sources hold zero charges, every lane runs Tc iterations, and
// a partial chunk costs only the iterations that it fills.
// Tile shape: a chunk is 64 SPL sources (SPL = 1 or 2 per lane): with SPL = 2
// each staging round (entry lookup, LDS writes, three wave syncs) serves
// twice the pair iterations.
constexpr int P2P_WAVES = 4;
// stage slots: splits * ceil(nchunk / splits) <= 84 for nt in [1, 64] and
// nchunk in [1, 64], <= 147 for nchunk in [1, 128]
template <int SPL>
constexpr int p2p_stage_slots() { return SPL == 1 ? 96 : 160; }

// SOFT: eps > 0 (each target meets itself at r = 0 with a finite softened
// potential); else the self pair and the empty slots are skipped by index.
// POT: the potential sums too (else the accelerations alone).
template <bool SOFT, int SPL, bool POT>
__device__ __forceinline__ void p2p_staged_body(const FmmParams& F, const State& S, uint32_t num_leaves) {
    // (num_leaves == DEV_COUNT: the leaf count of the device counters)
    constexpr int P2P_STAGE = p2p_stage_slots<SPL>();
    constexpr int CHUNK = WAVE * SPL;
    __shared__ float4 stage_s[P2P_WAVES][P2P_STAGE];
    __shared__ int stage_i[P2P_WAVES][P2P_STAGE];
    __shared__ int2 ent_s[P2P_WAVES][WAVE];      // list entry: first particle, flattened start
    __shared__ float4 entd_s[P2P_WAVES][WAVE];   // list entry: source centre - target centre
    __shared__ unsigned long long start_s[P2P_WAVES][SPL];  // entries starting in each 64 of the chunk
    // (left as a vector value: made uniform with readfirstlane, the leaf
    // records come through scalar loads, whose waits share the counter of
    // the LDS reads, and the kernel is slower)
    const int wslot = threadIdx.x / WAVE;
    const int lane = lane_id();
    float4* stage = stage_s[wslot];
    int* sidx = stage_i[wslot];
    int2* ent = ent_s[wslot];
    float4* entd = entd_s[wslot];
    if (num_leaves == DEV_COUNT) {
        if (F.C->overflow & OVF_WALK) return;
        num_leaves = F.C->num_leaves;
    }
    // slots past the chunk are never staged: zero charges for good
    static_assert(P2P_STAGE - CHUNK <= WAVE, "one slot per lane");
    if (lane < P2P_STAGE - CHUNK) {
        stage[CHUNK + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        sidx[CHUNK + lane] = -1;
    }
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
            // ceil(x / splits) = umulhi(x + splits - 1, magic) for x <= 64
            const uint32_t magic = splits == 1 ? 0u : 0xffffffffu / (uint32_t)splits + 1u;
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
                if (lane < ne) {
                    ent[lane] = make_int2(beg, excl);
                    entd[lane] = make_float4(dx, dy, dz, 0.f);
                }
                for (int f0 = 0; f0 < total; f0 += CHUNK) {
                    // entry of flattened source fidx: the last entry with excl <=
                    // fidx.  Every entry has at least one particle, so the starts
                    // are distinct: the entries before each 64 of the chunk, plus
                    // one bit per start inside it (set with one LDS atomic)
                    if (lane < SPL) start_s[wslot][lane] = 0ull;
                    wave_lds_sync();
#pragma unroll
                    for (int hh = 0; hh < SPL; ++hh) {
                        const int g0 = f0 + hh * WAVE;
                        if (lane < ne && excl >= g0 && excl < g0 + WAVE)
                            atomicOr(&start_s[wslot][hh], 1ull << (excl - g0));
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int hh = 0; hh < SPL; ++hh) {
                        const int g0 = f0 + hh * WAVE;
                        const int fidx = g0 + lane;
                        const int kb0 = __popcll(__ballot(lane < ne && excl < g0));
                        const unsigned long long starts = start_s[wslot][hh];
                        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);  // past the sources: no charge
                        int src = -1;
                        if (fidx < total) {
                            const int k = kb0 + __popcll(starts & ((2ull << lane) - 1ull)) - 1;
                            const int2 e = ent[k];
                            const float4 d = entd[k];
                            src = e.x + (fidx - e.y);
                            const float4 s = F.pf[src];
                            v = make_float4(d.x + s.x, d.y + s.y, d.z + s.z, s.w);
                        }
                        stage[hh * WAVE + lane] = v;
                        sidx[hh * WAVE + lane] = src;
                    }
                    const int nchunk = min(CHUNK, total - f0);
                    // uniform: counted on the scalar unit
                    const int Tc = __builtin_amdgcn_readfirstlane(
                        splits == 1 ? nchunk : (int)__umulhi((uint32_t)(nchunk + splits - 1), magic));
                    const float4* blk = stage + sp;
                    const int* blki = sidx + sp;
                    wave_lds_sync();
                    if (sp < splits) {
                        if (SOFT) {
                            // the self pair has r = 0 exactly: no force, potential
                            // q_i/eps (removed after the loop); a zero-charge
                            // slot at the target's own position gives q/eps = 0
#pragma unroll 4
                            for (int k = 0; k < Tc; ++k) {
                                const float4 s = blk[k * splits];
                                const float rx = s.x - me.x, ry = s.y - me.y, rz = s.z - me.z;
                                const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, eps2)));
                                // r2 >= eps2 is a normal float: the bare
                                // v_rsq_f32, without rsqrtf's denormal scaling
                                const float inv = __builtin_amdgcn_rsqf(r2);
                                const float qi = s.w * inv;
                                const float qi3 = qi * inv * inv;
                                ax = fmaf(rx, qi3, ax);
                                ay = fmaf(ry, qi3, ay);
                                az = fmaf(rz, qi3, az);
                                if (POT) ph += qi;
                            }
                        } else {
                            // zero-charge slots (index -1) are skipped like the self pair
                            for (int k = 0; k < Tc; ++k) {
                                const float4 s = blk[k * splits];
                                const float rx = s.x - me.x, ry = s.y - me.y, rz = s.z - me.z;
                                const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
                                const int sj = blki[k * splits];
                                const float inv = (sj == i || sj < 0) ? 0.f : rsqrtf(r2);
                                const float qi = s.w * inv;
                                const float qi3 = qi * inv * inv;
                                ax = fmaf(rx, qi3, ax);
                                ay = fmaf(ry, qi3, ay);
                                az = fmaf(rz, qi3, az);
                                if (POT) ph += qi;
                            }
                        }
                    }
                    wave_lds_sync();
                }
            }
            // the self term q_i/eps sits in exactly one stride's partial sum
            if (POT && SOFT && sp == 0) ph -= me.w * __builtin_amdgcn_rsqf(eps2);
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
                // raw near-field sums; l2p_kernel adds the far field
                S.ax[i] = ax;
                S.ay[i] = ay;
                S.az[i] = az;
                if (POT) S.pot[i] = ph;
            }
        }
    }
    inter = wave_sum(inter);
    if (lane == 0 && inter) atomicAdd(F.p2p_inter, inter);
}

template <int P, int SPL, bool POT>
__global__ __launch_bounds__(WAVE * P2P_WAVES) void p2p_staged_kernel(FmmParams F, State S, uint32_t num_leaves) {
    p2p_staged_body<true, SPL, POT>(F, S, num_leaves);
}

template <int P, int SPL, bool POT>
__global__ __launch_bounds__(WAVE * P2P_WAVES) void p2p_staged_unsoftened_kernel(FmmParams F, State S,
                                                                                 uint32_t num_leaves) {
    p2p_staged_body<false, SPL, POT>(F, S, num_leaves);
}

// ---------------------------------------------------------------------------
// M2L grouped by translation offset.  Same-level pairs of one system share
// their derivative tensor whenever they share the offset o = i_B - i_A (cells
// at level l, R = -o h_l).  Same-level pairs always have |o_i| <= K =
// 2 floor(sqrt(3)/theta) + 1 (their parents were split), and by the octant
// symmetry D_m(S R) = s^m D_m(R) only offsets with o_i >= 0 are tabled, and
// of [0, K]^3 only those the criterion accepts while rejecting the parent
// offset for some target position in its parent (the offsets that occur: 273
// of 512 at theta = 1/2), then a zero row; the host maps [0, K]^3 to rows.
// Built once per step in FP64 and stored FP32.  For theta = 1/2 and Q <= 5 a
// level's table fits in LDS twice per CU (274 x 60 floats = 64 KiB): the
// grouped kernel runs one launch per level, every block loads that level's
// table once, and the only global traffic per pair is the source moment row.
// (Per-lane 144-byte gathers of both the tensor and the moment rows made this
// contraction L2-bandwidth-bound.)
template <int Q>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 2))) void dtable_kernel(
    FmmParams F, float* out, int level0, int nlev, int E, int a, int b, int shift, int nsys, const uint32_t* offs,
    int nrows) {
    constexpr int NCQ = ncoef(Q);
    const int per = offs ? nrows : E * E * E;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * nlev * nsys) return;
    const int s = t / (per * nlev), lv = (t / per) % nlev, o = t % per;
    float* row = out + (size_t)t * F.dcp;
    int ex, ey, ez;
    if (offs) {
        const uint32_t w = offs[o];
        if (w == ~0u) {
            for (int k = 0; k < F.dcp; ++k) row[k] = 0.f;
            return;
        }
        ex = (int)(w & 255u);
        ey = (int)((w >> 8) & 255u);
        ez = (int)((w >> 16) & 255u);
    } else {
        ex = o / (E * E);
        ey = (o / E) % E;
        ez = o % E;
    }
    const double h = ldexp(F.box[s].w, -(level0 + lv + shift));
    double D[NCQ];
    derivs<Q, double>(-(a + b * ex) * h, -(a + b * ey) * h, -(a + b * ez) * h, (double)F.eps2_64, D);
#pragma unroll
    for (int k = 0; k < NCQ; ++k) row[k] = (float)D[k];
    for (int k = NCQ; k < F.dcp; ++k) row[k] = 0.f;
}

// The pipelined form (two entries of expansion and tensor in 256 VGPRs) up to
// p = 4, q = 5; above, the one-entry form with 256 threads and the table in
// global memory
constexpr bool m2l_pipelined(int p, int qc) { return ncoef(p) + ncoef(m2l_deg(qc)) <= 91; }

// Same-level part of the rows of the level's targets [off, off+count): one
// wavefront per target, lanes over its entries; writes L.  IN_LDS: the
// level's table and the offset -> row map are first copied to LDS (when they
// fit), else read from global memory.  Two forms:
//  - PIPE (up to p = 4, q = 5): 2 wavefronts per SIMD with 256 VGPRs each,
//    every wavefront holding two entries so that one entry's loads fly while
//    the other's products run.  256 threads where two workgroups' tables fit
//    one CU's LDS (the table holds only the offsets that occur: 273 rows,
//    64 KiB at theta = 1/2, q = 5; one workgroup computes while the other
//    copies its table or finishes its last list) or where the table stays in
//    global memory; 512 where one table fits;
//  - else 256 threads, one entry at a time with the next source id
//    prefetched, the table in global memory.
//  POT: the products into L[0] too (the potential; else accelerations only).
template <int P, int Q, bool IN_LDS, int NT, bool PIPE, bool POT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PIPE ? 2 : 1, PIPE ? 2 : 8)))
void m2l_lds_kernel(FmmParams F, uint32_t off, uint32_t count, int lvl) {
    constexpr int NC = ncoef(P), NCQ = ncoef(m2l_deg(Q));
    constexpr int NQD = (NCQ + 3) / 4;  // float4 per tensor row (the host sets dcp = 4 NQD)
    // LDS row stride in float4: odd, so that the rows that the lanes of a
    // wavefront read at once start in different bank groups (an even stride,
    // 14 at q = 5, maps every row to one of four groups of the 32 banks)
    constexpr int NQS = NQD | 1;
    extern __shared__ float4 dtab_lds[];
    // launched beside the walk: after an overflow at this depth or above the
    // level's rows are incomplete, and the host repeats the walk and M2L
    if (F.C->overflow & OVF_WALK) return;
    if (count == DEV_COUNT) {
        count = F.C->level_count[lvl + F.dlevel0];
        off = F.C->level_off[lvl + F.dlevel0];
    }
    const uint32_t wpb = NT / WAVE;
    if (blockIdx.x * wpb >= count) return;  // uniform: no target for this block
    const int K = F.dK, K1 = K + 1;
    const float4* src = reinterpret_cast<const float4*>(F.dtab + (size_t)lvl * F.dsz * F.dcp);
    uint16_t* map_lds = reinterpret_cast<uint16_t*>(dtab_lds + F.dsz * NQS);
    const uint16_t* map = IN_LDS ? map_lds : F.dmap;
    if (IN_LDS) {
        const int tot = F.dsz * NQD;
        for (int i = threadIdx.x; i < tot; i += NT) {
            const int r = i / NQD;
            dtab_lds[r * NQS + (i - r * NQD)] = src[i];
        }
        for (int i = threadIdx.x; i < K1 * K1 * K1; i += NT) map_lds[i] = F.dmap[i];
        __syncthreads();
    }
    constexpr int RS = IN_LDS ? NQS : NQD;
    // one system per launch when the table is in LDS; else each target reads
    // its own system's table from global memory
    const size_t sys_stride = (size_t)F.dnlev * F.dsz * NQD;
    const int lane = lane_id();
    int rs0, rsn;  // coefficients [rs0, rs0 + rsn) of this lane after the reduction
    rs_slot(NC, 6, lane, rs0, rsn);
    for (uint32_t t = blockIdx.x * wpb + threadIdx.x / WAVE; t < count; t += gridDim.x * wpb) {
        const uint32_t A = off + t;
        const int4 ga = F.nodes.geo[A];
        const float4* tab = IN_LDS ? dtab_lds : src + (size_t)F.nodes.sys[A] * sys_stride;
        float L[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = 0.f;
        // every same-level entry of a level-l target is emitted at depth l,
        // in its first row; the rows behind it (a leaf carried deeper) hold
        // other levels only and may still be being written by the walk
        const uint32_t r = F.lists.first[A];
        {
            const uint32_t s0 = F.lists.m_start[r];
            const uint32_t e1 = s0 + F.lists.m_cnt[r] - F.lists.mx_cnt[r];
            // The products run in tensor coefficient order (m2l_apply_by_d):
            // the tensor streams in and the registers it held take loads in
            // flight.
            if constexpr (PIPE) {
                // two entries in registers: while one entry's products run,
                // the next entry's tensor (LDS) and moments (global) are
                // loaded, with the cell coordinates one further and the
                // source id two further ahead.  Every load is issued on
                // every lane (a lane past its list reads pool entry 0, that
                // node and the zero row of the table, and computes nothing),
                // so that the compiler counts the loads in flight exactly and
                // waits only for the ones an instruction needs; the octant
                // signs are applied when the entry's products start.
                struct Ent {
                    float D[NCQ];
                    float M[NC];
                    uint32_t neg;
                };
                // (the zero row where gb = ga)
                auto fetch = [&](Ent& X, const int4 gb, const int B) {
                    const int ox = gb.x - ga.x, oy = gb.y - ga.y, oz = gb.z - ga.z;
                    X.neg = (uint32_t)(ox < 0) | ((uint32_t)(oy < 0) << 1) | ((uint32_t)(oz < 0) << 2);
                    const float4* row = tab + (uint32_t)map[(abs(ox) * K1 + abs(oy)) * K1 + abs(oz)] * RS;
#pragma unroll
                    for (int q = 0; q < NQD; ++q) {
                        const float4 v = row[q];
                        if (4 * q + 0 < NCQ) X.D[4 * q + 0] = v.x;
                        if (4 * q + 1 < NCQ) X.D[4 * q + 1] = v.y;
                        if (4 * q + 2 < NCQ) X.D[4 * q + 2] = v.z;
                        if (4 * q + 3 < NCQ) X.D[4 * q + 3] = v.w;
                    }
                    const float4* mr = reinterpret_cast<const float4*>(F.M + (size_t)B * F.ncp);
#pragma unroll
                    for (int q = 0; q < (NC + 3) / 4; ++q) {
                        const float4 v = mr[q];
                        if (4 * q + 0 < NC) X.M[4 * q + 0] = v.x;
                        if (4 * q + 1 < NC) X.M[4 * q + 1] = v.y;
                        if (4 * q + 2 < NC) X.M[4 * q + 2] = v.z;
                        if (4 * q + 3 < NC) X.M[4 * q + 3] = v.w;
                    }
                };
                auto apply = [&](Ent& X) {
                    octant_signs<m2l_deg(Q)>(X.neg, X.D);
                    m2l_apply_by_d<P, Q, POT>(X.D, X.M, L);
                };
                // source id of entry i; past the list, that of pool entry 0
                // (a node id, so its loads stay in bounds; the value itself
                // is not selected away, which would wait for the load here)
                auto source = [&](uint32_t i) { return F.lists.m_pool[i < e1 ? i : 0u]; };
                uint32_t e = s0 + lane;
                const int B0 = source(e);
                int Bn = source(e + WAVE);
                Ent X, Y;
                {
                    const int4 g0 = F.nodes.geo[B0];
                    fetch(X, e < e1 ? g0 : ga, B0);
                }
                int4 gn = F.nodes.geo[Bn];
                int Bnn = source(e + 2 * WAVE);
                // at the top: X holds entry e; gn, Bn are entry e+W's cell
                // and source, Bnn entry e+2W's source.  (The scheduling
                // barriers keep the loads ahead of the products: the
                // scheduler would otherwise sink them to the products that
                // use them, to save registers.)
                auto stage = [&](Ent& Z) {
                    fetch(Z, e + WAVE < e1 ? gn : ga, Bn);
                    gn = F.nodes.geo[Bnn];
                    Bn = Bnn;
                    Bnn = source(e + 3 * WAVE);
                    __builtin_amdgcn_sched_barrier(0);
                };
                while (e < e1) {
                    stage(Y);
                    apply(X);
                    e += WAVE;
                    if (e >= e1) break;
                    stage(X);
                    apply(Y);
                    e += WAVE;
                }
            } else {
            // one entry at a time; the lane's next source id is loaded before
            // this entry's products, so that of the dependent loads pool ->
            // geo -> moments only two wait per entry (with the tensor order,
            // 6% off the step of the time against source-coefficient order
            // and no prefetch; either alone gained 0-2%)
            uint32_t e = s0 + lane;
            int Bn = e < e1 ? F.lists.m_pool[e] : 0;
            for (; e < e1; e += WAVE) {
                const int B = Bn;
                Bn = e + WAVE < e1 ? F.lists.m_pool[e + WAVE] : 0;
                const int4 gb = F.nodes.geo[B];
                const int ox = gb.x - ga.x, oy = gb.y - ga.y, oz = gb.z - ga.z;
                const uint32_t neg = (uint32_t)(ox < 0) | ((uint32_t)(oy < 0) << 1) | ((uint32_t)(oz < 0) << 2);
                // (an offset outside the table maps to its zero row)
                const float4* row = tab + (uint32_t)map[(abs(ox) * K1 + abs(oy)) * K1 + abs(oz)] * RS;
                float D[NCQ];
#pragma unroll
                for (int q = 0; q < NQD; ++q) {
                    const float4 v = row[q];
                    if (4 * q + 0 < NCQ) D[4 * q + 0] = v.x;
                    if (4 * q + 1 < NCQ) D[4 * q + 1] = v.y;
                    if (4 * q + 2 < NCQ) D[4 * q + 2] = v.z;
                    if (4 * q + 3 < NCQ) D[4 * q + 3] = v.w;
                }
                octant_signs<m2l_deg(Q)>(neg, D);
                m2l_apply_by_d<P, Q, POT>(D, F.M + (size_t)B * F.ncp, L);
            }
            }
        }
        group_reduce_scatter<NC, 6>(L);
        float* out = F.L + (size_t)A * F.ncp + rs0;
#pragma unroll
        for (int j = 0; j < rs_len(NC, 6); ++j)
            if (j < rsn) out[j] = L[j];
    }
}

// Loads a tensor row (float4 x NQD) and applies the octant signs.
template <int Q>
__device__ __forceinline__ void load_tensor(const float4* row, uint32_t neg, float* D) {
    constexpr int NCQ = ncoef(Q), NQD = (NCQ + 3) / 4;
#pragma unroll
    for (int q = 0; q < NQD; ++q) {
        const float4 v = row[q];
        if (4 * q + 0 < NCQ) D[4 * q + 0] = v.x;
        if (4 * q + 1 < NCQ) D[4 * q + 1] = v.y;
        if (4 * q + 2 < NCQ) D[4 * q + 2] = v.z;
        if (4 * q + 3 < NCQ) D[4 * q + 3] = v.w;
    }
    octant_signs<Q>(neg, D);
}

// M2L with the tensor built per pair, every entry (no tables: ensembles of
// systems with different root sizes); writes L.  One wavefront per target,
// lanes over its entries.
template <int P, int Q>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2l_pairs_kernel(
    FmmParams F, uint32_t num_nodes) {
    constexpr int NC = ncoef(P);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const int lane = lane_id();
    int rs0, rsn;
    rs_slot(NC, 6, lane, rs0, rsn);
    for (uint32_t A = wid; A < num_nodes; A += nw) {
        const int4 ga = F.nodes.geo[A];
        const float unit = F.unit[F.nodes.sys[A]];
        float L[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = 0.f;
        for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r]) {
            const uint32_t e0 = F.lists.m_start[r], e1 = e0 + F.lists.m_cnt[r];
            for (uint32_t e = e0 + lane; e < e1; e += WAVE) {
                const int B = F.lists.m_pool[e];
                const int4 gb = F.nodes.geo[B];
                const float3 d = centre_diff(gb, ga, F.bits, unit);  // R = c_A - c_B
                float D[ncoef(m2l_deg(Q))];
                derivs<m2l_deg(Q)>(d.x, d.y, d.z, F.eps2, D);
                m2l_apply<P, Q>(D, F.M + (size_t)B * F.ncp, L);
            }
        }
        group_reduce_scatter<NC, 6>(L);
        float* out = F.L + (size_t)A * F.ncp + rs0;
#pragma unroll
        for (int j = 0; j < rs_len(NC, 6); ++j)
            if (j < rsn) out[j] = L[j];
    }
}

// Cross-level entry: its tensor row in F.xtab and octant signs, or -1 when
// the levels differ by more than one (or the offset lies beyond the table).
// Centres in units of half the finer cell: 2 i + 1 at the finer level,
// 4 i + 2 at the coarser; the difference is odd.
__device__ __forceinline__ int xtab_index(const FmmParams& F, int4 ga, int4 gb, int sys, uint32_t& neg) {
    const int la = ga.w & 0xff, lb = gb.w & 0xff;
    if (abs(la - lb) != 1) return -1;
    const int lf = la > lb ? la : lb;
    const int sa = la == lf ? 1 : 2, sb = lb == lf ? 1 : 2;  // centre = sx (2 i + 1)
    const int dx = sb * (2 * gb.x + 1) - sa * (2 * ga.x + 1);
    const int dy = sb * (2 * gb.y + 1) - sa * (2 * ga.y + 1);
    const int dz = sb * (2 * gb.z + 1) - sa * (2 * ga.z + 1);
    const int ex = (abs(dx) - 1) >> 1, ey = (abs(dy) - 1) >> 1, ez = (abs(dz) - 1) >> 1;
    const int E = F.xE;
    if (ex >= E || ey >= E || ez >= E || lf < F.xlevel0 || lf >= F.xlevel0 + F.xnlev) return -1;
    neg = (uint32_t)(dx < 0) | ((uint32_t)(dy < 0) << 1) | ((uint32_t)(dz < 0) << 2);
    return (((sys * F.xnlev + lf - F.xlevel0) * E + ex) * E + ey) * E + ez;
}

// The other-level tails of the rows (after the same-level kernel): pairs one
// level apart read their tensor from F.xtab; the rest are deferred to
// m2l_far_kernel as one contiguous block per target.  Adds to L.
// Sum over the 32 lanes of each half wavefront
template <typename T>
__device__ __forceinline__ T half_sum(T v) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

template <int P, int Q, bool POT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P <= 4 ? 4 : 1))) void m2l_cross_kernel(
    FmmParams F, uint32_t num_nodes) {
    constexpr int NC = ncoef(P), NCQ = ncoef(m2l_deg(Q)), NQD = (NCQ + 3) / 4;
    if (num_nodes == DEV_COUNT) {
        if (F.C->overflow & OVF_WALK) return;
        num_nodes = F.C->num_nodes;
    }
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const int lane = lane_id();
    // two targets per wavefront, one per half: a target has ~60 such pairs,
    // and the 35-coefficient reduction is shared by the two
    const int half = lane >> 5, sub = lane & 31;
    const uint64_t hmask = half ? 0xffffffff00000000ull : 0xffffffffull;
    const uint64_t lt = lanemask_lt();
    int rs0, rsn;  // coefficients of this lane after the reduction over the half
    rs_slot(NC, 5, lane, rs0, rsn);
    const float4* X4 = reinterpret_cast<const float4*>(F.xtab);
    for (uint32_t A0 = 2 * wid; A0 < num_nodes; A0 += 2 * nw) {  // uniform
        const uint32_t A = A0 + half;
        const bool act = A < num_nodes;
        const int4 ga = act ? F.nodes.geo[A] : make_int4(0, 0, 0, 0);
        const int sysA = act ? F.nodes.sys[A] : 0;
        float L[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) L[k] = 0.f;
        uint32_t nfar = 0, any = 0;
        if (act) {
            for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r]) {
                const uint32_t nx = F.lists.mx_cnt[r];
                any |= nx;
                const uint32_t e1 = F.lists.m_start[r] + F.lists.m_cnt[r];
                for (uint32_t e = e1 - nx + sub; e < e1; e += 32) {
                    const int B = F.lists.m_pool[e];
                    uint32_t neg = 0;
                    const int t = xtab_index(F, ga, F.nodes.geo[B], sysA, neg);
                    if (t < 0) {
                        ++nfar;
                        continue;
                    }
                    float D[NCQ];
                    load_tensor<m2l_deg(Q)>(X4 + (size_t)t * NQD, neg, D);
                    m2l_apply<P, Q, POT>(D, F.M + (size_t)B * F.ncp, L);
                }
            }
        }
        group_reduce_scatter<NC, 5>(L);
        nfar = half_sum(nfar);
        // a target without such entries has L = 0 here: Lx is still written
        // (it replaces the previous evaluation's), L is left alone
        if (act && F.Lx) {
            float* out = F.Lx + (size_t)A * F.ncp + rs0;
#pragma unroll
            for (int j = 0; j < rs_len(NC, 5); ++j)
                if (j < rsn) out[j] = L[j];
        } else if (act && any) {
            float* out = F.L + (size_t)A * F.ncp + rs0;
#pragma unroll
            for (int j = 0; j < rs_len(NC, 5); ++j)
                if (j < rsn) out[j] += L[j];
        }
        // deferred pairs: one contiguous run per target
        const bool defer = act && nfar > 0;
        if (!__ballot(defer)) continue;  // uniform
        uint32_t s0 = 0;
        if (defer && sub == 0) {
            atomicAdd(&F.far_cnt[0], 1u);
            s0 = atomicAdd(&F.far_cnt[1], nfar);
        }
        s0 = __shfl(s0, lane & 32, WAVE);
        // the capacity is the walk's count of other-level entries, which
        // bounds the deferred pairs (a captured step sizes it from the
        // previous step and flags a shortfall)
        if (defer && (unsigned long long)s0 + nfar > F.far_cap) {
            if (sub == 0) atomicOr(const_cast<uint32_t*>(&F.C->overflow), (uint32_t)OVF_FAR);
            continue;
        }
        if (!defer) continue;
        uint32_t pos = s0;
        for (uint32_t r = F.lists.first[A]; r != NO_ROW; r = F.lists.next[r]) {
            const uint32_t e1 = F.lists.m_start[r] + F.lists.m_cnt[r];
            const uint32_t e0 = e1 - F.lists.mx_cnt[r];
            for (uint32_t c0 = e0; c0 < e1; c0 += 32) {
                const uint32_t e = c0 + sub;
                int B = -1;
                bool far = false;
                if (e < e1) {
                    B = F.lists.m_pool[e];
                    uint32_t neg = 0;
                    far = xtab_index(F, ga, F.nodes.geo[B], sysA, neg) < 0;
                }
                const uint64_t bal = __ballot(far) & hmask;
                if (far) {
                    F.far_src[pos + __popcll(bal & lt)] = B;
                    F.far_tgt[pos + __popcll(bal & lt)] = (int)A;
                }
                pos += __popcll(bal);
            }
        }
    }
}

// Deferred pairs (levels two or more apart, ~0.1% of the pairs, a handful
// per target): tensor built per pair, one lane per pair (a wavefront per
// target would leave most lanes idle).  A target's pairs are contiguous, so
// the lanes of a wavefront form runs of equal targets: a segmented suffix sum
// over the lanes leaves each run's total in its first lane, which adds it to
// the target's L with float atomics (a run may continue in the next
// wavefront).
template <int P, int Q, bool POT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void m2l_far_kernel(FmmParams F,
                                                                                                uint32_t num_nodes) {
    constexpr int NC = ncoef(P);
    if (num_nodes == DEV_COUNT) {
        if (F.C->overflow & OVF_WALK) return;
        num_nodes = F.C->num_nodes;
    }
    const uint32_t npairs = min(F.far_cnt[1], F.far_cap);
    const int lane = lane_id();
    const uint32_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE * WAVE;
    for (uint32_t base = w0; base < npairs; base += gridDim.x * blockDim.x) {  // uniform
        const uint32_t k = base + lane;
        int A = -1, B = -1;
        if (k < npairs) {
            A = F.far_tgt[k];
            B = F.far_src[k];
            if ((uint32_t)A >= num_nodes || (uint32_t)B >= num_nodes) A = B = -1;
        }
        float L[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) L[c] = 0.f;
        if (A >= 0) {
            const int4 ga = F.nodes.geo[A];
            const float unit = F.unit[F.nodes.sys[A]];
            const float3 d = centre_diff(F.nodes.geo[B], ga, F.bits, unit);  // R = c_A - c_B
            float D[ncoef(m2l_deg(Q))];
            derivs<m2l_deg(Q)>(d.x, d.y, d.z, F.eps2, D);
            m2l_apply<P, Q, POT>(D, F.M + (size_t)B * F.ncp, L);
        }
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const int An = __shfl_down(A, off, WAVE);
            const bool same = lane + off < WAVE && An == A;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float o = __shfl_down(L[c], off, WAVE);
                if (same) L[c] += o;
            }
        }
        const int Ap = __shfl_up(A, 1, WAVE);
        if (A >= 0 && (lane == 0 || Ap != A)) {
            float* out = (F.Lx ? F.Lx : F.L) + (size_t)A * F.ncp;
#pragma unroll
            for (int c = 0; c < NC; ++c) atomicAdd(out + c, L[c]);
        }
    }
}

__global__ void zero_u32x2_kernel(uint32_t* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0u;
}

__global__ void zero_f32_kernel(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// rows [0, level_off[level]) of L (the cells above the tabled levels), the
// count read from the device counters
__global__ void zero_top_kernel(float* L, const Counters* C, int level, int ncp) {
    if (C->overflow & OVF_WALK) return;
    const size_t n = (size_t)C->level_off[level] * ncp;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        L[i] = 0.f;
}

// parts: M2L_LEVELS runs the same-level kernel of level `level` (all tabled
// levels when level < 0); M2L_REST runs everything else (initialisation, the
// other-level and deferred pairs), after the same-level kernels.  POT: F.pot
// (the per-pair and matrix-core kernels compute the potential always).
template <int P, int Q, bool POT>
static void launch_m2l_pq_pot(const FmmParams& F, const M2LPlan& plan, int parts, int level, hipStream_t st) {
    const uint32_t nn = plan.num_nodes;
    const unsigned grid = std::min<unsigned>(ceil_div(nn, 4), 65536u);
    if (plan.mode == 0) {
        if (parts & M2L_REST)
            hipLaunchKernelGGL((m2l_plain_kernel<P, Q, POT>), dim3(ceil_div(nn, 256)), dim3(256), 0, st, F, nn);
        return;
    }
    if (!F.dtab) {
        if (parts & M2L_REST) hipLaunchKernelGGL((m2l_pairs_kernel<P, Q>), dim3(grid), dim3(256), 0, st, F, nn);
        return;
    }
    if (parts & M2L_REST) {
        hipLaunchKernelGGL(zero_u32x2_kernel, dim3(1), dim3(64), 0, st, F.far_cnt);
        // cells above the tabled levels have no same-level pairs: L starts at 0
        if (plan.dev) {
            hipLaunchKernelGGL(zero_top_kernel, dim3(64), dim3(256), 0, st, F.L, F.C, F.dlevel0, F.ncp);
        } else {
            const uint32_t top = plan.level_off[F.dlevel0];
            if (top) {
                const size_t nz = (size_t)top * F.ncp;
                hipLaunchKernelGGL(zero_f32_kernel, dim3((unsigned)ceil_div(nz, (size_t)256)), dim3(256), 0, st, F.L,
                                   nz);
            }
        }
    }
    constexpr bool PIPE = m2l_pipelined(P, Q);
    const int K1 = F.dK + 1;
    // rows padded to an odd float4 count, then the offset -> row map
    const size_t lds = (size_t)F.dsz * ((F.dcp / 4) | 1) * 16 + (((size_t)K1 * K1 * K1 * 2 + 15) & ~(size_t)15);
    const bool in_lds = PIPE && plan.nsys == 1 && lds <= plan.max_lds;
    // two workgroups of 256 per CU where two tables fit (the largest
    // workgroup LDS, max_lds, is all of a CU's LDS on MI355X)
    const bool two = in_lds && 2 * lds <= plan.max_lds;
    const int T = in_lds && !two ? 512 : 256;
    // (set by the uncaptured evaluation that precedes every capture: not a
    // stream operation, kept out of a capture)
    if constexpr (PIPE) {
        if (in_lds && (parts & M2L_LEVELS) && !plan.dev) {
            const void* k = two ? reinterpret_cast<const void*>(&m2l_lds_kernel<P, Q, true, 256, true, POT>)
                                : reinterpret_cast<const void*>(&m2l_lds_kernel<P, Q, true, 512, true, POT>);
            (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        }
    }
    for (int l = F.dlevel0; l < F.dlevel0 + F.dnlev; ++l) {
        if (!(parts & M2L_LEVELS) || (level >= 0 && l != level)) continue;
        const uint32_t cnt = plan.level_count[l];
        if (!cnt) continue;
        if (plan.mfma) {
            // by translation offset on the matrix cores: one workgroup per
            // level-(l-2) cell, its level-l descendants as the targets
            launch_m2l_offset_level(P, plan.mfma == 1, F, plan.off, plan.level_off[l - 2], plan.level_count[l - 2],
                                    l - 2, st);
            continue;
        }
        // one target per wavefront: the workgroup scheduler balances the
        // uneven list lengths better than a grid-stride loop or a resident
        // grid with a work queue (measured); with device counts the grid is
        // sized from plan.level_count and the kernel strides over the rest
        const unsigned g = std::min<unsigned>(ceil_div(cnt, T / WAVE), 65535u);
        const uint32_t koff = plan.dev ? 0u : plan.level_off[l], kcnt = plan.dev ? DEV_COUNT : cnt;
        const int lv = l - F.dlevel0;
        if constexpr (PIPE) {
            if (two)
                hipLaunchKernelGGL((m2l_lds_kernel<P, Q, true, 256, true, POT>), dim3(g), dim3(256), lds, st, F, koff,
                                   kcnt, lv);
            else if (in_lds)
                hipLaunchKernelGGL((m2l_lds_kernel<P, Q, true, 512, true, POT>), dim3(g), dim3(512), lds, st, F, koff,
                                   kcnt, lv);
            else
                hipLaunchKernelGGL((m2l_lds_kernel<P, Q, false, 256, true, POT>), dim3(g), dim3(256), 0, st, F, koff,
                                   kcnt, lv);
        } else {
            hipLaunchKernelGGL((m2l_lds_kernel<P, Q, false, 256, false, POT>), dim3(g), dim3(256), 0, st, F, koff,
                               kcnt, lv);
        }
    }
    if (!(parts & M2L_REST)) return;
    const uint32_t knn = plan.dev ? DEV_COUNT : nn;
    hipLaunchKernelGGL((m2l_cross_kernel<P, Q, POT>), dim3(std::min<unsigned>(ceil_div(nn, 8), 65536u)), dim3(256),
                       0, st, F, knn);
    // the deferred pairs are a subset of the other-level entries (far_cap)
    const unsigned gfar = std::max<unsigned>(1u, std::min<unsigned>(ceil_div(F.far_cap, 256u), 1024u));
    hipLaunchKernelGGL((m2l_far_kernel<P, Q, POT>), dim3(gfar), dim3(256), 0, st, F, knn);
}
template <int P, int Q>
static void launch_m2l_pq(const FmmParams& F, const M2LPlan& plan, int parts, int level, hipStream_t st) {
    if (F.pot) launch_m2l_pq_pot<P, Q, true>(F, plan, parts, level, st);
    else launch_m2l_pq_pot<P, Q, false>(F, plan, parts, level, st);
}

template <int Q>
static void launch_dtable_t(const FmmParams& F, float* out, int level0, int nlev, int E, int a, int b, int shift,
                            int nsys, const uint32_t* offs, int nrows, hipStream_t st) {
    const unsigned total = (unsigned)((offs ? nrows : E * E * E) * nlev * nsys);
    hipLaunchKernelGGL((dtable_kernel<Q>), dim3(ceil_div(total, 128)), dim3(128), 0, st, F, out, level0, nlev, E, a,
                       b, shift, nsys, offs, nrows);
}

// grid of the launches whose count is read on the device (DEV_COUNT): sized
// for grid_items items; the kernels loop over the actual count
static uint32_t grid_items_of(uint32_t count, uint32_t grid_items) {
    return count == DEV_COUNT ? std::max<uint32_t>(grid_items, 1u) : count;
}

template <int P>
static void launch_p2m_t(const FmmParams& F, const State& S, uint32_t num_leaves, uint32_t items, hipStream_t st) {
    hipLaunchKernelGGL((p2m_kernel<P>), dim3(ceil_div(items, 256)), dim3(256), 0, st, F, S, num_leaves);
}
template <int P>
static void launch_l2p_t(const FmmParams& F, const State& S, uint32_t n, double kick_dt, hipStream_t st) {
    hipLaunchKernelGGL((l2p_kernel<P>), dim3(ceil_div(n, 256)), dim3(256), 0, st, F, S, n, kick_dt);
}
template <int P>
static void launch_m2m_t(const FmmParams& F, uint32_t count, uint32_t off, uint32_t items, hipStream_t st) {
    hipLaunchKernelGGL((m2m_kernel<P>), dim3(ceil_div(items, 64)), dim3(64), 0, st, F, count, off);
}
template <int P>
static void launch_l2l_t(const FmmParams& F, uint32_t count, uint32_t off, uint32_t items, hipStream_t st) {
    hipLaunchKernelGGL((l2l_kernel<P>), dim3(ceil_div(items, 64)), dim3(64), 0, st, F, count, off);
}
template <int P>
static void launch_p2p_t(const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st) {
    if (F.pot) hipLaunchKernelGGL((p2p_kernel<P, true>), dim3(ceil_div(num_leaves, 4)), dim3(256), 0, st, F, S, num_leaves);
    else hipLaunchKernelGGL((p2p_kernel<P, false>), dim3(ceil_div(num_leaves, 4)), dim3(256), 0, st, F, S, num_leaves);
}

template <int P, int SPL, bool POT>
static void launch_p2p_staged_pot(const FmmParams& F, const State& S, uint32_t num_leaves, unsigned grid,
                                  hipStream_t st) {
    if (F.eps2 > 0.f)
        hipLaunchKernelGGL((p2p_staged_kernel<P, SPL, POT>), dim3(grid), dim3(WAVE * P2P_WAVES), 0, st, F, S,
                           num_leaves);
    else
        hipLaunchKernelGGL((p2p_staged_unsoftened_kernel<P, SPL, POT>), dim3(grid), dim3(WAVE * P2P_WAVES), 0, st,
                           F, S, num_leaves);
}
template <int P, int SPL>
static void launch_p2p_staged_t(const FmmParams& F, const State& S, uint32_t num_leaves, uint32_t items,
                                hipStream_t st) {
    const unsigned grid = std::min<unsigned>(ceil_div(items, P2P_WAVES), 65536u);
    if (F.pot) launch_p2p_staged_pot<P, SPL, true>(F, S, num_leaves, grid, st);
    else launch_p2p_staged_pot<P, SPL, false>(F, S, num_leaves, grid, st);
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

void launch_p2m(int p, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st, uint32_t grid_items) {
    if (!num_leaves) return;
    const uint32_t items = grid_items_of(num_leaves, grid_items);
    NBH_DISPATCH_P(p, launch_p2m_t<PP>(F, S, num_leaves, items, st));
}
void launch_l2p(int p, const FmmParams& F, const State& S, uint32_t n, double kick_dt, hipStream_t st) {
    if (!n) return;
    NBH_DISPATCH_P(p, launch_l2p_t<PP>(F, S, n, kick_dt, st));
}
void launch_m2m_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st,
                      uint32_t grid_items) {
    if (!count) return;
    const uint32_t items = grid_items_of(count, grid_items);
    if (count == DEV_COUNT) off = (uint32_t)level;
    NBH_DISPATCH_P(p, launch_m2m_t<PP>(F, count, off, items, st));
}
void launch_l2l_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st,
                      uint32_t grid_items) {
    if (!count) return;
    const uint32_t items = grid_items_of(count, grid_items);
    if (count == DEV_COUNT) off = (uint32_t)level;
    NBH_DISPATCH_P(p, launch_l2l_t<PP>(F, count, off, items, st));
}
void launch_p2p_staged(int p, int chunk, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st,
                       uint32_t grid_items) {
    if (!num_leaves) return;
    const uint32_t items = grid_items_of(num_leaves, grid_items);
    if (chunk == 128) NBH_DISPATCH_P(p, (launch_p2p_staged_t<PP, 2>(F, S, num_leaves, items, st)))
    else NBH_DISPATCH_P(p, (launch_p2p_staged_t<PP, 1>(F, S, num_leaves, items, st)))
}
void launch_dtable(int q, const FmmParams& F, float* out, int level0, int nlev, int E, int a, int b, int shift,
                   int nsys, const uint32_t* offs, int nrows, hipStream_t st) {
#define NBH_DT(QQ) launch_dtable_t<QQ>(F, out, level0, nlev, E, a, b, shift, nsys, offs, nrows, st)
    switch (q) {
        case 1: NBH_DT(1); break;
        case 2: NBH_DT(2); break;
        case 3: NBH_DT(3); break;
        case 4: NBH_DT(4); break;
        case 5: NBH_DT(5); break;
        case 6: NBH_DT(6); break;
        case 7: NBH_DT(7); break;
        default: break;
    }
#undef NBH_DT
}
void launch_m2l_parts(int p, int q, const FmmParams& F, const M2LPlan& plan, int parts, int level, hipStream_t st) {
    if (!plan.num_nodes) return;
    // q: truncation code (m2l_deg(q) = p or p + 1, M2L_ALL for every term of
    // degree p + 1; below 2p only, where the two term sets differ)
    const int ext = m2l_deg(q) == p + 1 ? ((q & M2L_ALL) && p + 1 < 2 * p ? 2 : 1) : 0;
#define NBH_M2L(PP, QQ) launch_m2l_pq<PP, QQ>(F, plan, parts, level, st)
#define NBH_M2L3(PP) \
    (ext == 2 ? NBH_M2L(PP, (PP + 1) | M2L_ALL) : ext == 1 ? NBH_M2L(PP, PP + 1) : NBH_M2L(PP, PP))
    switch (p) {
        case 1: ext ? NBH_M2L(1, 2) : NBH_M2L(1, 1); break;
        case 2: NBH_M2L3(2); break;
        case 3: NBH_M2L3(3); break;
        case 4: NBH_M2L3(4); break;
        case 5: NBH_M2L3(5); break;
        case 6: NBH_M2L3(6); break;
        default: break;
    }
#undef NBH_M2L3
#undef NBH_M2L
}
void launch_m2l(int p, int q, const FmmParams& F, const M2LPlan& plan, hipStream_t st) {
    launch_m2l_parts(p, q, F, plan, M2L_LEVELS | M2L_REST, -1, st);
}
void launch_p2p(int p, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st) {
    if (!num_leaves) return;
    NBH_DISPATCH_P(p, launch_p2p_t<PP>(F, S, num_leaves, st));
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
// naive force law: separation (dx, dx, dx), no softening (:10-17), with each
// pair's term formed as the reference forms it in FP32 (uncentred x
// difference, r = sqrt(d d + d d + d d), force = (k q_i q_j) d / (r r r),
// divided by m_i; no contraction into FMAs) and summed in FP64, so one
// evaluation differs from the reference only by its FP32 summation order.
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
    // compat: absolute FP32 coordinates, as the reference holds them
    const double cx = compat ? 0.0 : bx.x + h, cy = compat ? 0.0 : bx.y + h, cz = compat ? 0.0 : bx.z + h;
    const bool valid = i < e;
    float px = 0, py = 0, pz = 0;
    if (valid) { px = (float)(S.x[i] - cx); py = (float)(S.y[i] - cy); pz = (float)(S.z[i] - cz); }
    float ax = 0, ay = 0, az = 0, ph = 0;
    double acd = 0.0;  // compat: sum of the reference's per-pair kicks / dt
    const float qi = valid ? (float)S.charge[i] : 0.f, mi = valid ? (float)S.mass[i] : 1.f;
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
                // the self pair contributes nothing (r2 = 0 when unsoftened)
                const float inv = (j0 + k == i) ? 0.f : rsqrtf(r2);
                const float w = q.w * inv;
                const float w3 = w * inv * inv;
                ax = fmaf(rx, w3, ax); ay = fmaf(ry, w3, ay); az = fmaf(rz, w3, az);
                ph += w;
            }
        } else {
            for (uint32_t k = 0; k < jn; ++k) {
                // IEEE single operations in the reference's order: no FMA
                // contraction; sqrtf and '/' are correctly rounded in HIP
#pragma clang fp contract(off)
                if (j0 + k == i) continue;
                const float4 q = tile[k];
                const float d = q.x - px;
                const float r = sqrtf(d * d + d * d + d * d);
                const float f = G * (qi * q.w) * d / (r * r * r);
                acd += (double)(f / mi);
                ph += q.w / r;
            }
        }
    }
    if (!valid) return;
    if (compat) {
        const float a = (float)acd;
        S.ax[i] = a; S.ay[i] = a; S.az[i] = a;
        S.pot[i] = -G * ph;
        if (kick_dt != 0.0) {
            S.vx[i] += (double)a * kick_dt;
            S.vy[i] += (double)a * kick_dt;
            S.vz[i] += (double)a * kick_dt;
        }
        return;
    }
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
    if (i >= sys_off[nsys]) {  // no such particle: report NaN rather than read out of bounds
        if (lane_id() == 0) {
            acc[3 * w] = acc[3 * w + 1] = acc[3 * w + 2] = NAN;
            pot[w] = NAN;
        }
        return;
    }
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

// Sorted slot of each target given in input order, by a scan of the id array
// (one wavefront per target; an id array is a permutation, so every index in
// [0, n) is found once; anything else gets 0xffffffff).  Costs a read of the
// ids per target, an eighth of what the FP64 sum then reads per target.
__global__ __launch_bounds__(256) void find_slots_kernel(const uint32_t* __restrict__ id, uint32_t n,
                                                         const int64_t* __restrict__ targets, uint32_t nt,
                                                         uint32_t* slots) {
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (w >= nt) return;
    const int64_t t = targets[w];
    uint32_t found = 0xffffffffu;
    for (uint32_t j0 = 0; j0 < n && found == 0xffffffffu; j0 += WAVE) {
        const uint32_t j = j0 + lane_id();
        const uint64_t hit = __ballot(j < n && (int64_t)id[j] == t);
        if (hit) found = j0 + (uint32_t)__ffsll((unsigned long long)hit) - 1;
    }
    if (lane_id() == 0) slots[w] = found;
}
void launch_find_slots(const uint32_t* id, uint32_t n, const int64_t* targets, uint32_t nt, uint32_t* slots,
                       hipStream_t st) {
    hipLaunchKernelGGL(find_slots_kernel, dim3(ceil_div(nt, 4)), dim3(256), 0, st, id, n, targets, nt, slots);
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
