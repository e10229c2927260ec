// m2l_offset.hip -- same-level M2L grouped by translation offset, on the
// matrix cores.
//
// Replaces the far field of field.cl:153-211 (the order-0 cell field of
// node_moment_field, field.cl:35-47) for the cell pairs of one level, at
// order P with the dense operator
//   L_k(A) += sum_{|n| <= P} D_{k+n}(z_A - z_B) (-1)^|n| M_n(B),   |k| <= P,
// i.e. every term (k, n) of orders <= P, so a translation by the offset
// o = B - A is one dense NC x NC matrix T_o[k][n] = D_{k+n}(-o h).
//
// Which same-level pairs the walk emits is a function of geometry alone: for
// a target A of level l whose position in its parent is c = A mod 2 (its
// "class") and a cell B = A + o of the same level, B is in A's list iff
// the opening criterion accepts o and rejects the parent offset
// floor((c + o) / 2) (the walk splits a pair of equal cells iff the criterion
// rejects it, interaction.cl:63-82, and a rejection at one level implies a
// rejection of every coarser ancestor pair, whose offsets are no larger).  The
// host tabulates the candidate offsets with an 8-bit mask of the classes that
// accept them.  A workgroup takes the (at most 64) level-l cells below one
// level-(l-2) cell, a 4x4x4 block, gathers their same-level sources from
// their lists into an LDS window of the block's neighbourhood indexed by
// position, and walks the candidate offsets: for each, every target's source
// (or none) is one LDS lookup, and the contraction
//   C[k][A] += sum_n T_o[k][n] M'[n][src(A, o)],   M'_n = (-1)^|n| M_n
// is a (k x targets) GEMM with T_o shared by the 64 targets.  Four offsets
// are batched along the contraction dimension (4 x NG groups of 8 n-values =
// NG K-steps of 32), the accumulators of all targets stay in registers across
// the whole offset loop, and the four wavefronts of the workgroup take
// interleaved candidates and add their partial sums at the end.  The pairs
// each target visited are counted against its list as a check.
//
// Scaling.  The tensors are tabled for the unit cell, D~_m = h^(|m|+1) D_m
// (R/h, eps/h), the moments enter as M~_n = h^-|n| M'_n and the sums leave as
// L_k = h^-(|k|+1) C_k (h the level's cell size), so that no FP32 or bf16 value
// leaves the exponent range at any depth: D_m grows like |m|! / |R|^(|m|+1).
//
// Precisions.  BF3: bf16x3 split products with FP32 accumulation,
// x = hi + lo, hi = bf16(x), lo = bf16(x - hi) (round to nearest even),
// C += lo_T hi_M + hi_T lo_M + hi_T hi_M on v_mfma_f32_16x16x32_bf16; the
// tensors are split once per step in the table and the moments once per cell.
// !BF3: FP32 products on v_mfma_f32_16x16x4_f32.
// Operand layouts (lane l, r = l % 16, h = l / 16): A row r, B column r; the
// K-slots of a lane are 8h..8h+7 (16x16x32) or h (16x16x4); accumulator
// element i is C[4h + i][r].
#include <vector>

#include "common.h"
#include "expansion.h"
#include "launch.h"

namespace nbh {

typedef __attribute__((ext_vector_type(8))) __bf16 obf16x8_t;
typedef __attribute__((ext_vector_type(4))) float of32x4_t;
typedef __attribute__((ext_vector_type(4))) uint32_t ou32x4_t;

// index of the coefficient with exponents (x, y, z) in the total-degree order
// (degree, then descending x, then descending y)
__host__ __device__ inline int off_coef_index(int x, int y, int z) {
    const int d = x + y + z, a = d - x;
    return d * (d + 1) * (d + 2) / 6 + a * (a + 1) / 2 + z;
}
// its inverse
__host__ __device__ inline void off_coef_exps(int idx, int& x, int& y, int& z) {
    int d = 0;
    while ((d + 1) * (d + 2) * (d + 3) / 6 <= idx) ++d;
    const int rem = idx - d * (d + 1) * (d + 2) / 6;
    int a = 0;
    while ((a + 1) * (a + 2) / 2 <= rem) ++a;
    z = rem - a * (a + 1) / 2;
    x = d - a;
    y = a - z;
}

// bf16 bits of x rounded to nearest even (finite x)
__host__ __device__ inline uint32_t off_bf16_rne(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__host__ __device__ inline float off_bf16_value(uint32_t b) {
    const uint32_t u = b << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

constexpr uint32_t OFF_PAD = 0xffffffffu;  // batch slot without an offset

// the workgroup's dynamic LDS (no static LDS in this file, so it starts at address 0)
extern __shared__ __align__(16) unsigned char off_lds[];

// LDS writes of one lane made visible to the other lanes of its wavefront,
// ordering LDS only: the tensor rows of the next batch stay in flight from
// global memory across it (a fence over every address space would wait for
// them)
__device__ __forceinline__ void off_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

template <int P>
struct OffDims {
    static constexpr int NC = ncoef(P);
    static constexpr int NC2 = ncoef(2 * P);           // tensor coefficients, orders <= 2P
    static constexpr int DCP = (NC2 + 1 + 3) & ~3;     // tensor row in words; slot NC2 is zero
    static constexpr int MT = (NC + 15) / 16;          // M tiles over k
    static constexpr int NG = (NC + 7) / 8;            // bf16: K groups of 8 over n
    static constexpr int KR = MT * 16;
    static constexpr int KCB = NG * 8;                 // n extent, bf16 layout
    static constexpr int KCF = (NC + 3) / 4 * 4;       // n extent, FP32 layout
};

template <int P, bool BF3>
struct OffLayout {
    using D = OffDims<P>;
    static constexpr int KC = BF3 ? D::KCB : D::KCF;
    // byte offsets in the workgroup's LDS; everything but the window (whose
    // edge W is a run-time value) at compile-time offsets, so that the hot
    // loop addresses LDS with immediate offsets
    static constexpr size_t HDR = 1024;                                    // int cnt[64], s0[64], pre[65]
    static constexpr size_t DB = HDR;                                      // u32 [wave][OFF_J][DCP] tensor rows
    static constexpr size_t SB = DB + (size_t)OFF_WAVES * OFF_J * D::DCP * 4;  // int [wave][2][OFF_J][64] sources
    static constexpr size_t HK = SB + (size_t)OFF_WAVES * 2 * OFF_J * WAVE * 4;  // u16 [OFF_J][KR][KC] A indices
    static constexpr size_t WIN = HK + (((size_t)OFF_J * D::KR * KC * 2 + 15) & ~(size_t)15);  // int [W^3]
    // int [wave][ceil(ncand / 4)]: the wavefront's candidates that some target takes
    static __host__ __device__ size_t CL(int W) { return WIN + (size_t)W * W * W * 4; }
    static __host__ __device__ size_t total(int W, int ncand) {
        const size_t main = CL(W) + (size_t)OFF_WAVES * ((ncand + OFF_WAVES - 1) / OFF_WAVES) * 4;
        const size_t red = HDR + (size_t)(OFF_WAVES - 1) * D::KR * WAVE * 4;  // partial sums of waves 1..3
        return main > red ? main : red;
    }
};

// ---------------------------------------------------------------------------
// h^-d for the cell size h of level l (root cube edge h0)
__device__ __forceinline__ float off_inv_pow(const double4* box, int level, int d) {
    const float ih = (float)ldexp(1.0 / box[0].w, level);
    float f = 1.f;
    for (int i = 0; i < d; ++i) f *= ih;
    return f;
}

// Signed scaled split moments M~_n = (-1)^|n| h^-|n| M_n of every node, and a
// zero row at index nn.  BF3: per node NG groups of {8 hi bf16, 8 lo bf16}
// (32 bytes); FP32: KCF floats.
template <int P, bool BF3>
__global__ __launch_bounds__(256) void off_moments_kernel(const float* __restrict__ M, int ncp, uint32_t nn,
                                                          const int4* __restrict__ geo, const double4* box,
                                                          void* out) {
    using D = OffDims<P>;
    if constexpr (BF3) {
        const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= (nn + 1) * D::NG) return;
        const uint32_t node = i / D::NG, g = i - node * D::NG;
        const int level = node < nn ? geo_level(geo[node]) : 0;
        uint32_t hb[8], lb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int n = 8 * g + e;
            float x = 0.f;
            if (node < nn && n < D::NC) {
                int ex, ey, ez;
                off_coef_exps(n, ex, ey, ez);
                x = M[(size_t)node * ncp + n] * off_inv_pow(box, level, ex + ey + ez);
                if ((ex + ey + ez) & 1) x = -x;
            }
            hb[e] = off_bf16_rne(x);
            lb[e] = off_bf16_rne(x - off_bf16_value(hb[e]));
        }
        uint4* o = (uint4*)out + (size_t)i * 2;
        o[0] = make_uint4(hb[0] | hb[1] << 16, hb[2] | hb[3] << 16, hb[4] | hb[5] << 16, hb[6] | hb[7] << 16);
        o[1] = make_uint4(lb[0] | lb[1] << 16, lb[2] | lb[3] << 16, lb[4] | lb[5] << 16, lb[6] | lb[7] << 16);
    } else {
        const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= (nn + 1) * D::KCF) return;
        const uint32_t node = i / D::KCF, n = i - node * D::KCF;
        float x = 0.f;
        if (node < nn && n < (uint32_t)D::NC) {
            int ex, ey, ez;
            off_coef_exps((int)n, ex, ey, ez);
            x = M[(size_t)node * ncp + n] * off_inv_pow(box, geo_level(geo[node]), ex + ey + ez);
            if ((ex + ey + ez) & 1) x = -x;
        }
        ((float*)out)[i] = x;
    }
}

// Unit-cell derivative tensors D~_m = D_m(-e, eps / h_l), |m| <= Q2, of
// (|R|^2 + eps^2)^(-1/2) for e in [0, E)^3 and levels level0 .. level0 +
// nlev - 1, in FP64 by the recurrence of expansion.h (s D_n = -[...]),
// stored per row of dcp words: BF3 as (bf16(x) << 16 | bf16(x - bf16(x))),
// else FP32.  One thread per tensor; the order-12 tensor (455 terms) lives in
// scratch.
template <int Q2, bool BF3>
__global__ __launch_bounds__(64) void off_dtable_kernel(const double4* box, double eps2, int level0, int nlev, int E,
                                                        int dcp, uint32_t* out) {
    constexpr int NC2 = ncoef(Q2);
    const int per = E * E * E;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * nlev) return;
    const int lv = t / per, o = t - lv * per;
    const int ex = o / (E * E), ey = (o / E) % E, ez = o % E;
    const double h = ldexp(box[0].w, -(level0 + lv));
    const double R[3] = {(double)-ex, (double)-ey, (double)-ez};
    double D[NC2];
    const double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps2 / (h * h);
    const double inv_s = 1.0 / s;
    D[0] = 1.0 / sqrt(s);
    for (int m = 1; m < NC2; ++m) {
        int e[3];
        off_coef_exps(m, e[0], e[1], e[2]);
        const int i = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
        double acc = 0.0;
        for (int j = 0; j < 3; ++j) {
            if (e[j] < 1) continue;
            int a[3] = {e[0], e[1], e[2]};
            a[j] -= 1;
            const double c1 = j == i ? 2.0 * e[i] - 1.0 : 2.0 * e[j];
            acc += c1 * R[j] * D[off_coef_index(a[0], a[1], a[2])];
            if (e[j] >= 2) {
                a[j] -= 1;
                const double c2 = j == i ? (double)(e[i] - 1) * (e[i] - 1) : (double)e[j] * (e[j] - 1);
                acc += c2 * D[off_coef_index(a[0], a[1], a[2])];
            }
        }
        D[m] = -acc * inv_s;
    }
    uint32_t* row = out + (size_t)t * dcp;
    for (int m = 0; m < dcp; ++m) {
        uint32_t w = 0;
        if (m < NC2) {
            if (BF3) {
                const uint32_t hi = off_bf16_rne((float)D[m]);
                const uint32_t lo = off_bf16_rne((float)(D[m] - (double)off_bf16_value(hi)));
                w = hi << 16 | lo;
            } else {
                const float f = (float)D[m];
                __builtin_memcpy(&w, &f, 4);
            }
        }
        row[m] = w;
    }
}

// ---------------------------------------------------------------------------
// One batch of OFF_J offsets for wavefront WV (compile-time, so that its LDS
// rows are at immediate offsets).  K layout: K-step s is n-group s of all
// OFF_J offsets, lane group h = l / 16 taking offset j = h; the A operand of
// lane (r, h) in M-tile mt is row k = 16 mt + r of T_j, the B operand the
// moments of the source of target 16 t + r at offset j.
template <int P, bool BF3, int WV>
__device__ __forceinline__ void off_batch(of32x4_t (&acc)[OffDims<P>::MT][4], const OffsetParams& O, int b, int ntile,
                                          int r, int hh) {
    using D = OffDims<P>;
    using Lay = OffLayout<P, BF3>;
    constexpr int MT = D::MT, KC = Lay::KC;
    constexpr size_t DBW = Lay::DB + (size_t)WV * OFF_J * D::DCP * 4;  // this wavefront's tensor rows
    const int* sb = (const int*)(off_lds + Lay::SB) + (WV * 2 + b) * OFF_J * WAVE;
    // this lane's row of the index table (byte offsets into DBW, including the row of offset hh)
    const uint16_t* hk = (const uint16_t*)(off_lds + Lay::HK) + ((size_t)hh * D::KR + r) * KC;
    if constexpr (BF3) {
        const unsigned char* mb = (const unsigned char*)O.mb;
        // the moments of the four target tiles, empty slots and tiles past
        // ntile reading the zero row: the loads carry no branches, so that a
        // K-step's are issued one step ahead of its products
        const unsigned char* bp[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int src = t < ntile ? sb[hh * WAVE + 16 * t + r] : (int)O.zero_node;
            bp[t] = mb + (size_t)src * (D::NG * 32);
        }
        uint4 bq[2][4][2];
        obf16x8_t ahi[2][MT], alo[2][MT];
        // A fragments of K-step s (LDS gathers of the Hankel-indexed tensor words)
        auto build_a = [&](int s, obf16x8_t (&hi)[MT], obf16x8_t (&lo)[MT]) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const uint4 hx = *(const uint4*)(hk + (size_t)16 * mt * KC + 8 * s);
                const uint32_t ix[4] = {hx.x, hx.y, hx.z, hx.w};
                ou32x4_t hv, lv;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t w0 = *(const uint32_t*)(off_lds + DBW + (ix[q] & 0xffffu));
                    const uint32_t w1 = *(const uint32_t*)(off_lds + DBW + (ix[q] >> 16));
                    hv[q] = __builtin_amdgcn_perm(w1, w0, 0x07060302u);  // hi halves: w0.hi | w1.hi << 16
                    lv[q] = __builtin_amdgcn_perm(w1, w0, 0x05040100u);  // lo halves: w0.lo | w1.lo << 16
                }
                hi[mt] = __builtin_bit_cast(obf16x8_t, hv);
                lo[mt] = __builtin_bit_cast(obf16x8_t, lv);
            }
        };
        auto load_b = [&](int s, uint4 (&bv)[4][2]) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bv[t][0] = *(const uint4*)(bp[t] + 32 * s);
                bv[t][1] = *(const uint4*)(bp[t] + 32 * s + 16);
            }
        };
        // software pipeline: the operands of K-step s + 1 (moment loads, tensor
        // gathers) are issued ahead of the products of step s
        load_b(0, bq[0]);
        build_a(0, ahi[0], alo[0]);
#pragma unroll
        for (int s = 0; s < D::NG; ++s) {
            if (s + 1 < D::NG) {
                load_b(s + 1, bq[(s + 1) & 1]);
                build_a(s + 1, ahi[(s + 1) & 1], alo[(s + 1) & 1]);
            }
            // the three split products in turn over all tiles, so that
            // consecutive MFMAs write different accumulators
#pragma unroll
            for (int pr = 0; pr < 3; ++pr)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (t < ntile) {
                        const obf16x8_t bhi = __builtin_bit_cast(obf16x8_t, bq[s & 1][t][0]);
                        const obf16x8_t blo = __builtin_bit_cast(obf16x8_t, bq[s & 1][t][1]);
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
                            acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                pr == 0 ? alo[s & 1][mt] : ahi[s & 1][mt], pr == 1 ? blo : bhi, acc[mt][t], 0, 0, 0);
                    }
                }
        }
    } else {
        const float* mf = (const float*)O.mb;
        const float* bp[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int src = t < ntile ? sb[hh * WAVE + 16 * t + r] : (int)O.zero_node;
            bp[t] = mf + (size_t)src * KC;
        }
        // K-step s: coefficient n = s of the offset of lane group h
        for (int s = 0; s < KC; ++s) {
            float a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) a[mt] = *(const float*)(off_lds + DBW + hk[(size_t)16 * mt * KC + s]);
            float bb[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) bb[t] = bp[t][s];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t < ntile) {
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt], bb[t], acc[mt][t], 0, 0, 0);
                }
            }
        }
    }
}

template <int P, bool BF3>
__global__ __launch_bounds__(OFF_WAVES * WAVE) __attribute__((amdgpu_waves_per_eu(P <= 4 ? 2 : 1, 2)))
void m2l_offset_kernel(FmmParams F, OffsetParams O, uint32_t goff, int lvl) {
    using D = OffDims<P>;
    using Lay = OffLayout<P, BF3>;
    constexpr int MT = D::MT, KC = Lay::KC, DCP = D::DCP, NC = D::NC;
    constexpr int NQ = (DCP + WAVE - 1) / WAVE;  // tensor words per lane per row
    const int W = O.W, K = O.K;
    int* cnt = (int*)off_lds;         // [64] pairs per target, summed over the wavefronts
    // [64] start of each target's same-level entries: a pool position, which
    // exceeds 2^31 on large trees (unsigned throughout)
    uint32_t* ls0 = (uint32_t*)(cnt + 64);
    int* lpre = (int*)(ls0 + 64);     // [65] exclusive prefix sum of their counts
    int* win = (int*)(off_lds + Lay::WIN);
    const int tid = threadIdx.x, lane = lane_id(), wave = tid / WAVE;
    uint32_t* db = (uint32_t*)(off_lds + Lay::DB) + wave * OFF_J * DCP;
    int* sbw = (int*)(off_lds + Lay::SB) + wave * 2 * OFF_J * WAVE;  // [2][OFF_J][64]

    // launched beside the walk: after an overflow at this depth or above the
    // level's rows are incomplete, and the host repeats the walk and M2L
    if (F.C->overflow & (OVF_SL | OVF_ROWS | OVF_M2L | OVF_P2P)) return;
    const uint32_t G = goff + blockIdx.x;
    const int4 gg = F.nodes.geo[G];
    const int nch = geo_nchild(gg);
    if (nch == 0) return;
    const int4 gl = F.nodes.link[G];
    int gc0 = -1, gc1 = -1;  // the level-l cells below G: [gc0, gc1)
    for (int c = 0; c < nch; ++c) {
        const int C = gl.z + c;
        const int cn = geo_nchild(F.nodes.geo[C]);
        if (cn) {
            const int fc = F.nodes.link[C].z;
            if (gc0 < 0) gc0 = fc;
            gc1 = fc + cn;
        }
    }
    if (gc0 < 0) return;
    const int nt = gc1 - gc0;  // <= 64
    const int ntile = (nt + 15) >> 4;

    // lane l of every wavefront stands for target gc0 + l
    const bool tv = lane < nt;
    int pl = 0, cls = 0;
    if (tv) {
        const int4 ga = F.nodes.geo[gc0 + lane];
        const int lx = ga.x - 4 * gg.x, ly = ga.y - 4 * gg.y, lz = ga.z - 4 * gg.z;
        pl = ((lx + K) * W + (ly + K)) * W + (lz + K);
        cls = ((ga.x & 1) << 2) | ((ga.y & 1) << 1) | (ga.z & 1);
    }
    // per tensor word m = lane + 64 q this lane stages: the octants neg whose
    // sign flips it (bit neg of fm[q]: an odd count of odd exponents on the
    // negative axes)
    uint32_t fm[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int m = lane + WAVE * q;
        const uint32_t pm = m < DCP ? O.par[m] : 0u;
        uint32_t f = 0;
#pragma unroll
        for (int ng = 0; ng < 8; ++ng) f |= (uint32_t)(__popc(pm & (uint32_t)ng) & 1) << ng;
        fm[q] = f;
    }

    // window cleared; index table copied in; the targets' entry ranges
    const int W3 = W * W * W;
    for (int i = tid; i < W3; i += OFF_WAVES * WAVE) win[i] = -1;
    for (int i = tid; i < OFF_J * D::KR * KC / 8; i += OFF_WAVES * WAVE)
        ((uint4*)(off_lds + Lay::HK))[i] = ((const uint4*)O.hank)[i];
    if (wave == 0) {
        int c = 0;
        if (tv) {
            const uint32_t r = F.lists.first[gc0 + lane];
            ls0[lane] = F.lists.m_start[r];
            c = (int)(F.lists.m_cnt[r] - F.lists.mx_cnt[r]);
        }
        // exclusive prefix sum over the wavefront
        int incl = c;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int v = __shfl_up(incl, o, WAVE);
            if (lane >= o) incl += v;
        }
        lpre[lane] = incl - c;
        if (lane == WAVE - 1) lpre[WAVE] = incl;
        cnt[lane] = 0;
    }
    __syncthreads();
    // window: every same-level source of the block's targets at its position,
    // the entries of all targets taken flat across the workgroup, eight per
    // thread at a time so that their list and geometry loads overlap
    const int total = lpre[WAVE];
    {
        constexpr int U = 8;
        int tcur = 0;  // target of this thread's current entry (entries only increase)
        for (int base = 0; base < total; base += U * OFF_WAVES * WAVE) {
            int Bv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = base + u * OFF_WAVES * WAVE + tid;
                Bv[u] = -1;
                if (e < total) {
                    while (lpre[tcur + 1] <= e) ++tcur;
                    Bv[u] = F.lists.m_pool[ls0[tcur] + (uint32_t)(e - lpre[tcur])];
                }
            }
            int4 gv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gv[u] = F.nodes.geo[Bv[u] >= 0 ? Bv[u] : 0];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (Bv[u] < 0) continue;
                const int wx = gv[u].x - 4 * gg.x + K, wy = gv[u].y - 4 * gg.y + K, wz = gv[u].z - 4 * gg.z + K;
                if ((unsigned)wx < (unsigned)W && (unsigned)wy < (unsigned)W && (unsigned)wz < (unsigned)W)
                    win[(wx * W + wy) * W + wz] = Bv[u];
                else
                    atomicOr(O.flags, OVF_GEOM);
            }
        }
    }
    __syncthreads();

    of32x4_t acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[mt][t] = of32x4_t{0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, hh = lane >> 4;
    const int E = O.E, E3 = E * E * E;
    const uint32_t* tab = O.tab + (size_t)lvl * E3 * DCP;
    int myc = 0;  // pairs of this lane's target taken by this wavefront

    // this wavefront's candidates (wave, wave + 4, ...) that some target of
    // the block takes, found 64 at a time (lane = candidate, independent
    // window reads over the targets) and listed in LDS
    uint32_t* clist = (uint32_t*)(off_lds + Lay::CL(W)) + wave * ((O.ncand + OFF_WAVES - 1) / OFF_WAVES);
    const int mine = O.ncand > wave ? (O.ncand - wave + OFF_WAVES - 1) / OFF_WAVES : 0;
    int ncl = 0;
    for (int g0 = 0; g0 < mine; g0 += WAVE) {
        const uint32_t c = g0 + lane < mine ? O.cand[wave + OFF_WAVES * (g0 + lane)] : OFF_PAD;
        const int ox = (int)(c & 31) - 16, oy = (int)((c >> 5) & 31) - 16, oz = (int)((c >> 10) & 31) - 16;
        const int lin = (ox * W + oy) * W + oz;
        const uint32_t cm = c == OFF_PAD ? 0u : c >> 16;
        bool any = false;
        for (int t = 0; t < nt; ++t) {
            const int tpl = __builtin_amdgcn_readlane(pl, t), tcl = __builtin_amdgcn_readlane(cls, t);
            if (((cm >> tcl) & 1u) && win[tpl + lin] >= 0) any = true;
        }
        const uint64_t bal = __ballot(any);
        if (any) clist[ncl + __popcll(bal & lanemask_lt())] = c;
        ncl += __popcll(bal);
    }
    int ci = 0;
    uint32_t ccur[OFF_J], cnxt[OFF_J];  // offset codes of the current and the next batch (uniform)
    // the next OFF_J listed offsets into batch buffer b: each target's source
    // (or the zero row) and the codes cd; returns how many
    auto collect = [&](int b, uint32_t (&cd)[OFF_J]) -> int {
        int* sb = sbw + b * OFF_J * WAVE;
        off_lds_sync();  // the list is written
        const int jn = ncl - ci < OFF_J ? ncl - ci : OFF_J;
#pragma unroll
        for (int j = 0; j < OFF_J; ++j) {
            uint32_t c = OFF_PAD;
            int src = -1;
            if (j < jn) {
                c = clist[ci + j];
                const int ox = (int)(c & 31) - 16, oy = (int)((c >> 5) & 31) - 16, oz = (int)((c >> 10) & 31) - 16;
                if (tv && ((c >> (16 + cls)) & 1)) src = win[pl + (ox * W + oy) * W + oz];
            }
            cd[j] = c;
            myc += src >= 0;
            sb[j * WAVE + lane] = src >= 0 ? src : (int)O.zero_node;
        }
        ci += jn;
        return jn;
    };
    // the batch's tensor rows from the table into registers (all loads in
    // flight at once) ...
    uint32_t dreg[OFF_J][NQ];
    auto stage_load = [&](const uint32_t (&cd)[OFF_J]) {
#pragma unroll
        for (int j = 0; j < OFF_J; ++j) {
            const uint32_t c = cd[j];
            const int ox = (int)(c & 31) - 16, oy = (int)((c >> 5) & 31) - 16, oz = (int)((c >> 10) & 31) - 16;
            const uint32_t* row = tab + (c != OFF_PAD ? (size_t)((abs(ox) * E + abs(oy)) * E + abs(oz)) * DCP : 0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int m = lane + WAVE * q;
                dreg[j][q] = row[m < DCP ? m : DCP - 1];  // masked at the write
            }
        }
    };
    // ... and from there, with their octant signs, to the wavefront's LDS rows
    auto stage_write = [&](const uint32_t (&cd)[OFF_J]) {
#pragma unroll
        for (int j = 0; j < OFF_J; ++j) {
            const uint32_t c = cd[j];
            const int ox = (int)(c & 31) - 16, oy = (int)((c >> 5) & 31) - 16, oz = (int)((c >> 10) & 31) - 16;
            const uint32_t neg = (uint32_t)(ox < 0) | ((uint32_t)(oy < 0) << 1) | ((uint32_t)(oz < 0) << 2);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int m = lane + WAVE * q;
                if (m < DCP) {
                    const uint32_t flip = (fm[q] >> neg) & 1u;
                    const uint32_t v = c != OFF_PAD ? dreg[j][q] : 0u;
                    db[j * DCP + m] = v ^ (flip ? (BF3 ? 0x80008000u : 0x80000000u) : 0u);
                }
            }
        }
    };
    auto process = [&](int b) {
        switch (wave) {
            case 0: off_batch<P, BF3, 0>(acc, O, b, ntile, r, hh); break;
            case 1: off_batch<P, BF3, 1>(acc, O, b, ntile, r, hh); break;
            case 2: off_batch<P, BF3, 2>(acc, O, b, ntile, r, hh); break;
            default: off_batch<P, BF3, 3>(acc, O, b, ntile, r, hh); break;
        }
    };

    // pipeline: the next batch is collected and its tensors are in flight
    // while the current one multiplies
    int cur = 0;
    if (collect(0, ccur) > 0) {
        stage_load(ccur);
        stage_write(ccur);
        off_lds_sync();
        for (;;) {
            const int jn = collect(cur ^ 1, cnxt);
            if (jn) stage_load(cnxt);
            off_lds_sync();
            process(cur);
            if (!jn) break;
            off_lds_sync();  // every lane is done with the current tensors
            stage_write(cnxt);
            off_lds_sync();
            cur ^= 1;
#pragma unroll
            for (int j = 0; j < OFF_J; ++j) ccur[j] = cnxt[j];
        }
    }

    // every target must have taken exactly the same-level entries of its list
    atomicAdd(&cnt[lane], myc);
    __syncthreads();
    if (wave == 0 && tv && cnt[lane] != lpre[lane + 1] - lpre[lane]) atomicOr(O.flags, OVF_GEOM);
    // the partial sums of wavefronts 1..3 to LDS (over the buffers), added by wavefront 0
    float* red = (float*)(off_lds + Lay::HDR);
    if (wave > 0) {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < ntile)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        red[((wave - 1) * D::KR + 16 * mt + 4 * hh + q) * WAVE + 16 * t + r] = acc[mt][t][q];
    }
    __syncthreads();
    if (wave == 0) {
        const int level = lvl + 2;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 16 * mt + 4 * hh + q;
                if (k >= NC) continue;
                int kx, ky, kz;
                off_coef_exps(k, kx, ky, kz);
                const float scale = off_inv_pow(F.box, level, 1 + kx + ky + kz);  // h^-(|k|+1)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int tg = 16 * t + r;
                    if (t >= ntile || tg >= nt) continue;
                    float v = acc[mt][t][q];
#pragma unroll
                    for (int w = 0; w < OFF_WAVES - 1; ++w) v += red[(w * D::KR + k) * WAVE + tg];
                    F.L[(size_t)(gc0 + tg) * F.ncp + k] = v * scale;
                }
            }
    }
}

// ---------------------------------------------------------------------------
template <int P, bool BF3>
static void launch_off_level_t(const FmmParams& F, const OffsetParams& O, uint32_t goff, uint32_t gcount, int lvl,
                               hipStream_t st) {
    const size_t lds = OffLayout<P, BF3>::total(O.W, O.ncand);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&m2l_offset_kernel<P, BF3>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((m2l_offset_kernel<P, BF3>), dim3(gcount), dim3(OFF_WAVES * WAVE), lds, st, F, O, goff, lvl);
}

template <int P, bool BF3>
static void launch_off_moments_t(const float* M, int ncp, uint32_t nn, const int4* geo, const double4* box, void* out,
                                 hipStream_t st) {
    using D = OffDims<P>;
    const uint32_t items = (nn + 1) * (BF3 ? D::NG : D::KCF);
    hipLaunchKernelGGL((off_moments_kernel<P, BF3>), dim3(ceil_div(items, 256u)), dim3(256), 0, st, M, ncp, nn, geo,
                       box, out);
}

template <int P, bool BF3>
static void launch_off_dtable_t(const double4* box, double eps2, int level0, int nlev, int E, uint32_t* out,
                                hipStream_t st) {
    const unsigned total = (unsigned)(E * E * E * nlev);
    hipLaunchKernelGGL((off_dtable_kernel<2 * P, BF3>), dim3(ceil_div(total, 64u)), dim3(64), 0, st, box, eps2, level0,
                       nlev, E, OffDims<P>::DCP, out);
}

#define NBH_OFF_DISPATCH(p, bf3, call)                                          \
    do {                                                                        \
        if (bf3) {                                                              \
            constexpr bool BB = true;                                           \
            switch (p) {                                                        \
                case 2: { constexpr int PP = 2; call; } break;                  \
                case 3: { constexpr int PP = 3; call; } break;                  \
                case 4: { constexpr int PP = 4; call; } break;                  \
                case 5: { constexpr int PP = 5; call; } break;                  \
                case 6: { constexpr int PP = 6; call; } break;                  \
                default: break;                                                 \
            }                                                                   \
        } else {                                                                \
            constexpr bool BB = false;                                          \
            switch (p) {                                                        \
                case 2: { constexpr int PP = 2; call; } break;                  \
                case 3: { constexpr int PP = 3; call; } break;                  \
                case 4: { constexpr int PP = 4; call; } break;                  \
                case 5: { constexpr int PP = 5; call; } break;                  \
                case 6: { constexpr int PP = 6; call; } break;                  \
                default: break;                                                 \
            }                                                                   \
        }                                                                       \
    } while (0)

bool off_supported(int p) { return p >= 2 && p <= 6; }

// Index table of the kernel's A operand: for the row j < OFF_J of a batch and
// k < KR, n < KC, the byte offset in the wavefront's staged rows of the
// tensor word D_{k+n} of row j, or of the row's zero word (index NC2) outside
// |k|, |n| <= P; and the octant parity bits of each tensor word.
template <int P, bool BF3>
static void off_tables_t(std::vector<uint16_t>& hank, std::vector<uint8_t>& par) {
    using D = OffDims<P>;
    constexpr int KC = OffLayout<P, BF3>::KC;
    hank.assign((size_t)OFF_J * D::KR * KC, 0);
    for (int j = 0; j < OFF_J; ++j)
        for (int k = 0; k < D::KR; ++k)
            for (int n = 0; n < KC; ++n) {
                int v = D::NC2;
                if (k < D::NC && n < D::NC) {
                    int kx, ky, kz, nx, ny, nz;
                    off_coef_exps(k, kx, ky, kz);
                    off_coef_exps(n, nx, ny, nz);
                    v = off_coef_index(kx + nx, ky + ny, kz + nz);
                }
                hank[((size_t)j * D::KR + k) * KC + n] = (uint16_t)(4 * (j * D::DCP + v));
            }
    par.assign(D::DCP, 0);
    for (int i = 0; i < D::NC2; ++i) {
        int x, y, z;
        off_coef_exps(i, x, y, z);
        par[i] = (uint8_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
    }
}
void off_tables(int p, bool bf3, std::vector<uint16_t>& hank, std::vector<uint8_t>& par) {
    NBH_OFF_DISPATCH(p, bf3, (off_tables_t<PP, BB>(hank, par)));
}

size_t off_lds_bytes(int p, bool bf3, int W, int ncand) {
    size_t b = 0;
    NBH_OFF_DISPATCH(p, bf3, (b = OffLayout<PP, BB>::total(W, ncand)));
    return b;
}
size_t off_moment_bytes(int p, bool bf3, uint32_t nn) {
    size_t b = 0;
    NBH_OFF_DISPATCH(p, bf3, (b = (size_t)(nn + 1) * (BB ? OffDims<PP>::NG * 32 : OffDims<PP>::KCF * 4)));
    return b;
}
int off_table_words(int p) { return (ncoef(2 * p) + 1 + 3) & ~3; }

void launch_off_moments(int p, bool bf3, const float* M, int ncp, uint32_t nn, const int4* geo, const double4* box,
                        void* out, hipStream_t st) {
    NBH_OFF_DISPATCH(p, bf3, (launch_off_moments_t<PP, BB>(M, ncp, nn, geo, box, out, st)));
}
void launch_off_dtable(int p, bool bf3, const double4* box, double eps2, int level0, int nlev, int E, uint32_t* out,
                       hipStream_t st) {
    NBH_OFF_DISPATCH(p, bf3, (launch_off_dtable_t<PP, BB>(box, eps2, level0, nlev, E, out, st)));
}
void launch_m2l_offset_level(int p, bool bf3, const FmmParams& F, const OffsetParams& O, uint32_t goff,
                             uint32_t gcount, int lvl, hipStream_t st) {
    if (!gcount) return;
    NBH_OFF_DISPATCH(p, bf3, (launch_off_level_t<PP, BB>(F, O, goff, gcount, lvl, st)));
}

}  // namespace nbh
