// walk.hip -- on-device dual-tree walk emitting compact M2L and P2P lists.
//
// Replaces the host-driven level loop of OpenClSimulation::step
// (open_cl_simulation.cpp:81-98): find_interactions (interaction.cl:10-100),
// the host classification (open_cl_simulation.cpp:242-267), the per-node
// interaction counters (interaction.cl:104-154) and the host prefix sums
// (open_cl_simulation.cpp:371-484).
//
// The reference splits unordered pairs (A,B) from (root, root): a side with
// children is replaced by its children, a childless side by itself
// (interaction.cl:42-49, 58-59); pairs are approximated when
//   A != B  and  0.75 (d_A + d_B)^2 / |c_B - c_A|^2 < theta^2   (:63-79)
// and split further when not approximated and either side has children
// (:80-82).  Unsplit, unapproximated pairs are near-field (P2P).  Read as
// directed pairs (target <- source), the set of pairs is the same, and the
// candidates for target A at recursion depth s are
//   - A at level s:       expand(SL(parent(A)))
//   - leaf A above s:     expand(SL(A))        (A stays, the source splits)
// where SL(X) are the sources that X's pairs were split against and
// expand(B) = children(B) if B has children, else {B}.  The walk runs one
// depth at a time, one wavefront per target row; rows of depth s are the
// level-s cells followed by the leaves carried from depth s-1.
//
// Each target's lists come out in ascending source id without any sort:
// split lists are ascending, kept leaves (coarser levels) precede children,
// and deeper depths only add deeper (larger-id) sources.
//
// The MAC is evaluated exactly in integers (units of half the finest cell):
//   3 den^2 (d_A + d_B)^2 < 4 num^2 |c_B - c_A|^2,  theta = num/den.
#include "common.h"
#include "launch.h"

namespace nbh {



__device__ __forceinline__ bool mac_approx(int A, int4 ga, int B, int4 gb, int bits, uint32_t num4,
                                           uint32_t den3) {
    if (A == B) return false;
    const int la = ga.w & 0xff, lb = gb.w & 0xff;
    const int sa = bits - la, sb = bits - lb;
    const int dx = ((2 * gb.x + 1) << sb) - ((2 * ga.x + 1) << sa);
    const int dy = ((2 * gb.y + 1) << sb) - ((2 * ga.y + 1) << sa);
    const int dz = ((2 * gb.z + 1) << sb) - ((2 * ga.z + 1) << sa);
    const uint64_t dsq = (uint64_t)((int64_t)dx * dx) + (uint64_t)((int64_t)dy * dy) + (uint64_t)((int64_t)dz * dz);
    const uint64_t side = (1ull << (sa + 1)) + (1ull << (sb + 1));
    return (uint64_t)den3 * side * side < (uint64_t)num4 * dsq;
}

__device__ __forceinline__ uint32_t depth_base(const Counters* C, int s) {
    uint32_t b = 0;
    for (int q = 0; q < s; ++q) b += C->level_count[q] + C->carried[q];
    return b;
}

enum { CAT_M2L = 0, CAT_P2P = 1, CAT_SL = 2, CAT_NONE = 3 };

// LDS writes of one lane made visible to the other lanes of its wavefront
__device__ __forceinline__ void walk_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS cache of one split list (SL_CACHE entries at a time): the
// entry's node id, geometry and first child.  Loading it once, coalesced,
// leaves one global round trip (the children's geometry) per candidate chunk.
constexpr int SL_CACHE = 256;
struct SLCache {
    int4* geo;
    int* node;
    int* fc;
};

__device__ __forceinline__ void sl_cache_fill(const WalkParams& W, const SLCache& c, uint32_t st, uint32_t e0,
                                              uint32_t n) {
    const int lane = lane_id();
    for (uint32_t k = lane; k < n; k += WAVE) {
        const int B = W.sl[st + e0 + k];
        c.node[k] = B;
        c.geo[k] = W.nodes.geo[B];
        c.fc[k] = W.nodes.link[B].z;
    }
    walk_lds_sync();
}

// Enumerate the candidates of one row; f(lane_valid, candidate, geo, category)
// is called once per chunk of 64 lanes, in ascending candidate order: first
// the split-list entries without children (kept as themselves), then the
// children of the others, entry by entry.
template <typename F>
__device__ __forceinline__ void enumerate(const WalkParams& W, const SLCache& cache, int A, int4 ga, int src,
                                          F&& f) {
    const bool a_leaf = (ga.w >> 8) == 0;
    const int lane = lane_id();
    auto classify = [&](bool valid, int B, int4 gb) -> int {
        if (!valid) return CAT_NONE;
        if (mac_approx(A, ga, B, gb, W.bits, W.num4, W.den3)) return CAT_M2L;
        const bool b_leaf = (gb.w >> 8) == 0;
        if (a_leaf && b_leaf) return CAT_P2P;
        return CAT_SL;
    };
    if (src < 0) {
        const bool valid = lane == 0;
        f(valid, A, ga, classify(valid, A, ga));
        return;
    }
    const uint32_t st = W.row_sl_start[src];
    uint32_t cnt = W.row_sl_cnt[src];
    if ((unsigned long long)st + cnt > W.sl_cap) cnt = 0;  // only after an overflow
    // phase a: sources kept as themselves (childless)
    for (uint32_t p0 = 0; p0 < cnt; p0 += SL_CACHE) {
        const uint32_t pn = min((uint32_t)SL_CACHE, cnt - p0);
        sl_cache_fill(W, cache, st, p0, pn);
        for (uint32_t e0 = 0; e0 < pn; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            int B = -1;
            int4 gb = make_int4(0, 0, 0, 0);
            if (e < pn) {
                B = cache.node[e];
                gb = cache.geo[e];
            }
            const bool valid = B >= 0 && (gb.w >> 8) == 0;
            if (__ballot(valid) == 0) continue;
            f(valid, B, gb, classify(valid, B, gb));
        }
        walk_lds_sync();
    }
    // phase b: children of the sources that have children, 8 entries x 8
    // slots per chunk, two chunks in flight
    for (uint32_t p0 = 0; p0 < cnt; p0 += SL_CACHE) {
        const uint32_t pn = min((uint32_t)SL_CACHE, cnt - p0);
        if (cnt > SL_CACHE) sl_cache_fill(W, cache, st, p0, pn);  // else still cached from phase a
        const int c = lane & 7;
        for (uint32_t e0 = 0; e0 < pn; e0 += 16) {
            bool v0 = false, v1 = false;
            int C0 = -1, C1 = -1;
            int4 g0 = make_int4(0, 0, 0, 0), g1 = make_int4(0, 0, 0, 0);
            const uint32_t ea = e0 + (lane >> 3), eb = ea + 8;
            if (ea < pn && c < (cache.geo[ea].w >> 8)) { C0 = cache.fc[ea] + c; v0 = true; }
            if (eb < pn && c < (cache.geo[eb].w >> 8)) { C1 = cache.fc[eb] + c; v1 = true; }
            if (v0) g0 = W.nodes.geo[C0];
            if (v1) g1 = W.nodes.geo[C1];
            if (__ballot(v0)) f(v0, C0, g0, classify(v0, C0, g0));
            if (__ballot(v1)) f(v1, C1, g1, classify(v1, C1, g1));
        }
        walk_lds_sync();
    }
}

// Phase 1: one enumeration per row.  The row's M2L, P2P and split entries
// are buffered in LDS (per-wave regions), then written as three contiguous
// segments carved from per-wave pool chunks; a row that overflows its LDS
// region is enumerated a second time straight into the pools.  Per-target
// counts (for the CSR offsets) and leaf carries as before.
constexpr int CAP_M = 1536, CAP_P = 256, CAP_S = 256;

// Take n entries from a wave's chunk {next, end}, reserving a new chunk of
// max(n, chunk) entries from the pool when the current one is exhausted.
// Returns the first index, or 0xffffffff when the pool is full.
__device__ __forceinline__ uint32_t pool_take(uint2& ch, uint32_t n, uint32_t chunk, uint32_t* used,
                                              unsigned long long cap) {
    if (ch.x + n > ch.y || ch.y == 0) {
        const uint32_t want = n > chunk ? n : chunk;
        const uint32_t s0 = atomicAdd(used, want);
        if ((unsigned long long)s0 + want > cap) {
            ch = make_uint2(0, 0);
            return 0xffffffffu;
        }
        ch = make_uint2(s0, s0 + want);
    }
    const uint32_t r = ch.x;
    ch.x += n;
    return r;
}

__global__ __launch_bounds__(WAVE * WALK_WAVES) void walk_phase1_kernel(WalkParams W, int s) {
    __shared__ int lbuf[WALK_WAVES][CAP_M + CAP_P + CAP_S];
    __shared__ int4 cgeo[WALK_WAVES][SL_CACHE];
    __shared__ int cnode[WALK_WAVES][SL_CACHE];
    __shared__ int cfc[WALK_WAVES][SL_CACHE];
    Counters* C = W.C;
    const uint32_t n_level = C->level_count[s];
    const uint32_t n_rows = n_level + C->carried[s];
    const uint32_t base = depth_base(C, s);
    const uint32_t base_prev = s > 0 ? base - (C->level_count[s - 1] + C->carried[s - 1]) : 0;
    const uint32_t base_next = base + n_rows;
    const int lane = lane_id();
    const int wslot = threadIdx.x / WAVE;
    int* bm = lbuf[wslot];
    int* bp = bm + CAP_M;
    int* bs = bp + CAP_P;
    const SLCache cache{cgeo[wslot], cnode[wslot], cfc[wslot]};
    const uint32_t wid = blockIdx.x * WALK_WAVES + wslot;
    const uint32_t nw = gridDim.x * WALK_WAVES;
    const uint64_t lt = lanemask_lt();
    uint2 mchunk = make_uint2(0, 0), pchunk = make_uint2(0, 0), schunk = make_uint2(0, 0);  // lane 0 only
    unsigned long long split_acc = 0;
    for (uint32_t rl = wid; rl < n_rows; rl += nw) {
        const uint32_t r = base + rl;
        if (r >= W.row_cap) {  // rows beyond capacity: flag, the host retries larger
            if (lane == 0) atomicOr(&C->overflow, (uint32_t)OVF_ROWS);
            continue;
        }
        int A, src;
        if (rl < n_level) {
            A = (int)(C->level_off[s] + rl);
            src = s == 0 ? -1 : (int)(base_prev + (uint32_t)(W.nodes.link[A].w - (int)C->level_off[s - 1]));
        } else {
            A = W.row_node[r];
            src = W.row_src[r];
        }
        const int4 ga = W.nodes.geo[A];
        uint32_t nm = 0, np = 0, ns = 0;
        enumerate(W, cache, A, ga, src, [&](bool valid, int B, int4 gb, int cat) {
            const uint64_t mm = __ballot(cat == CAT_M2L);
            const uint64_t mp = __ballot(cat == CAT_P2P);
            const uint64_t ms = __ballot(cat == CAT_SL);
            if (cat == CAT_M2L) { const uint32_t q = nm + __popcll(mm & lt); if (q < CAP_M) bm[q] = B; }
            if (cat == CAT_P2P) { const uint32_t q = np + __popcll(mp & lt); if (q < CAP_P) bp[q] = B; }
            if (cat == CAT_SL) { const uint32_t q = ns + __popcll(ms & lt); if (q < CAP_S) bs[q] = B; }
            nm += __popcll(mm);
            np += __popcll(mp);
            ns += __popcll(ms);
        });
        uint32_t ms0 = 0, ps0 = 0, ss0 = 0;
        if (lane == 0) {
            W.m2l_cnt[A] += nm;
            W.p2p_cnt[A] += np;
            // segments come from per-wave chunks (one global atomic per chunk,
            // not per row: same-address atomics serialise in L2)
            if (nm) {
                ms0 = pool_take(mchunk, nm, POOL_CHUNK_M, &C->m2l_pool_used, W.m2l_pool_cap);
                if (ms0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_M2L); nm = 0; ms0 = 0; }
            }
            if (np) {
                ps0 = pool_take(pchunk, np, POOL_CHUNK_P, &C->p2p_pool_used, W.p2p_pool_cap);
                if (ps0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_P2P); np = 0; ps0 = 0; }
            }
            if (ns) {
                ss0 = pool_take(schunk, ns, POOL_CHUNK_S, &C->sl_used, W.sl_cap);
                if (ss0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_SL); ns = 0; ss0 = 0; }
                split_acc += ns;
            }
            W.row_m_start[r] = ms0;
            W.row_m_cnt[r] = nm;
            W.row_p_start[r] = ps0;
            W.row_p_cnt[r] = np;
            W.row_sl_start[r] = ss0;
            W.row_sl_cnt[r] = ns;
            const bool a_leaf = (ga.w >> 8) == 0;
            if (a_leaf && ns) {
                const uint32_t k = atomicAdd(&C->carried[s + 1], 1u);
                const uint32_t rr = base_next + C->level_count[s + 1] + k;
                if (rr < W.row_cap) {
                    W.row_node[rr] = A;
                    W.row_src[rr] = (int)r;
                } else {
                    atomicOr(&C->overflow, (uint32_t)OVF_ROWS);
                }
            }
        }
        ms0 = __shfl(ms0, 0, WAVE); nm = __shfl(nm, 0, WAVE);
        ps0 = __shfl(ps0, 0, WAVE); np = __shfl(np, 0, WAVE);
        ss0 = __shfl(ss0, 0, WAVE); ns = __shfl(ns, 0, WAVE);
        walk_lds_sync();
        if (nm <= CAP_M && np <= CAP_P && ns <= CAP_S) {
            for (uint32_t k = lane; k < nm; k += WAVE) W.m2l_pool[ms0 + k] = bm[k];
            for (uint32_t k = lane; k < np; k += WAVE) W.p2p_pool[ps0 + k] = bp[k];
            for (uint32_t k = lane; k < ns; k += WAVE) W.sl[ss0 + k] = bs[k];
        } else {
            uint32_t cm = 0, cp = 0, cs = 0;
            enumerate(W, cache, A, ga, src, [&](bool valid, int B, int4 gb, int cat) {
                const uint64_t mm = __ballot(cat == CAT_M2L);
                const uint64_t mp = __ballot(cat == CAT_P2P);
                const uint64_t ms = __ballot(cat == CAT_SL);
                if (cat == CAT_M2L) { const uint32_t q = cm + __popcll(mm & lt); if (q < nm) W.m2l_pool[ms0 + q] = B; }
                if (cat == CAT_P2P) { const uint32_t q = cp + __popcll(mp & lt); if (q < np) W.p2p_pool[ps0 + q] = B; }
                if (cat == CAT_SL) { const uint32_t q = cs + __popcll(ms & lt); if (q < ns) W.sl[ss0 + q] = B; }
                cm += __popcll(mm);
                cp += __popcll(mp);
                cs += __popcll(ms);
            });
        }
        walk_lds_sync();
    }
    if (lane == 0 && split_acc) atomicAdd(&C->split_total, split_acc);
}

// Phase 2 (per depth, in depth order): copy each row's M2L/P2P segments to
// its target's CSR cursor.  Depth order makes a leaf's entries from deeper
// recursion depths follow its earlier ones, so every list stays ascending.
__global__ __launch_bounds__(256) void walk_phase2_kernel(WalkParams W, int s) {
    Counters* C = W.C;
    const uint32_t n_level = C->level_count[s];
    const uint32_t n_rows = n_level + C->carried[s];
    const uint32_t base = depth_base(C, s);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const int lane = lane_id();
    unsigned long long inter_acc = 0;
    for (uint32_t rl = wid; rl < n_rows; rl += nw) {
        const uint32_t r = base + rl;
        if (r >= W.row_cap) continue;  // cannot happen after a clean phase 1
        const int A = rl < n_level ? (int)(C->level_off[s] + rl) : W.row_node[r];
        const uint32_t nm = W.row_m_cnt[r], np = W.row_p_cnt[r];
        const uint32_t ms0 = W.row_m_start[r], ps0 = W.row_p_start[r];
        const unsigned long long cm = W.m2l_cur[A], cp = W.p2p_cur[A];
        // eight loads in flight per lane before their stores
        for (uint32_t k0 = 0; k0 < nm; k0 += 8 * WAVE) {
            int v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t k = k0 + u * WAVE + lane;
                v[u] = k < nm ? W.m2l_pool[ms0 + k] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t k = k0 + u * WAVE + lane;
                if (k < nm && cm + k < W.m2l_cap) W.m2l_src[cm + k] = v[u];
            }
        }
        unsigned long long inter = 0;
        const int na = W.nodes.link[A].y;
        for (uint32_t k = lane; k < np; k += WAVE) {
            const int B = W.p2p_pool[ps0 + k];
            if (cp + k < W.p2p_cap) W.p2p_src[cp + k] = B;
            const int nb = W.nodes.link[B].y;
            inter += (unsigned long long)na * (unsigned long long)(B == A ? nb - 1 : nb);
        }
        inter_acc += inter;
        if (lane == 0) {
            W.m2l_cur[A] = cm + nm;
            W.p2p_cur[A] = cp + np;
        }
    }
    inter_acc = wave_sum(inter_acc);
    if (lane == 0 && inter_acc) atomicAdd(&C->p2p_inter, inter_acc);
}

void launch_walk_phase1(const WalkParams& W, int s, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(walk_phase1_kernel, dim3(grid), dim3(WAVE * WALK_WAVES), 0, st, W, s);
}
void launch_walk_phase2(const WalkParams& W, int s, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(walk_phase2_kernel, dim3(grid), dim3(256), 0, st, W, s);
}

}  // namespace nbh
