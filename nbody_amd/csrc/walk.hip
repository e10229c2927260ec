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
// expand(B) = children(B) if B has children, else {B}.
//
// Rows and groups.  The walk runs one depth at a time.  A row is one target at
// one depth: rows of depth s are the level-s cells (in id order) followed by
// the leaves carried from depth s-1.  Every row of depth s-1 is a group: its
// candidate set expand(SL(row)) is shared by all of its targets at depth s --
// the cell's children, or the carried leaf itself.  One wavefront takes one
// group, loads each candidate once and classifies it against every target
// (up to 8), in two passes: count, reserve, write.  Each row's M2L, P2P and
// split entries land as contiguous segments carved from per-wave chunks of
// three pools; a node's rows are chained in depth order (first row = its
// level row, next = its carried row one depth deeper), and the force kernels
// follow the chain, so no compaction pass is needed.
//
// Each target's entries come out in ascending source id, chain order included:
// split lists are ascending, kept leaves (coarser levels) precede children,
// and deeper depths only add deeper (larger-id) sources.
//
// The MAC is evaluated exactly in integers (units of half the finest cell):
//   3 den^2 (d_A + d_B)^2 < 4 num^2 |c_B - c_A|^2,  theta = num/den.
#include "common.h"
#include "launch.h"

namespace nbh {

__device__ __forceinline__ uint32_t depth_base(const Counters* C, int s) {
    uint32_t b = 0;
    for (int q = 0; q < s; ++q) b += C->level_count[q] + C->carried[q];
    return b;
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
    wave_lds_sync();
}

// Enumerate expand(SL) for the split list at sl[st, st+cnt), one wavefront
// with the cache to itself; f(phase, valid, B, geo) is called once per chunk
// of 64 lanes (uniformly), in ascending candidate order: phase 0 the entries
// without children (kept as themselves), then phase 1 the children of the
// others, entry by entry.  The cache is refilled SL_CACHE entries at a time.
template <typename F>
__device__ __forceinline__ void enumerate_sl(const WalkParams& W, const SLCache& cache, uint32_t st, uint32_t cnt,
                                             F&& f) {
    const int lane = lane_id();
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
            f(0, valid, B, gb);
        }
        wave_lds_sync();
    }
    for (uint32_t p0 = 0; p0 < cnt; p0 += SL_CACHE) {
        const uint32_t pn = min((uint32_t)SL_CACHE, cnt - p0);
        sl_cache_fill(W, cache, st, p0, pn);
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
            if (__ballot(v0)) f(1, v0, C0, g0);
            if (__ballot(v1)) f(1, v1, C1, g1);
        }
        wave_lds_sync();
    }
}

// The same for the cached entries [lo, hi) of a list that the cache holds
// whole (one wavefront's slice of a workgroup-shared cache).
template <typename F>
__device__ __forceinline__ void enumerate_slice(const WalkParams& W, const SLCache& cache, uint32_t lo, uint32_t hi,
                                                F&& f) {
    const int lane = lane_id();
    for (uint32_t e0 = lo; e0 < hi; e0 += WAVE) {
        const uint32_t e = e0 + lane;
        int B = -1;
        int4 gb = make_int4(0, 0, 0, 0);
        if (e < hi) {
            B = cache.node[e];
            gb = cache.geo[e];
        }
        const bool valid = B >= 0 && (gb.w >> 8) == 0;
        if (__ballot(valid) == 0) continue;
        f(0, valid, B, gb);
    }
    // children: 8 entries x 8 slots per chunk, two chunks in flight
    const int c = lane & 7;
    for (uint32_t e0 = lo; e0 < hi; e0 += 16) {
        bool v0 = false, v1 = false;
        int C0 = -1, C1 = -1;
        int4 g0 = make_int4(0, 0, 0, 0), g1 = make_int4(0, 0, 0, 0);
        const uint32_t ea = e0 + (lane >> 3), eb = ea + 8;
        if (ea < hi && c < (cache.geo[ea].w >> 8)) { C0 = cache.fc[ea] + c; v0 = true; }
        if (eb < hi && c < (cache.geo[eb].w >> 8)) { C1 = cache.fc[eb] + c; v1 = true; }
        if (v0) g0 = W.nodes.geo[C0];
        if (v1) g1 = W.nodes.geo[C1];
        if (__ballot(v0)) f(1, v0, C0, g0);
        if (__ballot(v1)) f(1, v1, C1, g1);
    }
}

// The call sequence of enumerate_slice, with the candidate ids only (no
// geometry loads): f(phase, valid, B).  The validity tests read the cache
// alone, so the calls match enumerate_slice's one for one.
template <typename F>
__device__ __forceinline__ void enumerate_slice_ids(const SLCache& cache, uint32_t lo, uint32_t hi, F&& f) {
    const int lane = lane_id();
    for (uint32_t e0 = lo; e0 < hi; e0 += WAVE) {
        const uint32_t e = e0 + lane;
        int B = -1;
        bool valid = false;
        if (e < hi) {
            B = cache.node[e];
            valid = B >= 0 && (cache.geo[e].w >> 8) == 0;
        }
        if (__ballot(valid) == 0) continue;
        f(0, valid, B);
    }
    const int c = lane & 7;
    for (uint32_t e0 = lo; e0 < hi; e0 += 16) {
        bool v0 = false, v1 = false;
        int C0 = -1, C1 = -1;
        const uint32_t ea = e0 + (lane >> 3), eb = ea + 8;
        if (ea < hi && c < (cache.geo[ea].w >> 8)) { C0 = cache.fc[ea] + c; v0 = true; }
        if (eb < hi && c < (cache.geo[eb].w >> 8)) { C1 = cache.fc[eb] + c; v1 = true; }
        if (__ballot(v0)) f(1, v0, C0);
        if (__ballot(v1)) f(1, v1, C1);
    }
}

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

// Depth 0: one row per root (one per system), the candidate being the root
// itself: a childless root is its own P2P source, otherwise (root, root) is
// split.  One thread per root.
__global__ __launch_bounds__(64) void walk_root_kernel(WalkParams W) {
    Counters* C = W.C;
    if (C->overflow & OVF_NODES) return;  // a captured step's tree outgrew its node arrays
    const uint32_t n_roots = C->level_count[0];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n_roots; r += gridDim.x * blockDim.x) {
        if (r >= W.row_cap) {
            atomicOr(&C->overflow, (uint32_t)OVF_ROWS);
            continue;
        }
        const int A = (int)(C->level_off[0] + r);
        const bool leaf = (W.nodes.geo[A].w >> 8) == 0;
        W.row_first[A] = r;
        W.row_next[r] = NO_ROW;
        W.row_node[r] = A;
        W.row_m_start[r] = 0;
        W.row_m_cnt[r] = 0;
        W.row_mx_cnt[r] = 0;
        uint32_t pc = 0, ps = 0, sc = 0, ss = 0;
        if (leaf) {
            ps = atomicAdd(&C->p2p_pool_used, 1u);
            if (ps < W.p2p_pool_cap) {
                W.p2p_pool[ps] = A;
                pc = 1;
                atomicAdd(&C->p2p_total, 1ull);
            } else {
                atomicOr(&C->overflow, (uint32_t)OVF_P2P);
                ps = 0;
            }
        } else {
            ss = atomicAdd(&C->sl_used, 1u);
            if (ss < W.sl_cap) {
                W.sl[ss] = A;
                sc = 1;
                atomicAdd(&C->split_total, 1ull);
            } else {
                atomicOr(&C->overflow, (uint32_t)OVF_SL);
                ss = 0;
            }
        }
        W.row_p_start[r] = ps;
        W.row_p_cnt[r] = pc;
        W.row_sl_start[r] = ss;
        W.row_sl_cnt[r] = sc;
    }
}

constexpr int MAXT = 8;  // targets per group: the children of one cell
// Calls of enumerate_slice per wavefront and group: a slice of at most
// SL_CACHE / WALK_WAVES = 64 entries gives one chunk of kept entries and at
// most 4 x 2 chunks of children.
constexpr int CAT_CALLS = 1 + 2 * (SL_CACHE / WALK_WAVES / 16);

// Depth s >= 1: one workgroup of WALK_WAVES wavefronts per group (row of
// depth s-1).  The group's split list is cached once for the workgroup and
// each wavefront classifies the candidates of a contiguous slice of it
// against all targets.  Pass 1 counts per (wavefront, phase, target,
// category); thread 0 reserves one block per pool for the group; every
// wavefront then writes its entries at offsets ordered phase-major,
// wavefront-minor, which is the order of one wavefront walking the whole
// list.  Lists longer than the cache are walked by wavefront 0 alone.  Rows
// are taken with a fixed stride over the workgroups; their headers are read
// 64 at a time so rows without work (leaves with nothing left to split) cost
// one batched read.
__global__ __launch_bounds__(WAVE * WALK_WAVES) void walk_group_kernel(WalkParams W, int s) {
    __shared__ int4 cgeo[SL_CACHE];
    __shared__ int cnode[SL_CACHE];
    __shared__ int cfc[SL_CACHE];
    __shared__ uint32_t cnt_s[WALK_WAVES][2][4][MAXT];  // [wave][phase][same-level, other-level, P2P, split][target]
    __shared__ uint32_t bcast[8];                       // carried row, pool starts, failure
    // pass 1's categories of a cached group, 3 bits per target, per call and lane
    __shared__ uint32_t cat_s[WALK_WAVES][CAT_CALLS][WAVE];
    Counters* C = W.C;
    if (C->overflow & OVF_NODES) return;  // a captured step's tree outgrew its node arrays
    const uint32_t n_prev_level = C->level_count[s - 1];
    const uint32_t n_groups = n_prev_level + C->carried[s - 1];
    const uint32_t base_prev = depth_base(C, s - 1);
    const uint32_t base = base_prev + n_groups;
    const uint32_t n_level = C->level_count[s];
    const uint32_t loff = C->level_off[s], loff_prev = C->level_off[s - 1];
    const int lane = lane_id();
    const int wv = threadIdx.x / WAVE;
    const SLCache cache{cgeo, cnode, cfc};
    const uint64_t lt = lanemask_lt();
    const int bits = W.bits;
    const uint64_t num4 = W.num4, den3 = W.den3;
    const uint32_t nblk = gridDim.x;
    uint2 mchunk = make_uint2(0, 0), pchunk = make_uint2(0, 0), schunk = make_uint2(0, 0);  // thread 0 only
    unsigned long long acc_m = 0, acc_p = 0, acc_s = 0, acc_x = 0;                           // thread 0 only
    for (uint64_t j0 = 0; (uint64_t)blockIdx.x + j0 * nblk < n_groups; j0 += WAVE) {
        // headers of the next 64 rows of this workgroup (identical in every wavefront)
        const uint64_t gl = (uint64_t)blockIdx.x + (j0 + lane) * nblk;
        int P_l = 0;
        int4 gp_l = make_int4(0, 0, 0, 0);
        uint32_t st_l = 0, sn_l = 0;
        bool live = false;
        if (gl < n_groups && base_prev + gl < W.row_cap) {  // rows past the cap were flagged when made
            const uint32_t g = (uint32_t)gl, rg = base_prev + g;
            P_l = g < n_prev_level ? (int)(loff_prev + g) : W.row_node[rg];
            gp_l = W.nodes.geo[P_l];
            st_l = W.row_sl_start[rg];
            sn_l = W.row_sl_cnt[rg];
            if ((unsigned long long)st_l + sn_l > W.sl_cap) sn_l = 0;  // only after an overflow
            live = (gp_l.w >> 8) > 0 || sn_l > 0;
        }
        for (uint64_t todo = __ballot(live); todo; todo &= todo - 1) {  // uniform over the workgroup
            const int jl = __ffsll((unsigned long long)todo) - 1;
            const uint32_t g = (uint32_t)(blockIdx.x + (j0 + jl) * nblk);
            const uint32_t rg = base_prev + g;
            const int P = __builtin_amdgcn_readlane(P_l, jl);
            const int gpw = __builtin_amdgcn_readlane(gp_l.w, jl);
            const uint32_t sl_st = __builtin_amdgcn_readlane(st_l, jl);
            const uint32_t sl_n = __builtin_amdgcn_readlane(sn_l, jl);
            const int nch = gpw >> 8;
            __syncthreads();  // the previous group is done with the cache, counts and broadcasts
            if (nch == 0 && threadIdx.x == 0) bcast[0] = atomicAdd(&C->carried[s], 1u);
            __syncthreads();
            int nt, A0;
            uint32_t row0;
            if (nch > 0) {  // the cell's children: level-s rows in id order
                nt = nch;
                A0 = W.nodes.link[P].z;
                row0 = base + (uint32_t)(A0 - (int)loff);
            } else {  // a leaf with sources left to split: carried to depth s
                nt = 1;
                A0 = P;
                row0 = base + n_level + bcast[0];
            }
            if ((unsigned long long)row0 + nt > W.row_cap) {
                // the host retries with more rows; rows that fit are left empty so
                // that the next depth reads only rows of this evaluation
                if (threadIdx.x == 0) atomicOr(&C->overflow, (uint32_t)OVF_ROWS);
                const uint32_t r = row0 + lane;
                if (wv == 0 && lane < nt && r < W.row_cap) {
                    W.row_node[r] = A0 + lane;
                    W.row_next[r] = NO_ROW;
                    W.row_m_start[r] = W.row_m_cnt[r] = W.row_mx_cnt[r] = 0;
                    W.row_p_start[r] = W.row_p_cnt[r] = 0;
                    W.row_sl_start[r] = W.row_sl_cnt[r] = 0;
                }
                continue;
            }
            // targets (all of one level): lane t < nt holds target t; per-target
            // state lives in lane t and is read with readlane where the loop
            // over targets needs it
            int4 gt_l = make_int4(0, 0, 0, 0);
            if (lane < nt) gt_l = W.nodes.geo[A0 + lane];
            const int la = gpw & 0xff;
            const int lt_level = nch > 0 ? la + 1 : la;
            const int sa = bits - lt_level;
            const int side_a = 1 << (sa + 1);
            const int cax_l = (2 * gt_l.x + 1) << sa, cay_l = (2 * gt_l.y + 1) << sa, caz_l = (2 * gt_l.z + 1) << sa;
            const int leaf_l = (gt_l.w >> 8) == 0;
            // the split list: cached for the workgroup and sliced over the
            // wavefronts, or (longer than the cache) walked by wavefront 0
            const bool solo = sl_n > (uint32_t)SL_CACHE;
            if (!solo) {
                for (uint32_t k = threadIdx.x; k < sl_n; k += blockDim.x) {
                    const int B = W.sl[sl_st + k];
                    cache.node[k] = B;
                    cache.geo[k] = W.nodes.geo[B];
                    cache.fc[k] = W.nodes.link[B].z;
                }
            }
            __syncthreads();
            const uint32_t lo = solo ? 0u : (uint32_t)(((uint64_t)sl_n * wv) / WALK_WAVES);
            const uint32_t hi = solo ? (wv == 0 ? sl_n : 0u) : (uint32_t)(((uint64_t)sl_n * (wv + 1)) / WALK_WAVES);
            auto each = [&](auto&& f) {
                if (solo) {
                    if (wv == 0 && sl_n) enumerate_sl(W, cache, sl_st, sl_n, f);
                } else if (lo < hi) {
                    enumerate_slice(W, cache, lo, hi, f);
                }
            };
            // per-candidate part of the MAC, then the per-target categories
            // (0 = M2L same level, 4 = M2L other level, 1 = P2P, 2 = split, 3 = none)
            auto classify_all = [&](bool valid, int B, int4 gb, auto&& per_target) {
                const int lb = gb.w & 0xff;
                const int sb = bits - lb;
                const int cbx = (2 * gb.x + 1) << sb, cby = (2 * gb.y + 1) << sb, cbz = (2 * gb.z + 1) << sb;
                const uint64_t side = (uint64_t)side_a + (1ull << (sb + 1));
                const uint64_t lhs = den3 * side * side;
                const bool b_leaf = (gb.w >> 8) == 0;
#pragma unroll
                for (int t = 0; t < MAXT; ++t) {
                    if (t < nt) {
                        const int64_t dx = cbx - __builtin_amdgcn_readlane(cax_l, t);
                        const int64_t dy = cby - __builtin_amdgcn_readlane(cay_l, t);
                        const int64_t dz = cbz - __builtin_amdgcn_readlane(caz_l, t);
                        const uint64_t dsq = (uint64_t)(dx * dx + dy * dy + dz * dz);
                        // evaluated without short cuts: selects, not branches
                        const bool approx = (B != A0 + t) & (lhs < num4 * dsq);
                        const bool both_leaf = (__builtin_amdgcn_readlane(leaf_l, t) != 0) & b_leaf;
                        const int near = both_leaf ? 1 : 2;
                        const int far = lb == lt_level ? 0 : 4;
                        const int cat = valid ? (approx ? far : near) : 3;
                        per_target(t, cat);
                    }
                }
            };
            // pass 1: counts of this wavefront's slice per phase, category, target
            // (lane t); a cached group also keeps every call's categories for pass 2
            uint32_t k0m = 0, k0x = 0, k0p = 0, k0s = 0, k1m = 0, k1x = 0, k1p = 0, k1s = 0;
            int ncall = 0;  // uniform
            each([&](int ph, bool valid, int B, int4 gb) {
                uint32_t word = 0;
                classify_all(valid, B, gb, [&](int t, int cat) {
                    word |= (uint32_t)cat << (3 * t);
                    const uint32_t cm = __popcll(__ballot(cat == 0)), cx = __popcll(__ballot(cat == 4));
                    const uint32_t cp = __popcll(__ballot(cat == 1)), cs = __popcll(__ballot(cat == 2));
                    if (lane == t) {
                        if (ph == 0) { k0m += cm; k0x += cx; k0p += cp; k0s += cs; }
                        else { k1m += cm; k1x += cx; k1p += cp; k1s += cs; }
                    }
                });
                if (!solo && ncall < CAT_CALLS) cat_s[wv][ncall][lane] = word;
                ++ncall;
            });
            if (lane < MAXT) {
                cnt_s[wv][0][0][lane] = k0m; cnt_s[wv][0][1][lane] = k0x;
                cnt_s[wv][0][2][lane] = k0p; cnt_s[wv][0][3][lane] = k0s;
                cnt_s[wv][1][0][lane] = k1m; cnt_s[wv][1][1][lane] = k1x;
                cnt_s[wv][1][2][lane] = k1p; cnt_s[wv][1][3][lane] = k1s;
            }
            __syncthreads();
            // per target (lane t): totals, and this wavefront's start inside the
            // target's segments per phase (counts of earlier (phase, wavefront))
            uint32_t Tm = 0, Tx = 0, Tp = 0, Ts = 0;
            uint32_t o0m = 0, o0x = 0, o0p = 0, o0s = 0, o1m = 0, o1x = 0, o1p = 0, o1s = 0;
            if (lane < nt) {
#pragma unroll
                for (int ph = 0; ph < 2; ++ph)
#pragma unroll
                    for (int w = 0; w < (int)WALK_WAVES; ++w) {
                        const uint32_t m = cnt_s[w][ph][0][lane], x = cnt_s[w][ph][1][lane];
                        const uint32_t pp = cnt_s[w][ph][2][lane], ss = cnt_s[w][ph][3][lane];
                        if (ph == 0 && w == wv) { o0m = Tm; o0x = Tx; o0p = Tp; o0s = Ts; }
                        if (ph == 1 && w == wv) { o1m = Tm; o1x = Tx; o1p = Tp; o1s = Ts; }
                        Tm += m; Tx += x; Tp += pp; Ts += ss;
                    }
            }
            const uint32_t tm = wave_sum(Tm + Tx), tp = wave_sum(Tp), ts = wave_sum(Ts), tx = wave_sum(Tx);
            // reserve one block per pool for all targets of the group
            if (threadIdx.x == 0) {
                uint32_t ms0 = 0, ps0 = 0, ss0 = 0, fail = 0;
                if (tm) {
                    ms0 = pool_take(mchunk, tm, POOL_CHUNK_M, &C->m2l_pool_used, W.m2l_pool_cap);
                    if (ms0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_M2L); fail = 1; }
                }
                if (tp) {
                    ps0 = pool_take(pchunk, tp, POOL_CHUNK_P, &C->p2p_pool_used, W.p2p_pool_cap);
                    if (ps0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_P2P); fail = 1; }
                }
                if (ts) {
                    ss0 = pool_take(schunk, ts, POOL_CHUNK_S, &C->sl_used, W.sl_cap);
                    if (ss0 == 0xffffffffu) { atomicOr(&C->overflow, (uint32_t)OVF_SL); fail = 1; }
                }
                if (!fail) { acc_m += tm; acc_p += tp; acc_s += ts; acc_x += tx; }
                bcast[1] = ms0; bcast[2] = ps0; bcast[3] = ss0; bcast[4] = fail;
            }
            __syncthreads();
            const uint32_t fail = bcast[4];
            // segment starts of lane t's target: exclusive prefix over lanes
            // 0..7; a row's M2L segment holds its same-level sources, then the others
            uint32_t em = fail ? 0u : Tm + Tx, ep = fail ? 0u : Tp, es = fail ? 0u : Ts;
            const uint32_t nm_l = em, np_l = ep, ns_l = es;
#pragma unroll
            for (int o = 1; o < MAXT; o <<= 1) {
                const uint32_t ym = __shfl_up(em, o, WAVE), yp = __shfl_up(ep, o, WAVE), ys = __shfl_up(es, o, WAVE);
                if (lane >= o) { em += ym; ep += yp; es += ys; }
            }
            const uint32_t m0 = bcast[1] + em - nm_l, p0 = bcast[2] + ep - np_l, s0 = bcast[3] + es - ns_l;
            if (wv == 0 && lane < nt) {
                const uint32_t r = row0 + lane;
                W.row_node[r] = A0 + lane;
                W.row_next[r] = NO_ROW;
                W.row_m_start[r] = m0;
                W.row_m_cnt[r] = nm_l;
                W.row_mx_cnt[r] = fail ? 0u : Tx;
                W.row_p_start[r] = p0;
                W.row_p_cnt[r] = np_l;
                W.row_sl_start[r] = s0;
                W.row_sl_cnt[r] = ns_l;
                if (nch > 0) W.row_first[A0 + lane] = r;
            }
            if (nch == 0 && threadIdx.x == 0) W.row_next[rg] = row0;
            if (fail || !(tm | tp | ts)) continue;  // uniform; the host retries after a failure
            // pass 2: write this wavefront's entries
            uint32_t c0m = m0 + o0m, c0x = m0 + Tm + o0x, c0p = p0 + o0p, c0s = s0 + o0s;
            uint32_t c1m = m0 + o1m, c1x = m0 + Tm + o1x, c1p = p0 + o1p, c1s = s0 + o1s;
            auto write_entries = [&](int ph, int B, int t, int cat) {
                    const uint64_t bm = __ballot(cat == 0), bp = __ballot(cat == 1), bs = __ballot(cat == 2);
                    const uint64_t bx = __ballot(cat == 4);
                    const uint32_t cm = __builtin_amdgcn_readlane(ph == 0 ? c0m : c1m, t);
                    const uint32_t cx = __builtin_amdgcn_readlane(ph == 0 ? c0x : c1x, t);
                    const uint32_t cp = __builtin_amdgcn_readlane(ph == 0 ? c0p : c1p, t);
                    const uint32_t cs = __builtin_amdgcn_readlane(ph == 0 ? c0s : c1s, t);
                    if (cat == 0) W.m2l_pool[cm + __popcll(bm & lt)] = B;
                    if (cat == 4) W.m2l_pool[cx + __popcll(bx & lt)] = B;
                    if (cat == 1) W.p2p_pool[cp + __popcll(bp & lt)] = B;
                    if (cat == 2) W.sl[cs + __popcll(bs & lt)] = B;
                    if (lane == t) {
                        if (ph == 0) {
                            c0m += __popcll(bm); c0x += __popcll(bx); c0p += __popcll(bp); c0s += __popcll(bs);
                        } else {
                            c1m += __popcll(bm); c1x += __popcll(bx); c1p += __popcll(bp); c1s += __popcll(bs);
                        }
                    }
            };
            if (!solo && ncall <= CAT_CALLS) {
                // the categories of pass 1, without the children's geometry or the MAC
                int k = 0;
                if (lo < hi)
                    enumerate_slice_ids(cache, lo, hi, [&](int ph, bool valid, int B) {
                        const uint32_t word = cat_s[wv][k][lane];
                        ++k;
#pragma unroll
                        for (int t = 0; t < MAXT; ++t)
                            if (t < nt) write_entries(ph, B, t, valid ? (int)((word >> (3 * t)) & 7u) : 3);
                    });
            } else {
                each([&](int ph, bool valid, int B, int4 gb) {
                    classify_all(valid, B, gb, [&](int t, int cat) { write_entries(ph, B, t, cat); });
                });
            }
        }
    }
    if (threadIdx.x == 0) {
        if (acc_m) atomicAdd(&C->m2l_total, acc_m);
        if (acc_p) atomicAdd(&C->p2p_total, acc_p);
        if (acc_s) atomicAdd(&C->split_total, acc_s);
        if (acc_x) atomicAdd(&C->m2lx_total, acc_x);
    }
}

void launch_walk_root(const WalkParams& W, hipStream_t st) {
    hipLaunchKernelGGL(walk_root_kernel, dim3(1), dim3(64), 0, st, W);
}
uint32_t walk_group_resident_blocks(int num_cu) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&walk_group_kernel),
                                                     WAVE * WALK_WAVES, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)(num_cu > 0 ? num_cu : 1);
}
void launch_walk_group(const WalkParams& W, int s, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(walk_group_kernel, dim3(grid), dim3(WAVE * WALK_WAVES), 0, st, W, s);
}

}  // namespace nbh
