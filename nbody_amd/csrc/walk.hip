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

// Enumerate expand(SL) for the split list at sl[st, st+cnt); f(valid, B, geo)
// is called once per chunk of 64 lanes (uniformly), in ascending candidate
// order: first the entries without children (kept as themselves), then the
// children of the others, entry by entry.  `filled`: the cache already holds
// the whole list (cnt <= SL_CACHE).
template <typename F>
__device__ __forceinline__ void enumerate_sl(const WalkParams& W, const SLCache& cache, uint32_t st, uint32_t cnt,
                                             bool filled, F&& f) {
    const int lane = lane_id();
    const bool one_block = cnt <= (uint32_t)SL_CACHE;
    // phase a: sources kept as themselves (childless)
    for (uint32_t p0 = 0; p0 < cnt; p0 += SL_CACHE) {
        const uint32_t pn = min((uint32_t)SL_CACHE, cnt - p0);
        if (!(filled && one_block)) sl_cache_fill(W, cache, st, p0, pn);
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
            f(valid, B, gb);
        }
        if (!one_block) wave_lds_sync();
    }
    // phase b: children of the sources that have children, 8 entries x 8
    // slots per chunk, two chunks in flight
    for (uint32_t p0 = 0; p0 < cnt; p0 += SL_CACHE) {
        const uint32_t pn = min((uint32_t)SL_CACHE, cnt - p0);
        if (!one_block) sl_cache_fill(W, cache, st, p0, pn);
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
            if (__ballot(v0)) f(v0, C0, g0);
            if (__ballot(v1)) f(v1, C1, g1);
        }
        if (!one_block) wave_lds_sync();
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

// Depth s >= 1: one wavefront per group (row of depth s-1).
__global__ __launch_bounds__(WAVE * WALK_WAVES) void walk_group_kernel(WalkParams W, int s) {
    __shared__ int4 cgeo[WALK_WAVES][SL_CACHE];
    __shared__ int cnode[WALK_WAVES][SL_CACHE];
    __shared__ int cfc[WALK_WAVES][SL_CACHE];
    Counters* C = W.C;
    const uint32_t n_prev_level = C->level_count[s - 1];
    const uint32_t n_groups = n_prev_level + C->carried[s - 1];
    const uint32_t base_prev = depth_base(C, s - 1);
    const uint32_t base = base_prev + n_groups;
    const uint32_t n_level = C->level_count[s];
    const uint32_t loff = C->level_off[s], loff_prev = C->level_off[s - 1];
    const int lane = lane_id();
    const int wslot = threadIdx.x / WAVE;
    const SLCache cache{cgeo[wslot], cnode[wslot], cfc[wslot]};
    const uint32_t wid = blockIdx.x * WALK_WAVES + wslot;
    const uint32_t nw = gridDim.x * WALK_WAVES;
    const uint64_t lt = lanemask_lt();
    const int bits = W.bits;
    const uint64_t num4 = W.num4, den3 = W.den3;
    uint2 mchunk = make_uint2(0, 0), pchunk = make_uint2(0, 0), schunk = make_uint2(0, 0);  // lane 0 only
    unsigned long long acc_m = 0, acc_p = 0, acc_s = 0, acc_x = 0;                           // lane 0 only
    for (uint32_t g = wid; g < n_groups; g += nw) {
        const uint32_t rg = base_prev + g;
        if (rg >= W.row_cap) continue;  // flagged when the row was made
        const int P = g < n_prev_level ? (int)(loff_prev + g) : W.row_node[rg];
        const int4 gp = W.nodes.geo[P];
        const int nch = gp.w >> 8;
        const uint32_t sl_st = W.row_sl_start[rg];
        uint32_t sl_n = W.row_sl_cnt[rg];
        if ((unsigned long long)sl_st + sl_n > W.sl_cap) sl_n = 0;  // only after an overflow
        int nt, A0;
        uint32_t row0;
        if (nch > 0) {  // the cell's children: level-s rows in id order
            nt = nch;
            A0 = W.nodes.link[P].z;
            row0 = base + (uint32_t)(A0 - (int)loff);
        } else {  // a leaf with sources left to split: carried to depth s
            if (sl_n == 0) continue;
            nt = 1;
            A0 = P;
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&C->carried[s], 1u);
            k = __shfl(k, 0, WAVE);
            row0 = base + n_level + k;
        }
        if ((unsigned long long)row0 + nt > W.row_cap) {
            // the host retries with more rows; rows that fit are left empty so
            // that the next depth reads only rows of this evaluation
            if (lane == 0) atomicOr(&C->overflow, (uint32_t)OVF_ROWS);
            const uint32_t r = row0 + lane;
            if (lane < nt && r < W.row_cap) {
                W.row_node[r] = A0 + lane;
                W.row_next[r] = NO_ROW;
                W.row_m_start[r] = W.row_m_cnt[r] = W.row_mx_cnt[r] = 0;
                W.row_p_start[r] = W.row_p_cnt[r] = 0;
                W.row_sl_start[r] = W.row_sl_cnt[r] = 0;
            }
            continue;
        }
        // targets (all of one level): lane t < nt holds target t
        int4 gt_l = make_int4(0, 0, 0, 0);
        if (lane < nt) gt_l = W.nodes.geo[A0 + lane];
        const int la = gp.w & 0xff;
        const int lt_level = nch > 0 ? la + 1 : la;
        const int sa = bits - lt_level;
        const int side_a = 1 << (sa + 1);
        // per-target state lives in lane t (centres, leaf flag, counters,
        // cursors) and is read with readlane where the loop over targets
        // needs it, which keeps it out of the scalar registers
        const int cax_l = (2 * gt_l.x + 1) << sa, cay_l = (2 * gt_l.y + 1) << sa, caz_l = (2 * gt_l.z + 1) << sa;
        const int leaf_l = (gt_l.w >> 8) == 0;
        const bool filled = sl_n <= (uint32_t)SL_CACHE;
        if (filled && sl_n) sl_cache_fill(W, cache, sl_st, 0, sl_n);
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
                    const bool approx = B != A0 + t && lhs < num4 * dsq;
                    const bool both_leaf = __builtin_amdgcn_readlane(leaf_l, t) && b_leaf;
                    const int cat = !valid ? 3 : approx ? (lb == lt_level ? 0 : 4) : both_leaf ? 1 : 2;
                    per_target(t, cat);
                }
            }
        };
        // pass 1: counts per target (lane t)
        uint32_t nm_l = 0, nx_l = 0, np_l = 0, ns_l = 0;
        if (sl_n) {
            enumerate_sl(W, cache, sl_st, sl_n, filled, [&](bool valid, int B, int4 gb) {
                classify_all(valid, B, gb, [&](int t, int cat) {
                    const uint32_t cm = __popcll(__ballot(cat == 0)), cx = __popcll(__ballot(cat == 4));
                    const uint32_t cp = __popcll(__ballot(cat == 1)), cs = __popcll(__ballot(cat == 2));
                    if (lane == t) {
                        nm_l += cm;
                        nx_l += cx;
                        np_l += cp;
                        ns_l += cs;
                    }
                });
            });
        }
        const uint32_t tm = wave_sum(nm_l + nx_l), tp = wave_sum(np_l), ts = wave_sum(ns_l);
        const uint32_t tx = wave_sum(nx_l);
        // reserve one block per pool for all targets of the group
        uint32_t ms0 = 0, ps0 = 0, ss0 = 0, fail = 0;
        if (lane == 0) {
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
            if (!fail) {
                acc_m += tm; acc_p += tp; acc_s += ts; acc_x += tx;
            }
        }
        fail = __shfl(fail, 0, WAVE);
        ms0 = __shfl(ms0, 0, WAVE);
        ps0 = __shfl(ps0, 0, WAVE);
        ss0 = __shfl(ss0, 0, WAVE);
        if (fail) {  // the host retries with larger pools; leave empty rows
            nm_l = nx_l = np_l = ns_l = 0;
        }
        // segment starts of lane t's target: exclusive prefix over the
        // targets (lanes 0..7); a row's M2L segment holds its same-level
        // sources, then the others
        uint32_t em = nm_l + nx_l, ep = np_l, es = ns_l;
#pragma unroll
        for (int o = 1; o < MAXT; o <<= 1) {
            const uint32_t ym = __shfl_up(em, o, WAVE), yp = __shfl_up(ep, o, WAVE), ys = __shfl_up(es, o, WAVE);
            if (lane >= o) { em += ym; ep += yp; es += ys; }
        }
        uint32_t cm_l = ms0 + em - (nm_l + nx_l), cp_l = ps0 + ep - np_l, cs_l = ss0 + es - ns_l;
        uint32_t cx_l = cm_l + nm_l;
        if (lane < nt) {
            const uint32_t r = row0 + lane;
            W.row_node[r] = A0 + lane;
            W.row_next[r] = NO_ROW;
            W.row_m_start[r] = cm_l;
            W.row_m_cnt[r] = nm_l + nx_l;
            W.row_mx_cnt[r] = nx_l;
            W.row_p_start[r] = cp_l;
            W.row_p_cnt[r] = np_l;
            W.row_sl_start[r] = cs_l;
            W.row_sl_cnt[r] = ns_l;
            if (nch > 0) W.row_first[A0 + lane] = r;
        }
        if (nch == 0 && lane == 0) W.row_next[rg] = row0;
        // pass 2: write every target's segments
        if (!fail && (tm | tp | ts)) {
            enumerate_sl(W, cache, sl_st, sl_n, filled, [&](bool valid, int B, int4 gb) {
                classify_all(valid, B, gb, [&](int t, int cat) {
                    const uint64_t bm = __ballot(cat == 0), bp = __ballot(cat == 1), bs = __ballot(cat == 2);
                    const uint64_t bx = __ballot(cat == 4);
                    if (cat == 0) W.m2l_pool[__builtin_amdgcn_readlane(cm_l, t) + __popcll(bm & lt)] = B;
                    if (cat == 4) W.m2l_pool[__builtin_amdgcn_readlane(cx_l, t) + __popcll(bx & lt)] = B;
                    if (cat == 1) W.p2p_pool[__builtin_amdgcn_readlane(cp_l, t) + __popcll(bp & lt)] = B;
                    if (cat == 2) W.sl[__builtin_amdgcn_readlane(cs_l, t) + __popcll(bs & lt)] = B;
                    if (lane == t) {
                        cm_l += __popcll(bm);
                        cx_l += __popcll(bx);
                        cp_l += __popcll(bp);
                        cs_l += __popcll(bs);
                    }
                });
            });
        }
        wave_lds_sync();  // the cache is refilled by the next group
    }
    if (lane == 0) {
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
