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

// Enumerate the candidates of one row; f(lane_valid, candidate, geo, category)
// is called once per chunk of 64 lanes, in ascending candidate order.
template <typename F>
__device__ __forceinline__ void enumerate(const WalkParams& W, int A, int4 ga, int src, F&& f) {
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
    for (uint32_t e0 = 0; e0 < cnt; e0 += WAVE) {
        const uint32_t e = e0 + lane;
        int B = -1;
        int4 gb = make_int4(0, 0, 0, 0);
        if (e < cnt) {
            B = W.sl[st + e];
            gb = W.nodes.geo[B];
        }
        const bool valid = B >= 0 && (gb.w >> 8) == 0;
        if (__ballot(valid) == 0) continue;
        f(valid, B, gb, classify(valid, B, gb));
    }
    // phase b: children of sources that have children, 8 entries x 8 slots
    for (uint32_t e0 = 0; e0 < cnt; e0 += 8) {
        const uint32_t e = e0 + (lane >> 3);
        const int c = lane & 7;
        bool valid = false;
        int Cn = -1;
        int4 gc = make_int4(0, 0, 0, 0);
        if (e < cnt) {
            const int B = W.sl[st + e];
            const int4 gb = W.nodes.geo[B];
            const int nch = gb.w >> 8;
            if (c < nch) {
                Cn = W.nodes.link[B].z + c;
                gc = W.nodes.geo[Cn];
                valid = true;
            }
        }
        if (__ballot(valid) == 0) continue;
        f(valid, Cn, gc, classify(valid, Cn, gc));
    }
}

// Phase 1: count M2L/P2P per target, write split lists, carry leaf targets.
__global__ __launch_bounds__(256) void walk_phase1_kernel(WalkParams W, int s) {
    Counters* C = W.C;
    const uint32_t n_level = C->level_count[s];
    const uint32_t n_rows = n_level + C->carried[s];
    const uint32_t base = depth_base(C, s);
    const uint32_t base_prev = s > 0 ? base - (C->level_count[s - 1] + C->carried[s - 1]) : 0;
    const uint32_t base_next = base + n_rows;
    const int lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const uint64_t lt = lanemask_lt();
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
        enumerate(W, A, ga, src, [&](bool valid, int B, int4 gb, int cat) {
            nm += __popcll(__ballot(cat == CAT_M2L));
            np += __popcll(__ballot(cat == CAT_P2P));
            ns += __popcll(__ballot(cat == CAT_SL));
        });
        uint32_t start = 0;
        if (lane == 0) {
            W.m2l_cnt[A] += nm;
            W.p2p_cnt[A] += np;
            if (ns) {
                start = atomicAdd(&C->sl_used, ns);
                if ((unsigned long long)start + ns > W.sl_cap) {
                    atomicOr(&C->overflow, (uint32_t)OVF_SL);
                    ns = 0;
                }
                atomicAdd(&C->split_total, (unsigned long long)ns);
            }
            W.row_sl_start[r] = start;
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
        start = __shfl(start, 0, WAVE);
        ns = __shfl(ns, 0, WAVE);
        if (ns) {
            uint32_t cur = start;
            enumerate(W, A, ga, src, [&](bool valid, int B, int4 gb, int cat) {
                const uint64_t m = __ballot(cat == CAT_SL);
                if (cat == CAT_SL) W.sl[cur + __popcll(m & lt)] = B;
                cur += __popcll(m);
            });
        }
    }
}

// Phase 2: write the M2L/P2P entries at each target's cursor.
__global__ __launch_bounds__(256) void walk_phase2_kernel(WalkParams W, int s) {
    Counters* C = W.C;
    const uint32_t n_level = C->level_count[s];
    const uint32_t n_rows = n_level + C->carried[s];
    const uint32_t base = depth_base(C, s);
    const uint32_t base_prev = s > 0 ? base - (C->level_count[s - 1] + C->carried[s - 1]) : 0;
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / WAVE;
    const uint64_t lt = lanemask_lt();
    for (uint32_t rl = wid; rl < n_rows; rl += nw) {
        const uint32_t r = base + rl;
        if (r >= W.row_cap) continue;  // cannot happen after a clean phase 1
        int A, src;
        if (rl < n_level) {
            A = (int)(C->level_off[s] + rl);
            src = s == 0 ? -1 : (int)(base_prev + (uint32_t)(W.nodes.link[A].w - (int)C->level_off[s - 1]));
        } else {
            A = W.row_node[r];
            src = W.row_src[r];
        }
        const int4 ga = W.nodes.geo[A];
        const int na = W.nodes.link[A].y;
        unsigned long long cm = W.m2l_cur[A], cp = W.p2p_cur[A];
        unsigned long long inter = 0;
        enumerate(W, A, ga, src, [&](bool valid, int B, int4 gb, int cat) {
            const uint64_t mm = __ballot(cat == CAT_M2L);
            const uint64_t mp = __ballot(cat == CAT_P2P);
            if (cat == CAT_M2L) {
                const unsigned long long at = cm + __popcll(mm & lt);
                if (at < W.m2l_cap) W.m2l_src[at] = B;
            }
            if (cat == CAT_P2P) {
                const unsigned long long at = cp + __popcll(mp & lt);
                if (at < W.p2p_cap) W.p2p_src[at] = B;
                const int nb = W.nodes.link[B].y;
                inter += (unsigned long long)na * (unsigned long long)(B == A ? nb - 1 : nb);
            }
            cm += __popcll(mm);
            cp += __popcll(mp);
        });
        inter = wave_sum(inter);
        if (lane_id() == 0) {
            W.m2l_cur[A] = cm;
            W.p2p_cur[A] = cp;
            if (inter) atomicAdd(&C->p2p_inter, inter);
        }
    }
}

void launch_walk_phase1(const WalkParams& W, int s, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(walk_phase1_kernel, dim3(grid), dim3(256), 0, st, W, s);
}
void launch_walk_phase2(const WalkParams& W, int s, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(walk_phase2_kernel, dim3(grid), dim3(256), 0, st, W, s);
}

}  // namespace nbh
