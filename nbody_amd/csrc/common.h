// common.h -- shared device types and helpers of libnbody_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbh {

constexpr int WAVE = 64;
constexpr int MAX_LEVELS = 22;   // levels 0..21

// Node records (canonical BFS order: level, system, Morton prefix).
//   geo  = {ix, iy, iz, level | nchild << 8}   integer cell coords at `level`
//   link = {begin, count, first_child, parent} particle range [begin, begin+count)
struct NodeView {
    int4* geo;
    int4* link;
    int* sys;
};

__device__ __forceinline__ int geo_level(int4 g) { return g.w & 0xff; }
__device__ __forceinline__ int geo_nchild(int4 g) { return g.w >> 8; }

// Device-side counters of one force evaluation.  Written by kernels, read by
// the host once after the tree build and once after the walk's first phase.
struct Counters {
    uint32_t num_nodes;
    uint32_t num_leaves;
    uint32_t depth;
    uint32_t overflow;                 // bit flags, see OVF_*
    uint32_t level_count[MAX_LEVELS + 1];
    uint32_t level_off[MAX_LEVELS + 2];
    uint32_t carried[MAX_LEVELS + 2];  // walk: carried leaf rows per depth
    uint32_t sl_used;                  // walk: split-list entries used
    uint32_t rows_used;
    uint32_t m2l_pool_used;            // walk: phase-1 segment pools used
    uint32_t p2p_pool_used;
    unsigned long long m2l_total, p2p_total, p2p_inter, split_total;
    unsigned long long m2lx_total;     // walk: M2L entries between different levels
    uint32_t far_blocks, far_pairs;    // M2L: deferred pairs (level gap >= 2)
    uint32_t nonfinite;
    uint32_t outside;                  // fixed root box: a particle lies outside it
};

// OVF_GEOM: a same-level M2L block found a pair count or a source position
// that differs from its list (an internal consistency check)
// OVF_FAR: the deferred-pair buffers were too small (captured steps size
// them from the previous step)
enum : uint32_t { OVF_NODES = 1, OVF_SL = 2, OVF_ROWS = 4, OVF_M2L = 8, OVF_P2P = 16, OVF_GEOM = 32, OVF_FAR = 64 };
constexpr uint32_t OVF_WALK = OVF_NODES | OVF_SL | OVF_ROWS | OVF_M2L | OVF_P2P;

// A count argument equal to DEV_COUNT means "read the count from the device
// counters": the launch shapes of a captured step do not depend on values
// read back by the host, and the kernels loop over what the tree and the
// walk actually produced.
constexpr uint32_t DEV_COUNT = 0xffffffffu;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// LDS writes of one lane made visible to the other lanes of its wavefront
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Values per lane after s halvings of n (rounding up)
constexpr int rs_len(int n, int s) {
    for (; s > 0; --s) n = (n + 1) / 2;
    return n;
}

// Butterfly reduce-scatter of v[0, N) over each aligned group of 2^S lanes:
// at the step over lane bit b the pair (l, l ^ 2^b) splits the values in two
// halves, each lane keeps one half and adds its partner's copy of it.  After
// the S steps v[0, rs_len(N, S)) of a lane holds group sums; rs_slot names
// which.  It takes sum(rs_len(N, 1..S)) shuffles where a sum per value
// (wave_sum) takes N S.
template <int N, int S>
__device__ __forceinline__ void group_reduce_scatter(float* v) {
    const int lane = __lane_id();
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int b = S - 1 - t;
        const int n = rs_len(N, t), h = (n + 1) / 2;
        const bool up = (lane >> b) & 1;
#pragma unroll
        for (int i = 0; i < h; ++i) {
            const float lo = v[i], hi = h + i < n ? v[h + i] : 0.f;
            v[i] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, 1 << b, WAVE);
        }
    }
}

// After group_reduce_scatter<N, S>: v[j] of this lane, j < cnt, is the group
// sum of value start + j (cnt may be 0).
__device__ __forceinline__ void rs_slot(int N, int S, int lane, int& start, int& cnt) {
    int n = N;
    start = 0;
    cnt = N;
    for (int b = S - 1; b >= 0; --b) {
        const int h = (n + 1) / 2;
        if ((lane >> b) & 1) {
            start += h;
            cnt = cnt > h ? cnt - h : 0;
        } else {
            cnt = cnt < h ? cnt : h;
        }
        n = h;
    }
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, WAVE);
        v = w > v ? w : v;
    }
    return v;
}

// centre difference c_B - c_A of two cells of one system, FP32
__device__ __forceinline__ float3 centre_diff(int4 ga, int4 gb, int bits, float unit) {
    const int sa = bits - (ga.w & 0xff), sb = bits - (gb.w & 0xff);
    const int dx = ((2 * gb.x + 1) << sb) - ((2 * ga.x + 1) << sa);
    const int dy = ((2 * gb.y + 1) << sb) - ((2 * ga.y + 1) << sa);
    const int dz = ((2 * gb.z + 1) << sb) - ((2 * ga.z + 1) << sa);
    return make_float3((float)dx * unit, (float)dy * unit, (float)dz * unit);
}

// FP64 centre of a cell
__device__ __forceinline__ void centre64(int4 g, double4 box, int bits, double c[3]) {
    const int l = g.w & 0xff;
    const double h = ldexp(box.w, -l);
    c[0] = box.x + ((double)g.x + 0.5) * h;
    c[1] = box.y + ((double)g.y + 0.5) * h;
    c[2] = box.z + ((double)g.z + 0.5) * h;
}

__host__ __device__ __forceinline__ unsigned ceil_div(unsigned a, unsigned b) { return (a + b - 1) / b; }

}  // namespace nbh
