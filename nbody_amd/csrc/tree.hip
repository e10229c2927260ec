// tree.hip -- on-device octree over 63-bit Morton keys.
//
// Replaces the host glade::Orthtree build and per-step move
// (open_cl_simulation.cpp:43-47, :108-124) and the per-step re-upload of the
// tree (computeOctreeBuffers, :126-135).  The octree is a pure function of
// the sorted keys, the leaf capacity C and the depth limit D: a cell splits
// iff it holds more than C particles and its level is below D.
//
// Particle-parallel emission.  For sorted keys k[0..n) of one system, let
//   h(i)      = 0 for the first particle, else 1 + (#leading equal levels of
//               k[i-1], k[i])  -- the first level at which i starts a cell;
//   lam(j)    = #leading equal levels of k[j], k[j+C] (window of C+1 keys);
//   Lam(i)    = max lam(j) over windows [j, j+C] that contain i (-1 if none);
//   L(i)      = min(D, Lam(i)+1)  -- the level of i's leaf.
// The level-l cell containing i holds > C particles iff Lam(i) >= l, so node
// (i, l) exists iff h(i) <= l <= L(i).  Nodes of one level, in particle
// order, are that level's cells in Morton order: the BFS index of (i, l) is
// level_off[l] + #{j < i : h(j) <= l <= L(j)}, computed with block counts,
// one scan, and ballot ranks.
#include "common.h"
#include "launch.h"

namespace nbh {

__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    uint64_t x = v & 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t compact3(uint64_t v) {
    uint64_t x = v & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}

__device__ __forceinline__ int sys_of(const uint32_t* sys_off, int nsys, uint32_t i) {
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sys_off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Bounding cube per system (FP64 min/max are exact, so the result does not
// depend on the reduction order).
constexpr int BB_THREADS = 256;
constexpr int BB_TILE = 4096;

__global__ __launch_bounds__(BB_THREADS) void bbox_partial_kernel(const double* __restrict__ x,
                                                                  const double* __restrict__ y,
                                                                  const double* __restrict__ z,
                                                                  const uint32_t* sys_off, int chunks,
                                                                  double* part, uint32_t* nonfinite) {
    const int s = blockIdx.y, c = blockIdx.x;
    const uint32_t b = sys_off[s], e = sys_off[s + 1];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    for (uint32_t i = b + c * BB_TILE + threadIdx.x; i < e && i < b + (c + 1) * (uint32_t)BB_TILE; i += BB_THREADS) {
        double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            bad |= !isfinite(v[a]);
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
    if (bad) atomicOr(nonfinite, 1u);
    __shared__ double sh[6][BB_THREADS / WAVE];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double l = lo[a], h = hi[a];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            l = fmin(l, __shfl_xor(l, o, WAVE));
            h = fmax(h, __shfl_xor(h, o, WAVE));
        }
        if (lane_id() == 0) { sh[a][threadIdx.x / WAVE] = l; sh[3 + a][threadIdx.x / WAVE] = h; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double r = threadIdx.x < 3 ? INFINITY : -INFINITY;
        for (int w = 0; w < BB_THREADS / WAVE; ++w)
            r = threadIdx.x < 3 ? fmin(r, sh[threadIdx.x][w]) : fmax(r, sh[threadIdx.x][w]);
        part[((size_t)s * chunks + c) * 6 + threadIdx.x] = r;
    }
}

// one wavefront per system: fold its partials, write box = {lo, size} and unit.
__global__ __launch_bounds__(WAVE) void bbox_final_kernel(const double* part, int chunks, int nsys,
                                                          const uint32_t* sys_off, int bits, int fixed, double flo0,
                                                          double flo1, double flo2, double fsize, double4* box,
                                                          double* inv, float* unit) {
    const int s = blockIdx.x;
    if (s >= nsys) return;
    double lo[3], size;
    if (fixed) {
        lo[0] = flo0; lo[1] = flo1; lo[2] = flo2;
        size = fsize;
    } else {
        const int nch = (int)ceil_div(sys_off[s + 1] - sys_off[s], BB_TILE);
        double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = lane_id(); c < nch; c += WAVE)
            for (int a = 0; a < 3; ++a) {
                l[a] = fmin(l[a], part[((size_t)s * chunks + c) * 6 + a]);
                h[a] = fmax(h[a], part[((size_t)s * chunks + c) * 6 + 3 + a]);
            }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                l[a] = fmin(l[a], __shfl_xor(l[a], o, WAVE));
                h[a] = fmax(h[a], __shfl_xor(h[a], o, WAVE));
            }
        }
        size = fmax(h[0] - l[0], fmax(h[1] - l[1], h[2] - l[2]));
        if (!(size > 0.0)) size = 1.0;
        for (int a = 0; a < 3; ++a) lo[a] = l[a];
    }
    if (lane_id() == 0) {
        box[s] = make_double4(lo[0], lo[1], lo[2], size);
        inv[s] = ldexp(1.0, bits) / size;
        unit[s] = (float)ldexp(size, -(bits + 1));  // half of the finest cell
    }
}

// key = sys << 3B | interleave(qx, qy, qz), q = clamp(floor((x - lo) * 2^B / size))
// With a fixed root box (flags != nullptr) the bounding-cube pass does not
// run, so this kernel reports what it would have caught: a non-finite
// coordinate (flags[0]) and a particle outside the closed box [lo, lo + size]
// (flags[1]), which the clamp would otherwise file into a boundary cell whose
// centre-relative offsets and opening criterion no longer hold for it.  A
// coordinate equal to lo + size (the reference's inputs may reach the upper
// bound, main.cpp:118-122) stays inside and lands in the last cell.
__global__ void keys_kernel(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, uint32_t n, const uint32_t* sys_off, int nsys,
                            int bits, const double4* box, const double* inv, uint64_t* keys, uint32_t* perm,
                            uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = nsys == 1 ? 0 : sys_of(sys_off, nsys, i);
    const double4 b = box[s];
    const double iv = inv[s];
    const double qn = ldexp(1.0, bits), qmax = qn - 1.0;
    const double v[3] = {x[i], y[i], z[i]};
    const double lo[3] = {b.x, b.y, b.z};
    uint64_t q[3];
    bool bad = false, out = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double t = (v[a] - lo[a]) * iv;
        bad |= !isfinite(v[a]);
        out |= !(t >= 0.0 && t <= qn);
        double f = floor(t);
        f = fmin(fmax(f, 0.0), qmax);
        q[a] = (uint64_t)f;
    }
    if (flags && (bad || out)) atomicOr(&flags[bad ? 0 : 1], 1u);
    keys[i] = ((uint64_t)s << (3 * bits)) | (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    perm[i] = i;
}

// ---------------------------------------------------------------------------
constexpr int TB_THREADS = 256;

__device__ __forceinline__ int common_levels(uint64_t a, uint64_t b, int bits) {
    const uint64_t mask = (bits >= 21) ? 0x7fffffffffffffffull : ((1ull << (3 * bits)) - 1);
    if ((a >> (3 * bits)) != (b >> (3 * bits))) return -1;
    const uint64_t x = (a ^ b) & mask;
    if (x == 0) return bits;
    const int lz = __clzll(x) - (64 - 3 * bits);
    return lz / 3;
}

struct TreeParams {
    const uint64_t* keys;
    uint32_t n;
    const uint32_t* sys_off;
    int nsys, bits, cap, max_depth;
};

// h(i) and L(i) for the particles of this block; lam staged in LDS with a
// halo of C entries to the left.
__device__ void compute_hl(const TreeParams& P, uint32_t i, int* lam_sh, int& h, int& L) {
    const uint32_t t0 = blockIdx.x * TB_THREADS;
    // fill lam for j in [t0 - C, t0 + TB_THREADS)
    const int halo = P.cap;
    for (int t = threadIdx.x; t < TB_THREADS + halo; t += TB_THREADS) {
        long long j = (long long)t0 - halo + t;
        int v = -1;
        if (j >= 0 && j + P.cap < (long long)P.n) {
            v = common_levels(P.keys[j], P.keys[j + P.cap], P.bits);  // -1 across systems
        }
        lam_sh[t] = v;
    }
    __syncthreads();
    h = 0;
    L = 0;
    if (i < P.n) {
        const uint64_t k = P.keys[i];
        int prev = -1;
        if (i > 0) prev = common_levels(P.keys[i - 1], k, P.bits);
        h = prev < 0 ? 0 : prev + 1;  // new system (or first particle) starts the root
        int Lam = -1;
        const int base = (int)(i - t0) + halo;  // lam_sh index of j = i
        for (int d = 0; d <= P.cap; ++d) Lam = max(Lam, lam_sh[base - d]);
        L = min(P.max_depth, Lam + 1);
    }
}

__global__ __launch_bounds__(TB_THREADS) void tree_count_kernel(TreeParams P, uint32_t* cnt, uint32_t nblk,
                                                                uint16_t* hl) {
    extern __shared__ int lam_sh[];
    __shared__ uint32_t lc[MAX_LEVELS + 2];
    if (threadIdx.x < MAX_LEVELS + 2) lc[threadIdx.x] = 0;
    const uint32_t i = blockIdx.x * TB_THREADS + threadIdx.x;
    int h, L;
    compute_hl(P, i, lam_sh, h, L);
    if (i < P.n) hl[i] = (uint16_t)(h | (L << 8));
    const bool valid = i < P.n;
    for (int l = 0; l <= P.max_depth + 1; ++l) {
        bool pred = l <= P.max_depth ? (valid && h <= l && l <= L) : (valid && h <= L);
        uint64_t m = __ballot(pred);
        if (lane_id() == 0 && m) atomicAdd(&lc[l], (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x <= P.max_depth + 1) cnt[(size_t)threadIdx.x * nblk + blockIdx.x] = lc[threadIdx.x];
}

// level offsets / counts from the scanned block counts
__global__ void tree_counts_final_kernel(const uint32_t* off, const uint32_t* total, uint32_t nblk, int max_depth,
                                         Counters* C) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t depth = 0;
    for (int l = 0; l <= max_depth + 1; ++l) C->level_off[l] = off[(size_t)l * nblk];
    for (int l = 0; l <= max_depth; ++l) {
        C->level_count[l] = C->level_off[l + 1] - C->level_off[l];
        if (C->level_count[l]) depth = l;
    }
    for (int l = max_depth + 1; l <= MAX_LEVELS; ++l) C->level_count[l] = 0;
    C->num_nodes = C->level_off[max_depth + 1];
    C->num_leaves = *total - C->num_nodes;
    C->depth = depth;
}

__global__ __launch_bounds__(TB_THREADS) void tree_emit_kernel(TreeParams P, const uint32_t* off, uint32_t nblk,
                                                               const uint16_t* hl, NodeView nodes,
                                                               uint32_t* node_begin, uint32_t* leaf_ids,
                                                               uint32_t node_cap, uint32_t* overflow) {
    __shared__ uint32_t wt[TB_THREADS / WAVE];
    const uint32_t i = blockIdx.x * TB_THREADS + threadIdx.x;
    const bool valid = i < P.n;
    int h = 0, L = -1;
    uint64_t k = 0;
    int s = 0;
    if (valid) {
        h = hl[i] & 0xff;
        L = hl[i] >> 8;
        k = P.keys[i];
        s = (int)(k >> (3 * P.bits));
    }
    const int wave = threadIdx.x / WAVE;
    const uint64_t lt = lanemask_lt();
    const uint64_t mmask = (P.bits >= 21) ? 0x7fffffffffffffffull : ((1ull << (3 * P.bits)) - 1);
    uint32_t leaf_node = 0;
    for (int l = 0; l <= P.max_depth + 1; ++l) {
        const bool pred = l <= P.max_depth ? (valid && h <= l && l <= L) : (valid && h <= L);
        const uint64_t m = __ballot(pred);
        if (lane_id() == 0) wt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t rank = (uint32_t)__popcll(m & lt);
        for (int w = 0; w < wave; ++w) rank += wt[w];
        const uint32_t id = off[(size_t)l * nblk + blockIdx.x] + rank;
        if (pred) {
            if (l <= P.max_depth) {
                if (id < node_cap) {
                    const uint64_t prefix = (k & mmask) >> (3 * (P.bits - l));
                    const bool leaf = (l == L);
                    nodes.geo[id] = make_int4((int)compact3(prefix >> 2), (int)compact3(prefix >> 1),
                                              (int)compact3(prefix), l | (leaf ? 0 : (0xff << 8)));
                    nodes.sys[id] = s;
                    node_begin[id] = i;
                    if (leaf) leaf_node = id;
                } else {
                    atomicOr(overflow, (uint32_t)OVF_NODES);
                }
            } else {
                const uint32_t slot = id - off[(size_t)l * nblk];  // leaves in particle order
                leaf_ids[slot] = leaf_node;
            }
        }
        __syncthreads();
    }
}

// Per node: particle count (end of the key prefix), parent (last node of the
// previous level with begin <= mine) and children (next-level nodes whose
// begin lies in [begin, end)).
__global__ void tree_link_kernel(TreeParams P, const Counters* C, const uint32_t* node_begin, NodeView nodes) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    // past the node arrays (a captured step sized them from the previous step)
    if (id >= C->num_nodes || (C->overflow & OVF_NODES)) return;
    int4 g = nodes.geo[id];
    const int l = g.w & 0xff;
    const bool leaf = (g.w >> 8) == 0;
    const uint32_t b = node_begin[id];
    const int s = nodes.sys[id];
    const uint32_t se = P.sys_off[s + 1];
    // end: first j in (b, se) whose level-l prefix differs
    uint32_t end;
    if (l == 0) {
        end = se;
    } else {
        const int shift = 3 * (P.bits - l);
        const uint64_t pk = P.keys[b] >> shift;
        uint32_t lo = b + 1, hi = se;  // answer in [lo, hi]
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if ((P.keys[mid] >> shift) == pk) lo = mid + 1; else hi = mid;
        }
        end = lo;
    }
    int parent = -1;
    if (l > 0) {
        uint32_t lo = C->level_off[l - 1], hi = C->level_off[l] - 1;  // last with begin <= b
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo + 1) / 2;
            if (node_begin[mid] <= b) lo = mid; else hi = mid - 1;
        }
        parent = (int)lo;
    }
    int first = -1, nch = 0;
    if (!leaf) {
        uint32_t lo = C->level_off[l + 1], hi = C->level_off[l + 2];
        // first child has begin == b
        uint32_t a = lo, z = hi;
        while (a < z) { uint32_t mid = a + (z - a) / 2; if (node_begin[mid] < b) a = mid + 1; else z = mid; }
        first = (int)a;
        uint32_t c = a;
        z = hi;
        while (c < z) { uint32_t mid = c + (z - c) / 2; if (node_begin[mid] < end) c = mid + 1; else z = mid; }
        nch = (int)(c - a);
    }
    nodes.link[id] = make_int4((int)b, (int)(end - b), first, parent);
    g.w = l | (nch << 8);
    nodes.geo[id] = g;
}

// ---------------------------------------------------------------------------
void launch_bbox(const double* x, const double* y, const double* z, const uint32_t* sys_off, int nsys,
                 uint32_t max_sys_n, int bits, int fixed, const double* flo, double fsize, double* part,
                 double4* box, double* inv, float* unit, uint32_t* nonfinite, hipStream_t st) {
    const int chunks = (int)ceil_div(max_sys_n ? max_sys_n : 1, BB_TILE);
    if (!fixed)
        hipLaunchKernelGGL(bbox_partial_kernel, dim3(chunks, nsys), dim3(BB_THREADS), 0, st, x, y, z, sys_off,
                           chunks, part, nonfinite);
    hipLaunchKernelGGL(bbox_final_kernel, dim3(nsys), dim3(WAVE), 0, st, part, chunks, nsys, sys_off,
                       bits, fixed, flo[0], flo[1], flo[2], fsize, box, inv, unit);
}

void launch_keys(const double* x, const double* y, const double* z, uint32_t n, const uint32_t* sys_off,
                 int nsys, int bits, const double4* box, const double* inv, uint64_t* keys, uint32_t* perm,
                 uint32_t* fixed_flags, hipStream_t st) {
    hipLaunchKernelGGL(keys_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, x, y, z, n, sys_off, nsys, bits,
                       box, inv, keys, perm, fixed_flags);
}

size_t tree_tmp_bytes(uint32_t n, int max_depth) {
    const uint32_t nblk = ceil_div(n ? n : 1, TB_THREADS);
    const size_t cnt = (size_t)(max_depth + 2) * nblk;
    return cnt * 4 * 2 + scan_tmp_bytes((uint32_t)cnt) + (size_t)n * 2 + 1024;
}

void launch_tree_build(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                       int max_depth, NodeView nodes, uint32_t* node_begin, uint32_t* leaf_ids, uint32_t node_cap,
                       Counters* C, void* tmp, hipStream_t st) {
    const uint32_t nblk = ceil_div(n, TB_THREADS);
    const uint32_t rows = (uint32_t)(max_depth + 2) * nblk;
    uint32_t* cnt = (uint32_t*)tmp;
    uint32_t* off = cnt + rows;
    uint32_t* total = off + rows;
    uint16_t* hl = (uint16_t*)(total + 2);
    void* stmp = (void*)(((uintptr_t)(hl + n) + 255) & ~(uintptr_t)255);
    TreeParams P{keys, n, sys_off, nsys, bits, cap, max_depth};
    const size_t lds = (size_t)(TB_THREADS + cap) * sizeof(int);
    hipLaunchKernelGGL(tree_count_kernel, dim3(nblk), dim3(TB_THREADS), lds, st, P, cnt, nblk, hl);
    exclusive_scan<uint32_t, uint32_t>(cnt, off, nullptr, rows, total, stmp, st);
    hipLaunchKernelGGL(tree_counts_final_kernel, dim3(1), dim3(1), 0, st, off, total, nblk, max_depth, C);
}

// second half of the build, after the host has sized the node arrays
void launch_tree_emit(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                      int max_depth, NodeView nodes, uint32_t* node_begin, uint32_t* leaf_ids, uint32_t node_cap,
                      Counters* C, void* tmp, hipStream_t st) {
    const uint32_t nblk = ceil_div(n, TB_THREADS);
    const uint32_t rows = (uint32_t)(max_depth + 2) * nblk;
    uint32_t* off = (uint32_t*)tmp + rows;
    uint32_t* total = off + rows;
    uint16_t* hl = (uint16_t*)(total + 2);
    TreeParams P{keys, n, sys_off, nsys, bits, cap, max_depth};
    hipLaunchKernelGGL(tree_emit_kernel, dim3(nblk), dim3(TB_THREADS), 0, st, P, off, nblk, hl, nodes, node_begin,
                       leaf_ids, node_cap, &C->overflow);
}

void launch_tree_link(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                      int max_depth, const Counters* C, uint32_t num_nodes, const uint32_t* node_begin,
                      NodeView nodes, hipStream_t st) {
    TreeParams P{keys, n, sys_off, nsys, bits, cap, max_depth};
    hipLaunchKernelGGL(tree_link_kernel, dim3(ceil_div(num_nodes ? num_nodes : 1, 256)), dim3(256), 0, st, P, C,
                       node_begin, nodes);
}

}  // namespace nbh
