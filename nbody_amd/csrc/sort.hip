// sort.hip -- device-wide exclusive scan and stable LSD radix sort of
// (u64 key, u32 value) pairs, written for 64-lane CDNA wavefronts.
//
// Replaces the host octree's ordering step (glade::Orthtree build,
// open_cl_simulation.cpp:43-47): particles are ordered by 63-bit Morton key
// on the device, with ties kept in their previous order (stable), so the
// octree topology is a pure function of the keys.
#include "common.h"
#include "launch.h"

namespace nbh {

// ---------------------------------------------------------------------------
// Exclusive scan, three phases.  Tiles of SCAN_TILE elements per block.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_THREADS) void scan_reduce_kernel(const TI* in, const uint32_t* n_dev,
                                                                   uint32_t n_host, TO* partial) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    TO s = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        uint32_t i = base + k * SCAN_THREADS + threadIdx.x;
        if (i < n) s += (TO)in[i];
    }
    s = wave_sum(s);
    __shared__ TO ws[SCAN_THREADS / WAVE];
    if (lane_id() == 0) ws[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        TO t = 0;
        for (int w = 0; w < SCAN_THREADS / WAVE; ++w) t += ws[w];
        partial[blockIdx.x] = t;
    }
}

// single block: exclusive scan of the partials in place, total to *total
template <typename TO>
__global__ __launch_bounds__(1024) void scan_partials_kernel(TO* partial, const uint32_t* n_dev, uint32_t n_host,
                                                             TO* total) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    __shared__ TO ws[16];
    __shared__ TO carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t i = base + threadIdx.x;
        TO v = i < nb ? partial[i] : 0;
        // inclusive wave scan
        TO x = v;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            TO y = __shfl_up(x, o, WAVE);
            if (lane_id() >= o) x += y;
        }
        if (lane_id() == WAVE - 1) ws[threadIdx.x / WAVE] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            TO r = 0;
            for (int w = 0; w < 16; ++w) { TO t = ws[w]; ws[w] = r; r += t; }
        }
        __syncthreads();
        TO excl = carry + ws[threadIdx.x / WAVE] + x - v;
        if (i < nb) partial[i] = excl;
        __syncthreads();
        if (threadIdx.x == 1023) carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_THREADS) void scan_down_kernel(const TI* in, TO* out, const uint32_t* n_dev,
                                                                 uint32_t n_host, const TO* partial) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    // each thread owns SCAN_PER_THREAD consecutive elements
    const uint32_t my = base + threadIdx.x * SCAN_PER_THREAD;
    TO v[SCAN_PER_THREAD];
    TO s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        uint32_t i = my + k;
        v[k] = i < n ? (TO)in[i] : 0;
        s += v[k];
    }
    TO x = s;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        TO y = __shfl_up(x, o, WAVE);
        if (lane_id() >= o) x += y;
    }
    __shared__ TO ws[SCAN_THREADS / WAVE];
    if (lane_id() == WAVE - 1) ws[threadIdx.x / WAVE] = x;
    __syncthreads();
    TO wpre = 0;
    for (int w = 0; w < (int)(threadIdx.x / WAVE); ++w) wpre += ws[w];
    TO run = partial[blockIdx.x] + wpre + x - s;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        uint32_t i = my + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

template <typename TI, typename TO>
void exclusive_scan(const TI* in, TO* out, const uint32_t* n_dev, uint32_t n_max, TO* total, void* tmp,
                    hipStream_t st) {
    const unsigned nb = ceil_div(n_max ? n_max : 1, SCAN_TILE);
    TO* partial = (TO*)tmp;
    hipLaunchKernelGGL((scan_reduce_kernel<TI, TO>), dim3(nb), dim3(SCAN_THREADS), 0, st, in, n_dev, n_max,
                       partial);
    hipLaunchKernelGGL((scan_partials_kernel<TO>), dim3(1), dim3(1024), 0, st, partial, n_dev, n_max, total);
    hipLaunchKernelGGL((scan_down_kernel<TI, TO>), dim3(nb), dim3(SCAN_THREADS), 0, st, in, out, n_dev, n_max,
                       partial);
}

size_t scan_tmp_bytes(uint32_t n_max) { return (size_t)(ceil_div(n_max ? n_max : 1, SCAN_TILE) + 1) * 8; }

template void exclusive_scan<uint32_t, uint32_t>(const uint32_t*, uint32_t*, const uint32_t*, uint32_t,
                                                 uint32_t*, void*, hipStream_t);
template void exclusive_scan<uint32_t, unsigned long long>(const uint32_t*, unsigned long long*,
                                                           const uint32_t*, uint32_t, unsigned long long*,
                                                           void*, hipStream_t);

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: block histograms (digit-major),
// exclusive scan, then a stable scatter.  A block owns RS_TILE consecutive
// keys, processed in rounds of 256 (one per thread) so that equal digits keep
// their input order: rank = block base + earlier rounds + earlier waves +
// earlier lanes of the same digit (found with 8 ballots).
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;

// Per block: the digit histogram of its tile (digit-major counts), and the
// digit totals of the pass accumulated in `totals` (cleared beforehand), in
// RS_TOT_COPIES copies taken by block index to spread the atomics.
constexpr int RS_TOT_COPIES = 8;
__global__ __launch_bounds__(RS_THREADS) void radix_hist_kernel(const uint64_t* keys, uint32_t n, int shift,
                                                                uint32_t* counts, uint32_t nblocks,
                                                                uint32_t* totals) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    // every key of the thread is loaded before the first is counted: the
    // loads are in flight together (counted one by one, each waited for the
    // one before: 16 memory round trips in a row)
    const uint32_t base = blockIdx.x * RS_TILE;
    uint32_t d[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        // (unconditional loads, clamped into the array: a load under a
        // branch is waited for at the branch's end)
        const uint32_t i = base + r * RS_THREADS + threadIdx.x;
        const uint64_t k = keys[min(i, n - 1)];
        d[r] = i < n ? (uint32_t)(k >> shift) & 255u : 256u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r)
        if (d[r] < 256u) atomicAdd(&h[d[r]], 1u);
    __syncthreads();
    const uint32_t c = h[threadIdx.x];
    counts[threadIdx.x * nblocks + blockIdx.x] = c;
    if (c) atomicAdd(&totals[(blockIdx.x % RS_TOT_COPIES) * 256 + threadIdx.x], c);
}

// Exclusive scan of the digit-major counts of one pass in one launch: block
// j takes the G whole digits [j G, j G + G) (G nblocks counts), starts from
// the totals of the digits before them and scans its counts in tiles of
// 16 per thread.  (The three-phase device scan needs three dependent
// launches, and each costs ~10 us of idle GPU between these short kernels.)
__global__ __launch_bounds__(RS_THREADS) void radix_offsets_kernel(const uint32_t* counts, uint32_t* offs,
                                                                   const uint32_t* totals, uint32_t nblocks,
                                                                   uint32_t G) {
    constexpr int PER = 16;
    __shared__ uint32_t ws[RS_THREADS / WAVE];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x, wave = tid / WAVE;
    const uint32_t d0 = blockIdx.x * G;
    // base: the counts of every digit below d0
    uint32_t b = 0;
    if ((uint32_t)tid < d0)
        for (int c = 0; c < RS_TOT_COPIES; ++c) b += totals[c * 256 + tid];
    b = wave_sum(b);
    if (lane_id() == 0) ws[wave] = b;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < RS_THREADS / WAVE; ++w) t += ws[w];
        carry = t;
    }
    __syncthreads();
    const uint32_t e0 = d0 * nblocks, e1 = min(d0 + G, 256u) * nblocks;
    for (uint32_t t0 = e0; t0 < e1; t0 += RS_THREADS * PER) {
        const uint32_t my = t0 + tid * PER;
        uint32_t v[PER], s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            v[k] = my + k < e1 ? counts[my + k] : 0u;
            s += v[k];
        }
        uint32_t x = s;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, WAVE);
            if (lane_id() >= o) x += y;
        }
        __syncthreads();  // ws and carry of the previous tile are read
        if (lane_id() == WAVE - 1) ws[wave] = x;
        __syncthreads();
        uint32_t wpre = 0, tile = 0;
        for (int w = 0; w < RS_THREADS / WAVE; ++w) {
            wpre += w < wave ? ws[w] : 0u;
            tile += ws[w];
        }
        uint32_t run = carry + wpre + x - s;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (my + k < e1) offs[my + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == 0) carry += tile;
    }
}

__global__ void radix_zero_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ __launch_bounds__(RS_THREADS) void radix_scatter_kernel(const uint64_t* kin, const uint32_t* vin,
                                                                   uint64_t* kout, uint32_t* vout,
                                                                   const uint32_t* offsets, uint32_t n,
                                                                   int shift, uint32_t nblocks) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_THREADS / WAVE][256];
    const int tid = threadIdx.x, wave = tid / WAVE;
    run[tid] = offsets[tid * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_THREADS / WAVE; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    const uint64_t lt = lanemask_lt();
    // the tile's keys and values loaded up front (all in flight together)
    uint64_t kr[RS_ROUNDS];
    uint32_t vr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t i = base + r * RS_THREADS + tid;
        kr[r] = i < n ? kin[i] : 0;
        vr[r] = i < n ? vin[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t i = base + r * RS_THREADS + tid;
        const bool valid = i < n;
        const uint64_t k = kr[r];
        const uint32_t v = vr[r];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t lrank = __popcll(m & lt);
        if (valid && lrank == 0) wcnt[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + lrank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_THREADS / WAVE; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

size_t radix_tmp_bytes(uint32_t n) {
    const uint32_t nb = ceil_div(n ? n : 1, RS_TILE);
    // counts, offsets, digit totals per pass (RS_TOT_COPIES copies, up to 16 passes)
    return (size_t)256 * nb * 4 * 2 + (size_t)16 * RS_TOT_COPIES * 256 * 4 + 64;
}

// Sorts (keys, vals) in place over bits [0, nbits); alt buffers are scratch.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                      int nbits, void* tmp, hipStream_t st) {
    if (n <= 1) return;
    const uint32_t nb = ceil_div(n, RS_TILE);
    uint32_t* counts = (uint32_t*)tmp;
    uint32_t* offs = counts + (size_t)256 * nb;
    uint32_t* totals = offs + (size_t)256 * nb;  // [pass][RS_TOT_COPIES][256]
    int passes = (nbits + 7) / 8;
    if (passes & 1) ++passes;  // even number of passes: result ends in (keys, vals)
    // digits per offsets block: whole digits, about RS_TILE counts per block
    uint32_t G = 1;
    while (G < 256 && G * 2 * nb <= (uint32_t)RS_TILE) G *= 2;
    hipLaunchKernelGGL(radix_zero_kernel, dim3(1), dim3(1024), 0, st, totals,
                       (uint32_t)(passes * RS_TOT_COPIES * 256));
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        uint32_t* tot = totals + RS_TOT_COPIES * 256 * p;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ka, n, shift, counts, nb, tot);
        hipLaunchKernelGGL(radix_offsets_kernel, dim3(256 / G), dim3(RS_THREADS), 0, st, counts, offs, tot, nb, G);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ka, va, kb, vb, offs, n,
                           shift, nb);
        uint64_t* t = ka; ka = kb; kb = t;
        uint32_t* u = va; va = vb; vb = u;
    }
}

}  // namespace nbh
