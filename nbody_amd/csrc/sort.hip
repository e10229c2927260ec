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

__global__ __launch_bounds__(RS_THREADS) void radix_hist_kernel(const uint64_t* keys, uint32_t n, int shift,
                                                                uint32_t* counts, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        uint32_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
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
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t i = base + r * RS_THREADS + tid;
        const bool valid = i < n;
        uint64_t k = valid ? kin[i] : 0;
        uint32_t v = valid ? vin[i] : 0;
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

// Exclusive scan of the digit counts of one pass in a single workgroup (one
// launch where the three-phase scan takes three, and a dependent launch costs
// ~10 us of idle GPU between short kernels): thread t owns the consecutive
// elements [t per, t per + per), per a multiple of 4 (16-byte loads and
// stores; the buffers are 16-byte aligned), read once for its total and once
// more to write.  For n <= SCAN1_MAX.
constexpr uint32_t SCAN1_THREADS = 1024, SCAN1_MAX = 1024u * 256u;
__global__ __launch_bounds__(SCAN1_THREADS) void scan_counts_kernel(const uint32_t* __restrict__ in,
                                                                    uint32_t* __restrict__ out, uint32_t n) {
    const uint32_t per = ((n + SCAN1_THREADS - 1) / SCAN1_THREADS + 3u) & ~3u;
    const uint32_t b0 = min(threadIdx.x * per, n), b1 = min(b0 + per, n);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; i += 4) {
        if (i + 4 <= b1) {
            const uint4 x = *reinterpret_cast<const uint4*>(in + i);
            s += x.x + x.y + x.z + x.w;
        } else {
            for (uint32_t j = i; j < b1; ++j) s += in[j];
        }
    }
    // exclusive scan of the thread totals: within the wave, then over the 16 waves
    uint32_t x = s;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if (lane_id() >= o) x += y;
    }
    __shared__ uint32_t ws[SCAN1_THREADS / WAVE];
    const int wave = threadIdx.x / WAVE;
    if (lane_id() == WAVE - 1) ws[wave] = x;
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < wave; ++w) wpre += ws[w];
    uint32_t run = wpre + x - s;
    for (uint32_t i = b0; i < b1; i += 4) {
        if (i + 4 <= b1) {
            const uint4 v = *reinterpret_cast<const uint4*>(in + i);
            uint4 o;
            o.x = run; run += v.x;
            o.y = run; run += v.y;
            o.z = run; run += v.z;
            o.w = run; run += v.w;
            *reinterpret_cast<uint4*>(out + i) = o;
        } else {
            for (uint32_t j = i; j < b1; ++j) {
                out[j] = run;
                run += in[j];
            }
        }
    }
}

size_t radix_tmp_bytes(uint32_t n) {
    const uint32_t nb = ceil_div(n ? n : 1, RS_TILE);
    return (size_t)256 * nb * 4 * 2 + scan_tmp_bytes(256 * nb) + 64;
}

// Sorts (keys, vals) in place over bits [0, nbits); alt buffers are scratch.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                      int nbits, void* tmp, hipStream_t st) {
    if (n <= 1) return;
    const uint32_t nb = ceil_div(n, RS_TILE);
    uint32_t* counts = (uint32_t*)tmp;
    uint32_t* offs = counts + (size_t)256 * nb;
    void* stmp = (void*)(offs + (size_t)256 * nb);
    int passes = (nbits + 7) / 8;
    if (passes & 1) ++passes;  // even number of passes: result ends in (keys, vals)
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ka, n, shift, counts, nb);
        if (256u * nb <= SCAN1_MAX)
            hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(SCAN1_THREADS), 0, st, counts, offs, 256u * nb);
        else
            exclusive_scan<uint32_t, uint32_t>(counts, offs, nullptr, 256 * nb, (uint32_t*)nullptr, stmp, st);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ka, va, kb, vb, offs, n,
                           shift, nb);
        uint64_t* t = ka; ka = kb; kb = t;
        uint32_t* u = va; va = vb; vb = u;
    }
}

}  // namespace nbh
