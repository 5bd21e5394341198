// Device exclusive prefix sum and stable LSD radix sort of (key, value) pairs.
//
// The radix sort ranks elements inside an LDS tile with wave64 ballots, so it is
// stable: equal keys keep their input order.  The sort-and-sweep broadphase
// relies on that, because the reference resolves ties between equal FP32
// endpoint values by the previous frame's order (an insertion sort,
// src/rbc/BroadPhase.cpp:80-94).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rb {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;     // 2048 elements per block
constexpr int SCAN_TOP_THREADS = 1024;
constexpr int SCAN_TOP_ITEMS = 8;                         // top level handles <= 8192 tiles
constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;           // 1024 elements per block
constexpr int RS_BINS = 256;

// Block-wide exclusive scan of one value per thread (256 threads); returns the
// prefix, writes the block total to *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    // inclusive scan within the wave (wave64 shuffles)
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    if (t == 0) {
        uint32_t run = 0;
        for (int w = 0; w < nw; ++w) { uint32_t s = sh[w]; sh[w] = run; run += s; }
        sh[nw] = run;
    }
    __syncthreads();
    uint32_t r = sh[wave] + x - v;
    *total = sh[nw];
    __syncthreads();
    return r;
}

// Kernel 1: per-tile sums.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const uint32_t *__restrict__ in, const uint32_t *n_dev,
                                                              uint32_t n_host, uint32_t *__restrict__ tile_sums) {
    __shared__ uint32_t sh[SCAN_THREADS / 64 + 1];
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Kernel 2: exclusive scan of the tile sums in one block; total -> *total.
__global__ void __launch_bounds__(SCAN_TOP_THREADS) k_scan_top(uint32_t *__restrict__ tile_sums, uint32_t ntiles,
                                                                uint32_t *__restrict__ total) {
    __shared__ uint32_t sh[SCAN_TOP_THREADS / 64 + 1];
    const uint32_t base = threadIdx.x * SCAN_TOP_ITEMS;
    uint32_t v[SCAN_TOP_ITEMS];
    uint32_t s = 0;
    for (int k = 0; k < SCAN_TOP_ITEMS; ++k) {
        v[k] = (base + k < ntiles) ? tile_sums[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, sh, &tot);
    for (int k = 0; k < SCAN_TOP_ITEMS; ++k) {
        if (base + k < ntiles) tile_sums[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// Kernel 3: per-tile exclusive scan plus the tile offset.  out may alias in.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t *in, const uint32_t *n_dev, uint32_t n_host,
                                                             const uint32_t *__restrict__ tile_sums, uint32_t *out) {
    __shared__ uint32_t sh[SCAN_THREADS / 64 + 1];
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, sh, &tot) + tile_sums[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// ---------------------------------------------------------------- radix sort
// Order-preserving map of an FP32 value to uint32.  +0.0 and -0.0 map to the same
// key because they compare equal as floats.
__host__ __device__ __forceinline__ uint32_t float_key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) == 0u) return 0x80000000u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint32_t rs_digit(uint32_t key, int shift) { return (key >> shift) & 0xffu; }

// Per-tile digit histograms, digit-major: hist[d * ntiles + tile].
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint32_t *__restrict__ keys, const uint32_t *n_dev, uint32_t n_host,
                                                        int shift, uint32_t ntiles, uint32_t *__restrict__ hist,
                                                        const int *skip) {
    if (skip && *skip == 0) return;
    __shared__ uint32_t h[RS_BINS];
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    for (int i = threadIdx.x; i < RS_BINS; i += RS_THREADS) h[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; ++k) {
        uint32_t i = base + k * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[rs_digit(keys[i], shift)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RS_BINS; d += RS_THREADS) hist[d * ntiles + blockIdx.x] = h[d];
}

// Stable scatter of one tile using the scanned histogram offsets.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                           const uint32_t *n_dev, uint32_t n_host, int shift,
                                                           uint32_t ntiles, const uint32_t *__restrict__ hoff,
                                                           const int *skip) {
    if (skip && *skip == 0) return;
    __shared__ uint32_t base_d[RS_BINS];           // global offset of this tile's next element per digit
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];   // this round's per-wave digit counts
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int d = t; d < RS_BINS; d += RS_THREADS) {
        base_d[d] = hoff[d * ntiles + blockIdx.x];
        for (int w = 0; w < RS_WAVES; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t tile = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint32_t i = tile + r * RS_THREADS + t;
        const bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = rs_digit(key, shift);
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank_w = (uint32_t)__popcll(peers & lt);
        if (valid && rank_w == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t dst = base_d[d] + rank_w;
            for (int w = 0; w < wave; ++w) dst += wcnt[w][d];
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        for (int dd = t; dd < RS_BINS; dd += RS_THREADS) {
            uint32_t s = 0;
            for (int w = 0; w < RS_WAVES; ++w) { s += wcnt[w][dd]; wcnt[w][dd] = 0; }
            base_d[dd] += s;
        }
        __syncthreads();
    }
}

}  // namespace rb
