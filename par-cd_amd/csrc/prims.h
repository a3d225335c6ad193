// Device-wide primitives written for CDNA4 wave64: exclusive scan with
// user load/store functors, and wave/block reductions.
#ifndef PARCD_PRIMS_H
#define PARCD_PRIMS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace pcd {

// Community statistics update: size and int_vol are adjacent 32-bit fields,
// so one 64-bit atomic adds both (neither field can carry past 2^31).
// Adds (size, same, ext) to record l; size and same share one 64-bit
// atomic.  The deltas may be negative: the packed word changes by
// same * 2^32 + size mod 2^64, so a transient borrow between the halves is
// undone by the later additions and the final halves are exact.
__device__ inline void add_stats(CommStats *st, int l, int size, int same, int ext) {
    atomicAdd((unsigned long long *)&st[l].size,
              (unsigned long long)((long long)same * 4294967296LL + (long long)size));
    if (ext) atomicAdd(&st[l].ext_vol, ext);
}

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ROUNDS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ROUNDS;

__device__ inline int lane_id() { return __lane_id(); }

__device__ inline int wave_incl_scan(int x) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int y = __shfl_up(x, d, WAVE);
        if (lane_id() >= d) x += y;
    }
    return x;
}

__device__ inline int wave_sum(int x) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

__device__ inline long long wave_sum64(long long x) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

__device__ inline double wave_sum_f64(double x) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

// Block-wide exclusive scan of one value per thread (blockDim == SCAN_THREADS);
// returns the exclusive prefix, *total receives the block sum.
__device__ inline int block_excl_scan(int x, int *total, int *lds /* >= SCAN_THREADS/WAVE + 1 */) {
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    int inc = wave_incl_scan(x);
    if (lane == WAVE - 1) lds[w] = inc;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_THREADS / WAVE; i++) {
        int t = lds[i];
        if (i < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

// Block-wide exclusive scan for a block of NT threads (all threads call).
template <int NT>
__device__ inline int block_excl_scan_n(int x, int *total, int *lds /* >= NT/WAVE */) {
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    int inc = wave_incl_scan(x);
    __syncthreads();
    if (lane == WAVE - 1) lds[w] = inc;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / WAVE; i++) {
        int t = lds[i];
        if (i < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

// Phase 1: per-tile sums (a tile is `rounds` x SCAN_THREADS values).
template <class Load>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(int64_t n, Load load, int32_t *parts, int rounds) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * rounds * SCAN_THREADS;
    int acc = 0;
#pragma unroll 4
    for (int r = 0; r < rounds; r++) {
        int64_t i = base + (int64_t)r * SCAN_THREADS + threadIdx.x;
        if (i < n) acc += load(i);
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < SCAN_THREADS / WAVE; i++) s += lds[i];
        parts[blockIdx.x] = s;
    }
}

// Phase 2: exclusive scan of the tile sums by one block; parts[nparts] = total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_parts(int nparts, int32_t *parts);

// Phase 3: rescan each tile with its offset; store(i, exclusive, value).
template <class Load, class Store>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(int64_t n, Load load, Store store, const int32_t *parts,
                                                             int rounds) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * rounds * SCAN_THREADS;
    int run = parts[blockIdx.x];
    for (int r = 0; r < rounds; r++) {
        int64_t i = base + (int64_t)r * SCAN_THREADS + threadIdx.x;
        int x = i < n ? load(i) : 0;
        int tot;
        int ex = block_excl_scan(x, &tot, lds);
        if (i < n) store(i, run + ex, x);
        run += tot;
    }
}

// Rounds per scan tile for n values: of 16, 8, 4, 2 and 1, the one whose
// tile count is nearest (by ratio) to scan_tiles_per_cu tiles per CU, the
// smaller on a tie (a small n is one tile whatever the rounds).  A scan is then a few short rounds on many workgroups
// rather than 16 serial rounds on a few (its time is latency, not work),
// without so many tiles that the one-workgroup scan of their sums grows.
inline int scan_rounds(const pcd_ctx *ctx, int64_t n) {
    const double want = (double)ctx->tune.scan_tiles_per_cu * ctx->cu_count;
    int best = SCAN_ROUNDS;
    double best_d = 1e300;
    for (int r = SCAN_ROUNDS; r >= 1; r /= 2) {
        const int64_t tiles = (n + (int64_t)r * SCAN_THREADS - 1) / ((int64_t)r * SCAN_THREADS);
        const double ratio = tiles > 0 ? (double)tiles / want : 1e-300;
        const double d = ratio >= 1.0 ? ratio : 1.0 / ratio;
        if (d <= best_d) {
            best_d = d;
            best = r;
        }
    }
    return best;
}

// Exclusive scan over [0, n): load(i) gives the value, store(i, prefix,
// value) consumes it.  The total lands in ctx->d_scan_part[num_tiles].
// Returns the device address of the total.
template <class Load, class Store>
int device_scan(pcd_ctx *ctx, int64_t n, Load load, Store store, int32_t **total_dev) {
    PCD_TRY(ensure_scan_parts(ctx, n));
    const int rounds = scan_rounds(ctx, n);
    int64_t tiles = (n + (int64_t)rounds * SCAN_THREADS - 1) / ((int64_t)rounds * SCAN_THREADS);
    if (tiles < 1) tiles = 1;
    if (tiles + 1 > ctx->scan_part_cap) return fail(PCD_E_STATE, "scan tile sums exceed their buffer");
    hipLaunchKernelGGL(k_scan_tiles<Load>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, ctx->stream, n, load,
                       ctx->d_scan_part, rounds);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, (int)tiles, ctx->d_scan_part);
    hipLaunchKernelGGL((k_scan_apply<Load, Store>), dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, ctx->stream, n,
                       load, store, ctx->d_scan_part, rounds);
    PCD_HIP(hipGetLastError());
    if (total_dev) *total_dev = ctx->d_scan_part + tiles;
    return PCD_OK;
}

}  // namespace pcd

#endif
