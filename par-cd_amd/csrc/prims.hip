// Primitives: fills, scan support and an LSD radix sort for 64-bit keys.
// No vendor primitive library is used; everything is wave64 code.
#include "prims.h"

namespace pcd {

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_parts(int nparts, int32_t *parts) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    int run = 0;
    for (int base = 0; base < nparts; base += SCAN_THREADS) {
        int i = base + threadIdx.x;
        int x = i < nparts ? parts[i] : 0;
        int tot;
        int ex = block_excl_scan(x, &tot, lds);
        if (i < nparts) parts[i] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) parts[nparts] = run;
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t *p, uint64_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int fill_i32(pcd_ctx *ctx, int32_t *p, int32_t v, int64_t count) {
    if (count <= 0) return PCD_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(ctx, count, 256 * 4)), dim3(256), 0, ctx->stream, p, v, count);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

int fill_u64(pcd_ctx *ctx, uint64_t *p, uint64_t v, int64_t count) {
    if (count <= 0) return PCD_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, count, 256 * 4)), dim3(256), 0, ctx->stream, p, v, count);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

// Streaming copy with 16-byte accesses: the attainable HBM bandwidth, and a
// kernel of known traffic for calibrating memory counters.
__global__ __launch_bounds__(256) void k_stream_copy(const int4 *__restrict__ src, int4 *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int ensure_scan_parts(pcd_ctx *ctx, int64_t n) {
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles < 1) tiles = 1;
    if (tiles + 1 <= ctx->scan_part_cap) return PCD_OK;
    if (ctx->d_scan_part) dev_free(ctx, ctx->d_scan_part);
    ctx->d_scan_part = nullptr;
    int cap = (int)(tiles + 1 + 1024);
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_scan_part, cap));
    ctx->scan_part_cap = cap;
    return PCD_OK;
}

// ------------------------------------------------------------ radix sort
// LSD, 8-bit digits.  Per pass: tile histograms (digit-major, kept twice:
// raw counts and, after one exclusive scan over all (digit, tile)
// counters, global offsets), then a scatter that first sorts its tile
// stably by digit in LDS -- ranks from wave ballots (8 ballots give the
// lanes that share a digit) and per-round digit counters -- and then writes
// each digit's run to its global offset with consecutive threads on
// consecutive addresses.
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *keys, int64_t n, int shift, int ntiles,
                                                        int32_t *hist, int32_t *raw) {
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    uint64_t kv[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = base + (int64_t)r * RS_THREADS + threadIdx.x;
        kv[r] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = base + (int64_t)r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(int)((kv[r] >> shift) & 0xFF)], 1);
    }
    __syncthreads();
    const int64_t o = (int64_t)threadIdx.x * ntiles + blockIdx.x;
    hist[o] = cnt[threadIdx.x];
    raw[o] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *keys, uint64_t *out, int64_t n, int shift,
                                                           int ntiles, const int32_t *hist, const int32_t *raw) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ int gbase[256];
    __shared__ int lstart[256];
    __shared__ int run[256];
    __shared__ int wcnt[RS_THREADS / WAVE][256];
    __shared__ int scan_lds[RS_THREADS / WAVE + 1];
    const int t = threadIdx.x, lane = __lane_id(), w = t / WAVE;
    gbase[t] = hist[(int64_t)t * ntiles + blockIdx.x];
    int tot;
    lstart[t] = block_excl_scan(raw[(int64_t)t * ntiles + blockIdx.x], &tot, scan_lds);
    run[t] = 0;
    for (int i = 0; i < RS_THREADS / WAVE; i++) wcnt[i][t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int count = (int)min((int64_t)RS_TILE, n - base);
    const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    uint64_t kv[RS_ROUNDS];  // all of the thread's keys in flight at once
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = base + (int64_t)r * RS_THREADS + t;
        kv[r] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = base + (int64_t)r * RS_THREADS + t;
        bool valid = i < n;
        uint64_t key = kv[r];
        int d = (int)((key >> shift) & 0xFF);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int off = run[d] + rank;
            for (int j = 0; j < w; j++) off += wcnt[j][d];
            sk[lstart[d] + off] = key;
        }
        __syncthreads();
        int add = 0;
        for (int j = 0; j < RS_THREADS / WAVE; j++) { add += wcnt[j][t]; wcnt[j][t] = 0; }
        run[t] += add;
        __syncthreads();
    }
    // the tile is now sorted by digit: write each run to its global place
    for (int i = t; i < count; i += RS_THREADS) {
        uint64_t key = sk[i];
        int d = (int)((key >> shift) & 0xFF);
        out[gbase[d] + (i - lstart[d])] = key;
    }
}

struct RsLoad {
    const int32_t *h;
    __device__ int operator()(int64_t i) const { return h[i]; }
};
struct RsStore {
    int32_t *h;
    __device__ void operator()(int64_t i, int ex, int) const { h[i] = ex; }
};

// Sorts n keys whose significant bits are [0, bits).  Ping-pongs between
// keys and tmp; *out receives the buffer holding the result.
int radix_sort_u64(pcd_ctx *ctx, uint64_t *keys, uint64_t *tmp, int64_t n, int bits, uint64_t **out) {
    *out = keys;
    if (n <= 1) return PCD_OK;
    int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    size_t need = (size_t)ntiles * 256;
    if (2 * need > ctx->rs_hist_cap) {
        if (ctx->d_rs_hist) dev_free(ctx, ctx->d_rs_hist);
        ctx->d_rs_hist = nullptr;
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_rs_hist, 2 * need));
        ctx->rs_hist_cap = 2 * need;
    }
    int32_t *raw = ctx->d_rs_hist + need;
    uint64_t *src = keys, *dst = tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, ctx->stream, src, n, shift, ntiles,
                           ctx->d_rs_hist, raw);
        PCD_TRY(device_scan(ctx, (int64_t)need, RsLoad{ctx->d_rs_hist}, RsStore{ctx->d_rs_hist}, nullptr));
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, ctx->stream, src, dst, n, shift, ntiles,
                           ctx->d_rs_hist, raw);
        PCD_HIP(hipGetLastError());
        uint64_t *t = src; src = dst; dst = t;
    }
    *out = src;
    return PCD_OK;
}

}  // namespace pcd
