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

// Random gathers over a graph's own adjacency: every entry's neighbour id
// (coalesced, PF in flight per lane as the propagation kernels keep them),
// its label, and with two levels that label's label (a dependent gather into
// an array indexed by community id, as the gains' community totals are).
// Labels outside [0, n) are not followed.  The sum only keeps the gathers
// from being dropped.
__global__ __launch_bounds__(256) void k_gather_probe(int64_t E, const int32_t *adj, const int32_t *labels, int n,
                                                      int levels, unsigned long long *sink) {
    constexpr int PF = 8;
    unsigned acc = 0;
    const int64_t step = (int64_t)gridDim.x * blockDim.x * PF;
    for (int64_t e0 = (int64_t)blockIdx.x * blockDim.x * PF + threadIdx.x; e0 < E; e0 += step) {
        int x[PF];
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int64_t e = e0 + (int64_t)k * blockDim.x;
            x[k] = e < E ? adj[e] : -1;
        }
#pragma unroll
        for (int k = 0; k < PF; k++) x[k] = (x[k] >= 0 && x[k] < n) ? labels[x[k]] : -1;
        if (levels == 2) {
#pragma unroll
            for (int k = 0; k < PF; k++) x[k] = (x[k] >= 0 && x[k] < n) ? labels[x[k]] : -1;
        }
#pragma unroll
        for (int k = 0; k < PF; k++) acc += (unsigned)x[k];
    }
    acc = (unsigned)wave_sum((int)acc);
    if (lane_id() == 0) atomicAdd(sink, (unsigned long long)acc);
}

int ensure_scan_parts(pcd_ctx *ctx, int64_t n) {
    // room for the tile sums at any tile size scan_rounds picks for up to n
    int64_t tiles = (n + SCAN_THREADS - 1) / SCAN_THREADS;
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

// the tile histograms of a three-launch sort of n keys
int ensure_sort_space(pcd_ctx *ctx, int64_t n) {
    const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    const size_t need = (size_t)ntiles * 256;
    if (2 * need > ctx->rs_hist_cap) {
        if (ctx->d_rs_hist) dev_free(ctx, ctx->d_rs_hist);
        ctx->d_rs_hist = nullptr;
        ctx->rs_hist_cap = 0;
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_rs_hist, 2 * need));
        ctx->rs_hist_cap = 2 * need;
    }
    return PCD_OK;
}

// Sorts n keys whose significant bits are [0, bits) with the three-launch
// passes above (tile histograms, a scan of all (digit, tile) counters, the
// scatter).  Ping-pongs between src and dst from pass `shift0` on.
static int lsd_passes(pcd_ctx *ctx, uint64_t *src, uint64_t *dst, int64_t n, int shift0, int bits, uint64_t **out) {
    int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    size_t need = (size_t)ntiles * 256;
    PCD_TRY(ensure_sort_space(ctx, n));
    int32_t *raw = ctx->d_rs_hist + need;
    for (int shift = shift0; shift < bits; shift += 8) {
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

// ------------------------------------------------- one-sweep radix sort
// LSD with 8-bit digits, one launch per pass.  A first kernel reads every
// key once and builds the global histogram of every pass's digit; one
// workgroup scans them into per-pass digit bases.  Each pass's scatter then
// takes a tile ticket (tiles are numbered in the order workgroups start, so
// every lower ticket belongs to a workgroup that is running or done),
// counts its digits, publishes them as an aggregate, sorts its tile by
// digit in LDS (the ranking of k_rs_scatter), and finds each digit's
// offset among the earlier tiles by walking back over their published
// words until an inclusive prefix (decoupled look-back); it publishes its
// own inclusive prefix and writes each digit's run contiguously.  Per pass
// the keys are read once and written once.
//
// Status word per (tile, digit): bits 31-30 the state (0 none, 1
// aggregate, 2 inclusive prefix), bits 29-0 the count, so n < 2^30.  A
// wait polls with a bounded number of sleeps; a workgroup that runs out
// records its pass (plus one) in the error flag and moves on, and the
// host, which reads the flag after each pass, redoes the sort from that
// pass with the three-launch kernels, so every wait ends and the result
// stays exact.  (Reading the flag once after the last pass instead, with
// later passes skipping their work, measured slower: the scatter launches
// took twice as long back to back.)
constexpr int OS_MAX_PASSES = 8;
constexpr uint32_t OS_AGG = 1u << 30, OS_INC = 2u << 30, OS_VAL = (1u << 30) - 1;
constexpr int OS_WIN = 16;  // look-back words read at once per digit

__global__ __launch_bounds__(RS_THREADS) void k_os_hist(const uint64_t *keys, int64_t n, int passes,
                                                        uint32_t *ghist) {
    __shared__ uint32_t h[RS_THREADS / WAVE][OS_MAX_PASSES * 256];  // one copy per wave
    const int t = threadIdx.x, w = t / WAVE;
    for (int i = t; i < RS_THREADS / WAVE * OS_MAX_PASSES * 256; i += RS_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    // Keys of neighbouring entries often share their high digits (pair keys
    // of one community row share the community), and 64 lanes adding to one
    // LDS counter serialise.  So per pass, up to four times, the lanes with
    // the digit of the first lane still to count add it with one atomic;
    // the lanes left (digits spread out) add one each.  Whole waves iterate
    // together, so every lane takes part in the ballots and the shuffle.
    const int lane = __lane_id();
    for (int64_t b0 = (int64_t)blockIdx.x * RS_THREADS; b0 < n; b0 += (int64_t)gridDim.x * RS_THREADS) {
        const int64_t i = b0 + t;
        const bool valid = i < n;
        const uint64_t k = valid ? keys[i] : 0;
        for (int p = 0; p < passes; p++) {
            const int d = (int)((k >> (8 * p)) & 0xFF);
            bool todo = valid;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint64_t rem = __ballot(todo);
                if (rem == 0) break;
                const int leader = __ffsll((unsigned long long)rem) - 1;
                const int dl = __shfl(d, leader, WAVE);
                const bool mine = todo && d == dl;
                const uint64_t same = __ballot(mine);
                if (lane == leader) atomicAdd(&h[w][p * 256 + dl], (uint32_t)__popcll(same));
                todo = todo && !mine;
            }
            if (todo) atomicAdd(&h[w][p * 256 + d], 1u);
        }
    }
    __syncthreads();
    for (int i = t; i < passes * 256; i += RS_THREADS) {
        uint32_t c = 0;
        for (int j = 0; j < RS_THREADS / WAVE; j++) c += h[j][i];
        if (c) atomicAdd(&ghist[i], c);
    }
}

// one workgroup: exclusive scan of each pass's 256 digit counts
__global__ __launch_bounds__(RS_THREADS) void k_os_bases(int passes, const uint32_t *ghist, uint32_t *base) {
    __shared__ int lds[RS_THREADS / WAVE + 1];
    for (int p = 0; p < passes; p++) {
        int tot;
        const int ex = block_excl_scan((int)ghist[p * 256 + threadIdx.x], &tot, lds);
        base[p * 256 + threadIdx.x] = (uint32_t)ex;
    }
}

__global__ __launch_bounds__(RS_THREADS) void k_os_scatter(const uint64_t *keys, uint64_t *out, int64_t n, int shift,
                                                           const uint32_t *digit_base, uint32_t *status, int *ticket,
                                                           int *err, int pass, int spin_limit) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ int cnt[256];
    __shared__ int lstart[256];
    __shared__ int run[256];
    __shared__ int gout[256];
    __shared__ int wcnt[RS_THREADS / WAVE][256];
    __shared__ int scan_lds[RS_THREADS / WAVE + 1];
    __shared__ int s_tile;
    const int t = threadIdx.x, lane = __lane_id(), w = t / WAVE;
    if (t == 0) s_tile = atomicAdd(ticket, 1);
    cnt[t] = 0;
    run[t] = 0;
    for (int i = 0; i < RS_THREADS / WAVE; i++) wcnt[i][t] = 0;
    __syncthreads();
    const int tile = s_tile;
    const int64_t base = (int64_t)tile * RS_TILE;
    const int count = (int)min((int64_t)RS_TILE, n - base);
    uint64_t kv[RS_ROUNDS];  // all of the thread's keys in flight at once
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = base + (int64_t)r * RS_THREADS + t;
        kv[r] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        if (r * RS_THREADS + t < count) atomicAdd(&cnt[(int)((kv[r] >> shift) & 0xFF)], 1);
    __syncthreads();
    // publish this tile's counts as early as possible
    uint32_t *my = status + (size_t)tile * 256;
    __hip_atomic_store(&my[t], (tile == 0 ? OS_INC : OS_AGG) | (uint32_t)cnt[t], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    int tot;
    lstart[t] = block_excl_scan(cnt[t], &tot, scan_lds);
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const bool valid = r * RS_THREADS + t < count;
        const uint64_t key = kv[r];
        const int d = (int)((key >> shift) & 0xFF);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
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
    // decoupled look-back: thread t owns digit t and reads the words of
    // OS_WIN earlier tiles at once (nearest first), adding aggregates until
    // an inclusive prefix; a word not yet published is read again
    uint32_t prefix = 0;
    // spin limit 0 (a host setting for tests of the fallback): in pass 1
    // every tile after the first gives up as a look-back that ran out of
    // patience would
    if (spin_limit == 0 && pass == 1 && tile > 0) atomicCAS(err, 0, pass + 1);
    if (tile > 0) {
        int j = tile - 1, spins = 0;
        while (j >= 0) {
            uint32_t s[OS_WIN];
#pragma unroll
            for (int e = 0; e < OS_WIN; e++)
                s[e] = j - e >= 0 ? __hip_atomic_load(status + (size_t)(j - e) * 256 + t, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : OS_INC;  // never reached: tile 0 publishes an inclusive prefix
            int used = 0;
            bool found = false, waiting = false;
#pragma unroll
            for (int e = 0; e < OS_WIN; e++) {
                if (found || waiting) continue;
                const uint32_t fl = s[e] & ~OS_VAL;
                if (fl == 0) {
                    waiting = true;
                } else {
                    prefix += s[e] & OS_VAL;
                    used++;
                    if (fl == OS_INC) found = true;
                }
            }
            if (found) break;
            j -= used;
            if (waiting) {
                if (++spins > spin_limit) {
                    atomicCAS(err, 0, pass + 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __hip_atomic_store(&my[t], OS_INC | (prefix + (uint32_t)cnt[t]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    gout[t] = (int)(digit_base[t] + prefix) - lstart[t];
    __syncthreads();
    // the tile is sorted by digit: each digit's run goes to its global place
    for (int i = t; i < count; i += RS_THREADS) {
        const uint64_t key = sk[i];
        out[gout[(int)((key >> shift) & 0xFF)] + i] = key;
    }
}

// Sorts n keys whose significant bits are [0, bits).  Ping-pongs between
// keys and tmp; *out receives the buffer holding the result.  Uses the
// one-sweep passes when tuned so (tunable "onesweep") and n < 2^30.
int radix_sort_u64(pcd_ctx *ctx, uint64_t *keys, uint64_t *tmp, int64_t n, int bits, uint64_t **out) {
    *out = keys;
    if (n <= 1) return PCD_OK;
    const int passes = (bits + 7) / 8;
    // (a captured sort takes the three-launch passes: the one-sweep passes
    // read an error flag back after each pass)
    if (!ctx->tune.onesweep || ctx->capturing || n >= (int64_t)OS_VAL || passes > OS_MAX_PASSES)
        return lsd_passes(ctx, keys, tmp, n, 0, bits, out);
    const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    // status words, then 2 x 256 x passes histogram / bases, tickets, error flag
    const size_t words = (size_t)ntiles * 256 + 2 * 256 * OS_MAX_PASSES + OS_MAX_PASSES + 64;
    if (words > ctx->os_cap) {
        if (ctx->d_os) dev_free(ctx, ctx->d_os);
        ctx->d_os = nullptr;
        ctx->os_cap = 0;
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_os, words));
        ctx->os_cap = words;
    }
    uint32_t *status = ctx->d_os;
    uint32_t *ghist = status + (size_t)ntiles * 256;
    uint32_t *dbase = ghist + 256 * OS_MAX_PASSES;
    int *tickets = (int *)(dbase + 256 * OS_MAX_PASSES);
    int *err = tickets + OS_MAX_PASSES;
    PCD_TRY(fill_i32(ctx, (int32_t *)ghist, 0, 256 * OS_MAX_PASSES));
    PCD_TRY(fill_i32(ctx, tickets, 0, OS_MAX_PASSES + 1));
    hipLaunchKernelGGL(k_os_hist, dim3(grid_for(ctx, n, RS_THREADS * 16, 4)), dim3(RS_THREADS), 0, ctx->stream, keys,
                       n, passes, ghist);
    hipLaunchKernelGGL(k_os_bases, dim3(1), dim3(RS_THREADS), 0, ctx->stream, passes, ghist, dbase);
    uint64_t *src = keys, *dst = tmp;
    for (int p = 0; p < passes; p++) {
        PCD_TRY(fill_i32(ctx, (int32_t *)status, 0, (int64_t)ntiles * 256));
        hipLaunchKernelGGL(k_os_scatter, dim3(ntiles), dim3(RS_THREADS), 0, ctx->stream, src, dst, n, 8 * p,
                           dbase + 256 * p, status, tickets + p, err, p, ctx->tune.os_spin_limit);
        PCD_HIP(hipGetLastError());
        // a look-back that ran out of patience: redo this pass the other way
        PCD_HIP(hipMemcpyAsync(ctx->h_small + 12, err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->h_small[12] != 0) {
            if (ctx->h_small[12] != p + 1) return fail(PCD_E_STATE, "one-sweep sort: bad failed-pass record");
            ctx->os_fallbacks++;
            PCD_TRY(fill_i32(ctx, err, 0, 1));
            return lsd_passes(ctx, src, dst, n, 8 * p, bits, out);
        }
        uint64_t *t = src; src = dst; dst = t;
    }
    *out = src;
    return PCD_OK;
}

}  // namespace pcd
