// Partition phases on the device: label compression, community statistics
// and the metric evaluation of modularity / CPM / conductance.
#include "prims.h"

namespace pcd {

// ------------------------------------------------------------ compression
// framework.cpp:146-170 numbers labels in order of first occurrence over v.
// Here: first[l] = min v with label l (atomicMin), a vertex is a "first" if
// it is the minimum of its label or has NEW_LABEL, and the exclusive scan of
// those flags over v is exactly the reference's running counter.

// A plain read of first[l] skips the atomic when it already holds a smaller
// vertex: the value only decreases, so a stale read costs at most an
// atomic, and large communities do not serialise every member's update.
// The table's start: the vertex the last compression found first for label
// c, if it still has label c (an upper bound of the new first occurrence, so
// most members' reads skip their atomic, large communities included), else
// INT32_MAX.  firstv holds vertex ids or -1.
__global__ void k_first_init(int bound, const int32_t *labels, const int32_t *firstv, int32_t *first) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < bound; c += gridDim.x * blockDim.x) {
        const int w = firstv[c];
        first[c] = (w >= 0 && labels[w] == c) ? w : INT32_MAX;
    }
}

__global__ void k_first_occurrence(int n, const int32_t *labels, int32_t *first) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int l = labels[v];
        if (l >= 0 && first[l] > v) atomicMin(&first[l], v);
    }
}

struct FirstFlag {
    const int32_t *labels, *first;
    __device__ int operator()(int64_t v) const {
        int l = labels[v];
        return (l < 0 || first[l] == (int)v) ? 1 : 0;
    }
};
// Each first occurrence gets its new id.  With st_in set, the statistics
// move to the new ids as well: st_out[new id] = st_in[old label], and a
// NEW_LABEL vertex opens its singleton record {1, 0, deg}.
struct FirstStore {
    const int32_t *labels;
    int32_t *out, *newid;
    const CommStats *st_in;
    CommStats *st_out;
    const int32_t *indices;
    int32_t *firstv;  // firstv[new id] = its first vertex
    __device__ void operator()(int64_t v, int ex, int flag) const {
        if (!flag) return;
        int l = labels[v];
        firstv[ex] = (int)v;
        if (l < 0) out[v] = ex;
        else newid[l] = ex;
        if (st_in) st_out[ex] = l < 0 ? CommStats{1, 0, indices[v + 1] - indices[v], 0} : st_in[l];
    }
};

__global__ void k_relabel(int n, const int32_t *labels, const int32_t *newid, int32_t *out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int l = labels[v];
        if (l >= 0) out[v] = newid[l];
    }
}

__global__ void k_copy_i32(int32_t *dst, const int32_t *src) { *dst = *src; }

// the first *kp community records of src into dst
__global__ void k_copy_stats(const int32_t *kp, const CommStats *src, CommStats *dst) {
    const int k = *kp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x) dst[c] = src[c];
}

// slot of the pinned scratch ints that receives a deferred k
constexpr int K_SLOT = 8;

void resolve_k(pcd_ctx *ctx) {
    if (!ctx->k_deferred) return;
    ctx->k = ctx->h_small[K_SLOT];
    ctx->label_bound = ctx->k;
    ctx->k_deferred = false;
}

int compress_labels_dev(pcd_ctx *ctx, bool defer_k) {
    const int n = ctx->n;
    kt_begin(ctx, KS_COMPRESS);
    if (ctx->label_bound > 0 && ctx->label_bound <= n + 1)
        hipLaunchKernelGGL(k_first_init, dim3(grid_for(ctx, ctx->label_bound, 256)), dim3(256), 0, ctx->stream,
                           ctx->label_bound, ctx->d_labels, ctx->d_firstv, ctx->d_first);
    else if (ctx->label_bound > 0)
        PCD_TRY(fill_i32(ctx, ctx->d_first, INT32_MAX, ctx->label_bound));
    hipLaunchKernelGGL(k_first_occurrence, dim3(grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_labels,
                       ctx->d_first);
    // statistics of these labels (full count, or updated by the last
    // propagation step) are renumbered with them
    const bool renumber = ctx->stats_pending || ctx->has_stats;
    int32_t *total = nullptr;
    PCD_TRY(device_scan(ctx, n, FirstFlag{ctx->d_labels, ctx->d_first},
                        FirstStore{ctx->d_labels, ctx->d_labels_alt, ctx->d_newid, renumber ? ctx->d_stats : nullptr,
                                   ctx->d_stats_alt, ctx->d_indices, ctx->d_firstv},
                        &total));
    hipLaunchKernelGGL(k_relabel, dim3(grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_labels,
                       ctx->d_newid, ctx->d_labels_alt);
    hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(1), 0, ctx->stream, ctx->d_k, total);
    kt_end(ctx, KS_COMPRESS);
    PCD_HIP(hipGetLastError());
    if (ctx->capturing) {
        // a captured pass: k stays on the device, and the renumbered
        // statistics move back in place so that the replayed pass always
        // reads and writes the same buffers
        ctx->k_deferred = true;
        if (renumber)
            hipLaunchKernelGGL(k_copy_stats, dim3(grid_for(ctx, n + 1, 256, 8)), dim3(256), 0, ctx->stream,
                               ctx->d_k, ctx->d_stats_alt, ctx->d_stats);
    } else if (defer_k) {
        PCD_HIP(hipMemcpyAsync(ctx->h_small + K_SLOT, ctx->d_k, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        ctx->k_deferred = true;
    } else {
        PCD_HIP(hipMemcpyAsync(ctx->h_small, ctx->d_k, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        ctx->k = ctx->h_small[0];
        ctx->label_bound = ctx->k;
        ctx->k_deferred = false;
    }
    std::swap(ctx->d_labels, ctx->d_labels_alt);
    if (renumber && !ctx->capturing) std::swap(ctx->d_stats, ctx->d_stats_alt);
    ctx->labels_compressed = true;
    ctx->has_stats = renumber;
    ctx->stats_pending = false;
    return PCD_OK;
}

// ------------------------------------------------------------ statistics
// framework.cpp:69-143: per vertex, count neighbours with the same label;
// sizes[l] += 1, int_volume[l] += same, ext_volume[l] += deg - same.


__global__ void k_clear_stats(CommStats *st, const int32_t *kp) {
    const int k = *kp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x)
        st[c] = CommStats{0, 0, 0, 0};
}

// G lanes per vertex, vertices at positions [p0, p1) (all with degree <= G
// is the fast case; larger degrees are still correct).
template <int G>
__global__ __launch_bounds__(256) void k_stats_small(int p0, int p1, const int32_t *order, const int32_t *indices,
                                                     const int32_t *adj, const int32_t *labels, CommStats *st) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int per_wave = WAVE / G;
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int count = p1 - p0;
    for (int base = wave_g * per_wave; base < count; base += nwaves * per_wave) {
        int gi = base + lane / G;
        bool has = gi < count;
        int v = 0, start = 0, deg = 0, l = 0;
        if (has) {
            int p = p0 + gi;
            v = order ? order[p] : p;
            start = indices[v];
            deg = indices[v + 1] - start;
            l = labels[v];
        }
        int same = 0;
        for (int j = glane; j < deg; j += G) same += (labels[adj[start + j]] == l) ? 1 : 0;
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) same += __shfl_xor(same, d, WAVE);
        if (has && glane == 0) add_stats(st, l, 1, same, deg - same);
    }
}

// one wave per work item {v, begin, end, first} (rows longer than 64); each
// lane keeps PF gathers in flight
__global__ __launch_bounds__(256) void k_stats_items(int nitems, const int4 *items, const int32_t *adj,
                                                     const int32_t *labels, CommStats *st) {
    constexpr int PF = 8;
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int it = wave_g; it < nitems; it += nwaves) {
        int4 w = items[it];
        int l = labels[w.x];
        int same = 0;
        for (int j0 = w.y + lane; j0 < w.z; j0 += PF * WAVE) {
            int nb[PF];
#pragma unroll
            for (int e = 0; e < PF; e++) {
                int j = j0 + e * WAVE;
                nb[e] = j < w.z ? adj[j] : -1;
            }
#pragma unroll
            for (int e = 0; e < PF; e++) same += (nb[e] >= 0 && labels[nb[e]] == l) ? 1 : 0;
        }
        same = wave_sum(same);
        if (lane == 0) add_stats(st, l, w.w ? 1 : 0, same, (w.z - w.y) - same);
    }
}

// Same work with the labels of vertices [0, ncache) staged in LDS: ids are
// degree-descending, so those are the most frequent gather targets.  One
// 1024-thread workgroup per CU pair of slots; each wave takes items as above.
template <int NT>
__global__ __launch_bounds__(NT) void k_stats_items_cached(int nitems, const int4 *items, const int32_t *adj,
                                                           const int32_t *labels, CommStats *st, int ncache) {
    constexpr int PF = 8;
    __shared__ int32_t hot[STATS_CACHE];
    for (int i = threadIdx.x; i < ncache; i += NT) hot[i] = labels[i];
    __syncthreads();
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int it = wave_g; it < nitems; it += nwaves) {
        int4 w = items[it];
        int l = labels[w.x];
        int same = 0;
        for (int j0 = w.y + lane; j0 < w.z; j0 += PF * WAVE) {
            int nb[PF];
#pragma unroll
            for (int e = 0; e < PF; e++) {
                int j = j0 + e * WAVE;
                nb[e] = j < w.z ? adj[j] : -1;
            }
            int lab[PF];
#pragma unroll
            for (int e = 0; e < PF; e++) lab[e] = nb[e] < 0 ? -2 : (nb[e] < ncache ? hot[nb[e]] : labels[nb[e]]);
#pragma unroll
            for (int e = 0; e < PF; e++) same += lab[e] == l ? 1 : 0;
        }
        same = wave_sum(same);
        if (lane == 0) add_stats(st, l, w.w ? 1 : 0, same, (w.z - w.y) - same);
    }
}

template <int G>
static void launch_stats_small(pcd_ctx *ctx, int p0, int p1) {
    if (p1 <= p0) return;
    int per_block = 256 / G;
    hipLaunchKernelGGL(k_stats_small<G>, dim3(grid_for(ctx, p1 - p0, per_block, 16)), dim3(256), 0, ctx->stream, p0,
                       p1, ctx->d_order, ctx->d_indices, ctx->d_adj, ctx->d_labels, ctx->d_stats);
}

int collect_statistics_dev(pcd_ctx *ctx) {
    const BinLayout &b = ctx->bins;
    hipLaunchKernelGGL(k_clear_stats, dim3(grid_for(ctx, ctx->k, 256)), dim3(256), 0, ctx->stream, ctx->d_stats,
                       ctx->d_k);
    if (ctx->num_stat_items > 0) {
        kt_begin(ctx, KS_STATS_ITEMS);
        if (ctx->tune.stats_cache)
            hipLaunchKernelGGL(k_stats_items_cached<1024>, dim3(std::min(ctx->cu_count * 2, ctx->num_stat_items)),
                               dim3(1024), 0, ctx->stream, ctx->num_stat_items, ctx->d_stat_items, ctx->d_adj,
                               ctx->d_labels, ctx->d_stats, std::min(ctx->n, STATS_CACHE));
        else
            hipLaunchKernelGGL(k_stats_items, dim3(grid_for(ctx, ctx->num_stat_items, 4, 16)), dim3(256), 0,
                               ctx->stream, ctx->num_stat_items, ctx->d_stat_items, ctx->d_adj, ctx->d_labels,
                               ctx->d_stats);
        kt_end(ctx, KS_STATS_ITEMS);
    }
    kt_begin(ctx, KS_STATS_SMALL);
    launch_stats_small<64>(ctx, b.start[BIN_G64], b.start[BIN_G64 + 1]);
    launch_stats_small<32>(ctx, b.start[BIN_G32], b.start[BIN_G32 + 1]);
    launch_stats_small<16>(ctx, b.start[BIN_G16], b.start[BIN_G16 + 1]);
    launch_stats_small<8>(ctx, b.start[BIN_G8], b.start[BIN_G8 + 1]);
    launch_stats_small<4>(ctx, b.start[BIN_G4], b.start[BIN_G4 + 1]);
    kt_end(ctx, KS_STATS_SMALL);
    PCD_HIP(hipGetLastError());
    ctx->has_stats = true;
    return PCD_OK;
}

// ------------------------------------------------------------ statistics after a propagation step
// The counts of framework.cpp:69-143, brought from the labels before a
// propagation step (L) to those after it (N) through the rows of the
// vertices that moved.  Moving v from lo = L(v) to ln = N(v) touches the
// records of lo and ln alone.  For each neighbour w of v, the ordered pair
// (v, w) -- and (w, v) too when w kept its label, weight f = 2 -- was
// internal to lo when L(w) = lo and is internal to ln when N(w) = ln; a
// pair whose ends both moved is counted once from each end's row.  The
// volume moves by deg(v), and ext = volume - int.  A vertex with NEW_LABEL
// only leaves lo here: compression opens its singleton record.  The sums are
// exact integer additions, so the records equal a full recount.

// One list of the moved vertices, built by an exclusive scan of "moved"
// flags over index t: t < nsmall is small-bin position s[0] + t (its entry
// is the vertex), t in [nsmall, nsmall + nitems) is long-row item
// t - nsmall (its entry is the item index), and a last index with flag 0
// closes the list.  The scan keeps position order, so the small bins'
// moved vertices come out grouped by bin; bounds[b] (b < 5) is where bin b
// starts, bounds[5] where the items start and bounds[6] the total.
struct MovedArgs {
    int s[6];  // small bins G64..G4: positions [s[b], s[b+1])
    int nitems;
    const int32_t *order;
    const int4 *items;
    const int32_t *L, *N;
};

struct MovedFlag {
    MovedArgs a;
    __device__ int operator()(int64_t t) const {
        const int nsmall = a.s[5] - a.s[0];
        int v;
        if (t < nsmall) {
            const int p = a.s[0] + (int)t;
            v = a.order ? a.order[p] : p;
        } else if (t < nsmall + a.nitems) {
            v = a.items[t - nsmall].x;
        } else {
            return 0;
        }
        return a.L[v] != a.N[v] ? 1 : 0;
    }
};

struct MovedStore {
    MovedArgs a;
    int32_t *bounds, *list;
    __device__ void operator()(int64_t t, int ex, int flag) const {
        const int nsmall = a.s[5] - a.s[0];
#pragma unroll
        for (int b = 0; b < 5; b++)
            if (t == a.s[b] - a.s[0]) bounds[b] = ex;
        if (t == nsmall) bounds[5] = ex;
        if (t == nsmall + a.nitems) bounds[6] = ex;
        if (!flag) return;
        if (t < nsmall) {
            const int p = a.s[0] + (int)t;
            list[ex] = a.order ? a.order[p] : p;
        } else {
            list[ex] = (int)(t - nsmall);
        }
    }
};

// G lanes per moved vertex of one small bin (degree <= 64)
// Moved bitmap: bit v of M is set when vertex v's label changed in the step
// (L[v] != N[v]).  One wave writes one 64-bit word.
__global__ __launch_bounds__(256) void k_moved_bits(int n, const int32_t *L, const int32_t *N, uint64_t *M) {
    const int lane = __lane_id();
    const int nwords = (n + WAVE - 1) / WAVE;
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int wd = wave_g; wd < nwords; wd += nwaves) {
        const int v = wd * WAVE + lane;
        const uint64_t bits = __ballot(v < n && L[v] != N[v]);
        if (lane == 0) M[wd] = bits;
    }
}

// A neighbour's label before the step: its label after the step (already
// gathered) unless the bitmap says it moved, so unmoved neighbours cost one
// random gather instead of two (the bitmap itself stays in L2).
__device__ inline int label_before(const uint64_t *M, const int32_t *L, int w, int after) {
    return ((M[w >> 6] >> (w & 63)) & 1) ? L[w] : after;
}

template <int G>
__device__ inline void delta_small(int wave_g, int nwaves, const int32_t *list_all, const int32_t *bounds,
                                   const int32_t *indices, const int32_t *adj, const int32_t *L, const int32_t *N,
                                   const uint64_t *M, CommStats *st) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int per_wave = WAVE / G;
    const int32_t *list = list_all + bounds[0];
    const int cnt = bounds[1] - bounds[0];
    for (int base = wave_g * per_wave; base < cnt; base += nwaves * per_wave) {
        const int gi = base + lane / G;
        const bool has = gi < cnt;
        int start = 0, deg = 0, lo = 0, ln = 0;
        if (has) {
            const int v = list[gi];
            lo = L[v];
            ln = N[v];
            start = indices[v];
            deg = indices[v + 1] - start;
        }
        int a = 0, b = 0;
        for (int j = glane; j < deg; j += G) {
            const int w = adj[start + j];
            const int nw = N[w];
            const int lw = label_before(M, L, w, nw);
            const int f = lw != nw ? 1 : 2;
            a += lw == lo ? f : 0;
            b += nw == ln ? f : 0;
        }
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) {
            a += __shfl_xor(a, d, WAVE);
            b += __shfl_xor(b, d, WAVE);
        }
        if (has && glane == 0) {
            add_stats(st, lo, -1, -a, a - deg);
            if (ln >= 0) add_stats(st, ln, 1, b, deg - b);
        }
    }
}

// Moved items {v, begin, end, first} (rows longer than 64): each wave reads
// 64 list entries at once, lane j taking entry wave + (64 r + j) * waves in
// round r, and runs their row chunks one after another; one long row's
// chunks are spread over many waves.
__device__ inline void delta_items(int wave_g, int nwaves, const int32_t *list_all, const int32_t *bounds,
                                   const int4 *items, const int32_t *adj, const int32_t *L, const int32_t *N,
                                   const uint64_t *M, CommStats *st) {
    constexpr int PF = 4;
    const int lane = __lane_id();
    const int32_t *list = list_all + bounds[0];
    const int cnt = bounds[1] - bounds[0];
    for (int64_t r = 0; wave_g + r * WAVE * nwaves < cnt; r++) {
        const int64_t i = wave_g + (r * WAVE + lane) * nwaves;
        int4 w = make_int4(0, 0, 0, 0);
        int lo = 0, ln = 0;
        if (i < cnt) {
            w = items[list[i]];
            lo = L[w.x];
            ln = N[w.x];
        }
        uint64_t mv = __ballot(i < cnt);
        while (mv) {
            const int src = __builtin_ctzll(mv);
            mv &= mv - 1;
            const int xb = __builtin_amdgcn_readlane(w.y, src), xe = __builtin_amdgcn_readlane(w.z, src);
            const int xf = __builtin_amdgcn_readlane(w.w, src);
            const int xlo = __builtin_amdgcn_readlane(lo, src), xln = __builtin_amdgcn_readlane(ln, src);
            int a = 0, b = 0;
            for (int j0 = xb + lane; j0 < xe; j0 += PF * WAVE) {
                int nb[PF];
#pragma unroll
                for (int e = 0; e < PF; e++) {
                    const int j = j0 + e * WAVE;
                    nb[e] = j < xe ? adj[j] : -1;
                }
                int lw[PF], nw[PF];
#pragma unroll
                for (int e = 0; e < PF; e++) nw[e] = nb[e] < 0 ? -2 : N[nb[e]];
#pragma unroll
                for (int e = 0; e < PF; e++) lw[e] = nb[e] < 0 ? -2 : label_before(M, L, nb[e], nw[e]);
#pragma unroll
                for (int e = 0; e < PF; e++) {
                    if (nb[e] < 0) continue;
                    const int f = lw[e] != nw[e] ? 1 : 2;
                    a += lw[e] == xlo ? f : 0;
                    b += nw[e] == xln ? f : 0;
                }
            }
            a = wave_sum(a);
            b = wave_sum(b);
            if (lane == 0) {
                const int len = xe - xb;
                add_stats(st, xlo, xf ? -1 : 0, -a, a - len);
                if (xln >= 0) add_stats(st, xln, xf ? 1 : 0, b, len - b);
            }
        }
    }
}

// All six moved lists in one launch, so the short dependent-load chains of
// the lists overlap instead of running one kernel after another.  Every
// workgroup reads the six list lengths and computes the same split of the
// grid into contiguous ranges, one per list, proportional to the list's
// entries (at most G per small-bin vertex, DELTA_ITEM_COST per long-row
// item), with at least one workgroup for a list that is not empty.  The
// lists are disjoint and the updates are atomic adds.
constexpr int DELTA_ITEM_COST = 512;
__global__ __launch_bounds__(256) void k_delta_all(const int32_t *list_all, const int32_t *bounds, const int4 *items,
                                                   const int32_t *indices, const int32_t *adj, const int32_t *L,
                                                   const int32_t *N, const uint64_t *M, CommStats *st) {
    __shared__ int s_lo[7];
    if (threadIdx.x == 0) {
        constexpr int cost[6] = {64, 32, 16, 8, 4, DELTA_ITEM_COST};
        double w[6], tot = 0.0;
        int nonempty = 0;
        for (int l = 0; l < 6; l++) {
            const int c = bounds[l + 1] - bounds[l];
            w[l] = (double)c * cost[l];
            tot += w[l];
            nonempty += c > 0;
        }
        // one workgroup reserved per non-empty list, the rest by weight
        const int spare = (int)gridDim.x - nonempty;
        int at = 0;
        double acc = 0.0;
        for (int l = 0; l < 6; l++) {
            s_lo[l] = at;
            if (w[l] > 0) {
                const int before = tot > 0 ? (int)(spare * (acc / tot)) : 0;
                acc += w[l];
                const int after = tot > 0 ? (int)(spare * (acc / tot)) : 0;
                at += 1 + (after - before);
            }
        }
        s_lo[6] = at;  // <= gridDim.x
    }
    __syncthreads();
    int l = 0;
    while (l < 6 && !(s_lo[l] <= (int)blockIdx.x && (int)blockIdx.x < s_lo[l + 1])) l++;
    if (l == 6) return;  // beyond the last range (rounding) or no moved rows
    const int wpb = blockDim.x / WAVE;
    const int wave_g = ((int)blockIdx.x - s_lo[l]) * wpb + (int)threadIdx.x / WAVE;
    const int nwaves = (s_lo[l + 1] - s_lo[l]) * wpb;
    switch (l) {
    case 0: delta_small<64>(wave_g, nwaves, list_all, bounds + 0, indices, adj, L, N, M, st); break;
    case 1: delta_small<32>(wave_g, nwaves, list_all, bounds + 1, indices, adj, L, N, M, st); break;
    case 2: delta_small<16>(wave_g, nwaves, list_all, bounds + 2, indices, adj, L, N, M, st); break;
    case 3: delta_small<8>(wave_g, nwaves, list_all, bounds + 3, indices, adj, L, N, M, st); break;
    case 4: delta_small<4>(wave_g, nwaves, list_all, bounds + 4, indices, adj, L, N, M, st); break;
    default: delta_items(wave_g, nwaves, list_all, bounds + 5, items, adj, L, N, M, st); break;
    }
}

int statistics_delta_dev(pcd_ctx *ctx, const int32_t *before, const int32_t *after) {
    const BinLayout &b = ctx->bins;
    kt_begin(ctx, KS_STATS_DELTA);
    MovedArgs m;
    for (int i = 0; i < 6; i++) m.s[i] = b.start[BIN_G64 + i];
    m.nitems = ctx->num_stat_items;
    m.order = ctx->d_order;
    m.items = ctx->d_stat_items;
    m.L = before;
    m.N = after;
    int32_t *bounds = ctx->d_moved;
    int32_t *list = ctx->d_moved + 8;
    const int64_t len = (int64_t)(m.s[5] - m.s[0]) + m.nitems + 1;
    PCD_TRY(device_scan(ctx, len, MovedFlag{m}, MovedStore{m, bounds, list}, nullptr));
    uint64_t *M = ctx->d_moved_bits;
    hipLaunchKernelGGL(k_moved_bits, dim3(grid_for(ctx, (ctx->n + WAVE - 1) / WAVE, 4, 8)), dim3(256), 0, ctx->stream,
                       ctx->n, before, after, M);
    // the lists of empty bins (and no items) are empty: their workgroups exit
    hipLaunchKernelGGL(k_delta_all, dim3(ctx->cu_count * 24), dim3(256), 0, ctx->stream, list, bounds,
                       ctx->d_stat_items, ctx->d_indices, ctx->d_adj, before, after, M, ctx->d_stats);
    kt_end(ctx, KS_STATS_DELTA);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

// ------------------------------------------------------------ metric
// Modularity (rules.cpp:9-22) and CPM (rules.cpp:121-129) depend on the
// statistics only through sum(int), sum(vol^2) and sum(size*(size-1)),
// which are accumulated exactly in int64 (order independent); the final
// double arithmetic runs on the device.  Conductance (rules.cpp:58-70) is
// an FP64 sum in a fixed two-level order.

__global__ __launch_bounds__(256) void k_metric_sums(const CommStats *st, const int32_t *kp,
                                                     unsigned long long *acc) {
    __shared__ long long lds[3][256 / WAVE];
    const int k = *kp;
    long long a = 0, b = 0, c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        CommStats s = st[i];
        long long vol = (long long)s.int_vol + s.ext_vol;
        a += s.int_vol;
        b += vol * vol;
        c += (long long)s.size * (s.size - 1);
    }
    a = wave_sum64(a); b = wave_sum64(b); c = wave_sum64(c);
    const int w = threadIdx.x / WAVE;
    if (__lane_id() == 0) { lds[0][w] = a; lds[1][w] = b; lds[2][w] = c; }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long t = 0;
        for (int i = 0; i < 256 / WAVE; i++) t += lds[threadIdx.x][i];
        atomicAdd(&acc[threadIdx.x], (unsigned long long)t);
    }
}

__global__ void k_metric_final(const unsigned long long *acc, int kind, double lambda, int m, double *out) {
    long long a = (long long)acc[0], b = (long long)acc[1], c = (long long)acc[2];
    double two_m = 2.0 * m;
    if (kind == PCD_RULE_MODULARITY) *out = a / two_m - (double)b / (two_m * two_m);
    else *out = (a - lambda * (double)c) / (double)(2 * m);
}

constexpr int FSUM_BLOCKS = 1024;

__global__ __launch_bounds__(256) void k_cond_partial(const CommStats *st, const int32_t *kp, double *part) {
    __shared__ double lds[256 / WAVE];
    const int k = *kp;
    double acc = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x) {
        CommStats s = st[c];
        acc += s.size * (s.int_vol / (double)(s.int_vol + s.ext_vol));
    }
    acc = wave_sum_f64(acc);
    if (__lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < 256 / WAVE; i++) t += lds[i];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *part, int np, double div, double *out) {
    __shared__ double lds[256 / WAVE];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    acc = wave_sum_f64(acc);
    if (__lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < 256 / WAVE; i++) t += lds[i];
        *out = t / div;
    }
}

int metric_dev(pcd_ctx *ctx, pcd_rule rule, double *out) {
    if (rule.kind == PCD_RULE_WCC) return wcc_metric_dev(ctx, out);
    PCD_TRY(metric_enqueue(ctx, rule));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_result[0];
    return PCD_OK;
}

int metric_enqueue(pcd_ctx *ctx, pcd_rule rule) {
    if (rule.kind == PCD_RULE_WCC) return wcc_metric_enqueue(ctx);
    kt_begin(ctx, KS_METRIC);
    if (rule.kind == PCD_RULE_CONDUCTANCE) {
        hipLaunchKernelGGL(k_cond_partial, dim3(FSUM_BLOCKS), dim3(256), 0, ctx->stream, ctx->d_stats, ctx->d_k,
                           ctx->d_dacc);
        // rules.cpp:69 divides the sum by n
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->d_dacc, FSUM_BLOCKS,
                           (double)ctx->n, ctx->d_result);
    } else {
        PCD_TRY(fill_u64(ctx, (uint64_t *)ctx->d_acc, 0, 4));
        // the kernel reads k from d_k; a deferred k is bounded by n + 1
        const int kg = ctx->k_deferred ? ctx->n + 1 : ctx->k;
        hipLaunchKernelGGL(k_metric_sums, dim3(grid_for(ctx, kg, 256 * 8, 2)), dim3(256), 0, ctx->stream,
                           ctx->d_stats, ctx->d_k, (unsigned long long *)ctx->d_acc);
        hipLaunchKernelGGL(k_metric_final, dim3(1), dim3(1), 0, ctx->stream, (const unsigned long long *)ctx->d_acc,
                           rule.kind, rule.lambda, ctx->m, ctx->d_result);
    }
    kt_end(ctx, KS_METRIC);
    PCD_HIP(hipGetLastError());
    if (!ctx->capturing)
        PCD_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return PCD_OK;
}

}  // namespace pcd
