// Triangle-based initial partition on the device (framework.cpp:27-59,
// create_clustering_partition).
//
// The reference visits the vertices in ascending (-cc, -deg, v) order; a
// vertex still unlabeled when visited opens a community with itself and its
// unlabeled neighbours.  That is the lexicographically first maximal
// independent set of that order: v opens a community (is a "center") iff no
// neighbour before it in the order is a center, the centers are numbered in
// order, and every other vertex joins its first center neighbour in the
// order.  Here:
//  * the order is one 64-bit key per vertex, (rank of (-cc, -deg)) << 32 | v,
//    from two radix sorts; cc is the reference's double, with its int32
//    product deg * (deg - 1), and -0.0 ranks with 0.0 as in its comparison;
//  * rounds decide in place every undecided vertex whose earlier neighbours
//    are all decided: a center if none of them is one.  Decisions are final,
//    so a wave reads either a neighbour's final state or "undecided" (one
//    round late), never a wrong one;
//  * a dependency chain longer than the round cap (say a path numbered in
//    order) is finished on the host, in order;
//  * a scan over the order numbers the centers, and every other vertex takes
//    the id of its earliest center neighbour (the smallest key among them,
//    whose low 32 bits are that neighbour).
#include <vector>

#include "prims.h"

namespace pcd {

constexpr uint8_t ST_OPEN = 0, ST_CENTER = 1, ST_MEMBER = 2;

// order-preserving map of a double to an unsigned integer; -0.0 maps with 0.0
__device__ inline uint64_t sortable_f64(double x) {
    if (x == 0.0) x = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ULL << 63));
}

__device__ inline int lower_bound_u64(const uint64_t *a, int n, uint64_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// framework.cpp:31-34: key of -cc (two copies: one is sorted)
__global__ void k_cc_key(int n, const int32_t *indices, const int32_t *tri, uint64_t *k, uint64_t *ks) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const int deg = indices[v + 1] - indices[v];
        const double cc = deg > 1 ? tri[v] / (double)(int)((unsigned)deg * (unsigned)(deg - 1)) : 0.0;
        const uint64_t key = sortable_f64(-cc);
        k[v] = key;
        ks[v] = key;
    }
}

// (rank of -cc among all vertices) << 32 | (INT32_MAX - deg)
__global__ void k_cc_deg_key(int n, const int32_t *indices, const uint64_t *k1, const uint64_t *sorted1, uint64_t *k,
                             uint64_t *ks) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const int deg = indices[v + 1] - indices[v];
        const uint64_t key = ((uint64_t)lower_bound_u64(sorted1, n, k1[v]) << 32) | (uint32_t)(INT32_MAX - deg);
        k[v] = key;
        ks[v] = key;
    }
}

// final key: (rank of (-cc, -deg)) << 32 | v
__global__ void k_order_key(int n, const uint64_t *k2, const uint64_t *sorted2, uint64_t *key, uint64_t *ks) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint64_t k = ((uint64_t)lower_bound_u64(sorted2, n, k2[v]) << 32) | (uint32_t)v;
        key[v] = k;
        ks[v] = k;
    }
}

__global__ void k_order_of(int n, const uint64_t *sorted, int32_t *order, uint8_t *state) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        order[t] = (int32_t)(uint32_t)sorted[t];
        state[t] = ST_OPEN;
    }
}

// One round, one wave per vertex in order.  *open is set when a vertex stays
// undecided.
__global__ __launch_bounds__(256) void k_center_round(int n, const int32_t *order, const int32_t *indices,
                                                      const int32_t *adj, const uint64_t *key, uint8_t *state,
                                                      int32_t *open) {
    const int lane = __lane_id();
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    bool left = false;
    for (int t = wave_g; t < n; t += nwaves) {
        const int v = order[t];
        if (state[v] != ST_OPEN) continue;
        const uint64_t kv = key[v];
        const int b = indices[v], e = indices[v + 1];
        bool center_before = false, open_before = false;
        for (int j0 = b; j0 < e; j0 += WAVE) {
            const int j = j0 + lane;
            if (j < e) {
                const int u = adj[j];
                if (key[u] < kv) {
                    const uint8_t s = state[u];
                    center_before |= s == ST_CENTER;
                    open_before |= s == ST_OPEN;
                }
            }
            if (__ballot(center_before)) break;
        }
        const bool any_center = __ballot(center_before) != 0;
        const bool any_open = __ballot(open_before) != 0;
        if (lane == 0) {
            if (any_center) state[v] = ST_MEMBER;
            else if (!any_open) state[v] = ST_CENTER;
        }
        left |= !any_center && any_open;
    }
    if (lane == 0 && left) *open = 1;
}

// center ids: exclusive scan of "is a center" over the order
struct CenterFlag {
    const int32_t *order;
    const uint8_t *state;
    __device__ int operator()(int64_t t) const { return state[order[t]] == ST_CENTER ? 1 : 0; }
};
struct CenterStore {
    const int32_t *order;
    int32_t *labels;
    __device__ void operator()(int64_t t, int ex, int flag) const {
        if (flag) labels[order[t]] = ex;
    }
};

// every other vertex: the id of its center neighbour with the smallest key
__global__ __launch_bounds__(256) void k_member_labels(int n, const int32_t *indices, const int32_t *adj,
                                                       const uint64_t *key, const uint8_t *state, int32_t *labels) {
    const int lane = __lane_id();
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int v = wave_g; v < n; v += nwaves) {
        if (state[v] != ST_MEMBER) continue;
        uint64_t best = ~0ULL;
        for (int j = indices[v] + lane; j < indices[v + 1]; j += WAVE) {
            const int u = adj[j];
            if (state[u] == ST_CENTER) best = min(best, key[u]);
        }
#pragma unroll
        for (int d = WAVE / 2; d > 0; d >>= 1) best = min(best, (uint64_t)__shfl_xor((unsigned long long)best, d, WAVE));
        // a member always has an earlier center neighbour; the test keeps a
        // corrupt state from writing out of bounds
        if (lane == 0 && best != ~0ULL) labels[v] = labels[(int)(uint32_t)best];
    }
}

int clustering_labels_dev(pcd_ctx *ctx, int *num_communities) {
    const int n = ctx->n;
    uint64_t *a = nullptr, *b = nullptr, *c = nullptr, *key = nullptr;
    int32_t *order = nullptr, *open = nullptr;
    uint8_t *state = nullptr;
    int rc = PCD_OK;
    auto cleanup = [&] {
        for (void *p : {(void *)a, (void *)b, (void *)c, (void *)key, (void *)order, (void *)open, (void *)state})
            if (p) dev_free(ctx, p);
    };
#define PCD_CL(expr)            \
    do {                        \
        rc = (expr);            \
        if (rc != PCD_OK) {     \
            cleanup();          \
            return rc;          \
        }                       \
    } while (0)
    PCD_CL(dev_alloc_t(ctx, &a, n));
    PCD_CL(dev_alloc_t(ctx, &b, n));
    PCD_CL(dev_alloc_t(ctx, &c, n));
    PCD_CL(dev_alloc_t(ctx, &key, n));
    PCD_CL(dev_alloc_t(ctx, &order, n));
    PCD_CL(dev_alloc_t(ctx, &open, 1));
    PCD_CL(dev_alloc_t(ctx, &state, n));
    const int grid = grid_for(ctx, n, 256);
    int rank_bits = 1;
    while (rank_bits < 31 && (1LL << rank_bits) < n) rank_bits++;
    const int bits2 = 32 + rank_bits;
    uint64_t *sorted = nullptr;
    // -cc, then (-cc, -deg), then the final key
    hipLaunchKernelGGL(k_cc_key, dim3(grid), dim3(256), 0, ctx->stream, n, ctx->d_indices, ctx->d_tri, a, b);
    PCD_CL(radix_sort_u64(ctx, b, c, n, 64, &sorted));
    hipLaunchKernelGGL(k_cc_deg_key, dim3(grid), dim3(256), 0, ctx->stream, n, ctx->d_indices, a, sorted, key,
                       sorted == b ? c : b);
    uint64_t *tmp = sorted == b ? b : c;  // the buffer the sort leaves free
    uint64_t *k2s = sorted == b ? c : b;
    PCD_CL(radix_sort_u64(ctx, k2s, tmp, n, bits2, &sorted));
    hipLaunchKernelGGL(k_order_key, dim3(grid), dim3(256), 0, ctx->stream, n, key, sorted, a,
                       sorted == k2s ? tmp : k2s);
    uint64_t *k3s = sorted == k2s ? tmp : k2s;
    uint64_t *free3 = sorted == k2s ? k2s : tmp;
    PCD_CL(radix_sort_u64(ctx, k3s, free3, n, bits2, &sorted));
    hipLaunchKernelGGL(k_order_of, dim3(grid), dim3(256), 0, ctx->stream, n, sorted, order, state);
    // key now lives in `a`
    const uint64_t *fkey = a;
    bool done = false;
    for (int r = 0; r < ctx->tune.cluster_rounds; r++) {
        PCD_CL(fill_i32(ctx, open, 0, 1));
        hipLaunchKernelGGL(k_center_round, dim3(grid_for(ctx, n, 4, 8)), dim3(256), 0, ctx->stream, n, order,
                           ctx->d_indices, ctx->d_adj, fkey, state, open);
        if (hipGetLastError() != hipSuccess) PCD_CL(fail(PCD_E_HIP, "k_center_round launch"));
        if (hipMemcpyAsync(ctx->h_small, open, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            PCD_CL(fail(PCD_E_HIP, "center round"));
        if (ctx->h_small[0] == 0) {
            done = true;
            break;
        }
    }
    if (!done) {
        // the chain outlasted the rounds: finish in order on the host
        std::vector<uint8_t> hs((size_t)n);
        std::vector<int32_t> ho((size_t)n), hi((size_t)n + 1), ha((size_t)ctx->m * 2);
        std::vector<uint64_t> hk((size_t)n);
        if (hipMemcpyAsync(hs.data(), state, (size_t)n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(ho.data(), order, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(hk.data(), fkey, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(hi.data(), ctx->d_indices, 4 * ((size_t)n + 1), hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipMemcpyAsync(ha.data(), ctx->d_adj, 8 * (size_t)ctx->m, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            PCD_CL(fail(PCD_E_HIP, "center download"));
        for (int t = 0; t < n; t++) {
            const int v = ho[t];
            if (hs[v] != ST_OPEN) continue;
            bool center_before = false;
            for (int j = hi[v]; j < hi[v + 1] && !center_before; j++) {
                const int u = ha[j];
                center_before = hk[u] < hk[v] && hs[u] == ST_CENTER;
            }
            hs[v] = center_before ? ST_MEMBER : ST_CENTER;
        }
        if (hipMemcpyAsync(state, hs.data(), (size_t)n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            PCD_CL(fail(PCD_E_HIP, "center upload"));
    }
    int32_t *total = nullptr;
    PCD_CL(device_scan(ctx, n, CenterFlag{order, state}, CenterStore{order, ctx->d_labels}, &total));
    hipLaunchKernelGGL(k_member_labels, dim3(grid_for(ctx, n, 4, 8)), dim3(256), 0, ctx->stream, n, ctx->d_indices,
                       ctx->d_adj, fkey, state, ctx->d_labels);
    if (hipGetLastError() != hipSuccess) PCD_CL(fail(PCD_E_HIP, "k_member_labels launch"));
    if (hipMemcpyAsync(ctx->h_small, total, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        PCD_CL(fail(PCD_E_HIP, "center count"));
#undef PCD_CL
    *num_communities = ctx->h_small[0];
    cleanup();
    return PCD_OK;
}

}  // namespace pcd
