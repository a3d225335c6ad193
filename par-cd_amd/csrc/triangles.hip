// Triangle counting (Graph::count_triangles, graph.cpp:107-175) and the
// internal-triangle pass of WCCRule::calculate (rules.cpp:150-235).
//
// For each vertex v and each neighbour u < v at row position r, the
// reference intersects N(v) ∩ [0,u) -- which, rows being ascending, is the
// row prefix of v before r -- with N(u) ∩ [0,u), and for every common w adds
// one to the entry counters of (v,w), (u,w) and (v,u).  Here one wave owns
// a vertex v: the row prefix of v (ids < v) and its entry counters are
// staged in LDS, the wave walks u over that prefix, and for each u its lanes
// take the entries of N(u) ∩ [0,u) and binary-search them in the staged
// prefix.  Counters of row v stay in LDS and are added to memory once;
// counters of (u,w) go to memory with atomics.  With labels given, only
// edges and triangles inside one community count (rules.cpp:172,185).
#include "prims.h"

namespace pcd {

constexpr int TRI_STAGE = 2048;  // staged prefix entries per wave

// lowpos[v] = number of neighbours of v with a smaller id
__global__ void k_lowpos(int n, const int32_t *indices, const int32_t *adj, int32_t *lowpos) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int lo = indices[v], hi = indices[v + 1];
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (adj[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        lowpos[v] = lo - indices[v];
    }
}

__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int find_sorted(const int32_t *a, int len, int x) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return (lo < len && a[lo] == x) ? lo : -1;
}

// Vertices whose prefix exceeds FLAT_TS: one wave walks u over the prefix
// and its lanes take the entries of N(u) ∩ [0,u).
__global__ __launch_bounds__(256) void k_triangles(int n, const int32_t *indices, const int32_t *adj,
                                                   const int32_t *lowpos, const int32_t *labels, int32_t *et,
                                                   int min_len) {
    __shared__ int32_t s_val[4][TRI_STAGE];
    __shared__ int32_t s_cnt[4][TRI_STAGE];
    const int lane = __lane_id(), w = threadIdx.x / WAVE;
    int32_t *sval = s_val[w], *scnt = s_cnt[w];
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int v = wave_g; v < n; v += nwaves) {
        const int lenv = lowpos[v];
        if (lenv < 2 || lenv <= min_len) continue;  // a triangle needs two smaller neighbours
        const int sv = indices[v];
        const int l = labels ? labels[v] : 0;
        const bool staged = lenv <= TRI_STAGE;
        const int32_t *vals = staged ? sval : adj + sv;
        if (staged) {
            for (int i = lane; i < lenv; i += WAVE) { sval[i] = adj[sv + i]; scnt[i] = 0; }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int r = 1; r < lenv; r++) {
            const int u = vals[r];
            if (labels && labels[u] != l) continue;
            const int su = indices[u], lenu = lowpos[u];
            int found = 0;
            for (int y = lane; y < lenu; y += WAVE) {
                int wv = adj[su + y];
                int x = find_sorted(vals, r, wv);
                if (x >= 0 && (!labels || labels[wv] == l)) {
                    found++;
                    if (staged) atomicAdd(&scnt[x], 1);
                    else atomicAdd(&et[sv + x], 1);
                    atomicAdd(&et[su + y], 1);
                }
            }
            found = wave_sum(found);
            if (lane == 0 && found) {
                if (staged) scnt[r] += found;
                else atomicAdd(&et[sv + r], found);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (staged) {
            for (int i = lane; i < lenv; i += WAVE)
                if (scnt[i]) atomicAdd(&et[sv + i], scnt[i]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Vertices with prefixes up to TS: the wedge checks (r, y) of vertex v --
// u = the r-th prefix entry, w = the y-th entry of N(u) ∩ [0,u) -- are
// flattened over the lanes (an exclusive scan of the per-r lengths in LDS,
// a binary search of the lane's item in it), so a wave stays busy however
// short the rows are.  Same checks and counts as k_triangles.
constexpr int FLAT_TS = 256;  // 16 KB of LDS per workgroup: eight workgroups (32 waves) per CU

template <int TS>
__global__ __launch_bounds__(256) void k_triangles_flat(int n, const int32_t *indices, const int32_t *adj,
                                                        const int32_t *lowpos, const int32_t *labels, int32_t *et) {
    __shared__ int32_t s_val[4][TS];
    __shared__ int32_t s_cnt[4][TS];
    __shared__ int32_t s_su[4][TS];
    __shared__ int32_t s_pre[4][TS + 1];
    const int lane = __lane_id(), w = threadIdx.x / WAVE;
    int32_t *sval = s_val[w], *scnt = s_cnt[w], *ssu = s_su[w], *spre = s_pre[w];
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int v = wave_g; v < n; v += nwaves) {
        const int lenv = lowpos[v];
        if (lenv < 2 || lenv > TS) continue;
        const int sv = indices[v];
        const int l = labels ? labels[v] : 0;
        int run = 0;
        for (int base = 0; base < lenv; base += WAVE) {
            const int i = base + lane;
            int len = 0;
            if (i < lenv) {
                const int u = adj[sv + i];
                int su = 0;
                if (i >= 1 && (!labels || labels[u] == l)) {
                    su = indices[u];
                    len = lowpos[u];
                }
                sval[i] = u;
                scnt[i] = 0;
                ssu[i] = su;
            }
            const int inc = wave_incl_scan(len);
            if (i < lenv) spre[i] = run + inc - len;
            run += __shfl(inc, WAVE - 1, WAVE);
        }
        if (lane == 0) spre[lenv] = run;
        wave_sync_lds();
        for (int t = lane; t < run; t += WAVE) {
            int lo = 0, hi = lenv;  // largest r with spre[r] <= t; its range holds t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (spre[mid] <= t) lo = mid;
                else hi = mid;
            }
            const int r = lo, y = t - spre[r];
            const int su = ssu[r];
            const int wv = adj[su + y];
            const int x = find_sorted(sval, r, wv);
            if (x >= 0 && (!labels || labels[wv] == l)) {
                atomicAdd(&scnt[x], 1);
                atomicAdd(&scnt[r], 1);
                atomicAdd(&et[su + y], 1);
            }
        }
        wave_sync_lds();
        for (int i = lane; i < lenv; i += WAVE)
            if (scnt[i]) atomicAdd(&et[sv + i], scnt[i]);
        wave_sync_lds();
    }
}

// graph.cpp:146-159 (rules.cpp:197-211): per entry (v,u) with t > 0,
// t is added to both endpoints and each endpoint counts one edge.
// Prefixes of at most ACC_G * 8 entries: ACC_G lanes per vertex, 64 / ACC_G
// vertices per wave; longer prefixes: k_tri_accumulate, one wave each.
constexpr int ACC_G = 8, ACC_SMALL = ACC_G * 8;
__global__ __launch_bounds__(256) void k_tri_accumulate_small(int n, const int32_t *indices, const int32_t *adj,
                                                              const int32_t *lowpos, const int32_t *et, int32_t *tri,
                                                              int32_t *tdeg) {
    const int lane = __lane_id(), glane = lane & (ACC_G - 1);
    const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / ACC_G;
    const int ngroups = gridDim.x * blockDim.x / ACC_G;
    // the lanes of a group share v, so the group's shuffles stay inside
    // lanes that are active together
    for (int v = gid; v < n; v += ngroups) {
        const int lenv = lowpos[v];
        const bool mine = lenv <= ACC_SMALL;
        const int sv = mine ? indices[v] : 0;
        int ts = 0, tc = 0;
        for (int i = glane; mine && i < lenv; i += ACC_G) {
            int t = et[sv + i];
            if (t > 0) {
                int u = adj[sv + i];
                ts += t;
                tc++;
                atomicAdd(&tri[u], t);
                atomicAdd(&tdeg[u], 1);
            }
        }
#pragma unroll
        for (int d = ACC_G / 2; d > 0; d >>= 1) {
            ts += __shfl_xor(ts, d, WAVE);
            tc += __shfl_xor(tc, d, WAVE);
        }
        if (glane == 0 && tc) {
            atomicAdd(&tri[v], ts);
            atomicAdd(&tdeg[v], tc);
        }
    }
}

__global__ __launch_bounds__(256) void k_tri_accumulate(int n, const int32_t *indices, const int32_t *adj,
                                                        const int32_t *lowpos, const int32_t *et, int32_t *tri,
                                                        int32_t *tdeg) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int v = wave_g; v < n; v += nwaves) {
        const int lenv = lowpos[v];
        if (lenv <= ACC_SMALL) continue;  // k_tri_accumulate_small
        const int sv = indices[v];
        int ts = 0, tc = 0;
        for (int i = lane; i < lenv; i += WAVE) {
            int t = et[sv + i];
            if (t > 0) {
                int u = adj[sv + i];
                ts += t;
                tc++;
                atomicAdd(&tri[u], t);
                atomicAdd(&tdeg[u], 1);
            }
        }
        ts = wave_sum(ts);
        tc = wave_sum(tc);
        if (lane == 0 && tc) {
            atomicAdd(&tri[v], ts);
            atomicAdd(&tdeg[v], tc);
        }
    }
}

constexpr int TSUM_BLOCKS = 1024;

// graph.cpp:161-169: cc = sum over v with deg > 1 of (tri/(deg(deg-1)))/n
__global__ __launch_bounds__(256) void k_cc_partial(int n, const int32_t *indices, const int32_t *tri, double *part) {
    __shared__ double lds[256 / WAVE];
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int deg = indices[v + 1] - indices[v];
        if (deg > 1) acc += (tri[v] / (deg * (deg - 1.0))) / n;
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

// rules.cpp:213-228
__global__ __launch_bounds__(256) void k_wcc_partial(int n, const int32_t *labels, const CommStats *st,
                                                     const int32_t *tri, const int32_t *tdeg, const int32_t *it,
                                                     const int32_t *itd, double *part) {
    __shared__ double lds[256 / WAVE];
    double acc = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int t = tri[v];
        if (t > 0) {
            int size = st[labels[v]].size;
            int dt = tdeg[v];
            double a = it[v] / (double)t;
            double b = dt / (double)(size - 1 + dt - itd[v]);
            acc += (a * b) / n;
        }
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

__global__ __launch_bounds__(256) void k_sum_partials_plain(const double *part, int np, double *out) {
    __shared__ double lds[256 / WAVE];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    acc = wave_sum_f64(acc);
    if (__lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < 256 / WAVE; i++) t += lds[i];
        *out = t;
    }
}

static int ensure_tri_buffers(pcd_ctx *ctx) {
    if (ctx->d_tri) return PCD_OK;
    const int64_t E = (int64_t)2 * ctx->m;
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_tri, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_tri_deg, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_et, E + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_lowpos, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_it, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_itd, ctx->n + 1));
    hipLaunchKernelGGL(k_lowpos, dim3(grid_for(ctx, ctx->n, 256)), dim3(256), 0, ctx->stream, ctx->n, ctx->d_indices,
                       ctx->d_adj, ctx->d_lowpos);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

// Runs the counting pass (labels: nullptr for all triangles) into et and
// accumulates per-vertex totals into (t, td).
static int count_pass(pcd_ctx *ctx, const int32_t *labels, int32_t *t, int32_t *td) {
    const int n = ctx->n;
    const int64_t E = (int64_t)2 * ctx->m;
    kt_begin(ctx, KS_TRIANGLES);
    PCD_TRY(fill_i32(ctx, ctx->d_et, 0, E));
    PCD_TRY(fill_i32(ctx, t, 0, n));
    PCD_TRY(fill_i32(ctx, td, 0, n));
    hipLaunchKernelGGL(k_triangles_flat<FLAT_TS>, dim3(grid_for(ctx, n, 4, 8)), dim3(256), 0, ctx->stream, n,
                       ctx->d_indices, ctx->d_adj, ctx->d_lowpos, labels, ctx->d_et);
    hipLaunchKernelGGL(k_triangles, dim3(grid_for(ctx, n, 4, 2)), dim3(256), 0, ctx->stream, n, ctx->d_indices,
                       ctx->d_adj, ctx->d_lowpos, labels, ctx->d_et, FLAT_TS);
    hipLaunchKernelGGL(k_tri_accumulate_small, dim3(grid_for(ctx, n, 256 / ACC_G, 8)), dim3(256), 0, ctx->stream, n,
                       ctx->d_indices, ctx->d_adj, ctx->d_lowpos, ctx->d_et, t, td);
    hipLaunchKernelGGL(k_tri_accumulate, dim3(grid_for(ctx, n, 4, 8)), dim3(256), 0, ctx->stream, n, ctx->d_indices,
                       ctx->d_adj, ctx->d_lowpos, ctx->d_et, t, td);
    kt_end(ctx, KS_TRIANGLES);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

int triangles_dev(pcd_ctx *ctx) {
    PCD_TRY(ensure_tri_buffers(ctx));
    PCD_TRY(count_pass(ctx, nullptr, ctx->d_tri, ctx->d_tri_deg));
    hipLaunchKernelGGL(k_cc_partial, dim3(TSUM_BLOCKS), dim3(256), 0, ctx->stream, ctx->n, ctx->d_indices, ctx->d_tri,
                       ctx->d_dacc);
    hipLaunchKernelGGL(k_sum_partials_plain, dim3(1), dim3(256), 0, ctx->stream, ctx->d_dacc, TSUM_BLOCKS,
                       ctx->d_result);
    PCD_HIP(hipGetLastError());
    PCD_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->cc = ctx->h_result[0];
    ctx->has_triangles = true;
    return PCD_OK;
}

// WCCRule::calculate (rules.cpp:150-235) into d_result; the triangles of
// the graph must be counted.  The score is copied to the host unless a pass
// is being captured.
int wcc_metric_enqueue(pcd_ctx *ctx) {
    if (!ctx->has_triangles) return fail(PCD_E_STATE, "WCC metric needs the triangle counts");
    PCD_TRY(count_pass(ctx, ctx->d_labels, ctx->d_it, ctx->d_itd));
    hipLaunchKernelGGL(k_wcc_partial, dim3(TSUM_BLOCKS), dim3(256), 0, ctx->stream, ctx->n, ctx->d_labels,
                       ctx->d_stats, ctx->d_tri, ctx->d_tri_deg, ctx->d_it, ctx->d_itd, ctx->d_dacc);
    hipLaunchKernelGGL(k_sum_partials_plain, dim3(1), dim3(256), 0, ctx->stream, ctx->d_dacc, TSUM_BLOCKS,
                       ctx->d_result);
    PCD_HIP(hipGetLastError());
    if (!ctx->capturing)
        PCD_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return PCD_OK;
}

int wcc_metric_dev(pcd_ctx *ctx, double *out) {
    if (!ctx->has_triangles) PCD_TRY(triangles_dev(ctx));
    PCD_TRY(wcc_metric_enqueue(ctx));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_result[0];
    return PCD_OK;
}

}  // namespace pcd
