// Label propagation (PropagationRefinement::refine, refinement.cpp:156-224)
// on gfx950.
//
// Per active vertex v the step needs, for every distinct neighbour label c,
// the count of neighbours in c, then score_join(v, c, count) for c != own,
// the best of those (score, then tie key, then label), score_leave(v, own
// count) and the move decision.  Vertices are processed by degree bin:
//
//   G4..G64   deg <= G: G lanes per vertex, one neighbour label per lane,
//             bitonic sort across the lanes (shuffles), run heads found by
//             ballot, butterfly argmax.  No LDS.
//   W256      one wave per vertex, LDS hash table of nextpow2(2 deg) slots.
//   W1024     as W256 with a larger per-wave table.
//   BLOCK     one workgroup per vertex, LDS hash table (deg <= block_max).
//   HUB       rows split into chunks; each chunk is pre-aggregated in an LDS
//             table and flushed into a per-hub global hash table; one
//             workgroup per hub then scores the table.
//
// Inactive vertices (counter hash > frac) copy their label in the select
// pass, which also builds the per-bin active lists.
#include "prims.h"

namespace pcd {

struct Cand {
    double s;
    uint32_t key;
    int label;
};

__device__ inline Cand cand_none() { return Cand{-__builtin_inf(), 0xFFFFFFFFu, INT32_MAX}; }

__device__ inline bool cand_better(const Cand &a, const Cand &b) {
    if (a.s > b.s) return true;
    if (a.s < b.s) return false;
    if (a.key != b.key) return a.key < b.key;
    return a.label < b.label;
}

__device__ inline Cand cand_shfl_xor(const Cand &c, int d) {
    Cand o;
    o.s = __shfl_xor(c.s, d, WAVE);
    o.key = (uint32_t)__shfl_xor((int)c.key, d, WAVE);
    o.label = __shfl_xor(c.label, d, WAVE);
    return o;
}

struct LPArgs {
    const int32_t *indices, *adj, *labels;
    const CommStats *stats;
    int32_t *new_labels;
    RuleParams rule;
    uint64_t seed;
    int step;
    int tie;
    unsigned long long *counters;  // {active, entries, distinct} or nullptr
};

// refinement.cpp:211-221: best_score (0 if no join candidate beat 0) plus
// score_leave decides between moving, leaving to a new singleton, staying.
__device__ inline int lp_decide(const LPArgs &a, int deg, int own, int own_cnt, const Cand &best) {
    double leave = score_leave(a.rule, deg, a.stats[own], own_cnt);
    double bs = best.label != INT32_MAX ? best.s : 0.0;
    bs += leave;
    if (bs > 0 && bs > leave) return best.label;
    if (leave > 0) return PCD_NEW_LABEL;
    return own;
}

// A candidate for neighbour label c with count cnt (c != own).
__device__ inline Cand lp_candidate(const LPArgs &a, int deg, uint64_t tb, int c, int cnt) {
    double s = score_join(a.rule, deg, a.stats[c], cnt);
    if (s > 0) return Cand{s, tie_key(a.tie, tb, c), c};
    return cand_none();
}

// ------------------------------------------------------------- select
__global__ __launch_bounds__(256) void k_lp_select(int n, const int32_t *order, const int32_t *labels,
                                                   int32_t *new_labels, uint64_t seed, int step, int f, BinLayout bins,
                                                   int32_t *active, int32_t *counts, int32_t *hub_active) {
    const int lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    for (int64_t p0 = (int64_t)blockIdx.x * blockDim.x; p0 < n; p0 += (int64_t)gridDim.x * blockDim.x) {
        int p = (int)(p0 + threadIdx.x);
        bool valid = p < n;
        int v = 0, b = NUM_BINS;
        bool act = false;
        if (valid) {
            v = order ? order[p] : p;
            act = selected(seed, step, v, f);
            if (!act) new_labels[v] = labels[v];
            b = 0;
            while (p >= bins.start[b + 1]) b++;
            if (b == BIN_HUB) {
                hub_active[p - bins.start[BIN_HUB]] = act ? 1 : 0;
                act = false;
            }
        }
        uint64_t pend = __ballot(act);
        while (pend) {
            int leader = __builtin_ctzll(pend);
            int b0 = __shfl(b, leader, WAVE);
            uint64_t mine = __ballot(act && b == b0);
            int base = 0;
            if (lane == leader) base = atomicAdd(&counts[b0], __popcll(mine));
            base = __shfl(base, leader, WAVE);
            if (act && b == b0) active[bins.start[b0] + base + __popcll(mine & lt)] = v;
            pend &= ~mine;
        }
    }
}

// ------------------------------------------------------------- G lanes / vertex
template <int G>
__global__ __launch_bounds__(256) void k_lp_small(const int32_t *list, const int32_t *count_ptr, LPArgs a) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int per_wave = WAVE / G;
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int count = *count_ptr;
    for (int base = wave_g * per_wave; base < count; base += nwaves * per_wave) {
        int gi = base + lane / G;
        bool has = gi < count;
        int v = 0, start = 0, deg = 0, own = 0;
        if (has) {
            v = list[gi];
            start = a.indices[v];
            deg = a.indices[v + 1] - start;
            own = a.labels[v];
        }
        int x = (has && glane < deg) ? a.labels[a.adj[start + glane]] : INT32_MAX;
        // bitonic sort (ascending) across the G lanes of the group
#pragma unroll
        for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                int y = __shfl_xor(x, j, WAVE);
                bool up = (glane & k) == 0;
                bool lower = (glane & j) == 0;
                x = (lower == up) ? min(x, y) : max(x, y);
            }
        }
        int prev = __shfl(x, lane == 0 ? 0 : lane - 1, WAVE);
        bool head = x != INT32_MAX && (glane == 0 || x != prev);
        uint64_t heads = __ballot(head);
        Cand c = cand_none();
        int own_cnt = 0;
        if (head) {
            uint64_t above = lane == 63 ? 0ULL : (heads >> (lane + 1));
            int nxt = above ? lane + 1 + __builtin_ctzll(above) : WAVE;
            int gend = (lane & ~(G - 1)) + deg;
            int cnt = min(nxt, gend) - lane;
            if (x == own) own_cnt = cnt;
            else c = lp_candidate(a, deg, a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, a.step, v), x, cnt);
        }
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) {
            Cand o = cand_shfl_xor(c, d);
            if (cand_better(o, c)) c = o;
            own_cnt = max(own_cnt, __shfl_xor(own_cnt, d, WAVE));
        }
        if (has && glane == 0) a.new_labels[v] = lp_decide(a, deg, own, own_cnt, c);
        if (a.counters) {
            int nh = __popcll(heads);
            int nact = __popcll(__ballot(has && glane == 0));
            int ent = (has && glane < deg) ? 1 : 0;
            ent = wave_sum(ent);
            if (lane == 0) {
                atomicAdd(&a.counters[0], (unsigned long long)nact);
                atomicAdd(&a.counters[1], (unsigned long long)ent);
                atomicAdd(&a.counters[2], (unsigned long long)nh);
            }
        }
    }
}

// ------------------------------------------------------------- LDS hash helpers
__device__ inline uint32_t hslot(int key, int log_t) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32 - log_t);
}

__device__ inline void lds_insert(int *keys, int *cnts, int mask, int log_t, int key, int add) {
    uint32_t h = hslot(key, log_t);
    while (true) {
        int old = atomicCAS(&keys[h], EMPTY_KEY, key);
        if (old == EMPTY_KEY || old == key) {
            atomicAdd(&cnts[h], add);
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int ceil_log2(int x) { return x <= 1 ? 0 : 32 - __clz(x - 1); }

// Reduce (Cand, own_cnt, distinct) across a wave; result valid in all lanes.
__device__ inline void wave_reduce_cand(Cand &c, int &own_cnt) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        Cand o = cand_shfl_xor(c, d);
        if (cand_better(o, c)) c = o;
        own_cnt = max(own_cnt, __shfl_xor(own_cnt, d, WAVE));
    }
}

// ------------------------------------------------------------- wave / vertex
template <int TW>
__global__ __launch_bounds__(256) void k_lp_wave(const int32_t *list, const int32_t *count_ptr, LPArgs a) {
    __shared__ int s_keys[4][TW];
    __shared__ int s_cnts[4][TW];
    const int lane = __lane_id(), w = threadIdx.x / WAVE;
    int *keys = s_keys[w], *cnts = s_cnts[w];
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int count = *count_ptr;
    for (int gi = wave_g; gi < count; gi += nwaves) {
        int v = list[gi];
        int start = a.indices[v], deg = a.indices[v + 1] - start, own = a.labels[v];
        int log_t = max(6, ceil_log2(2 * deg));
        int T = 1 << log_t, mask = T - 1;
        for (int s = lane; s < T; s += WAVE) { keys[s] = EMPTY_KEY; cnts[s] = 0; }
        wave_sync();
        for (int j = lane; j < deg; j += WAVE) lds_insert(keys, cnts, mask, log_t, a.labels[a.adj[start + j]], 1);
        wave_sync();
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, a.step, v);
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        for (int s = lane; s < T; s += WAVE) {
            int key = keys[s];
            if (key == EMPTY_KEY) continue;
            distinct++;
            int cnt = cnts[s];
            if (key == own) own_cnt = cnt;
            else {
                Cand o = lp_candidate(a, deg, tb, key, cnt);
                if (cand_better(o, c)) c = o;
            }
        }
        wave_reduce_cand(c, own_cnt);
        if (lane == 0) a.new_labels[v] = lp_decide(a, deg, own, own_cnt, c);
        if (a.counters) {
            distinct = wave_sum(distinct);
            if (lane == 0) {
                atomicAdd(&a.counters[0], 1ULL);
                atomicAdd(&a.counters[1], (unsigned long long)deg);
                atomicAdd(&a.counters[2], (unsigned long long)distinct);
            }
        }
        wave_sync();
    }
}

// ------------------------------------------------------------- block reduce
template <int NT>
__device__ inline void block_reduce_cand(Cand &c, int &own_cnt, int &distinct, Cand *s_c, int *s_i) {
    const int lane = __lane_id(), w = threadIdx.x / WAVE;
    wave_reduce_cand(c, own_cnt);
    distinct = wave_sum(distinct);
    if (lane == 0) { s_c[w] = c; s_i[w] = own_cnt; s_i[NT / WAVE + w] = distinct; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < NT / WAVE; i++) {
            if (cand_better(s_c[i], c)) c = s_c[i];
            own_cnt = max(own_cnt, s_i[i]);
            distinct += s_i[NT / WAVE + i];
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------- block / vertex
template <int NT, int TB>
__global__ __launch_bounds__(NT) void k_lp_block(const int32_t *list, const int32_t *count_ptr, LPArgs a) {
    __shared__ int keys[TB];
    __shared__ int cnts[TB];
    __shared__ Cand s_c[NT / WAVE];
    __shared__ int s_i[2 * NT / WAVE];
    const int count = *count_ptr;
    for (int gi = blockIdx.x; gi < count; gi += gridDim.x) {
        int v = list[gi];
        int start = a.indices[v], deg = a.indices[v + 1] - start, own = a.labels[v];
        int log_t = max(6, ceil_log2(2 * deg));
        int T = 1 << log_t, mask = T - 1;
        for (int s = threadIdx.x; s < T; s += NT) { keys[s] = EMPTY_KEY; cnts[s] = 0; }
        __syncthreads();
        for (int j = threadIdx.x; j < deg; j += NT) lds_insert(keys, cnts, mask, log_t, a.labels[a.adj[start + j]], 1);
        __syncthreads();
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, a.step, v);
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        for (int s = threadIdx.x; s < T; s += NT) {
            int key = keys[s];
            if (key == EMPTY_KEY) continue;
            distinct++;
            int cnt = cnts[s];
            if (key == own) own_cnt = cnt;
            else {
                Cand o = lp_candidate(a, deg, tb, key, cnt);
                if (cand_better(o, c)) c = o;
            }
        }
        block_reduce_cand<NT>(c, own_cnt, distinct, s_c, s_i);
        if (threadIdx.x == 0) {
            a.new_labels[v] = lp_decide(a, deg, own, own_cnt, c);
            if (a.counters) {
                atomicAdd(&a.counters[0], 1ULL);
                atomicAdd(&a.counters[1], (unsigned long long)deg);
                atomicAdd(&a.counters[2], (unsigned long long)distinct);
            }
        }
    }
}

// ------------------------------------------------------------- hubs
// items: {hub index, v, begin, end}.  LDS pre-aggregation of one chunk,
// then flush into the hub's global table (open addressing, CAS on keys).
template <int NT, int TB>
__global__ __launch_bounds__(NT) void k_hub_insert(int nitems, const int4 *items, const int32_t *hub_active,
                                                   const int64_t *toff, const int32_t *tlog, int32_t *gkeys,
                                                   int32_t *gcnts, const int32_t *adj, const int32_t *labels) {
    __shared__ int keys[TB];
    __shared__ int cnts[TB];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        int4 w = items[it];
        if (!hub_active[w.x]) continue;
        const int len = w.w - w.z;
        const int log_t = max(6, ceil_log2(2 * len));
        const int T = 1 << log_t, mask = T - 1;
        for (int s = threadIdx.x; s < T; s += NT) { keys[s] = EMPTY_KEY; cnts[s] = 0; }
        __syncthreads();
        for (int j = w.z + threadIdx.x; j < w.w; j += NT) lds_insert(keys, cnts, mask, log_t, labels[adj[j]], 1);
        __syncthreads();
        const int64_t off = toff[w.x];
        const int glog = tlog[w.x];
        const int gmask = (1 << glog) - 1;
        for (int s = threadIdx.x; s < T; s += NT) {
            int key = keys[s];
            if (key == EMPTY_KEY) continue;
            int add = cnts[s];
            uint32_t h = hslot(key, glog);
            while (true) {
                int old = atomicCAS(&gkeys[off + h], EMPTY_KEY, key);
                if (old == EMPTY_KEY || old == key) {
                    atomicAdd(&gcnts[off + h], add);
                    break;
                }
                h = (h + 1) & gmask;
            }
        }
        __syncthreads();
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_hub_decide(int p0, int nhubs, const int32_t *order, const int32_t *hub_active,
                                                   const int64_t *toff, const int32_t *tlog, int32_t *gkeys,
                                                   int32_t *gcnts, LPArgs a) {
    __shared__ Cand s_c[NT / WAVE];
    __shared__ int s_i[2 * NT / WAVE];
    for (int h = blockIdx.x; h < nhubs; h += gridDim.x) {
        if (!hub_active[h]) continue;
        int p = p0 + h;
        int v = order ? order[p] : p;
        int start = a.indices[v], deg = a.indices[v + 1] - start, own = a.labels[v];
        const int64_t off = toff[h];
        const int T = 1 << tlog[h];
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, a.step, v);
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        for (int s = threadIdx.x; s < T; s += NT) {
            int key = gkeys[off + s];
            if (key == EMPTY_KEY) continue;
            int cnt = gcnts[off + s];
            gkeys[off + s] = EMPTY_KEY;  // leave the table clean for the next step
            gcnts[off + s] = 0;
            distinct++;
            if (key == own) own_cnt = cnt;
            else {
                Cand o = lp_candidate(a, deg, tb, key, cnt);
                if (cand_better(o, c)) c = o;
            }
        }
        block_reduce_cand<NT>(c, own_cnt, distinct, s_c, s_i);
        if (threadIdx.x == 0) {
            a.new_labels[v] = lp_decide(a, deg, own, own_cnt, c);
            if (a.counters) {
                atomicAdd(&a.counters[0], 1ULL);
                atomicAdd(&a.counters[1], (unsigned long long)deg);
                atomicAdd(&a.counters[2], (unsigned long long)distinct);
            }
        }
    }
}

// ------------------------------------------------------------- driver
template <int G>
static void launch_small(pcd_ctx *ctx, int bin, const LPArgs &a) {
    const BinLayout &b = ctx->bins;
    int cap = b.start[bin + 1] - b.start[bin];
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_lp_small<G>, dim3(grid_for(ctx, cap, 256 / G, 16)), dim3(256), 0, ctx->stream,
                       ctx->d_active + b.start[bin], ctx->d_counts + bin, a);
}

constexpr int HUB_NT = 512, HUB_TB = 8192;
constexpr int BLK_NT = 512, BLK_TB = 8192;

int propagation_dev(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie) {
    const int n = ctx->n;
    const BinLayout &b = ctx->bins;
    double fr = frac < 1.0 ? frac : 1.0;
    int f = (int)(fr * 2147483647.0);  // refinement.cpp:158
    LPArgs a;
    a.indices = ctx->d_indices;
    a.adj = ctx->d_adj;
    a.labels = ctx->d_labels;
    a.stats = ctx->d_stats;
    a.new_labels = ctx->d_labels_alt;
    a.rule = make_rule(ctx, rule);
    a.seed = seed;
    a.step = step;
    a.tie = tie;
    a.counters = nullptr;
    unsigned long long *cnt64 = (unsigned long long *)(ctx->d_acc + 4);
    if (ctx->counters) {
        PCD_TRY(fill_u64(ctx, (uint64_t *)cnt64, 0, 3));
        a.counters = cnt64;
    }
    PCD_TRY(fill_i32(ctx, ctx->d_counts, 0, NUM_BINS));
    hipLaunchKernelGGL(k_lp_select, dim3(grid_for(ctx, n, 256, 16)), dim3(256), 0, ctx->stream, n, ctx->d_order,
                       ctx->d_labels, ctx->d_labels_alt, seed, step, f, b, ctx->d_active, ctx->d_counts,
                       ctx->d_hub_active);
    if (ctx->num_hubs > 0) {
        hipLaunchKernelGGL((k_hub_insert<HUB_NT, HUB_TB>), dim3(grid_for(ctx, ctx->num_hub_items, 1, 2)),
                           dim3(HUB_NT), 0, ctx->stream, ctx->num_hub_items, ctx->d_hub_items, ctx->d_hub_active,
                           ctx->d_hub_table_off, ctx->d_hub_table_log, ctx->d_hub_keys, ctx->d_hub_cnts, ctx->d_adj,
                           ctx->d_labels);
        hipLaunchKernelGGL(k_hub_decide<HUB_NT>, dim3(grid_for(ctx, ctx->num_hubs, 1, 2)), dim3(HUB_NT), 0,
                           ctx->stream, b.start[BIN_HUB], ctx->num_hubs, ctx->d_order, ctx->d_hub_active,
                           ctx->d_hub_table_off, ctx->d_hub_table_log, ctx->d_hub_keys, ctx->d_hub_cnts, a);
    }
    {
        int cap = b.start[BIN_BLOCK + 1] - b.start[BIN_BLOCK];
        if (cap > 0)
            hipLaunchKernelGGL((k_lp_block<BLK_NT, BLK_TB>), dim3(grid_for(ctx, cap, 1, 2)), dim3(BLK_NT), 0,
                               ctx->stream, ctx->d_active + b.start[BIN_BLOCK], ctx->d_counts + BIN_BLOCK, a);
    }
    {
        int cap = b.start[BIN_W1024 + 1] - b.start[BIN_W1024];
        if (cap > 0)
            hipLaunchKernelGGL(k_lp_wave<2048>, dim3(grid_for(ctx, cap, 4, 2)), dim3(256), 0, ctx->stream,
                               ctx->d_active + b.start[BIN_W1024], ctx->d_counts + BIN_W1024, a);
    }
    {
        int cap = b.start[BIN_W256 + 1] - b.start[BIN_W256];
        if (cap > 0)
            hipLaunchKernelGGL(k_lp_wave<512>, dim3(grid_for(ctx, cap, 4, 8)), dim3(256), 0, ctx->stream,
                               ctx->d_active + b.start[BIN_W256], ctx->d_counts + BIN_W256, a);
    }
    launch_small<64>(ctx, BIN_G64, a);
    launch_small<32>(ctx, BIN_G32, a);
    launch_small<16>(ctx, BIN_G16, a);
    launch_small<8>(ctx, BIN_G8, a);
    launch_small<4>(ctx, BIN_G4, a);
    PCD_HIP(hipGetLastError());
    if (ctx->counters) {
        // active counts of hubs are included by k_hub_decide
        PCD_HIP(hipMemcpyAsync(ctx->h_small64, cnt64, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        ctx->cnt_active = ctx->h_small64[0];
        ctx->cnt_entries = ctx->h_small64[1];
        ctx->cnt_distinct = ctx->h_small64[2];
    }
    std::swap(ctx->d_labels, ctx->d_labels_alt);
    ctx->labels_compressed = false;
    ctx->has_stats = false;
    ctx->label_bound = ctx->k;
    return PCD_OK;
}

}  // namespace pcd
