// Parallel merging (MergingRefinement::refine, refinement.cpp:185-288) and
// contraction on gfx950.
//
// Contraction: every CSR entry (v, u) writes the 2b-bit key (l_v << b | l_u)
// (or an all-ones sentinel when l_v == l_u) into a buffer indexed by the
// entry itself, the buffer is radix-sorted, and a run-length reduce-by-key
// (scan of run heads) gives the community graph as unique (l, k, links)
// sorted by (l, k): the table refinement.cpp:228-238 hashes per community
// and framework-cuda.cu:744-807 sorts.
//
// Merging round: each unmatched, eligible community l takes the neighbour
// community k (unmatched) of best score_merge >= 0 (refinement.cpp:240-260;
// a zero score is accepted, as there); handshake (refinement.cpp:267-274);
// relabel to min(l, partner) (refinement.cpp:277-286).
//
// Two forms of the partner choice:
//  * community rows (the default): as refinement.cpp:222-238 states it,
//    each community counts the labels of its members' neighbours.  Once per
//    step every entry's label is gathered into its community's row (rows in
//    label order, members by a stable radix sort of the vertices by label),
//    and each round runs the propagation kernels in their merging form over
//    the rows, binned by length.  No sort of the entries.
//  * sorted pairs (tunable merge_sort): the contraction above, then one pass
//    over the unique pairs with two order-independent atomics per l (max of
//    the order-preserving score bits, then min of (tie key, k)).
// Both choose the same partners: (score, tie key, label) order is total.
#include <algorithm>
#include <type_traits>

#include "work.h"

namespace pcd {

// ---------------------------------------------------------------- keys
template <int G>
__global__ __launch_bounds__(256) void k_pair_keys_small(int p0, int p1, const int32_t *order, const int32_t *indices,
                                                         const int32_t *adj, const int32_t *labels, int bits,
                                                         uint64_t *keys) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int per_wave = WAVE / G;
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int count = p1 - p0;
    const uint64_t sentinel = (bits >= 64) ? ~0ULL : ((1ULL << bits) - 1);
    const int half = bits / 2;
    for (int base = wave_g * per_wave; base < count; base += nwaves * per_wave) {
        int gi = base + lane / G;
        if (gi >= count) continue;
        int p = p0 + gi;
        int v = order ? order[p] : p;
        int start = indices[v], deg = indices[v + 1] - start, l = labels[v];
        for (int j = glane; j < deg; j += G) {
            int kk = labels[adj[start + j]];
            keys[start + j] = kk == l ? sentinel : (((uint64_t)(uint32_t)l << half) | (uint32_t)kk);
        }
    }
}

__global__ __launch_bounds__(256) void k_pair_keys_items(int nitems, const int4 *items, const int32_t *adj,
                                                         const int32_t *labels, int bits, uint64_t *keys) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const uint64_t sentinel = (bits >= 64) ? ~0ULL : ((1ULL << bits) - 1);
    const int half = bits / 2;
    for (int it = wave_g; it < nitems; it += nwaves) {
        int4 w = items[it];
        int l = labels[w.x];
        for (int j = w.y + lane; j < w.z; j += WAVE) {
            int kk = labels[adj[j]];
            keys[j] = kk == l ? sentinel : (((uint64_t)(uint32_t)l << half) | (uint32_t)kk);
        }
    }
}

// run heads of the first P sorted keys
struct HeadFlag {
    const uint64_t *keys;
    __device__ int operator()(int64_t i) const { return (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0; }
};
struct HeadStore {
    const uint64_t *keys;
    uint64_t *ukeys;
    int32_t *hpos;
    __device__ void operator()(int64_t i, int ex, int flag) const {
        if (flag) {
            ukeys[ex] = keys[i];
            hpos[ex] = (int)i;
        }
    }
};

__global__ void k_run_lengths(const int32_t *Up, const int32_t *hpos, int P, int32_t *links) {
    const int U = *Up;
    for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < U; u += gridDim.x * blockDim.x)
        links[u] = (u + 1 < U ? hpos[u + 1] : P) - hpos[u];
}

__global__ void k_sum_int_vol(const CommStats *st, const int32_t *kp, unsigned long long *acc) {
    const int k = *kp;
    long long a = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x) a += st[c].int_vol;
    a = wave_sum64(a);
    if (__lane_id() == 0) atomicAdd(acc, (unsigned long long)a);
}

static int label_bits(int k) {
    int b = 1;
    while (((int64_t)1 << b) - 1 < (int64_t)k) b++;  // all-ones never a label
    return b;
}

template <int G>
static void launch_keys_small(pcd_ctx *ctx, int bin, int bits) {
    const BinLayout &b = ctx->bins;
    int p0 = b.start[bin], p1 = b.start[bin + 1];
    if (p1 <= p0) return;
    hipLaunchKernelGGL(k_pair_keys_small<G>, dim3(grid_for(ctx, p1 - p0, 256 / G, 16)), dim3(256), 0, ctx->stream, p0,
                       p1, ctx->d_order, ctx->d_indices, ctx->d_adj, ctx->d_labels, bits, ctx->d_keys0);
}

// Builds the sorted unique (l, k) keys in ctx->d_keys1[0..U) with link
// counts in ctx->d_pair_links.  Returns U and the key layout.
static int build_pairs(pcd_ctx *ctx, int64_t *U_out, int *half_out) {
    const int64_t E = (int64_t)2 * ctx->m;
    if ((size_t)E > ctx->keys_cap) {
        if (ctx->d_keys0) dev_free(ctx, ctx->d_keys0);
        if (ctx->d_keys1) dev_free(ctx, ctx->d_keys1);
        if (ctx->d_pair_links) dev_free(ctx, ctx->d_pair_links);
        if (ctx->d_pair_head) dev_free(ctx, ctx->d_pair_head);
        ctx->d_keys0 = ctx->d_keys1 = nullptr;
        ctx->d_pair_links = ctx->d_pair_head = nullptr;
        ctx->keys_cap = 0;
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_keys0, E + 1));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_keys1, E + 1));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_pair_links, E + 1));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_pair_head, E + 1));
        ctx->keys_cap = (size_t)E;
    }
    const int half = label_bits(ctx->k);
    const int bits = 2 * half;
    kt_begin(ctx, KS_MERGE_KEYS);
    if (ctx->num_stat_items > 0)
        hipLaunchKernelGGL(k_pair_keys_items, dim3(grid_for(ctx, ctx->num_stat_items, 4, 16)), dim3(256), 0,
                           ctx->stream, ctx->num_stat_items, ctx->d_stat_items, ctx->d_adj, ctx->d_labels, bits,
                           ctx->d_keys0);
    launch_keys_small<64>(ctx, BIN_G64, bits);
    launch_keys_small<32>(ctx, BIN_G32, bits);
    launch_keys_small<16>(ctx, BIN_G16, bits);
    launch_keys_small<8>(ctx, BIN_G8, bits);
    launch_keys_small<4>(ctx, BIN_G4, bits);
    // P = E - sum(int_volume): entries between different communities
    unsigned long long *acc = (unsigned long long *)ctx->d_acc;
    PCD_TRY(fill_u64(ctx, (uint64_t *)acc, 0, 1));
    hipLaunchKernelGGL(k_sum_int_vol, dim3(grid_for(ctx, ctx->k, 256, 4)), dim3(256), 0, ctx->stream, ctx->d_stats,
                       ctx->d_k, acc);
    kt_end(ctx, KS_MERGE_KEYS);
    PCD_HIP(hipGetLastError());
    uint64_t *sorted = nullptr;
    kt_begin(ctx, KS_MERGE_SORT);
    PCD_TRY(radix_sort_u64(ctx, ctx->d_keys0, ctx->d_keys1, E, bits, &sorted));
    kt_end(ctx, KS_MERGE_SORT);
    PCD_HIP(hipMemcpyAsync(ctx->h_small64, acc, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t P = E - ctx->h_small64[0];
    uint64_t *ukeys = sorted == ctx->d_keys0 ? ctx->d_keys1 : ctx->d_keys0;
    int32_t *total = nullptr;
    kt_begin(ctx, KS_MERGE_RLE);
    PCD_TRY(device_scan(ctx, P, HeadFlag{sorted}, HeadStore{sorted, ukeys, ctx->d_pair_head}, &total));
    hipLaunchKernelGGL(k_run_lengths, dim3(grid_for(ctx, P, 256)), dim3(256), 0, ctx->stream, total,
                       ctx->d_pair_head, (int)P, ctx->d_pair_links);
    kt_end(ctx, KS_MERGE_RLE);
    PCD_HIP(hipGetLastError());
    PCD_HIP(hipMemcpyAsync(ctx->h_small, total, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    int64_t U = P > 0 ? ctx->h_small[0] : 0;
    // keep the unique keys in d_keys1
    if (ukeys != ctx->d_keys1 && U > 0)
        PCD_HIP(hipMemcpyAsync(ctx->d_keys1, ukeys, U * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
    *U_out = U;
    *half_out = half;
    return PCD_OK;
}

__global__ void k_unpack_pairs(const uint64_t *ukeys, int64_t U, int half, int32_t *l, int32_t *k) {
    const uint64_t mask = (1ULL << half) - 1;
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < U; u += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key = ukeys[u];
        l[u] = (int32_t)(key >> half);
        k[u] = (int32_t)(key & mask);
    }
}

// ---------------------------------------------------------------- rounds
__device__ inline bool eligible(int l, const int32_t *found, const int32_t *bp) {
    // refinement.cpp:224: skip matched communities and ones waiting on an
    // unmatched partner
    if (found[l]) return false;
    int b = bp[l];
    return !(b != -1 && !found[b]);
}

// order-preserving bits of a score >= 0 (-0.0 counts as +0.0); 0 = none
__device__ inline uint64_t score_ord(double s) {
    if (s == 0.0) s = 0.0;
    return (uint64_t)__double_as_longlong(s) | 0x8000000000000000ULL;
}

struct MergeArgs {
    const uint64_t *ukeys;
    const int32_t *links;
    int64_t U;
    int half;
    const CommStats *stats;
    const int32_t *found, *bp;
    RuleParams rule;
    uint64_t seed;  // per-round seed
    int step, tie;
    unsigned long long *best_ord, *best_key;
};

// Unique pairs arrive sorted by l, so the pairs of one community are a
// contiguous run.  One pass: each wave reduces its runs lexicographically
// (best score, then smallest tie key) with a segmented shuffle scan from the
// run heads.  A run that lies inside the wave is complete: its last lane
// stores the result.  A run that touches the wave's first or last lane may
// continue in a neighbouring wave: its partial result goes to best_ord with
// atomicMax and to one of two candidate slots of its 64 pairs, and a second
// pass over the slots keeps the smallest key among the maxima.
__device__ inline uint64_t shfl_up_u64(uint64_t x, int d) {
    int lo = __shfl_up((int)(uint32_t)x, d, WAVE), hi = __shfl_up((int)(uint32_t)(x >> 32), d, WAVE);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

struct MergeCand {
    uint64_t ord, key;
    int l, pad;
};

__global__ __launch_bounds__(256) void k_merge_pass(MergeArgs a, MergeCand *cand) {
    const uint64_t mask = (1ULL << a.half) - 1;
    const int lane = __lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // whole waves iterate together (the scan needs every lane)
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < a.U; base += stride) {
        const int64_t u = base + lane;
        const bool in = u < a.U;
        int l = -1, kk = 0;
        uint64_t ord = 0, key = ~0ULL;  // "no candidate"
        if (in) {
            uint64_t k64 = a.ukeys[u];
            l = (int)(k64 >> a.half);
            kk = (int)(k64 & mask);
            if (eligible(l, a.found, a.bp) && !a.found[kk]) {
                double s = score_merge(a.rule, a.stats[l], a.stats[kk], a.links[u]);
                if (s >= 0.0) {  // negative or NaN never beats best = 0
                    ord = score_ord(s);
                    uint32_t tk = tie_key(a.tie, a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, a.step, l), kk);
                    key = ((uint64_t)tk << 32) | (uint32_t)kk;
                }
            }
        }
        const int prev = __shfl_up(l, 1, WAVE);
        const uint64_t heads = __ballot(lane == 0 || l != prev);
        const int seg_start = 63 - __clzll(heads & (lane == 63 ? ~0ULL : ((2ULL << lane) - 1)));
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            uint64_t o = shfl_up_u64(ord, d), ok = shfl_up_u64(key, d);
            if (lane - d >= seg_start && (o > ord || (o == ord && ok < key))) {
                ord = o;
                key = ok;
            }
        }
        const bool tail = lane == WAVE - 1 || ((heads >> (lane + 1)) & 1);
        const bool open = seg_start == 0 || lane == WAVE - 1;  // may continue in another wave
        const bool has = in && tail && ord != 0;
        if (has && !open) {
            a.best_ord[l] = ord;
            a.best_key[l] = key;
        }
        // slot 2w: the run holding lane 0; slot 2w+1: the run holding lane 63
        // when that is another run; a slot without a candidate is marked empty
        const int64_t w2 = 2 * (base / WAVE);
        const bool emit_a = has && open && seg_start == 0;
        const bool emit_b = has && open && seg_start > 0;
        const uint64_t ea = __ballot(emit_a), eb = __ballot(emit_b);
        if (emit_a || emit_b) {
            atomicMax(&a.best_ord[l], (unsigned long long)ord);
            cand[w2 + (emit_a ? 0 : 1)] = MergeCand{ord, key, l, 0};
        }
        if (lane == 0 && !ea) cand[w2].ord = 0;
        if (lane == 0 && !eb) cand[w2 + 1].ord = 0;
    }
}

// among the partial results of runs that cross waves: smallest key at the maximum score
__global__ void k_merge_resolve(MergeArgs a, const MergeCand *cand, int64_t nc) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
        MergeCand c = cand[i];
        if (c.ord != 0 && c.ord == a.best_ord[c.l]) atomicMin(&a.best_key[c.l], (unsigned long long)c.key);
    }
}

__global__ void k_merge_finalize(int k, const int32_t *found, int32_t *bp, unsigned long long *best_ord,
                                 unsigned long long *best_key) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        if (eligible(l, found, bp)) {
            unsigned long long bk = best_key[l];
            bp[l] = bk == ~0ULL ? -1 : (int32_t)(uint32_t)(bk & 0xFFFFFFFFULL);
        }
        best_ord[l] = 0;
        best_key[l] = ~0ULL;
    }
}

__global__ void k_handshake(int k, int32_t *found, const int32_t *bp, int32_t *unmatched) {
    int local = 0;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        if (!found[l] && (bp[l] == -1 || bp[bp[l]] == l)) found[l] = 1;
        if (!found[l]) local++;
    }
    local = wave_sum(local);
    if (__lane_id() == 0 && local) atomicAdd(unmatched, local);
}

__global__ void k_merge_relabel(int n, const int32_t *labels, const int32_t *found, const int32_t *bp, int32_t *out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int l = labels[v];
        int b = bp[l];
        out[v] = (found[l] && b != -1) ? min(l, b) : l;
    }
}

// ------------------------------------------------------ community rows
// A merging step over community rows is laid out by kernels: no size is
// read back, so the step can be queued behind others or captured in a
// device-resident pass.  Community ids lie below `bound` (the label bound:
// k, or n + 1 while k stays on the device, when ids from *kp on are empty).
__device__ inline int k_of(int bound, const int *kp) { return kp ? min(*kp, bound) : bound; }

__global__ void k_member_keys(int n, const int32_t *labels, uint64_t *keys) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        keys[v] = ((uint64_t)(uint32_t)v << 32) | (uint32_t)labels[v];
}

// Community l's row starts at the volume (int_volume + ext_volume, the sum
// of its members' degrees) of the communities below l; an empty community
// has an empty row.
struct CommVol {
    const CommStats *st;
    int bound;
    const int *kp;
    __device__ int operator()(int64_t l) const {
        return (int)l < k_of(bound, kp) ? st[l].int_vol + st[l].ext_vol : 0;
    }
};
struct CommStart {
    int32_t *cind;
    int64_t bound;
    __device__ void operator()(int64_t l, int ex, int vol) const {
        cind[l] = ex;
        if (l == bound - 1) cind[bound] = ex + vol;
    }
};

// Member i of the label-sorted vertex list: an exclusive scan of the
// members' degrees in that order puts each member's entries at its place in
// its community's row.
struct MemberDeg {
    const uint64_t *members;
    const int32_t *indices;
    __device__ int operator()(int64_t i) const {
        const int v = (int)(members[i] >> 32);
        return indices[v + 1] - indices[v];
    }
};
struct MemberDst {
    const uint64_t *members;
    int32_t *dst;
    __device__ void operator()(int64_t i, int ex, int) const { dst[(int)(members[i] >> 32)] = ex; }
};

// every entry's label, written into the row of its vertex's community
template <int G>
__global__ __launch_bounds__(256) void k_member_rows_small(int p0, int p1, const int32_t *order, const int32_t *indices,
                                                           const int32_t *adj, const int32_t *labels,
                                                           const int32_t *dst, int32_t *ent) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int per_wave = WAVE / G;
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int count = p1 - p0;
    for (int base = wave_g * per_wave; base < count; base += nwaves * per_wave) {
        const int gi = base + lane / G;
        if (gi >= count) continue;
        const int p = p0 + gi;
        const int v = order ? order[p] : p;
        const int start = indices[v], deg = indices[v + 1] - start, o = dst[v];
        for (int j = glane; j < deg; j += G) ent[o + j] = labels[adj[start + j]];
    }
}

__global__ __launch_bounds__(256) void k_member_rows_items(int nitems, const int4 *items, const int32_t *indices,
                                                           const int32_t *adj, const int32_t *labels,
                                                           const int32_t *dst, int32_t *ent) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    constexpr int PF = 8;  // neighbour-label gathers in flight per lane
    for (int it = wave_g; it < nitems; it += nwaves) {
        const int4 w = items[it];
        const int o = dst[w.x] - indices[w.x];
        for (int j0 = w.y + lane; j0 < w.z; j0 += PF * WAVE) {
            int nb[PF];
#pragma unroll
            for (int e = 0; e < PF; e++) {
                const int j = j0 + e * WAVE;
                nb[e] = j < w.z ? adj[j] : -1;
            }
            int lab[PF];
#pragma unroll
            for (int e = 0; e < PF; e++) lab[e] = nb[e] >= 0 ? labels[nb[e]] : 0;
#pragma unroll
            for (int e = 0; e < PF; e++)
                if (nb[e] >= 0) ent[o + j0 + e * WAVE] = lab[e];
        }
    }
}

// bin of a row of `len` entries (the vertex bins' degree bounds)
__device__ inline int row_bin(int len) {
    if (len > 4096) return BIN_HUB;
    if (len > 1024) return BIN_B4096;
    if (len > 512) return BIN_B1024;
    if (len > 256) return BIN_B512;
    if (len > 64) return BIN_W256;
    if (len > 32) return BIN_G64;
    if (len > 16) return BIN_G32;
    if (len > 8) return BIN_G16;
    if (len > 4) return BIN_G8;
    return BIN_G4;
}

// Per community id below `bound`: its bin as the sort key (the id above);
// the ids from k on get bin NUM_BINS and sort after every bin.  The rows
// and entries of each bin are counted.
__global__ __launch_bounds__(256) void k_comm_bins(int bound, const int *kp, const int32_t *cind, uint64_t *keys,
                                                   unsigned long long *bcount) {
    __shared__ unsigned long long rows[NUM_BINS], ents[NUM_BINS];
    if (threadIdx.x < NUM_BINS) {
        rows[threadIdx.x] = 0;
        ents[threadIdx.x] = 0;
    }
    __syncthreads();
    const int k = k_of(bound, kp);
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < bound; l += gridDim.x * blockDim.x) {
        const int len = cind[l + 1] - cind[l];
        const int b = l < k ? row_bin(len) : NUM_BINS;
        keys[l] = ((uint64_t)(uint32_t)l << 32) | (uint32_t)b;
        if (b < NUM_BINS) {
            atomicAdd(&rows[b], 1ULL);
            atomicAdd(&ents[b], (unsigned long long)len);
        }
    }
    __syncthreads();
    if (threadIdx.x < NUM_BINS) {
        if (rows[threadIdx.x]) atomicAdd(&bcount[threadIdx.x], rows[threadIdx.x]);
        if (ents[threadIdx.x]) atomicAdd(&bcount[NUM_BINS + threadIdx.x], ents[threadIdx.x]);
    }
}

__global__ void k_high_words(int k, const uint64_t *keys, int32_t *out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
        out[i] = (int32_t)(keys[i] >> 32);
}

// One thread: the bin starts (positions in the sorted order) and the
// fused small-row launch's split of cu16 workgroups (one per non-empty bin,
// the rest in proportion to the bins' entries), as the host computes them
// for the vertex bins.
__global__ void k_row_layout(const unsigned long long *bcount, int cu16, int *layout) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    long long at = 0;
    for (int b = 0; b < NUM_BINS; b++) {
        layout[b] = (int)at;
        at += (long long)bcount[b];
    }
    layout[NUM_BINS] = (int)at;
    int *lo = layout + NUM_BINS + 1;
    int nonempty = 0;
    long long tot = 0;
    for (int i = 0; i < 6; i++) {
        nonempty += bcount[BIN_W256 + i] > 0;
        tot += (long long)bcount[NUM_BINS + BIN_W256 + i];
    }
    const long long spare = cu16 - nonempty;
    int a = 0;
    long long acc = 0;
    for (int i = 0; i < 6; i++) {
        lo[i] = a;
        if (bcount[BIN_W256 + i] > 0) {
            const int before = tot > 0 ? (int)(spare * acc / tot) : 0;
            acc += (long long)bcount[NUM_BINS + BIN_W256 + i];
            const int after = tot > 0 ? (int)(spare * acc / tot) : 0;
            a += 1 + (after - before);
        }
    }
    lo[6] = a;
}

// The hub layout of the rows above 4096 entries (positions [0, nh) of the
// sorted order), built on the device the way build_hub_layout builds the
// vertex hubs' on the host: per hub ceil(len / HUB_CHUNK) chunks, P =
// ceil(len / part_entries) label partitions (at most HUB_MAX_PARTS), P + 1
// partition offsets per chunk, its entries in the partition buffer and a
// fallback table of nextpow2(2 len) slots (at least 64).
struct HubDev {
    int4 *chunks, *parts, *where, *info;
    int32_t *chunk_off, *tlog, *over;
    int32_t *hc, *hq, *hpo;  // per hub: first chunk, first partition item, first partition offset (nh + 1)
    int64_t *buf_base, *toff;
    int *counts;             // {chunks, partition items, hubs}
    int *err;                // set when a total would exceed its room (the hubs are then left out)
    int64_t cap_hubs, cap_chunks, cap_parts, cap_poff, cap_buf, cap_slots;
};

constexpr int HL_NT = 1024;

__device__ inline int ceil_log2_ll(long long x) {
    int b = 0;
    while ((1LL << b) < x) b++;
    return b;
}

// inclusive scan over the workgroup (Hillis-Steele in LDS)
__device__ inline long long hl_scan(long long x, long long *buf) {
    const int t = threadIdx.x;
    buf[t] = x;
    __syncthreads();
    for (int d = 1; d < HL_NT; d <<= 1) {
        const long long y = t >= d ? buf[t - d] : 0;
        __syncthreads();
        buf[t] += y;
        __syncthreads();
    }
    const long long r = buf[t];
    __syncthreads();
    return r;
}

struct HubSize {
    long long nch, P, poff, len, slots;
    int tl;
};
__device__ inline HubSize hub_size(long long len, int part_entries) {
    HubSize s;
    s.len = len;
    s.nch = (len + HUB_CHUNK - 1) / HUB_CHUNK;
    s.P = min((long long)HUB_MAX_PARTS, max(1LL, (len + part_entries - 1) / part_entries));
    s.poff = s.nch * (s.P + 1);
    s.tl = max(6, ceil_log2_ll(2 * len));
    s.slots = 1LL << s.tl;
    return s;
}

// one workgroup of HL_NT threads: the hubs' sizes, their prefix sums and the
// per-hub records
__global__ __launch_bounds__(HL_NT) void k_hub_sizes(const int *layout, const int32_t *corder, const int32_t *cind,
                                                     int part_entries, HubDev d) {
    __shared__ long long buf[HL_NT];
    __shared__ int s_fits;
    const int nh = layout[BIN_HUB + 1];
    const int t = threadIdx.x;
    const int per = (nh + HL_NT - 1) / HL_NT;
    const int h0 = min(nh, t * per), h1 = min(nh, h0 + per);
    long long s[5] = {0, 0, 0, 0, 0};
    for (int h = h0; h < h1; h++) {
        const int l = corder[h];
        const HubSize z = hub_size(cind[l + 1] - cind[l], part_entries);
        s[0] += z.nch;
        s[1] += z.P;
        s[2] += z.poff;
        s[3] += z.len;
        s[4] += z.slots;
    }
    long long ex[5], tot[5];
    for (int i = 0; i < 5; i++) {
        const long long inc = hl_scan(s[i], buf);
        ex[i] = inc - s[i];
        if (t == HL_NT - 1) buf[0] = inc;
        __syncthreads();
        tot[i] = buf[0];
        __syncthreads();
    }
    if (t == 0) {
        const bool fits = nh <= d.cap_hubs && tot[0] <= d.cap_chunks && tot[1] <= d.cap_parts && tot[2] <= d.cap_poff &&
                          tot[3] <= d.cap_buf && tot[4] <= d.cap_slots;
        s_fits = fits;
        d.counts[0] = fits ? (int)tot[0] : 0;
        d.counts[1] = fits ? (int)tot[1] : 0;
        d.counts[2] = fits ? nh : 0;
        if (!fits) *d.err = 1;
        if (fits) {
            d.hc[nh] = (int)tot[0];
            d.hq[nh] = (int)tot[1];
            d.hpo[nh] = (int)tot[2];
            d.buf_base[nh] = tot[3];
            d.toff[nh] = tot[4];
        }
    }
    __syncthreads();
    if (!s_fits) return;
    for (int h = h0; h < h1; h++) {
        const int l = corder[h];
        const HubSize z = hub_size(cind[l + 1] - cind[l], part_entries);
        d.hc[h] = (int)ex[0];
        d.hq[h] = (int)ex[1];
        d.hpo[h] = (int)ex[2];
        d.buf_base[h] = ex[3];
        d.toff[h] = ex[4];
        d.tlog[h] = z.tl;
        d.over[h] = 0;
        d.info[h] = make_int4(l, (int)z.P, (int)ex[1], (int)ex[0]);
        ex[0] += z.nch;
        ex[1] += z.P;
        ex[2] += z.poff;
        ex[3] += z.len;
        ex[4] += z.slots;
    }
}

// last h in [0, n) with a[h] <= x (a non-decreasing, a[0] <= x)
__device__ inline int last_le(const int32_t *a, int n, int x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// the chunk and partition items of the hubs
__global__ void k_hub_expand(const int32_t *cind, HubDev d) {
    const int nc = d.counts[0], np = d.counts[1], nh = d.counts[2];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        const int h = last_le(d.hc, nh, i);
        const int4 inf = d.info[h];
        const int c = i - d.hc[h];
        const int begin = cind[inf.x] + c * HUB_CHUNK;
        d.chunks[i] = make_int4(h, inf.x, begin, min(cind[inf.x + 1], begin + HUB_CHUNK));
        d.chunk_off[i] = d.hpo[h] + c * (inf.y + 1);
    }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < np; j += stride) {
        const int h = last_le(d.hq, nh, j);
        const int4 inf = d.info[h];
        d.parts[j] = make_int4(h, j - d.hq[h], d.hc[h], d.hc[h + 1] - d.hc[h]);
        d.where[j] = make_int4(inf.x, inf.y, d.hpo[h], (int)d.buf_base[h]);
    }
}

// A grow-only device buffer in the context's allocation list.
template <typename T>
struct GrowBuf {
    T *p = nullptr;
    size_t cap = 0;
    // holds at least `need` elements afterwards (contents not kept); *grew
    // tells whether it was reallocated
    int reserve(pcd_ctx *ctx, size_t need, bool *grew = nullptr) {
        if (grew) *grew = false;
        if (p && cap >= need) return PCD_OK;
        if (p) dev_free(ctx, p);
        p = nullptr;
        cap = 0;
        const size_t c = need + need / 8 + 16;
        PCD_TRY(dev_alloc_t(ctx, &p, c));
        cap = c;
        if (grew) *grew = true;
        return PCD_OK;
    }
};

// ------------------------------------------------ statistics after merging
// A matched pair (l, p), l < p, becomes the community l (refinement.cpp:
// 277-286): size_l + size_p, and the links between them turn from external
// into internal volume on both sides, so int_volume gains 2 links and
// ext_volume loses 2 links.  links(l, p) = the entries of l's row that hold
// p; the lower community of each pair counts them.
__device__ inline int pair_low(int l, const int32_t *found, const int32_t *bp) {
    if (!found[l]) return -1;
    const int p = bp[l];
    return p > l ? p : -1;  // the partner when l is the lower side of a pair
}

// rows of at most 64 entries (positions [layout[b0], layout[b1])), G lanes
// per row
template <int G>
__global__ __launch_bounds__(256) void k_pair_links_small(const int *layout, int b0, int b1, const int32_t *corder,
                                                          const int32_t *cind, const int32_t *ent, const int32_t *found,
                                                          const int32_t *bp, int32_t *links) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int nt = gridDim.x * blockDim.x / G;
    const int p0 = layout[b0], p1 = layout[b1];
    for (int base = t; base < p1 - p0; base += nt) {
        const int l = corder[p0 + base];
        const int p = pair_low(l, found, bp);
        int c = 0;
        if (p >= 0)
            for (int j = cind[l] + glane; j < cind[l + 1]; j += G) c += ent[j] == p;
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, WAVE);
        if (p >= 0 && glane == 0) links[l] = c;
    }
}

// longer rows (positions [layout[b0], layout[b1])): one wave per row (up to
// 4096 entries), or per 4096-entry chunk of a hub row (summed with atomics)
__global__ __launch_bounds__(256) void k_pair_links_rows(const int *layout, int b0, int b1, const int32_t *corder,
                                                         const int32_t *cind, const int32_t *ent, const int32_t *found,
                                                         const int32_t *bp, int32_t *links) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int p1 = layout[b1];
    for (int q = layout[b0] + wave_g; q < p1; q += nwaves) {
        const int l = corder[q];
        const int p = pair_low(l, found, bp);
        if (p < 0) continue;
        int c = 0;
        for (int j = cind[l] + lane; j < cind[l + 1]; j += WAVE) c += ent[j] == p;
        c = wave_sum(c);
        if (lane == 0) links[l] = c;
    }
}

__global__ __launch_bounds__(256) void k_pair_links_hub(const int *counts, const int4 *chunks, const int32_t *ent,
                                                        const int32_t *found, const int32_t *bp, int32_t *links) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int nchunks = counts[0];
    for (int it = wave_g; it < nchunks; it += nwaves) {
        const int4 ck = chunks[it];  // {hub, community, begin, end}
        const int p = pair_low(ck.y, found, bp);
        if (p < 0) continue;
        int c = 0;
        for (int j = ck.z + lane; j < ck.w; j += WAVE) c += ent[j] == p;
        c = wave_sum(c);
        if (lane == 0 && c) atomicAdd(&links[ck.y], c);
    }
}

// (also clears the records' matched flags, kept in their pad words during
// the step; the other records of a pair are never read again)
__global__ void k_merged_stats(int bound, const int *kp, const int32_t *found, const int32_t *bp, const int32_t *links,
                               CommStats *st) {
    const int k = k_of(bound, kp);
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        const int p = pair_low(l, found, bp);
        if (p < 0) {
            if (found[l]) st[l].pad = 0;
            continue;
        }
        const CommStats a = st[l], b = st[p];  // p's record is not written: p is no pair's lower side
        const int x = links[l];
        st[l] = CommStats{a.size + b.size, a.int_vol + b.int_vol + 2 * x, a.ext_vol + b.ext_vol - 2 * x, 0};
    }
}

// The matched flags of a merging step on community rows live in the
// records' pad words too (the merging kernels read the record of every
// candidate anyway): cleared at the start, set with `found`.
__global__ void k_clear_pads(int bound, const int *kp, CommStats *st) {
    const int k = k_of(bound, kp);
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) st[l].pad = 0;
}

// refinement.cpp:264-274 for the communities below k
__global__ void k_handshake_rec(int bound, const int *kp, int32_t *found, const int32_t *bp, int32_t *unmatched,
                                CommStats *st) {
    const int k = k_of(bound, kp);
    int local = 0;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        if (!found[l] && (bp[l] == -1 || bp[bp[l]] == l)) {
            found[l] = 1;
            st[l].pad = 1;
        }
        if (!found[l]) local++;
    }
    local = wave_sum(local);
    if (__lane_id() == 0 && local) atomicAdd(unmatched, local);
}

// Device state of the community rows and the row set of the merging step
// (every buffer sized for the largest step the graph allows).
struct MergeRows {
    GrowBuf<int32_t> ent, cind, corder, dst, links;
    GrowBuf<uint64_t> mk0, mk1;
    GrowBuf<unsigned long long> bcount;
    GrowBuf<int> layout, counts;
    // hub layout
    GrowBuf<int4> items, parts, info, where;
    GrowBuf<int32_t> item_off, tlog, poff, hbuf, over, hc, hq, hpo;
    GrowBuf<int64_t> buf_base, toff;
    GrowBuf<char> partial;
    GrowBuf<int32_t> gkeys, gcnts;  // fallback tables: all slots empty between uses
    HubDev hd{};
    RowSet rows;
};

void merge_release(pcd_ctx *ctx) {
    delete static_cast<MergeRows *>(ctx->mrows);
    ctx->mrows = nullptr;
}

// Reserves every buffer for the largest merging step of the graph (E
// entries, n + 1 communities): rows above 4096 entries number at most
// E / 4097; their chunks at most 2E / 4096 + 1; their partition items at
// most E / part_entries + hubs; their partition offsets at most
// (chunks of a row) x (P + 1) <= half its entries; their fallback tables at
// most 4 E + 64 per hub slots.
static int reserve_rows(pcd_ctx *ctx, MergeRows &R) {
    const size_t n = (size_t)ctx->n, E = (size_t)2 * ctx->m;
    const size_t pe = (size_t)std::max(64, ctx->tune.hub_part_entries);
    const size_t hubs = std::min(n + 1, E / 4097 + 1);
    const size_t chunks = 2 * E / HUB_CHUNK + 2, parts = E / pe + hubs + 1, poff = E / 2 + 1024;
    const size_t slots = 4 * E + 64 * hubs + 64;
    if (slots > (size_t)INT32_MAX * 4) return fail(PCD_E_NOMEM, "graph too large for the merging step's hub tables");
    PCD_TRY(R.ent.reserve(ctx, E + 1));
    PCD_TRY(R.cind.reserve(ctx, n + 3));
    PCD_TRY(R.corder.reserve(ctx, n + 2));
    PCD_TRY(R.dst.reserve(ctx, n + 2));
    PCD_TRY(R.links.reserve(ctx, n + 2));
    PCD_TRY(R.mk0.reserve(ctx, n + 2));
    PCD_TRY(R.mk1.reserve(ctx, n + 2));
    PCD_TRY(R.bcount.reserve(ctx, 2 * NUM_BINS));
    PCD_TRY(R.layout.reserve(ctx, ROW_LAYOUT_INTS));
    bool grew_n = false;
    PCD_TRY(R.counts.reserve(ctx, 4, &grew_n));
    if (grew_n) PCD_TRY(fill_i32(ctx, R.counts.p, 0, (int64_t)R.counts.cap));
    PCD_TRY(R.items.reserve(ctx, chunks));
    PCD_TRY(R.item_off.reserve(ctx, chunks));
    PCD_TRY(R.parts.reserve(ctx, parts));
    PCD_TRY(R.where.reserve(ctx, parts));
    PCD_TRY(R.partial.reserve(ctx, (size_t)HUB_PARTIAL_BYTES * parts));
    PCD_TRY(R.poff.reserve(ctx, poff));
    PCD_TRY(R.hbuf.reserve(ctx, E + 1));
    PCD_TRY(R.info.reserve(ctx, hubs + 1));
    PCD_TRY(R.tlog.reserve(ctx, hubs + 1));
    PCD_TRY(R.over.reserve(ctx, hubs + 1));
    PCD_TRY(R.hc.reserve(ctx, hubs + 1));
    PCD_TRY(R.hq.reserve(ctx, hubs + 1));
    PCD_TRY(R.hpo.reserve(ctx, hubs + 1));
    PCD_TRY(R.buf_base.reserve(ctx, hubs + 1));
    PCD_TRY(R.toff.reserve(ctx, hubs + 1));
    bool grew_k = false, grew_c = false;
    PCD_TRY(R.gkeys.reserve(ctx, slots, &grew_k));
    PCD_TRY(R.gcnts.reserve(ctx, slots, &grew_c));
    if (grew_k) PCD_TRY(fill_i32(ctx, R.gkeys.p, EMPTY_KEY, (int64_t)R.gkeys.cap));
    if (grew_c) PCD_TRY(fill_i32(ctx, R.gcnts.p, 0, (int64_t)R.gcnts.cap));
    HubDev &d = R.hd;
    d.chunks = R.items.p;
    d.parts = R.parts.p;
    d.where = R.where.p;
    d.info = R.info.p;
    d.chunk_off = R.item_off.p;
    d.tlog = R.tlog.p;
    d.over = R.over.p;
    d.hc = R.hc.p;
    d.hq = R.hq.p;
    d.hpo = R.hpo.p;
    d.buf_base = R.buf_base.p;
    d.toff = R.toff.p;
    d.counts = R.counts.p;
    d.err = R.counts.p + 3;
    d.cap_hubs = (int64_t)std::min({R.info.cap, R.tlog.cap, R.over.cap, R.hc.cap, R.hq.cap, R.hpo.cap,
                                     R.buf_base.cap, R.toff.cap}) - 1;
    d.cap_chunks = (int64_t)std::min(R.items.cap, R.item_off.cap);
    d.cap_parts = (int64_t)std::min(std::min(R.parts.cap, R.where.cap), R.partial.cap / HUB_PARTIAL_BYTES);
    d.cap_poff = (int64_t)R.poff.cap;
    d.cap_buf = (int64_t)R.hbuf.cap;
    d.cap_slots = (int64_t)std::min(R.gkeys.cap, R.gcnts.cap);
    RowSet &rs = R.rows;
    rs = RowSet{};
    rs.layout = R.layout.p;
    rs.hub_counts = R.counts.p;
    rs.hub_items = R.items.p;
    rs.hub_item_off = R.item_off.p;
    rs.hub_part_items = R.parts.p;
    rs.hub_info = R.info.p;
    rs.hub_part_where = R.where.p;
    rs.hub_buf_base = R.buf_base.p;
    rs.hub_poff = R.poff.p;
    rs.hub_buf = R.hbuf.p;
    rs.hub_partial = R.partial.p;
    rs.hub_over = R.over.p;
    rs.hub_table_off = R.toff.p;
    rs.hub_table_log = R.tlog.p;
    rs.hub_keys = R.gkeys.p;
    rs.hub_cnts = R.gcnts.p;
    return PCD_OK;
}

// Builds the community rows of the current labels (ids below `bound`, k on
// the device when kp is set): entries, offsets, bin order and hub layout,
// all on the stream, nothing read back.
static int build_community_rows(pcd_ctx *ctx, MergeRows &R, int bound, const int *kp) {
    const int n = ctx->n;
    const int half = label_bits(bound);
    // members by label (stable: ascending vertex id inside a community)
    kt_begin(ctx, KS_MERGE_SORT);
    hipLaunchKernelGGL(k_member_keys, dim3(grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_labels,
                       R.mk0.p);
    uint64_t *members = nullptr;
    PCD_TRY(radix_sort_u64(ctx, R.mk0.p, R.mk1.p, n, half, &members));
    PCD_TRY(device_scan(ctx, n, MemberDeg{members, ctx->d_indices}, MemberDst{members, R.dst.p}, nullptr));
    PCD_TRY(device_scan(ctx, bound, CommVol{ctx->d_stats, bound, kp}, CommStart{R.cind.p, bound}, nullptr));
    kt_end(ctx, KS_MERGE_SORT);
    // the rows: every entry's label at its member's place
    kt_begin(ctx, KS_MERGE_KEYS);
    if (ctx->num_stat_items > 0)
        hipLaunchKernelGGL(k_member_rows_items, dim3(grid_for(ctx, ctx->num_stat_items, 4, 16)), dim3(256), 0,
                           ctx->stream, ctx->num_stat_items, ctx->d_stat_items, ctx->d_indices, ctx->d_adj,
                           ctx->d_labels, R.dst.p, R.ent.p);
    const BinLayout &vb = ctx->bins;
    auto small = [&](auto g, int bin) {
        constexpr int G = decltype(g)::value;
        const int p0 = vb.start[bin], p1 = vb.start[bin + 1];
        if (p1 > p0)
            hipLaunchKernelGGL(k_member_rows_small<G>, dim3(grid_for(ctx, p1 - p0, 256 / G, 16)), dim3(256), 0,
                               ctx->stream, p0, p1, ctx->d_order, ctx->d_indices, ctx->d_adj, ctx->d_labels, R.dst.p,
                               R.ent.p);
    };
    small(std::integral_constant<int, 64>{}, BIN_G64);
    small(std::integral_constant<int, 32>{}, BIN_G32);
    small(std::integral_constant<int, 16>{}, BIN_G16);
    small(std::integral_constant<int, 8>{}, BIN_G8);
    small(std::integral_constant<int, 4>{}, BIN_G4);
    kt_end(ctx, KS_MERGE_KEYS);
    PCD_HIP(hipGetLastError());
    // bins of the community rows by length (one radix pass over 4-bit bin
    // keys), the launch layout, and the hub layout
    kt_begin(ctx, KS_MERGE_SORT);
    PCD_TRY(fill_u64(ctx, (uint64_t *)R.bcount.p, 0, 2 * NUM_BINS));
    hipLaunchKernelGGL(k_comm_bins, dim3(grid_for(ctx, bound, 256, 4)), dim3(256), 0, ctx->stream, bound, kp, R.cind.p,
                       R.mk0.p, R.bcount.p);
    uint64_t *cs = nullptr;
    PCD_TRY(radix_sort_u64(ctx, R.mk0.p, R.mk1.p, bound, 4, &cs));
    hipLaunchKernelGGL(k_high_words, dim3(grid_for(ctx, bound, 256)), dim3(256), 0, ctx->stream, bound, cs,
                       R.corder.p);
    hipLaunchKernelGGL(k_row_layout, dim3(1), dim3(64), 0, ctx->stream, R.bcount.p, ctx->cu_count * 16, R.layout.p);
    hipLaunchKernelGGL(k_hub_sizes, dim3(1), dim3(HL_NT), 0, ctx->stream, R.layout.p, R.corder.p, R.cind.p,
                       std::max(64, ctx->tune.hub_part_entries), R.hd);
    hipLaunchKernelGGL(k_hub_expand, dim3(ctx->cu_count * 2), dim3(256), 0, ctx->stream, R.cind.p, R.hd);
    kt_end(ctx, KS_MERGE_SORT);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

static int ensure_merge_state(pcd_ctx *ctx) {
    // sized for n communities at most (k <= n)
    if (ctx->d_bp) return PCD_OK;
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_bp, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_found, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_best_ord, ctx->n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_best_key, ctx->n + 1));
    return PCD_OK;
}

// refinement.cpp:277-286: every vertex of a matched community with a partner
// takes min(l, partner)
static int merge_relabel(pcd_ctx *ctx, int k) {
    hipLaunchKernelGGL(k_merge_relabel, dim3(grid_for(ctx, ctx->n, 256)), dim3(256), 0, ctx->stream, ctx->n,
                       ctx->d_labels, ctx->d_found, ctx->d_bp, ctx->d_labels_alt);
    PCD_HIP(hipGetLastError());
    std::swap(ctx->d_labels, ctx->d_labels_alt);
    ctx->labels_compressed = false;
    ctx->has_stats = false;
    ctx->stats_pending = false;
    ctx->label_bound = k;
    return PCD_OK;
}

// Buffers a merging step over community rows needs, reserved ahead (a
// captured step may not allocate).
int merge_rows_reserve(pcd_ctx *ctx) {
    PCD_TRY(ensure_merge_state(ctx));
    if (!ctx->mrows) ctx->mrows = new MergeRows();
    MergeRows &R = *static_cast<MergeRows *>(ctx->mrows);
    PCD_TRY(reserve_rows(ctx, R));
    PCD_TRY(ensure_scan_parts(ctx, (int64_t)ctx->n + 2));
    PCD_TRY(ensure_sort_space(ctx, (int64_t)ctx->n + 2));
    return PCD_OK;
}

// Fails if a merging step queued since the last check left hubs out (a hub
// layout larger than its reserved room; reached only beyond the bounds of
// reserve_rows).
int merge_rows_check(pcd_ctx *ctx) {
    MergeRows *R = static_cast<MergeRows *>(ctx->mrows);
    if (!R || !R->counts.p) return PCD_OK;
    PCD_HIP(hipMemcpyAsync(ctx->h_small + 9, R->counts.p + 3, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->h_small[9] != 0) {
        PCD_TRY(fill_i32(ctx, R->counts.p + 3, 0, 1));
        return fail(PCD_E_NOMEM, "merging step: hub layout exceeds its reserved room");
    }
    return PCD_OK;
}

// the partner choice over community rows (see the top of this file); k on
// the device while a pass is captured or k is deferred
static int merging_rows_dev(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie) {
    const bool dev_k = ctx->capturing || ctx->k_deferred;
    const int bound = dev_k ? ctx->n + 1 : ctx->k;
    const int *kp = dev_k ? ctx->d_k : nullptr;
    if (!ctx->capturing) PCD_TRY(merge_rows_reserve(ctx));
    if (!ctx->mrows) return fail(PCD_E_STATE, "merging rows not reserved before a capture");
    MergeRows &R = *static_cast<MergeRows *>(ctx->mrows);
    if (bound <= 0) return PCD_OK;
    PCD_TRY(build_community_rows(ctx, R, bound, kp));
    PCD_TRY(fill_i32(ctx, ctx->d_bp, -1, bound));
    PCD_TRY(fill_i32(ctx, ctx->d_found, 0, bound));
    hipLaunchKernelGGL(k_clear_pads, dim3(grid_for(ctx, bound, 256)), dim3(256), 0, ctx->stream, bound, kp,
                       ctx->d_stats);
    LPArgs a{};
    a.indices = R.cind.p;  // community l's row: entries [cind[l], cind[l+1]) of ent
    a.adj = R.ent.p;
    a.order = R.corder.p;
    a.stats = ctx->d_stats;
    a.new_labels = ctx->d_bp;
    a.found = ctx->d_found;
    a.rule = make_rule(ctx, rule);
    a.step = step;
    a.step_ptr = ctx->graph_step;  // a replayed pass reads its step (or -1) there
    a.tie = tie;
    int32_t *unmatched = ctx->d_counts + NUM_BINS;
    for (int rnd = 0; rnd < rounds; rnd++) {
        a.seed = merge_seed(seed, rnd);
        PCD_TRY(merge_rows_launch(ctx, R.rows, a));
        kt_begin(ctx, KS_MERGE_ROUNDS);
        PCD_TRY(fill_i32(ctx, unmatched, 0, 1));
        hipLaunchKernelGGL(k_handshake_rec, dim3(grid_for(ctx, bound, 256)), dim3(256), 0, ctx->stream, bound, kp,
                           ctx->d_found, ctx->d_bp, unmatched, ctx->d_stats);
        kt_end(ctx, KS_MERGE_ROUNDS);
        PCD_HIP(hipGetLastError());
        if (rnd + 1 < rounds && !ctx->capturing) {
            // refinement.cpp:220,264-274: stop once every community is matched
            // (a captured pass runs every round: once all are matched, a
            // round changes nothing)
            PCD_HIP(hipMemcpyAsync(ctx->h_small, unmatched, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            PCD_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->h_small[0] == 0) break;
        }
    }
    const bool follow = ctx->tune.inc_stats != 0;
    if (follow) {
        // the statistics of the merged labels, from the records and the
        // pairs' links (the next compression renumbers them)
        kt_begin(ctx, KS_MERGE_STATS);
        PCD_TRY(fill_i32(ctx, R.links.p, 0, bound));
        hipLaunchKernelGGL(k_pair_links_hub, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, R.counts.p, R.items.p,
                           R.ent.p, ctx->d_found, ctx->d_bp, R.links.p);
        hipLaunchKernelGGL(k_pair_links_rows, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, R.layout.p,
                           (int)BIN_B4096, (int)BIN_G64, R.corder.p, R.cind.p, R.ent.p, ctx->d_found, ctx->d_bp,
                           R.links.p);
        hipLaunchKernelGGL(k_pair_links_small<8>, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, R.layout.p,
                           (int)BIN_G64, (int)NUM_BINS, R.corder.p, R.cind.p, R.ent.p, ctx->d_found, ctx->d_bp,
                           R.links.p);
        hipLaunchKernelGGL(k_merged_stats, dim3(grid_for(ctx, bound, 256)), dim3(256), 0, ctx->stream, bound, kp,
                           ctx->d_found, ctx->d_bp, R.links.p, ctx->d_stats);
        kt_end(ctx, KS_MERGE_STATS);
        PCD_HIP(hipGetLastError());
    } else {  // the records stay as they are for a caller that reads them
        hipLaunchKernelGGL(k_clear_pads, dim3(grid_for(ctx, bound, 256)), dim3(256), 0, ctx->stream, bound, kp,
                           ctx->d_stats);
    }
    PCD_TRY(merge_relabel(ctx, bound));
    ctx->stats_pending = follow;
    // a hub layout that did not fit its room would have left hubs out
    if (!ctx->capturing) PCD_TRY(merge_rows_check(ctx));
    return PCD_OK;
}

// the partner choice over the sorted unique community pairs
static int merging_pairs_dev(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie) {
    const int k = ctx->k;
    int64_t U = 0;
    int half = 0;
    PCD_TRY(build_pairs(ctx, &U, &half));
    PCD_TRY(ensure_merge_state(ctx));
    // candidate slots: two per 64 pairs (the runs that may cross a wave boundary)
    const size_t ncap = (size_t)(2 * ((U + WAVE - 1) / WAVE) + 2);
    if (ncap > ctx->mcand_cap) {
        if (ctx->d_mcand) dev_free(ctx, ctx->d_mcand);
        ctx->d_mcand = nullptr;
        ctx->mcand_cap = 0;
        PCD_TRY(dev_alloc(ctx, &ctx->d_mcand, sizeof(MergeCand) * ncap));
        ctx->mcand_cap = ncap;
    }
    PCD_TRY(fill_i32(ctx, ctx->d_bp, -1, k));
    PCD_TRY(fill_i32(ctx, ctx->d_found, 0, k));
    PCD_TRY(fill_u64(ctx, ctx->d_best_ord, 0, k));
    PCD_TRY(fill_u64(ctx, ctx->d_best_key, ~0ULL, k));
    MergeArgs a;
    a.ukeys = ctx->d_keys1;
    a.links = ctx->d_pair_links;
    a.U = U;
    a.half = half;
    a.stats = ctx->d_stats;
    a.found = ctx->d_found;
    a.bp = ctx->d_bp;
    a.rule = make_rule(ctx, rule);
    a.step = step;
    a.tie = tie;
    a.best_ord = (unsigned long long *)ctx->d_best_ord;
    a.best_key = (unsigned long long *)ctx->d_best_key;
    int32_t *unmatched = ctx->d_counts + NUM_BINS;
    for (int rnd = 0; rnd < rounds; rnd++) {
        kt_begin(ctx, KS_MERGE_ROUNDS);
        a.seed = merge_seed(seed, rnd);
        if (U > 0) {
            const int64_t nslots = 2 * ((U + WAVE - 1) / WAVE);
            hipLaunchKernelGGL(k_merge_pass, dim3(grid_for(ctx, U, 256)), dim3(256), 0, ctx->stream, a,
                               (MergeCand *)ctx->d_mcand);
            hipLaunchKernelGGL(k_merge_resolve, dim3(grid_for(ctx, nslots, 256)), dim3(256), 0, ctx->stream, a,
                               (const MergeCand *)ctx->d_mcand, nslots);
        }
        hipLaunchKernelGGL(k_merge_finalize, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_found,
                           ctx->d_bp, a.best_ord, a.best_key);
        PCD_TRY(fill_i32(ctx, unmatched, 0, 1));
        hipLaunchKernelGGL(k_handshake, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_found,
                           ctx->d_bp, unmatched);
        kt_end(ctx, KS_MERGE_ROUNDS);
        PCD_HIP(hipGetLastError());
        if (rnd + 1 < rounds) {
            // refinement.cpp:220,264-274: stop once every community is matched
            PCD_HIP(hipMemcpyAsync(ctx->h_small, unmatched, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            PCD_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->h_small[0] == 0) break;
        }
    }
    return merge_relabel(ctx, k);
}

int merging_dev(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie) {
    if (ctx->tune.merge_sort) return merging_pairs_dev(ctx, rule, rounds, seed, step, tie);
    return merging_rows_dev(ctx, rule, rounds, seed, step, tie);
}

int contract_dev(pcd_ctx *ctx, int64_t *num_pairs) {
    int64_t U = 0;
    int half = 0;
    PCD_TRY(build_pairs(ctx, &U, &half));
    *num_pairs = U;
    return PCD_OK;
}

int contract_download(pcd_ctx *ctx, int64_t U, int32_t *l, int32_t *k, int32_t *links) {
    if (U <= 0) return PCD_OK;
    int half = label_bits(ctx->k);
    int32_t *dl = nullptr, *dk = nullptr;
    PCD_TRY(dev_alloc_t(ctx, &dl, U));
    PCD_TRY(dev_alloc_t(ctx, &dk, U));
    hipLaunchKernelGGL(k_unpack_pairs, dim3(grid_for(ctx, U, 256)), dim3(256), 0, ctx->stream, ctx->d_keys1, U, half,
                       dl, dk);
    PCD_HIP(hipGetLastError());
    if (l) PCD_HIP(hipMemcpyAsync(l, dl, U * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (k) PCD_HIP(hipMemcpyAsync(k, dk, U * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (links) PCD_HIP(hipMemcpyAsync(links, ctx->d_pair_links, U * 4, hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, dl);
    dev_free(ctx, dk);
    return PCD_OK;
}

}  // namespace pcd
