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
__global__ void k_member_keys(int n, const int32_t *labels, uint64_t *keys) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        keys[v] = ((uint64_t)(uint32_t)v << 32) | (uint32_t)labels[v];
}

// Community l's row starts at the volume (int_volume + ext_volume, the sum
// of its members' degrees) of the communities below l; an empty community
// has an empty row.
struct CommVol {
    const CommStats *st;
    __device__ int operator()(int64_t l) const { return st[l].int_vol + st[l].ext_vol; }
};
struct CommStart {
    int32_t *cind;
    int64_t k;
    __device__ void operator()(int64_t l, int ex, int vol) const {
        cind[l] = ex;
        if (l == k - 1) cind[k] = ex + vol;
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

// per community: its bin as the sort key (community id above), and the
// rows and entries of each bin
__global__ __launch_bounds__(256) void k_comm_bins(int k, const int32_t *cind, uint64_t *keys,
                                                   unsigned long long *bcount) {
    __shared__ unsigned long long rows[NUM_BINS], ents[NUM_BINS];
    if (threadIdx.x < NUM_BINS) {
        rows[threadIdx.x] = 0;
        ents[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        const int len = cind[l + 1] - cind[l];
        const int b = row_bin(len);
        keys[l] = ((uint64_t)(uint32_t)l << 32) | (uint32_t)b;
        atomicAdd(&rows[b], 1ULL);
        atomicAdd(&ents[b], (unsigned long long)len);
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

// the HUB-bin rows {community, first entry, end entry}
__global__ void k_hub_rows(int nh, const int32_t *corder, const int32_t *cind, int4 *out) {
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
        const int l = corder[h];
        out[h] = make_int4(l, cind[l], cind[l + 1], 0);
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
    int upload(pcd_ctx *ctx, const std::vector<T> &v) {
        PCD_TRY(reserve(ctx, v.size()));
        if (!v.empty())
            PCD_HIP(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, ctx->stream));
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

// rows of at most 64 entries, G lanes per row
template <int G>
__global__ __launch_bounds__(256) void k_pair_links_small(int p0, int p1, const int32_t *corder, const int32_t *cind,
                                                          const int32_t *ent, const int32_t *found, const int32_t *bp,
                                                          int32_t *links) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int nt = gridDim.x * blockDim.x / G;
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

// longer rows: one wave per row (up to 4096 entries), or per 4096-entry
// chunk of a hub row (summed with atomics)
__global__ __launch_bounds__(256) void k_pair_links_rows(int p0, int p1, const int32_t *corder, const int32_t *cind,
                                                         const int32_t *ent, const int32_t *found, const int32_t *bp,
                                                         int32_t *links) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int q = p0 + wave_g; q < p1; q += nwaves) {
        const int l = corder[q];
        const int p = pair_low(l, found, bp);
        if (p < 0) continue;
        int c = 0;
        for (int j = cind[l] + lane; j < cind[l + 1]; j += WAVE) c += ent[j] == p;
        c = wave_sum(c);
        if (lane == 0) links[l] = c;
    }
}

__global__ __launch_bounds__(256) void k_pair_links_hub(int nchunks, const int4 *chunks, const int32_t *ent,
                                                        const int32_t *found, const int32_t *bp, int32_t *links) {
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
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
__global__ void k_merged_stats(int k, const int32_t *found, const int32_t *bp, const int32_t *links, CommStats *st) {
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
__global__ void k_clear_pads(int k, CommStats *st) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) st[l].pad = 0;
}

__global__ void k_handshake_rec(int k, int32_t *found, const int32_t *bp, int32_t *unmatched, CommStats *st) {
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

// Device state of the community rows and the row set of the current step.
struct MergeRows {
    GrowBuf<int32_t> ent, cind, corder, dst, links;
    GrowBuf<uint64_t> mk0, mk1;
    GrowBuf<unsigned long long> bcount;
    GrowBuf<int4> hrows;
    // hub layout
    GrowBuf<int4> items, parts, info, where;
    GrowBuf<int32_t> item_off, tlog, poff, hbuf, over;
    GrowBuf<int64_t> buf_base, toff;
    GrowBuf<char> partial;
    GrowBuf<int32_t> gkeys, gcnts;  // fallback tables: all slots empty between uses
    RowSet rows;
};

void merge_release(pcd_ctx *ctx) {
    delete static_cast<MergeRows *>(ctx->mrows);
    ctx->mrows = nullptr;
}

// Builds the community rows of the current compressed labels (k
// communities): entries, offsets, bin order and hub layout.
static int build_community_rows(pcd_ctx *ctx, MergeRows &R) {
    const int n = ctx->n, k = ctx->k;
    const int64_t E = (int64_t)2 * ctx->m;
    PCD_TRY(R.ent.reserve(ctx, (size_t)E + 1));
    PCD_TRY(R.cind.reserve(ctx, (size_t)n + 2));
    PCD_TRY(R.corder.reserve(ctx, (size_t)n + 2));
    PCD_TRY(R.dst.reserve(ctx, (size_t)n + 2));
    PCD_TRY(R.mk0.reserve(ctx, (size_t)n + 2));
    PCD_TRY(R.mk1.reserve(ctx, (size_t)n + 2));
    PCD_TRY(R.bcount.reserve(ctx, 2 * NUM_BINS));
    const int half = label_bits(k);
    // members by label (stable: ascending vertex id inside a community)
    kt_begin(ctx, KS_MERGE_SORT);
    hipLaunchKernelGGL(k_member_keys, dim3(grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_labels,
                       R.mk0.p);
    uint64_t *members = nullptr;
    PCD_TRY(radix_sort_u64(ctx, R.mk0.p, R.mk1.p, n, half, &members));
    PCD_TRY(device_scan(ctx, n, MemberDeg{members, ctx->d_indices}, MemberDst{members, R.dst.p}, nullptr));
    PCD_TRY(device_scan(ctx, k, CommVol{ctx->d_stats}, CommStart{R.cind.p, k}, nullptr));
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
    // bins of the community rows by length: one radix pass over 4-bit bins
    kt_begin(ctx, KS_MERGE_SORT);
    PCD_TRY(fill_u64(ctx, (uint64_t *)R.bcount.p, 0, 2 * NUM_BINS));
    hipLaunchKernelGGL(k_comm_bins, dim3(grid_for(ctx, k, 256, 4)), dim3(256), 0, ctx->stream, k, R.cind.p, R.mk0.p,
                       R.bcount.p);
    uint64_t *cs = nullptr;
    PCD_TRY(radix_sort_u64(ctx, R.mk0.p, R.mk1.p, k, 4, &cs));
    hipLaunchKernelGGL(k_high_words, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, cs, R.corder.p);
    kt_end(ctx, KS_MERGE_SORT);
    PCD_HIP(hipGetLastError());
    PCD_HIP(hipMemcpyAsync(ctx->h_small64, R.bcount.p, sizeof(int64_t) * 2 * NUM_BINS, hipMemcpyDeviceToHost,
                           ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    RowSet &rs = R.rows;
    rs = RowSet{};
    int64_t at = 0;
    for (int b = 0; b < NUM_BINS; b++) {
        rs.bins.start[b] = (int)at;
        at += ctx->h_small64[b];
        rs.bin_entries[b] = ctx->h_small64[NUM_BINS + b];
    }
    rs.bins.start[NUM_BINS] = (int)at;
    rs.bins.split2048 = rs.bins.start[BIN_B4096 + 1];
    if (at != k) return fail(PCD_E_STATE, "community bins do not cover the k communities");
    // rows above 4096 entries: the hub layout, built on the host
    const int nh = rs.bins.start[BIN_HUB + 1];
    if (nh > 0) {
        PCD_TRY(R.hrows.reserve(ctx, (size_t)nh));
        hipLaunchKernelGGL(k_hub_rows, dim3(grid_for(ctx, nh, 256)), dim3(256), 0, ctx->stream, nh, R.corder.p,
                           R.cind.p, R.hrows.p);
        PCD_HIP(hipGetLastError());
        std::vector<int4> hr(nh);
        PCD_HIP(hipMemcpyAsync(hr.data(), R.hrows.p, sizeof(int4) * nh, hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        HubLayoutHost L;
        PCD_TRY(build_hub_layout(hr, ctx->tune.hub_part_entries, L));
        PCD_TRY(R.items.upload(ctx, L.chunks));
        PCD_TRY(R.item_off.upload(ctx, L.chunk_off));
        PCD_TRY(R.parts.upload(ctx, L.parts));
        PCD_TRY(R.where.upload(ctx, L.where));
        PCD_TRY(R.info.upload(ctx, L.info));
        PCD_TRY(R.buf_base.upload(ctx, L.buf_base));
        PCD_TRY(R.toff.upload(ctx, L.toff));
        PCD_TRY(R.tlog.upload(ctx, L.tlog));
        PCD_TRY(R.partial.reserve(ctx, (size_t)HUB_PARTIAL_BYTES * L.parts.size()));
        PCD_TRY(R.poff.reserve(ctx, (size_t)L.poff_total));
        PCD_TRY(R.hbuf.reserve(ctx, (size_t)L.buf_base[nh]));
        PCD_TRY(R.over.reserve(ctx, (size_t)nh));
        PCD_TRY(fill_i32(ctx, R.over.p, 0, nh));
        const size_t gslots = (size_t)L.toff[nh];
        bool grew_k = false, grew_c = false;
        PCD_TRY(R.gkeys.reserve(ctx, gslots, &grew_k));
        PCD_TRY(R.gcnts.reserve(ctx, gslots, &grew_c));
        if (grew_k) PCD_TRY(fill_i32(ctx, R.gkeys.p, EMPTY_KEY, (int64_t)R.gkeys.cap));
        if (grew_c) PCD_TRY(fill_i32(ctx, R.gcnts.p, 0, (int64_t)R.gcnts.cap));
        // the host vectors must outlive their copies
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        rs.num_hubs = nh;
        rs.num_hub_items = (int)L.chunks.size();
        rs.num_hub_parts = (int)L.parts.size();
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
    }
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

// the partner choice over community rows (see the top of this file)
static int merging_rows_dev(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie) {
    const int k = ctx->k;
    if (!ctx->mrows) ctx->mrows = new MergeRows();
    MergeRows &R = *static_cast<MergeRows *>(ctx->mrows);
    PCD_TRY(build_community_rows(ctx, R));
    PCD_TRY(ensure_merge_state(ctx));
    PCD_TRY(fill_i32(ctx, ctx->d_bp, -1, k));
    PCD_TRY(fill_i32(ctx, ctx->d_found, 0, k));
    hipLaunchKernelGGL(k_clear_pads, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_stats);
    LPArgs a{};
    a.indices = R.cind.p;  // community l's row: entries [cind[l], cind[l+1]) of ent
    a.adj = R.ent.p;
    a.order = R.corder.p;
    a.stats = ctx->d_stats;
    a.new_labels = ctx->d_bp;
    a.found = ctx->d_found;
    a.rule = make_rule(ctx, rule);
    a.step = step;
    a.tie = tie;
    int32_t *unmatched = ctx->d_counts + NUM_BINS;
    for (int rnd = 0; rnd < rounds; rnd++) {
        a.seed = merge_seed(seed, rnd);
        PCD_TRY(merge_rows_launch(ctx, R.rows, a));
        kt_begin(ctx, KS_MERGE_ROUNDS);
        PCD_TRY(fill_i32(ctx, unmatched, 0, 1));
        hipLaunchKernelGGL(k_handshake_rec, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_found,
                           ctx->d_bp, unmatched, ctx->d_stats);
        kt_end(ctx, KS_MERGE_ROUNDS);
        PCD_HIP(hipGetLastError());
        if (rnd + 1 < rounds) {
            // refinement.cpp:220,264-274: stop once every community is matched
            PCD_HIP(hipMemcpyAsync(ctx->h_small, unmatched, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            PCD_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->h_small[0] == 0) break;
        }
    }
    const bool follow = ctx->tune.inc_stats != 0;
    if (!follow)  // the records stay as they are for a caller that reads them
        hipLaunchKernelGGL(k_clear_pads, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_stats);
    if (follow) {
        // the statistics of the merged labels, from the records and the
        // pairs' links (the next compression renumbers them)
        kt_begin(ctx, KS_MERGE_STATS);
        const RowSet &rs = R.rows;
        PCD_TRY(R.links.reserve(ctx, (size_t)k + 1));
        if (rs.num_hubs > 0) {
            PCD_TRY(fill_i32(ctx, R.links.p, 0, k));
            hipLaunchKernelGGL(k_pair_links_hub, dim3(grid_for(ctx, rs.num_hub_items, 4, 8)), dim3(256), 0,
                               ctx->stream, rs.num_hub_items, rs.hub_items, R.ent.p, ctx->d_found, ctx->d_bp,
                               R.links.p);
        }
        const int q0 = rs.bins.start[BIN_B4096], q1 = rs.bins.start[BIN_G64];
        if (q1 > q0)
            hipLaunchKernelGGL(k_pair_links_rows, dim3(grid_for(ctx, q1 - q0, 4, 8)), dim3(256), 0, ctx->stream, q0,
                               q1, R.corder.p, R.cind.p, R.ent.p, ctx->d_found, ctx->d_bp, R.links.p);
        const int s0 = rs.bins.start[BIN_G64], s1 = rs.bins.start[NUM_BINS];
        if (s1 > s0)
            hipLaunchKernelGGL(k_pair_links_small<8>, dim3(grid_for(ctx, s1 - s0, 32, 8)), dim3(256), 0, ctx->stream,
                               s0, s1, R.corder.p, R.cind.p, R.ent.p, ctx->d_found, ctx->d_bp, R.links.p);
        hipLaunchKernelGGL(k_merged_stats, dim3(grid_for(ctx, k, 256)), dim3(256), 0, ctx->stream, k, ctx->d_found,
                           ctx->d_bp, R.links.p, ctx->d_stats);
        kt_end(ctx, KS_MERGE_STATS);
        PCD_HIP(hipGetLastError());
    }
    PCD_TRY(merge_relabel(ctx, k));
    ctx->stats_pending = follow;
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
