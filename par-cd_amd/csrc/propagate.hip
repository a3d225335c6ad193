// Label propagation (PropagationRefinement::refine, refinement.cpp:114-182)
// on gfx950.
//
// Per active vertex v the step needs, for every distinct neighbour label c,
// the count of neighbours in c, then score_join(v, c, count) for c != own,
// the best of those (score, then tie key, then label), score_leave(v, own
// count) and the move decision.
//
// Neighbour gathers read labels; the community record of each distinct
// label is read once per row when its gain is scored.  For modularity and
// CPM, whose gains read one community total (int_volume + ext_volume, or
// size), kernels can be tuned to gather an 8-byte {label, aux} record per
// entry instead (built at the start of the step), so that one gather brings
// both; WCC and conductance always read the whole record.
//
// Vertex selection (counter hash against frac) is evaluated inline.  The
// next labels start as a copy of the current ones (one bulk copy before the
// kernels), so kernels write active vertices only.  Vertices are processed by degree
// bin (positions are degree-descending, so every bin is a contiguous range):
//
//   G4..G64   deg <= G: each wave compacts the active vertices of CP
//             consecutive positions (ballot), then G lanes per vertex load
//             one neighbour label each, bitonic-sort them by label across
//             the lanes (shuffles), find run heads by ballot and reduce the
//             best candidate by butterfly.  No LDS table.
//   W256      one wave per vertex, LDS hash table of 2 deg slots.
//   B1024     one 256-thread workgroup per vertex (or, for degrees up to
//             512, one wave), 2 deg slots.
//   B4096     one 512-thread workgroup per vertex, 2 deg slots.
//   HUB       deg > 4096.  Pass 1: each 4096-entry chunk of a hub row
//             gathers its labels and writes them grouped by label-hash
//             partition (LDS counting sort).  Pass 2: one workgroup per
//             (hub, partition) counts the partition in an LDS table, scans
//             the slots it claimed and reduces its best candidate.  Pass 3:
//             one wave per hub combines the partitions and decides.  A
//             partition whose distinct labels overflow its table flags the
//             hub, which is then recounted in a global-memory hash table
//             (never full).
#include <type_traits>

#include "work.h"

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

// x of lane ^ J, every lane of the wave active: a DPP quad permute for
// J = 1 and 2 (no LDS access), ds_swizzle's bit mode (and 0x1f, xor J,
// within 32 lanes; no address register) for J = 4..16, ds_bpermute for 32.
template <int J>
__device__ inline int xor_shfl(int x) {
    static_assert(J >= 1 && J <= 32 && (J & (J - 1)) == 0, "lane distance");
    if constexpr (J == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);       // quad_perm [1,0,3,2]
    else if constexpr (J == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (J <= 16) return __builtin_amdgcn_ds_swizzle(x, 0x1F | (J << 10));
    else return __shfl_xor(x, J, WAVE);
}

// x of lane - 1 (DPP wave_shr:1, every lane active); lane 0's result is
// undefined
__device__ inline int prev_lane(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, false); }

template <int J>
__device__ inline Cand cand_xor(const Cand &c) {
    Cand o;
    const long long b = __double_as_longlong(c.s);
    const int lo = xor_shfl<J>((int)(uint32_t)b), hi = xor_shfl<J>((int)(b >> 32));
    o.s = __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
    o.key = (uint32_t)xor_shfl<J>((int)c.key);
    o.label = xor_shfl<J>(c.label);
    return o;
}

// best candidate and largest own count over aligned groups of 2D lanes
template <int D>
__device__ inline void group_reduce_cand(Cand &c, int &own_cnt) {
    const Cand o = cand_xor<D>(c);
    if (cand_better(o, c)) c = o;
    own_cnt = max(own_cnt, xor_shfl<D>(own_cnt));
    if constexpr (D > 1) group_reduce_cand<D / 2>(c, own_cnt);
}

// bitonic sort of each aligned group of G lanes, ascending: one
// compare-exchange per step, the lane keeping the smaller or the larger of
// its value and its partner's (lane ^ J)
template <int G, int K = 2, int J = 1>
__device__ inline int group_sort(int x, int glane) {
    const bool take_min = ((glane & K) == 0) == ((glane & J) == 0);
    const int y = xor_shfl<J>(x);
    x = take_min ? min(x, y) : max(x, y);
    if constexpr (J > 1) return group_sort<G, K, J / 2>(x, glane);
    else if constexpr (K < G) return group_sort<G, 2 * K, K>(x, glane);
    else return x;
}

// partial result of one hub partition
struct HubPartial {
    Cand c;
    int own_cnt;
    int distinct;
};
static_assert(sizeof(HubPartial) <= HUB_PARTIAL_BYTES, "hub partial record size");

// Community record as the gains see it.  With aux: modularity reads
// int_volume + ext_volume (= aux + 0), CPM reads size (= aux).
__device__ inline CommStats comm_of(const LPArgs &a, int label, int aux) {
    if (a.use_aux) return CommStats{aux, aux, 0, 0};
    return a.stats[label];
}

// Kernel builds for the batch read the 16-byte community records (a batch
// keeps no per-community totals); an extra bit of the rule-kind mask.
constexpr unsigned RK_REC = 1u << 8;
constexpr unsigned RK_BATCH = RK_MOD | RK_CPM | RK_REC;

// Builds for the merging step (MergingRefinement::refine,
// refinement.cpp:185-288): a row is a community l, its entries are the
// labels of its members' neighbours (gathered into the row once per step),
// its own label is l, a candidate is another community not yet matched,
// scored by score_merge and accepted from 0 up, and the row's result is its
// best partner or -1.  They read the community records, whose pad word
// holds the community's "matched" flag during a merging step (one gather
// per candidate brings both).
constexpr unsigned RK_MERGE = 1u << 9;
constexpr unsigned RK_MERGE_REC = RK_MERGE | RK_REC;

template <unsigned KM>
__device__ inline bool lp_active(const LPArgs &a, int v) {
    if constexpr ((KM & RK_MERGE) != 0) {
        // a replayed merging pass whose loop has stopped (step -1) merges nothing
        if (a.step_ptr && *a.step_ptr < 0) return false;
        // refinement.cpp:224-226: matched communities, and those whose chosen
        // partner is still unmatched, sit the round out
        if (a.found[v]) return false;
        const int b = a.new_labels[v];
        return b == -1 || a.found[b] != 0;
    } else {
        return lp_selected(a, v);
    }
}

// label of a neighbour entry (merging rows hold labels already) and of the row
template <unsigned KM>
__device__ inline int entry_label(const LPArgs &a, int nb) {
    if constexpr ((KM & RK_MERGE) != 0) return nb;
    else return a.labels[nb];
}
template <unsigned KM>
__device__ inline int own_label(const LPArgs &a, int v) {
    if constexpr ((KM & RK_MERGE) != 0) return v;
    else return a.labels[v];
}

// Score of a candidate label other than the row's own, and whether it is a
// candidate at all: score_join > 0 (refinement.cpp:152-166), or for a
// merging row score_merge >= 0 (refinement.cpp:247-260: a zero score ties
// with "no partner" and is taken).  A merging caller has already dropped
// matched communities.
template <unsigned KM>
__device__ inline bool cand_score(const LPArgs &a, int deg, const CommStats &own_rec, const CommStats &rec, int cnt,
                                  double &sc) {
    if constexpr ((KM & RK_MERGE) != 0) {
        sc = score_merge<KM & RK_ALL>(a.rule, own_rec, rec, cnt);
        return sc >= 0.0;
    } else {
        sc = score_join<KM>(a.rule, deg, rec, cnt);
        return sc > 0;
    }
}

// merging row's result (refinement.cpp:262): the best partner, or -1
__device__ inline int merge_partner(const Cand &c) { return c.label != INT32_MAX ? c.label : -1; }

// Community record of `label` as the gains read it.  Builds for modularity
// or CPM alone read the 4-byte per-community total the step built (one
// random read into a 4x smaller array than the records); the others read
// the record.
template <unsigned KM>
__device__ inline CommStats comm_rec(const LPArgs &a, int label) {
    if constexpr ((KM & (RK_WCC | RK_COND | RK_REC)) == 0) {
        const int x = a.tot[label];
        return CommStats{x, x, 0, 0};
    } else {
        return a.stats[label];
    }
}

// refinement.cpp:169-179: best_score (0 if no join candidate beat 0) plus
// score_leave decides between moving, leaving to a new singleton, staying.
template <unsigned KM>
__device__ inline int lp_decide(const LPArgs &a, int deg, int own, int own_aux, int own_cnt, const Cand &best) {
    double leave = score_leave<KM>(a.rule, deg, comm_of(a, own, own_aux), own_cnt);
    double bs = best.label != INT32_MAX ? best.s : 0.0;
    bs += leave;
    if (bs > 0 && bs > leave) return best.label;
    if (leave > 0) return PCD_NEW_LABEL;
    return own;
}

template <unsigned KM>
__device__ inline int lp_decide_rec(const LPArgs &a, int deg, int own, const CommStats &own_rec, int own_cnt,
                                    const Cand &best) {
    double leave = score_leave<KM>(a.rule, deg, own_rec, own_cnt);
    double bs = best.label != INT32_MAX ? best.s : 0.0;
    bs += leave;
    if (bs > 0 && bs > leave) return best.label;
    if (leave > 0) return PCD_NEW_LABEL;
    return own;
}

template <unsigned KM>
__device__ inline int lp_decide_stats(const LPArgs &a, int deg, int own, int own_cnt, const Cand &best) {
    return lp_decide_rec<KM>(a, deg, own, comm_rec<KM>(a, own), own_cnt, best);
}

__device__ inline void count3(const LPArgs &a, long long act, long long ent, long long dist) {
    atomicAdd(&a.counters[0], (unsigned long long)act);
    atomicAdd(&a.counters[1], (unsigned long long)ent);
    atomicAdd(&a.counters[2], (unsigned long long)dist);
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------- records
__global__ void k_lp_records(int n, const int32_t *labels, const CommStats *st, int kind, int2 *lv) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        int l = labels[v];
        CommStats c = st[l];
        int aux = kind == PCD_RULE_MODULARITY ? c.int_vol + c.ext_vol : c.size;
        lv[v] = make_int2(l, aux);
    }
}

// Per community the one total the gains of modularity (int_volume +
// ext_volume) or CPM (size) read: 4 bytes per community instead of the
// 16-byte record, so the random reads of the step hit a 4x smaller array.
__global__ void k_label_totals(const int32_t *kp, const CommStats *st, int kind, int32_t *tot) {
    const int k = *kp;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < k; l += gridDim.x * blockDim.x) {
        CommStats c = st[l];
        tot[l] = kind == PCD_RULE_MODULARITY ? c.int_vol + c.ext_vol : c.size;
    }
}

// ------------------------------------------------------------- G lanes / vertex
template <int G, int CP, bool AUX, unsigned KM, class Src>
__device__ inline void lp_small(const Src &src, int wave_g, int nwaves, int *list) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    const int per_wave = WAVE / G;
    const int nunits = src.units();
    for (int u = wave_g; u < nunits; u += nwaves) {
        LPArgs a;
        int q0, q1;
        if (!src.unit(u, a, q0, q1)) continue;
        // compaction of the active vertices among (at most CP) consecutive positions
        int p = q0 + lane;
        bool valid = lane < CP && p < q1;
        int v = valid ? vertex_at(a, p) : 0;
        bool act = valid && lp_active<KM>(a, v);
        uint64_t am = __ballot(act);
        int nact = __popcll(am);
        if (act) list[__popcll(am & lt)] = v;
        wave_sync();
        // the row, own label and own community record of every active vertex
        // of the unit, one lane each, loaded together (nact <= CP <= 64);
        // each group takes its vertices' values by shuffle
        int pu = 0, pstart = 0, pdeg = 0, pown = 0, paux = 0;
        CommStats pst{0, 0, 0, 0};
        uint64_t ptb = 0;  // the vertex's tie-key base (one hash per vertex, not per pass)
        if (lane < nact) {
            pu = list[lane];
            pstart = a.indices[pu];
            pdeg = a.indices[pu + 1] - pstart;
            if constexpr (AUX) {
                const int2 o = a.lv[pu];
                pown = o.x;
                paux = o.y;
            } else {
                pown = own_label<KM>(a, pu);
                pst = comm_rec<KM>(a, pown);
            }
            if (a.tie != PCD_TIE_LOWEST) ptb = tie_base(a.seed, step_of(a), pu);
        }
        for (int gb = 0; gb < nact; gb += per_wave) {
            int gi = gb + lane / G;
            bool has = gi < nact;
            const int sl = has ? gi : 0;
            int u = __shfl(pu, sl, WAVE), start = __shfl(pstart, sl, WAVE), deg = __shfl(pdeg, sl, WAVE);
            const uint64_t tb = ((uint64_t)(uint32_t)__shfl((int)(ptb >> 32), sl, WAVE) << 32) |
                                (uint32_t)__shfl((int)(uint32_t)ptb, sl, WAVE);
            int2 own = make_int2(__shfl(pown, sl, WAVE), __shfl(paux, sl, WAVE));
            const CommStats own_rec{__shfl(pst.size, sl, WAVE), __shfl(pst.int_vol, sl, WAVE),
                                    __shfl(pst.ext_vol, sl, WAVE), 0};
            if (!has) {
                u = start = deg = 0;
                own = make_int2(0, 0);
            }
            int2 r = make_int2(INT32_MAX, 0);
            if (has && glane < deg) {
                int nb = a.adj[start + glane];
                r = AUX ? a.lv[nb] : make_int2(entry_label<KM>(a, nb), 0);
            }
            int x = r.x, xa = r.y;
            if constexpr (AUX) {
#pragma unroll
                for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
                    for (int j = k >> 1; j > 0; j >>= 1) {
                        const bool take_min = ((glane & k) == 0) == ((glane & j) == 0);
                        const int y = __shfl_xor(x, j, WAVE), ya = __shfl_xor(xa, j, WAVE);
                        const bool other = take_min ? (y < x) : (y > x);
                        x = other ? y : x;
                        xa = other ? ya : xa;
                    }
                }
            } else {
                x = group_sort<G>(x, glane);
            }
            const int prev = prev_lane(x);  // read only where glane > 0
            bool head = x != INT32_MAX && (glane == 0 || x != prev);
            uint64_t heads = __ballot(head);
            Cand c = cand_none();
            int own_cnt = 0;
            if (head) {
                uint64_t above = lane == 63 ? 0ULL : (heads >> (lane + 1));
                int nxt = above ? lane + 1 + __builtin_ctzll(above) : WAVE;
                int gend = (lane & ~(G - 1)) + deg;
                int cnt = min(nxt, gend) - lane;
                if (x == own.x) {
                    own_cnt = cnt;
                } else {
                    const CommStats xr = AUX ? CommStats{xa, xa, 0, 0} : comm_rec<KM>(a, x);
                    if ((KM & RK_MERGE) == 0 || !xr.pad) {  // merging: matched communities are no candidates
                        double sc;
                        if (cand_score<KM>(a, deg, own_rec, xr, cnt, sc)) c = Cand{sc, tie_key(a.tie, tb, x), x};
                    }
                }
            }
            group_reduce_cand<G / 2>(c, own_cnt);
            if (has && glane == 0) {
                if constexpr ((KM & RK_MERGE) != 0) a.new_labels[u] = merge_partner(c);
                else a.new_labels[u] = AUX ? lp_decide<KM>(a, deg, own.x, own.y, own_cnt, c)
                                           : lp_decide_rec<KM>(a, deg, own.x, own_rec, own_cnt, c);
            }
            if (a.counters) {
                // every lane votes: a wave holds up to 64 / G active vertices
                const int act = __popcll(__ballot(has && glane == 0));
                int ent = wave_sum((has && glane < deg) ? 1 : 0);
                if (lane == 0) count3(a, act, ent, __popcll(heads));
            }
        }
        wave_sync();
    }
}

template <int G, int CP, bool AUX, unsigned KM, class Src>
__global__ __launch_bounds__(256) void k_lp_small(Src src) {
    __shared__ int s_list[256 / WAVE][WAVE];
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    lp_small<G, CP, AUX, KM>(src, wave_g, gridDim.x * blockDim.x / WAVE, s_list[threadIdx.x / WAVE]);
}

// ------------------------------------------------------------- LDS hash table
// Hub partition tables: T slots (any multiple of 64), linear probing from
// (hash * T) >> 32.
__device__ inline int tslot(int key, int T) {
    return (int)(((uint64_t)((uint32_t)key * 0x9E3779B1u) * (uint32_t)T) >> 32);
}

// Listed-slot scans read SB slots at a time so that their community-record
// gathers overlap.
constexpr int SB = 4;

__device__ inline void wave_reduce_cand(Cand &c, int &own_cnt) { group_reduce_cand<WAVE / 2>(c, own_cnt); }

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

__device__ inline int round64(int x) { return (x + 63) & ~63; }

// ------------------------------------------------------------- label tables
// Tables of the per-vertex kernels hold nextpow2(2 deg) slots, more than the
// row's distinct labels, so probing always ends at a free or matching slot.
__device__ inline uint32_t pslot(int key, int log_t) { return ((uint32_t)key * 0x9E3779B1u) >> (32 - log_t); }

// Counts of a table: 32-bit, or 16-bit where a row's degree bounds them
// below 2^16.  A 16-bit count is incremented by a 32-bit LDS atomic on the
// word holding it (the other half cannot carry into it); it is read and
// cleared with 16-bit accesses.
__device__ inline void count_add(int *cnts, uint32_t h) { atomicAdd(&cnts[h], 1); }
__device__ inline void count_add(uint16_t *cnts, uint32_t h) {
    atomicAdd((unsigned *)(cnts + (h & ~1u)), 1u << ((h & 1u) * 16));
}

// Inserts `key` (with AUX its aux, stored by the lane that claims the slot)
// with count 1; returns true when this lane claimed a free slot, whose
// index lands in `slot`.
template <bool AUX, class CT = int>
__device__ inline bool pinsert(int *keys, CT *cnts, int *auxs, int mask, int log_t, int key, int aux, int &slot) {
    uint32_t h = pslot(key, log_t);
    while (true) {
        int old = atomicCAS(&keys[h], EMPTY_KEY, key);
        if (AUX && old == EMPTY_KEY) auxs[h] = aux;
        if (old == EMPTY_KEY || old == key) {
            count_add(cnts, h);
            slot = (int)h;
            return old == EMPTY_KEY;
        }
        h = (h + 1) & mask;
    }
}

// Appends the slots this wave's lanes claimed to the table's slot list
// (one LDS atomic per wave).  Called by every lane of the wave.
__device__ inline void occ_append(bool claimed, int slot, uint16_t *occ, int *nocc) {
    const uint64_t cm = __ballot(claimed);
    if (!cm) return;
    const int lane = __lane_id();
    const int leader = __builtin_ctzll(cm);
    int base = 0;
    if (lane == leader) base = atomicAdd(nocc, __popcll(cm));
    base = __builtin_amdgcn_readlane(base, leader);
    if (claimed) occ[base + __popcll(cm & ((1ULL << lane) - 1))] = (uint16_t)slot;
}

// Gathers the labels (with AUX the {label, aux} records: one 8-byte gather
// brings what the gains read, and aux lands in the table) of row entries
// base + lo + e * stride, PF at a time (independent loads in flight), and
// inserts them; claimed slots go to the slot list.  A merging row (its
// entries are labels) leaves out its own label, which is never a
// candidate.  The loop bounds are the same for every lane of the workgroup.
template <int PF, unsigned KM, bool AUX = false, class CT = int>
__device__ inline void insert_labels(const LPArgs &a, int *keys, CT *cnts, int *auxs, uint16_t *occ, int *nocc,
                                     int mask, int log_t, int start, int deg, int lo, int stride, int own) {
    constexpr bool MERGE = (KM & RK_MERGE) != 0;
    for (int base = 0; base < deg; base += PF * stride) {
        int nb[PF];
#pragma unroll
        for (int e = 0; e < PF; e++) {
            int j = base + lo + e * stride;
            nb[e] = j < deg ? a.adj[start + j] : -1;
        }
        int2 rec[PF];
#pragma unroll
        for (int e = 0; e < PF; e++)
            rec[e] = nb[e] < 0 ? make_int2(EMPTY_KEY, 0)
                               : (AUX ? a.lv[nb[e]] : make_int2(entry_label<KM>(a, nb[e]), 0));
#pragma unroll
        for (int e = 0; e < PF; e++) {
            bool claimed = false;
            int slot = 0;
            if (rec[e].x != EMPTY_KEY && (!MERGE || rec[e].x != own))
                claimed = pinsert<AUX, CT>(keys, cnts, auxs, mask, log_t, rec[e].x, rec[e].y, slot);
            occ_append(claimed, slot, occ, nocc);
        }
    }
}

// Scans the D listed slots [lo, D) with stride `stride` for candidates, SB
// at a time so that their community-record gathers overlap, and clears each
// slot it read (the table is empty again for the next row).
template <bool AUX, unsigned KM, class CT = int>
__device__ inline void scan_listed(const LPArgs &a, int *keys, CT *cnts, const int *auxs, const uint16_t *occ, int D,
                                   int lo, int stride, int deg, int own, uint64_t tb, Cand &c, int &own_cnt) {
    constexpr bool MERGE = (KM & RK_MERGE) != 0;
    CommStats orec{0, 0, 0, 0};
    if constexpr (MERGE) orec = a.stats[own];
    for (int i0 = lo; i0 < D; i0 += SB * stride) {
        int key[SB], cnt[SB];
        CommStats st[SB];
#pragma unroll
        for (int e = 0; e < SB; e++) {
            const int i = i0 + e * stride;
            key[e] = EMPTY_KEY;
            cnt[e] = 0;
            if (i < D) {
                const int h = occ[i];
                key[e] = keys[h];
                cnt[e] = cnts[h];
                if (AUX) {
                    const int x = auxs[h];
                    st[e] = CommStats{x, x, 0, 0};
                }
                keys[h] = EMPTY_KEY;
                cnts[h] = 0;
            }
        }
        if (!AUX) {
#pragma unroll
            for (int e = 0; e < SB; e++)
                if (key[e] != EMPTY_KEY && key[e] != own) st[e] = comm_rec<KM>(a, key[e]);
        }
#pragma unroll
        for (int e = 0; e < SB; e++) {
            if (key[e] == EMPTY_KEY) continue;
            if (MERGE && key[e] != own && st[e].pad) continue;  // matched: no candidate
            if (key[e] == own) {
                own_cnt = cnt[e];
            } else {
                double sc;
                if (cand_score<KM>(a, deg, orec, st[e], cnt[e], sc)) {
                    Cand o{sc, tie_key(a.tie, tb, key[e]), key[e]};
                    if (cand_better(o, c)) c = o;
                }
            }
        }
    }
}

__device__ inline int ceil_log2(int x) { return x <= 1 ? 0 : 32 - __clz(x - 1); }

// ------------------------------------------------------------- wave / vertex
// The table starts empty and each row's scan clears the slots it used.
// One wave per vertex over the units of src, with this wave's LDS table
// (TW slots of keys, counts and aux, a slot list of TW / 2 entries).
template <int TW, unsigned KM, class Src, bool AUX, int PF, class CT = int>
__device__ inline void lp_wave(const Src &src, int wave_g, int nwaves, int *keys, CT *cnts, int *auxs, uint16_t *occ,
                               int *nocc) {
    const int lane = __lane_id();
    for (int i = lane; i < TW; i += WAVE) {
        keys[i] = EMPTY_KEY;
        cnts[i] = 0;
    }
    if (lane == 0) *nocc = 0;
    wave_sync();
    const int nunits = src.units();
    for (int u = wave_g; u < nunits; u += nwaves) {
      LPArgs a;
      int q0, q1;
      if (!src.unit(u, a, q0, q1)) continue;
      for (int p = q0; p < q1; p++) {
        int v = vertex_at(a, p);
        if (!lp_active<KM>(a, v)) continue;
        int start = a.indices[v], deg = a.indices[v + 1] - start;
        int2 ownr = AUX ? a.lv[v] : make_int2(own_label<KM>(a, v), 0);
        int own = ownr.x;
        int log_t = max(6, ceil_log2(2 * deg));
        insert_labels<PF, KM, AUX, CT>(a, keys, cnts, auxs, occ, nocc, (1 << log_t) - 1, log_t, start, deg, lane, WAVE,
                                       own);
        wave_sync();
        const int D = *nocc;
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, step_of(a), v);
        Cand c = cand_none();
        int own_cnt = 0;
        scan_listed<AUX, KM, CT>(a, keys, cnts, auxs, occ, D, lane, WAVE, deg, own, tb, c, own_cnt);
        wave_reduce_cand(c, own_cnt);
        if (lane == 0) {
            if constexpr ((KM & RK_MERGE) != 0) a.new_labels[v] = merge_partner(c);
            else a.new_labels[v] = AUX ? lp_decide<KM>(a, deg, own, ownr.y, own_cnt, c)
                                       : lp_decide_stats<KM>(a, deg, own, own_cnt, c);
            if (a.counters) count3(a, 1, deg, D);
            *nocc = 0;
        }
        wave_sync();
      }
    }
}

// CT = uint16_t needs degrees below 2^16 (the table's counts).
template <int TW, unsigned KM, class Src, bool AUX = false, int PF = 4, class CT = int>
__global__ __launch_bounds__(256) void k_lp_wave(Src src) {
    __shared__ int s_keys[256 / WAVE][TW];
    __shared__ __align__(16) CT s_cnts[256 / WAVE][TW];
    __shared__ int s_auxs[256 / WAVE][AUX ? TW : 1];
    __shared__ uint16_t s_occ[256 / WAVE][TW / 2];
    __shared__ int s_nocc[256 / WAVE];
    const int w = threadIdx.x / WAVE;
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    lp_wave<TW, KM, Src, AUX, PF, CT>(src, wave_g, gridDim.x * blockDim.x / WAVE, s_keys[w], s_cnts[w], s_auxs[w],
                                      s_occ[w], &s_nocc[w]);
}

// The W256 and G64..G4 bins in one launch: the grid is split into
// contiguous workgroup ranges, one per bin, in proportion to the bin's
// entries (at least one workgroup for a bin that is not empty), so the
// bins' ramps and tails overlap instead of following one another.  The
// W256 waves keep their 512-slot tables; the LDS of a workgroup is that
// of a W256 workgroup either way.
template <class Src>
struct SmallBinsT {
    Src src[6];       // W256, G64, G32, G16, G8, G4
    int lo[7];        // workgroup range of bin b: [lo[b], lo[b + 1])
    const int *dlo;   // with DevRangeSrc: the seven bounds in device memory (laid out on the device)
    __device__ int lo_of(int b) const {
        if constexpr (std::is_same<Src, DevRangeSrc>::value) return dlo[b];
        else return lo[b];
    }
};
using SmallBins = SmallBinsT<RangeSrc>;

// WPE: waves per SIMD the registers are allocated for (1: the compiler's
// choice).  At 8 the modularity and CPM builds fit 64 VGPRs without scratch;
// the scalar registers beyond 96 spill to VGPR lanes.  SAUX: the G bins
// gather {label, aux} records (one 8-byte gather brings a neighbour's label
// and its community's total, one dependent load fewer per row); W256 keeps
// gathering labels.
template <unsigned KM, int WPE = 1, bool SAUX = false, class Src = RangeSrc>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_lp_small_all(SmallBinsT<Src> sb) {
    __shared__ int s_keys[256 / WAVE][512];
    __shared__ int s_cnts[256 / WAVE][512];
    __shared__ uint16_t s_occ[256 / WAVE][256];
    __shared__ int s_nocc[256 / WAVE];
    __shared__ int s_list[256 / WAVE][WAVE];
    int b = 0;
    while (b < 6 && !(sb.lo_of(b) <= (int)blockIdx.x && (int)blockIdx.x < sb.lo_of(b + 1))) b++;
    if (b == 6) return;
    const int wpb = blockDim.x / WAVE, w = threadIdx.x / WAVE;
    const int b0 = sb.lo_of(b);
    const int wave_g = __builtin_amdgcn_readfirstlane(((int)blockIdx.x - b0) * wpb + w);
    const int nwaves = (sb.lo_of(b + 1) - b0) * wpb;
    int *list = s_list[w];
    switch (b) {
    case 0: lp_wave<512, KM, Src, false, 4>(sb.src[0], wave_g, nwaves, s_keys[w], s_cnts[w], nullptr, s_occ[w],
                                            &s_nocc[w]); break;
    case 1: lp_small<64, 8, SAUX, KM>(sb.src[1], wave_g, nwaves, list); break;
    case 2: lp_small<32, 16, SAUX, KM>(sb.src[2], wave_g, nwaves, list); break;
    case 3: lp_small<16, 32, SAUX, KM>(sb.src[3], wave_g, nwaves, list); break;
    case 4: lp_small<8, 64, SAUX, KM>(sb.src[4], wave_g, nwaves, list); break;
    default: lp_small<4, 64, SAUX, KM>(sb.src[5], wave_g, nwaves, list); break;
    }
}


// Table slots: nextpow2(2 deg), at most TB.
// CT = uint16_t needs rows below 2^16 entries (the table's counts).
template <int NT, int TB, unsigned KM, class Src, bool AUX = false, int PF = 4, class CT = int>
__global__ __launch_bounds__(NT) void k_lp_block(Src src) {
    __shared__ int keys[TB];
    __shared__ __align__(16) CT cnts[TB];
    __shared__ int auxs[AUX ? TB : 1];
    __shared__ uint16_t occ[TB / 2];
    __shared__ int s_nocc;
    __shared__ Cand s_c[NT / WAVE];
    __shared__ int s_i[2 * NT / WAVE];
    for (int i = threadIdx.x; i < TB; i += NT) {
        keys[i] = EMPTY_KEY;
        cnts[i] = 0;
    }
    if (threadIdx.x == 0) s_nocc = 0;
    __syncthreads();
    const int nunits = src.units();
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
      LPArgs a;
      int q0, q1;
      if (!src.unit(u, a, q0, q1)) continue;
      for (int p = q0; p < q1; p++) {
        int v = vertex_at(a, p);
        if (!lp_active<KM>(a, v)) continue;
        int start = a.indices[v], deg = a.indices[v + 1] - start;
        int2 ownr = AUX ? a.lv[v] : make_int2(own_label<KM>(a, v), 0);
        int own = ownr.x;
        int log_t = max(6, ceil_log2(2 * deg));
        insert_labels<PF, KM, AUX, CT>(a, keys, cnts, auxs, occ, &s_nocc, (1 << log_t) - 1, log_t, start, deg,
                                      threadIdx.x, NT, own);
        __syncthreads();
        const int D = s_nocc;
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, step_of(a), v);
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        scan_listed<AUX, KM, CT>(a, keys, cnts, auxs, occ, D, threadIdx.x, NT, deg, own, tb, c, own_cnt);
        block_reduce_cand<NT>(c, own_cnt, distinct, s_c, s_i);
        if (threadIdx.x == 0) {
            if constexpr ((KM & RK_MERGE) != 0) a.new_labels[v] = merge_partner(c);
            else a.new_labels[v] = AUX ? lp_decide<KM>(a, deg, own, ownr.y, own_cnt, c)
                                       : lp_decide_stats<KM>(a, deg, own, own_cnt, c);
            if (a.counters) count3(a, 1, deg, D);
            s_nocc = 0;
        }
        __syncthreads();
      }
    }
}

// ------------------------------------------------------------- hubs
struct HubArgs {
    const int4 *chunks;        // {hub, v, begin, end}
    const int32_t *chunk_off;  // chunk -> first of its P+1 offsets in poff
    const int4 *parts;         // {hub, part, first chunk, number of chunks}
    const int4 *where;         // per part item: {v, P, poff of the hub's first chunk, buffer base}
    const int4 *info;          // per hub {v, P, first part item, first chunk}
    const int64_t *buf_base;   // per hub
    int32_t *poff;
    int2 *buf;
    HubPartial *partial;
    int32_t *over;
    int32_t *over_count;  // partitions that overflowed (diagnostic)
    int nchunks, nparts, nhubs;
    int table_slots;  // LDS slots per partition (multiple of 64, <= HUB_TB)
    const int *counts;  // device {nchunks, nparts, nhubs} (a layout built on the device), or nullptr
};

__device__ inline HubArgs hub_args(HubArgs h) {
    if (h.counts) {
        h.nchunks = h.counts[0];
        h.nparts = h.counts[1];
        h.nhubs = h.counts[2];
    }
    return h;
}

// Partition of a label: murmur3's 32-bit finaliser, independent of the
// multiplicative slot hash, so one partition's keys still spread over the
// whole table.
__device__ inline int hub_part_of(int label, int P) {
    uint32_t x = (uint32_t)label;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return (int)(((uint64_t)x * (uint32_t)P) >> 32);
}

constexpr int HUB1_NT = 512, HUB1_PER = HUB_CHUNK / HUB1_NT;
static_assert(HUB1_NT * HUB1_PER == HUB_CHUNK, "hub chunk layout");

// pass 1: gather the chunk's labels (with AUX its {label, aux} records) and
// group them by partition; a merging row leaves its own label out
template <unsigned KM, bool AUX>
__global__ __launch_bounds__(HUB1_NT) void k_hub_partition(LPArgs a, HubArgs hin) {
    const HubArgs h = hub_args(hin);
    constexpr bool MERGE = (KM & RK_MERGE) != 0;
    using Rec = typename std::conditional<AUX, int2, int32_t>::type;
    Rec *out = reinterpret_cast<Rec *>(h.buf);
    __shared__ int cnt[HUB_MAX_PARTS + 1];
    for (int it = blockIdx.x; it < h.nchunks; it += gridDim.x) {
        const int4 ck = h.chunks[it];
        const int v = ck.y;
        if (!lp_active<KM>(a, v)) continue;
        const int P = h.info[ck.x].y;
        for (int i = threadIdx.x; i <= P; i += HUB1_NT) cnt[i] = 0;
        __syncthreads();
        int nb[HUB1_PER], part[HUB1_PER], rank[HUB1_PER];
        Rec rec[HUB1_PER];
#pragma unroll
        for (int e = 0; e < HUB1_PER; e++) {
            int j = ck.z + e * HUB1_NT + threadIdx.x;
            nb[e] = j < ck.w ? a.adj[j] : -1;
        }
#pragma unroll
        for (int e = 0; e < HUB1_PER; e++) {
            if constexpr (AUX) rec[e] = nb[e] >= 0 ? a.lv[nb[e]] : make_int2(0, 0);
            else rec[e] = nb[e] >= 0 ? entry_label<KM>(a, nb[e]) : 0;
        }
#pragma unroll
        for (int e = 0; e < HUB1_PER; e++) {
            part[e] = -1;
            if (nb[e] >= 0) {
                int key;
                if constexpr (AUX) key = rec[e].x;
                else key = rec[e];
                if (!MERGE || key != v) {
                    part[e] = hub_part_of(key, P);
                    rank[e] = atomicAdd(&cnt[part[e]], 1);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // exclusive scan of the partition counts
            int run = 0;
            for (int q = 0; q < P; q++) {
                int t = cnt[q];
                cnt[q] = run;
                run += t;
            }
            cnt[P] = run;
        }
        __syncthreads();
        const int rel = ck.z - a.indices[v];  // chunk position inside the hub row
        const int64_t base = h.buf_base[ck.x] + rel;
        int32_t *po = h.poff + h.chunk_off[it];
        for (int q = threadIdx.x; q <= P; q += HUB1_NT) po[q] = rel + cnt[q];
#pragma unroll
        for (int e = 0; e < HUB1_PER; e++)
            if (part[e] >= 0) out[base + cnt[part[e]] + rank[e]] = rec[e];
        __syncthreads();
    }
}

// Inserts (key, aux) with count 1 into a hub partition table of T slots,
// probing at most T slots; false when the table is full.  claimed: this
// lane took a free slot, whose index lands in `slot`.
template <bool AUX>
__device__ inline bool tclaim(int *keys, int *cnts, int *auxs, int T, int key, int aux, bool &claimed, int &slot) {
    int hs = tslot(key, T);
    for (int probe = 0; probe < T; probe++) {
        int old = atomicCAS(&keys[hs], EMPTY_KEY, key);
        if (AUX && old == EMPTY_KEY) auxs[hs] = aux;
        if (old == EMPTY_KEY || old == key) {
            atomicAdd(&cnts[hs], 1);
            claimed = old == EMPTY_KEY;
            slot = hs;
            return true;
        }
        if (++hs == T) hs = 0;
    }
    return false;
}

// pass 2: one workgroup per (hub, partition) counts the partition in an LDS
// table that lists the slots it claims, so the scan visits only the
// partition's distinct labels (reading each one's community record) and
// clears just those slots for the next partition.  A key that finds the
// table full flags the hub for the global-table path; the listed slots are
// cleared all the same.
// NT threads per workgroup, TB table slots (the tuned table size is capped
// at TB); the entries of a partition are gathered PF per thread at a time so
// that their loads are in flight together.  With AUX the records are
// {label, aux} and the table keeps aux (no community-record reads).
// A partition's entries lie in one segment per chunk of its hub; each wave
// builds its own copy of the segment table, 64 chunks at a time (wave scan,
// no workgroup barrier), and the next partition item is loaded while the
// current one is counted.
template <unsigned KM, int NT, int TB, bool AUX>
__global__ __launch_bounds__(NT) void k_hub_count_listed(LPArgs a, HubArgs hin) {
    const HubArgs h = hub_args(hin);
    constexpr int PF = 4;
    constexpr int NW = NT / WAVE;
    using Rec = typename std::conditional<AUX, int2, int32_t>::type;
    __shared__ int keys[TB];
    __shared__ int cnts[TB];
    __shared__ int auxs[AUX ? TB : 1];
    __shared__ uint16_t occ[TB];
    __shared__ int s_seg[NW][WAVE];
    __shared__ int s_segbase[NW][WAVE];
    __shared__ Cand s_c[NW];
    __shared__ int s_i[2 * NW];
    __shared__ int s_over, s_nocc;
    const int T = min(h.table_slots, TB);
    const int lane = __lane_id(), w = threadIdx.x / WAVE;
    int *seg = s_seg[w], *segbase = s_segbase[w];
    for (int i = threadIdx.x; i < TB; i += NT) {
        keys[i] = EMPTY_KEY;
        cnts[i] = 0;
    }
    if (threadIdx.x == 0) { s_over = 0; s_nocc = 0; }
    __syncthreads();
    const Rec *buf = reinterpret_cast<const Rec *>(h.buf);
    int it = blockIdx.x;
    int4 pi = make_int4(0, 0, 0, 0), pw = make_int4(0, 0, 0, 0);
    if (it < h.nparts) { pi = h.parts[it]; pw = h.where[it]; }
    while (it < h.nparts) {
        const int nxt = it + gridDim.x;
        int4 pi_n = pi, pw_n = pw;
        if (nxt < h.nparts) { pi_n = h.parts[nxt]; pw_n = h.where[nxt]; }
        const int v = pw.x, P = pw.y, part = pi.y;
        if (lp_active<KM>(a, v)) {
            for (int c0 = 0; c0 < pi.w; c0 += WAVE) {
                const int nc = min(WAVE, pi.w - c0);
                int len = 0, b0 = 0;
                if (lane < nc) {
                    const int32_t *po = h.poff + pw.z + (int64_t)(c0 + lane) * (P + 1);
                    b0 = po[part];
                    len = po[part + 1] - b0;
                }
                const int inc = wave_incl_scan(len);
                const int tot = __shfl(inc, WAVE - 1, WAVE);
                seg[lane] = inc - len;
                segbase[lane] = pw.w + b0;
                wave_sync();
                // the trip count is the same for every lane of the
                // workgroup: occ_append is called by whole waves
                for (int i0 = 0; i0 < tot; i0 += PF * NT) {
                    int key[PF], aux[PF];
#pragma unroll
                    for (int e = 0; e < PF; e++) {
                        const int i = i0 + e * NT + threadIdx.x;
                        key[e] = EMPTY_KEY;
                        aux[e] = 0;
                        if (i < tot) {
                            int lo = 0, hi = nc;  // last segment with seg[lo] <= i
                            while (hi - lo > 1) {
                                int mid = (lo + hi) >> 1;
                                if (seg[mid] <= i) lo = mid;
                                else hi = mid;
                            }
                            const Rec r = buf[(int64_t)segbase[lo] + (i - seg[lo])];
                            if constexpr (AUX) { key[e] = r.x; aux[e] = r.y; }
                            else key[e] = r;
                        }
                    }
#pragma unroll
                    for (int e = 0; e < PF; e++) {
                        bool claimed = false;
                        int slot = 0;
                        if (key[e] != EMPTY_KEY && !tclaim<AUX>(keys, cnts, auxs, T, key[e], aux[e], claimed, slot))
                            s_over = 1;
                        occ_append(claimed, slot, occ, &s_nocc);
                    }
                }
                wave_sync();
            }
            __syncthreads();
            const int D = s_nocc;
            const bool over = s_over != 0;
            const int deg = a.indices[v + 1] - a.indices[v], own = own_label<KM>(a, v);
            Cand c = cand_none();
            int own_cnt = 0, distinct = threadIdx.x == 0 ? D : 0;
            if (over) {
                for (int i = threadIdx.x; i < D; i += NT) {
                    const int hs = occ[i];
                    keys[hs] = EMPTY_KEY;
                    cnts[hs] = 0;
                }
            } else {
                uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, step_of(a), v);
                scan_listed<AUX, KM>(a, keys, cnts, auxs, occ, D, threadIdx.x, NT, deg, own, tb, c, own_cnt);
            }
            block_reduce_cand<NT>(c, own_cnt, distinct, s_c, s_i);
            if (threadIdx.x == 0) {
                if (over) {
                    h.over[pi.x] = 1;
                    atomicAdd(h.over_count, 1);
                } else {
                    h.partial[it] = HubPartial{c, own_cnt, distinct};
                }
                s_over = 0;
                s_nocc = 0;
            }
            __syncthreads();
        }
        it = nxt;
        pi = pi_n;
        pw = pw_n;
    }
}

// pass 3: one wave per hub combines its partitions and decides
template <unsigned KM, bool AUX>
__global__ __launch_bounds__(256) void k_hub_final(LPArgs a, HubArgs hin) {
    const HubArgs h = hub_args(hin);
    const int lane = __lane_id();
    const int wave_g = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int hub = wave_g; hub < h.nhubs; hub += nwaves) {
        const int4 inf = h.info[hub];
        const int v = inf.x, P = inf.y;
        if (!lp_active<KM>(a, v)) continue;
        if (h.over[hub]) continue;  // recounted by the global-table path
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        for (int q = lane; q < P; q += WAVE) {
            HubPartial hp = h.partial[inf.z + q];
            if (cand_better(hp.c, c)) c = hp.c;
            own_cnt = max(own_cnt, hp.own_cnt);
            distinct += hp.distinct;
        }
        wave_reduce_cand(c, own_cnt);
        distinct = wave_sum(distinct);
        if (lane == 0) {
            const int deg = a.indices[v + 1] - a.indices[v];
            if constexpr ((KM & RK_MERGE) != 0) {
                a.new_labels[v] = merge_partner(c);
            } else if constexpr (AUX) {
                int2 own = a.lv[v];
                a.new_labels[v] = lp_decide<KM>(a, deg, own.x, own.y, own_cnt, c);
            } else {
                a.new_labels[v] = lp_decide_stats<KM>(a, deg, a.labels[v], own_cnt, c);
            }
            if (a.counters) count3(a, 1, deg, distinct);
        }
    }
}

// Fallback for flagged hubs: the whole row counted in a per-hub global
// open-addressing table of nextpow2(2 deg) slots (never full); community
// records are read from the statistics.
__device__ inline uint32_t gslot(int key, int log_t) { return ((uint32_t)key * 0x9E3779B1u) >> (32 - log_t); }

template <unsigned KM>
__global__ __launch_bounds__(512) void k_hub_global_insert(LPArgs a, HubArgs hin, const int64_t *toff,
                                                           const int32_t *tlog, int32_t *gkeys, int32_t *gcnts) {
    const HubArgs h = hub_args(hin);
    for (int it = blockIdx.x; it < h.nchunks; it += gridDim.x) {
        const int4 ck = h.chunks[it];
        if (!h.over[ck.x]) continue;
        const int64_t off = toff[ck.x];
        const int glog = tlog[ck.x], gmask = (1 << glog) - 1;
        for (int j = ck.z + threadIdx.x; j < ck.w; j += blockDim.x) {
            int key = entry_label<KM>(a, a.adj[j]);
            uint32_t s = gslot(key, glog);
            while (true) {
                int old = atomicCAS(&gkeys[off + s], EMPTY_KEY, key);
                if (old == EMPTY_KEY || old == key) {
                    atomicAdd(&gcnts[off + s], 1);
                    break;
                }
                s = (s + 1) & gmask;
            }
        }
    }
}

template <unsigned KM>
__global__ __launch_bounds__(1024) void k_hub_global_decide(LPArgs a, HubArgs hin, const int64_t *toff,
                                                            const int32_t *tlog, int32_t *gkeys, int32_t *gcnts) {
    const HubArgs h = hub_args(hin);
    __shared__ Cand s_c[1024 / WAVE];
    __shared__ int s_i[2 * 1024 / WAVE];
    for (int hub = blockIdx.x; hub < h.nhubs; hub += gridDim.x) {
        if (!h.over[hub]) continue;
        const int v = h.info[hub].x;
        const int deg = a.indices[v + 1] - a.indices[v], own = own_label<KM>(a, v);
        CommStats orec{0, 0, 0, 0};
        if constexpr ((KM & RK_MERGE) != 0) orec = a.stats[own];
        const int64_t off = toff[hub];
        const int T = 1 << tlog[hub];
        uint64_t tb = a.tie == PCD_TIE_LOWEST ? 0 : tie_base(a.seed, step_of(a), v);
        Cand c = cand_none();
        int own_cnt = 0, distinct = 0;
        for (int s = threadIdx.x; s < T; s += 1024) {
            int key = gkeys[off + s];
            if (key == EMPTY_KEY) continue;
            int cnt = gcnts[off + s];
            gkeys[off + s] = EMPTY_KEY;
            gcnts[off + s] = 0;
            distinct++;
            if (key == own) own_cnt = cnt;
            else {
                const CommStats xr = comm_rec<KM>(a, key);
                double sc;
                if (((KM & RK_MERGE) == 0 || !xr.pad) && cand_score<KM>(a, deg, orec, xr, cnt, sc)) {
                    Cand o{sc, tie_key(a.tie, tb, key), key};
                    if (cand_better(o, c)) c = o;
                }
            }
        }
        block_reduce_cand<1024>(c, own_cnt, distinct, s_c, s_i);
        if (threadIdx.x == 0) {
            if constexpr ((KM & RK_MERGE) != 0) {
                a.new_labels[v] = merge_partner(c);
            } else {
                double leave = score_leave<KM>(a.rule, deg, comm_rec<KM>(a, own), own_cnt);
                double bs = c.label != INT32_MAX ? c.s : 0.0;
                bs += leave;
                a.new_labels[v] = (bs > 0 && bs > leave) ? c.label : (leave > 0 ? PCD_NEW_LABEL : own);
            }
            if (a.counters) count3(a, 1, deg, distinct);
            h.over[hub] = 0;
        }
    }
}

// ------------------------------------------------------------- driver
RowSet graph_rows(pcd_ctx *ctx) {
    RowSet r;
    r.bins = ctx->bins;
    for (int b = 0; b < NUM_BINS; b++) r.bin_entries[b] = ctx->bin_entries[b];
    r.num_hubs = ctx->num_hubs;
    r.num_hub_items = ctx->num_hub_items;
    r.num_hub_parts = ctx->num_hub_parts;
    r.hub_items = ctx->d_hub_items;
    r.hub_item_off = ctx->d_hub_item_off;
    r.hub_part_items = ctx->d_hub_part_items;
    r.hub_info = ctx->d_hub_info;
    r.hub_part_where = ctx->d_hub_part_where;
    r.hub_buf_base = ctx->d_hub_buf_base;
    r.hub_poff = ctx->d_hub_poff;
    r.hub_buf = ctx->d_hub_buf;
    r.hub_partial = ctx->d_hub_partial;
    r.hub_over = ctx->d_hub_over;
    r.hub_table_off = ctx->d_hub_table_off;
    r.hub_table_log = ctx->d_hub_table_log;
    r.hub_keys = ctx->d_hub_keys;
    r.hub_cnts = ctx->d_hub_cnts;
    return r;
}

// Launches the propagation kernels of one step over the rows `rs`, compiled
// for the rule kinds KM (with RK_MERGE: one merging round over community
// rows, which writes each active community's partner).
template <unsigned KM>
static int launch_propagation(pcd_ctx *ctx, const RowSet &rs, const LPArgs &a) {
    constexpr bool MERGE = (KM & RK_MERGE) != 0;
    const BinLayout &b = rs.bins;
    auto args_for = [&](int bin) {
        LPArgs x = a;
        if (ctx->counters && !MERGE) x.counters = ctx->d_bin_counters + 3 * bin;
        return x;
    };
    // Bins laid out on the device (the merging rows): every bin is launched
    // with a fixed grid and reads its range there; the propagation rows'
    // bins are host ranges.
    const bool dev = rs.layout != nullptr;
    if (dev != MERGE) return fail(PCD_E_STATE, "row set layout does not match the kernel build");
    using Src = typename std::conditional<MERGE, DevRangeSrc, RangeSrc>::type;
    auto range = [&](int bin, int &p0, int &p1) {
        p0 = dev ? 0 : b.start[bin];
        p1 = dev ? 0 : b.start[bin + 1];
        return dev || p1 > p0;
    };
    auto src = [&](int bin, int q0, int q1, int per) {
        if constexpr (MERGE) return DevRangeSrc{rs.layout + bin, per, args_for(bin)};
        else return RangeSrc{q0, q1, per, args_for(bin)};
    };
    auto grid = [&](int64_t items, int per_block, int max_per_cu) {
        return dev ? ctx->cu_count * max_per_cu : grid_for(ctx, items, per_block, max_per_cu);
    };
    // {label, aux} records are read by the hub path and, when tuned so, by
    // the small bins; rules without aux read the statistics instead
    // inactive vertices keep their label: copy all, the kernels overwrite the
    // active ones (a merging round writes the partners of active communities
    // into the partner array, which keeps the others)
    if (!MERGE)
        PCD_HIP(hipMemcpyAsync(a.new_labels, a.labels, sizeof(int32_t) * (size_t)ctx->n, hipMemcpyDeviceToDevice,
                               ctx->stream));
    const bool small_aux = a.use_aux && ctx->tune.small_aux;
    const bool block_aux = a.use_aux && ctx->tune.block_aux;
    // hubs count {label, aux} records only when tuned so (and the rule's
    // gains read aux alone); otherwise labels, with the community record read
    // per distinct label
    const bool hub_aux = a.use_aux && ctx->tune.hub_aux;
    if (a.use_aux && ((rs.num_hubs > 0 && hub_aux) || small_aux || block_aux)) {
        kt_begin(ctx, KS_LP_SELECT);
        hipLaunchKernelGGL(k_lp_records, dim3(grid_for(ctx, ctx->n, 256, 8)), dim3(256), 0, ctx->stream, ctx->n,
                           ctx->d_labels, ctx->d_stats, a.rule.kind, ctx->d_lv);
        kt_end(ctx, KS_LP_SELECT);
    }
    if (a.use_aux) {  // the modularity and CPM builds read the totals
        kt_begin(ctx, KS_LP_SELECT);
        // k is read from d_k (a deferred k is bounded by n + 1)
        const int kg = ctx->k_deferred ? ctx->n + 1 : ctx->k;
        hipLaunchKernelGGL(k_label_totals, dim3(grid_for(ctx, kg, 256, 8)), dim3(256), 0, ctx->stream, ctx->d_k,
                           ctx->d_stats, a.rule.kind, ctx->d_tot);
        kt_end(ctx, KS_LP_SELECT);
    }
    // Hub rows on one side stream and degrees 1025-4096 on another, beside
    // the other bins here (tuning bin_streams): bins write disjoint vertices
    // and read the same step inputs, so one kernel's tail overlaps another's
    // work (R-MAT-22 modularity 1.255 -> 1.083 ms per iteration; sending
    // degrees 513-1024 or 257-512 to the side streams as well, or to a third
    // one, measured slower).  Kept on one stream while a timer brackets a side
    // section (its events are recorded on `stream`).
    constexpr uint32_t SIDE_SLOTS = (1u << KS_LP_HUB_INSERT) | (1u << KS_LP_HUB_DECIDE) | (1u << KS_LP_HUB_FINAL) |
                                    (1u << KS_LP_B4096);
    // Only where there is something to overlap: hub rows or degrees
    // 1025-4096, and at least bin_streams_min_edges edges (below that the
    // fork's event calls cost more than the tails it fills: per-graph runs of
    // SBM graphs up to 20,000 vertices were 5 % slower with the fork).
    const bool side_work = rs.num_hubs > 0 || rs.hub_counts || b.start[BIN_B4096 + 1] > b.start[BIN_B4096];
    const bool fork = !MERGE && ctx->tune.bin_streams && ctx->side[1] && side_work &&
                      ctx->m >= ctx->tune.bin_streams_min_edges && !(ctx->timing && (ctx->timer_mask & SIDE_SLOTS));
    const int nside = fork ? 2 : 0;
    hipStream_t hs = fork ? ctx->side[0] : ctx->stream, bs = fork ? ctx->side[1] : ctx->stream;
    if (fork) {
        PCD_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
        for (int i = 0; i < nside; i++) PCD_HIP(hipStreamWaitEvent(ctx->side[i], ctx->ev_fork, 0));
    }
    int p0, p1;
    if (rs.num_hubs > 0 || rs.hub_counts) {
        HubArgs h;
        h.chunks = rs.hub_items;
        h.chunk_off = rs.hub_item_off;
        h.parts = rs.hub_part_items;
        h.info = rs.hub_info;
        h.where = rs.hub_part_where;
        h.buf_base = rs.hub_buf_base;
        h.poff = rs.hub_poff;
        h.buf = (int2 *)rs.hub_buf;
        h.partial = (HubPartial *)rs.hub_partial;
        h.over = rs.hub_over;
        h.over_count = ctx->d_counts + NUM_BINS + 2;
        h.nchunks = rs.num_hub_items;
        h.nparts = rs.num_hub_parts;
        h.nhubs = rs.num_hubs;
        h.counts = rs.hub_counts;
        const bool hdev = rs.hub_counts != nullptr;
        auto hgrid = [&](int64_t items, int per_block, int max_per_cu) {
            return hdev ? ctx->cu_count * max_per_cu : grid_for(ctx, items, per_block, max_per_cu);
        };
        // the LDS table holds HUB_TB slots: never index past it
        h.table_slots = std::min(HUB_TB, std::max(64, ctx->tune.hub_table_slots & ~63));
        LPArgs ah = args_for(BIN_HUB);
        kt_begin(ctx, KS_LP_HUB_INSERT);
        if (hub_aux)
            hipLaunchKernelGGL((k_hub_partition<KM, true>), dim3(hgrid(h.nchunks, 1, 4)), dim3(HUB1_NT), 0,
                               hs, ah, h);
        else
            hipLaunchKernelGGL((k_hub_partition<KM, false>), dim3(hgrid(h.nchunks, 1, 4)), dim3(HUB1_NT), 0,
                               hs, ah, h);
        kt_end(ctx, KS_LP_HUB_INSERT);
        kt_begin(ctx, KS_LP_HUB_DECIDE);
        // 512 threads with 4096 slots (three workgroups per CU) or 256
        // threads with 2048 slots (seven per CU; five with aux records)
        if (hub_aux) {
            if (ctx->tune.hub_nt == 256)
                hipLaunchKernelGGL((k_hub_count_listed<KM, 256, 2048, true>), dim3(hgrid(h.nparts, 1, 5)),
                                   dim3(256), 0, hs, ah, h);
            else
                hipLaunchKernelGGL((k_hub_count_listed<KM, 512, HUB_TB, true>), dim3(hgrid(h.nparts, 1, 2)),
                                   dim3(512), 0, hs, ah, h);
            kt_end(ctx, KS_LP_HUB_DECIDE);
            kt_begin(ctx, KS_LP_HUB_FINAL);
            hipLaunchKernelGGL((k_hub_final<KM, true>), dim3(hgrid(h.nhubs, 4, 4)), dim3(256), 0, hs,
                               ah, h);
        } else {
            if (ctx->tune.hub_nt == 256)
                hipLaunchKernelGGL((k_hub_count_listed<KM, 256, 2048, false>), dim3(hgrid(h.nparts, 1, 7)),
                                   dim3(256), 0, hs, ah, h);
            else
                hipLaunchKernelGGL((k_hub_count_listed<KM, 512, HUB_TB, false>), dim3(hgrid(h.nparts, 1, 3)),
                                   dim3(512), 0, hs, ah, h);
            kt_end(ctx, KS_LP_HUB_DECIDE);
            kt_begin(ctx, KS_LP_HUB_FINAL);
            hipLaunchKernelGGL((k_hub_final<KM, false>), dim3(hgrid(h.nhubs, 4, 4)), dim3(256), 0,
                               hs, ah, h);
        }
        hipLaunchKernelGGL(k_hub_global_insert<KM>, dim3(hgrid(h.nchunks, 1, 4)), dim3(512), 0, hs,
                           ah, h, rs.hub_table_off, rs.hub_table_log, rs.hub_keys, rs.hub_cnts);
        hipLaunchKernelGGL(k_hub_global_decide<KM>, dim3(hgrid(h.nhubs, 1, 1)), dim3(1024), 0, hs, ah,
                           h, rs.hub_table_off, rs.hub_table_log, rs.hub_keys, rs.hub_cnts);
        kt_end(ctx, KS_LP_HUB_FINAL);
    }
    if (range(BIN_B4096, p0, p1)) {
        // degrees 2049-4096: 512-thread workgroups with 8192-slot tables (two
        // per CU; a 12-byte-slot table of 8192 entries would leave one);
        // 1025-2048: 256-thread workgroups with 4096-slot tables (four per CU)
        const int ps = ctx->tune.b4096_split && !MERGE && !dev ? std::min(std::max(b.split2048, p0), p1) : p1;
        kt_begin(ctx, KS_LP_B4096);
        // 16-bit counts (rows of at most 4096 entries): 56 instead of 72 KB of
        // LDS, room on a CU for workgroups of the bins running beside it
        if ((ps > p0 || dev) && ctx->tune.b4096_u16)
            hipLaunchKernelGGL((k_lp_block<512, 8192, KM, Src, false, 4, uint16_t>), dim3(grid(ps - p0, 1, 2)),
                               dim3(512), 0, bs, src(BIN_B4096, p0, ps, 1));
        else if (ps > p0 || dev)
            hipLaunchKernelGGL((k_lp_block<512, 8192, KM, Src>), dim3(grid(ps - p0, 1, 2)), dim3(512), 0,
                               bs, src(BIN_B4096, p0, ps, 1));
        if constexpr (!MERGE) {
            if (p1 > ps)
                hipLaunchKernelGGL((k_lp_block<256, 4096, KM, RangeSrc>), dim3(grid_for(ctx, p1 - ps, 1, 4)),
                                   dim3(256), 0, bs, RangeSrc{ps, p1, 1, args_for(BIN_B4096)});
        }
        kt_end(ctx, KS_LP_B4096);
    }
    // degrees 513-1024: one workgroup per vertex (2048-slot table)
    if (range(BIN_B1024, p0, p1)) {
        kt_begin(ctx, KS_LP_B1024);
        if (block_aux)
            hipLaunchKernelGGL((k_lp_block<256, 2048, KM, Src, true>), dim3(grid(p1 - p0, 1, 8)), dim3(256), 0,
                               ctx->stream, src(BIN_B1024, p0, p1, 1));
        else
            hipLaunchKernelGGL((k_lp_block<256, 2048, KM, Src>), dim3(grid(p1 - p0, 1, 8)), dim3(256), 0,
                               ctx->stream, src(BIN_B1024, p0, p1, 1));
        kt_end(ctx, KS_LP_B1024);
    }
    // degrees 257-512: one wave per vertex with a 1024-slot table, so a CU
    // holds 16 vertices in flight instead of 8 (or, tuned off, workgroups as
    // for 513-1024)
    if (range(BIN_B512, p0, p1)) {
        kt_begin(ctx, KS_LP_B512);
        if (!ctx->tune.b1024_wave)
            hipLaunchKernelGGL((k_lp_block<256, 2048, KM, Src>), dim3(grid(p1 - p0, 1, 8)), dim3(256), 0,
                               ctx->stream, src(BIN_B512, p0, p1, 1));
        else if (block_aux)
            hipLaunchKernelGGL((k_lp_wave<1024, KM, Src, true>), dim3(grid(p1 - p0, 4, 3)), dim3(256), 0,
                               ctx->stream, src(BIN_B512, p0, p1, 1));
        else if (ctx->tune.b1024_u16)  // 16-bit counts (degree <= 512): 28.7 KB, five workgroups per CU
            hipLaunchKernelGGL((k_lp_wave<1024, KM, Src, false, 8, uint16_t>), dim3(grid(p1 - p0, 4, 5)), dim3(256),
                               0, ctx->stream, src(BIN_B512, p0, p1, 1));
        else  // 4-8 entries per lane: all eight gathers in flight at once
            hipLaunchKernelGGL((k_lp_wave<1024, KM, Src, false, 8>), dim3(grid(p1 - p0, 4, 4)), dim3(256), 0,
                               ctx->stream, src(BIN_B512, p0, p1, 1));
        kt_end(ctx, KS_LP_B512);
    }
    // W256 and G64..G4 in one launch unless the wave / workgroup bins gather
    // records (small_aux: the launch's G bins gather records)
    const bool fused = !block_aux && ctx->tune.small_fused;
    if (!fused && range(BIN_W256, p0, p1)) {
        kt_begin(ctx, KS_LP_W256);
        if (block_aux)
            hipLaunchKernelGGL((k_lp_wave<512, KM, Src, true>), dim3(grid(p1 - p0, 4, 8)), dim3(256), 0,
                               ctx->stream, src(BIN_W256, p0, p1, 1));
        else
            hipLaunchKernelGGL((k_lp_wave<512, KM, Src>), dim3(grid(p1 - p0, 4, 8)), dim3(256), 0, ctx->stream,
                               src(BIN_W256, p0, p1, 1));
        kt_end(ctx, KS_LP_W256);
    }
#define PCD_SMALL(G, CP, BIN, SLOT)                                                                            \
    if (range(BIN, p0, p1)) {                                                                                  \
        kt_begin(ctx, SLOT);                                                                                         \
        if (small_aux)                                                                                         \
            hipLaunchKernelGGL((k_lp_small<G, CP, true, KM, Src>), dim3(grid(p1 - p0, 4 * CP, 16)),                \
                               dim3(256), 0, ctx->stream, src(BIN, p0, p1, CP));                                \
        else                                                                                                   \
            hipLaunchKernelGGL((k_lp_small<G, CP, false, KM, Src>), dim3(grid(p1 - p0, 4 * CP, 16)),               \
                               dim3(256), 0, ctx->stream, src(BIN, p0, p1, CP));                                \
        kt_end(ctx, SLOT);                                                                                     \
    }
    if (fused) {
        SmallBinsT<Src> sb;
        // bins W256, G64..G4 (consecutive bin numbers) and positions per unit
        const int cps[6] = {1, 8, 16, 32, 64, 64};
        static_assert(BIN_G64 == BIN_W256 + 1 && BIN_G4 == BIN_W256 + 5 && BIN_W256 == BIN_B512 + 1, "bin order");
        int64_t tot = 0;
        int nonempty = 0;
        bool has[6];
        for (int i = 0; i < 6; i++) {
            range(BIN_W256 + i, p0, p1);
            sb.src[i] = src(BIN_W256 + i, p0, p1, cps[i]);
            tot += rs.bin_entries[BIN_W256 + i];
            has[i] = p1 > p0;
            nonempty += has[i];
        }
        sb.dlo = MERGE ? rs.layout + NUM_BINS + 1 : nullptr;
        if (nonempty > 0 || dev) {
            // one workgroup per non-empty bin, the rest by entries: up to
            // 16 per CU, and no more than one per 2048 entries, so that a
            // small graph's launch is not mostly empty workgroups
            const int spare = (int)std::min<int64_t>(ctx->cu_count * 16 - nonempty, tot / 2048);
            int at = 0;
            int64_t acc = 0;
            for (int i = 0; i < 6; i++) {
                sb.lo[i] = at;
                if (has[i]) {
                    const int64_t e = rs.bin_entries[BIN_W256 + i];
                    const int before = tot > 0 ? (int)(spare * acc / tot) : 0;
                    acc += e;
                    const int after = tot > 0 ? (int)(spare * acc / tot) : 0;
                    at += 1 + (after - before);
                }
            }
            sb.lo[6] = at;
            if (dev) at = ctx->cu_count * 16;  // the split the device computed covers this grid
            kt_begin(ctx, KS_LP_SMALL);
            // modularity and CPM builds (propagation or merging) at 8 waves per SIMD
            constexpr bool ONE_KIND = (KM & RK_ALL) == RK_MOD || (KM & RK_ALL) == RK_CPM;
            constexpr int W8 = ONE_KIND ? 8 : 1;
            bool launched = false;
            if constexpr (KM == RK_MOD || KM == RK_CPM) {
                if (small_aux) {  // records of the modularity / CPM totals
                    if (ctx->tune.small_occ8)
                        hipLaunchKernelGGL((k_lp_small_all<KM, W8, true, Src>), dim3(at), dim3(256), 0, ctx->stream,
                                           sb);
                    else
                        hipLaunchKernelGGL((k_lp_small_all<KM, 1, true, Src>), dim3(at), dim3(256), 0, ctx->stream,
                                           sb);
                    launched = true;
                }
            }
            if (!launched) {
                if (ctx->tune.small_occ8 && ONE_KIND)
                    hipLaunchKernelGGL((k_lp_small_all<KM, W8, false, Src>), dim3(at), dim3(256), 0, ctx->stream, sb);
                else
                    hipLaunchKernelGGL((k_lp_small_all<KM, 1, false, Src>), dim3(at), dim3(256), 0, ctx->stream, sb);
            }
            kt_end(ctx, KS_LP_SMALL);
        }
    } else {
        PCD_SMALL(64, 8, BIN_G64, KS_LP_G64)
        PCD_SMALL(32, 16, BIN_G32, KS_LP_G32)
        PCD_SMALL(16, 32, BIN_G16, KS_LP_G16)
        PCD_SMALL(8, 64, BIN_G8, KS_LP_G8)
        PCD_SMALL(4, 64, BIN_G4, KS_LP_G4)
    }
#undef PCD_SMALL
    if (fork) {
        for (int i = 0; i < nside; i++) {
            PCD_HIP(hipEventRecord(ctx->ev_join[i], ctx->side[i]));
            PCD_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join[i], 0));
        }
    }
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

int propagation_dev(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie) {
    PCD_TRY(propagation_launch(ctx, rule, frac, seed, step, tie));
    return propagation_commit(ctx);
}

int propagation_launch(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie) {
    double fr = frac < 1.0 ? frac : 1.0;
    LPArgs a;
    a.indices = ctx->d_indices;
    a.adj = ctx->d_adj;
    a.labels = ctx->d_labels;
    a.order = ctx->d_order;
    a.lv = ctx->d_lv;
    a.stats = ctx->d_stats;
    a.new_labels = ctx->d_labels_alt;
    a.rule = make_rule(ctx, rule);
    a.seed = seed;
    a.step = step;
    a.tie = tie;
    a.f = (int)(fr * 2147483647.0);  // refinement.cpp:116
    a.use_aux = rule.kind == PCD_RULE_MODULARITY || rule.kind == PCD_RULE_CPM;
    a.tot = a.use_aux ? ctx->d_tot : nullptr;
    a.counters = nullptr;
    a.step_ptr = ctx->graph_step;  // set while a pass is captured for replay
    a.found = nullptr;
    // kernels specialised for modularity or CPM hold fewer registers (more
    // waves per SIMD); WCC and conductance use the general build
    const RowSet rs = graph_rows(ctx);
    if (rule.kind == PCD_RULE_MODULARITY) PCD_TRY(launch_propagation<RK_MOD>(ctx, rs, a));
    else if (rule.kind == PCD_RULE_CPM) PCD_TRY(launch_propagation<RK_CPM>(ctx, rs, a));
    else PCD_TRY(launch_propagation<RK_ALL>(ctx, rs, a));
    return PCD_OK;
}

// builds per rule kind, as for propagation (modularity and CPM hold fewer
// registers without the WCC and conductance formulas)
int merge_rows_launch(pcd_ctx *ctx, const RowSet &rows, const LPArgs &a) {
    if (a.rule.kind == PCD_RULE_MODULARITY) return launch_propagation<RK_MERGE_REC | RK_MOD>(ctx, rows, a);
    if (a.rule.kind == PCD_RULE_CPM) return launch_propagation<RK_MERGE_REC | RK_CPM>(ctx, rows, a);
    return launch_propagation<RK_MERGE_REC | RK_ALL>(ctx, rows, a);
}

int propagation_commit(pcd_ctx *ctx) {
    if (ctx->k_deferred && !ctx->capturing) {  // label_bound below needs k on the host
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        resolve_k(ctx);
    }
    // the statistics follow the moved vertices; the next compression
    // renumbers them
    const bool follow = ctx->tune.inc_stats != 0;
    if (follow) PCD_TRY(statistics_delta_dev(ctx, ctx->d_labels, ctx->d_labels_alt));
    std::swap(ctx->d_labels, ctx->d_labels_alt);
    ctx->labels_compressed = false;
    ctx->has_stats = false;
    ctx->stats_pending = follow;
    // a captured pass keeps k on the device: labels stay below n
    ctx->label_bound = ctx->capturing ? ctx->n : ctx->k;
    return PCD_OK;
}

// ------------------------------------------------------------- batches
// One persistent launch per variant in use: a fixed grid whose waves or
// workgroups stride over the variant's work items (the item count is read
// on the device, so the launches can be captured once and replayed while
// the lists change).  Small bins read labels and statistics (no records).
int batch_propagation_launch(hipStream_t stream, const BatchLPLaunch &l) {
    const int cu = l.cu_count;
    auto src = [&](int v) { return ItemSrc{l.items[v], l.count[v], l.graphs, l.state}; };
    if (l.used[LPV_B512X8192])
        hipLaunchKernelGGL((k_lp_block<512, 8192, RK_BATCH, ItemSrc>), dim3(cu * 2), dim3(512), 0, stream, src(LPV_B512X8192));
    if (l.used[LPV_B512])
        hipLaunchKernelGGL((k_lp_block<512, 2048, RK_BATCH, ItemSrc>), dim3(cu * 4), dim3(512), 0, stream, src(LPV_B512));
    if (l.used[LPV_B256])
        hipLaunchKernelGGL((k_lp_block<256, 2048, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_B256));
    // the plain path's form of the degree 257-512 wave kernel: all of a
    // lane's entries gathered at once, 16-bit counts (batch degrees are at
    // most 4096)
    if (l.used[LPV_W1024])
        hipLaunchKernelGGL((k_lp_wave<1024, RK_BATCH, ItemSrc, false, 8, uint16_t>), dim3(cu * 8), dim3(256), 0, stream,
                           src(LPV_W1024));
    if (l.used[LPV_W512])
        hipLaunchKernelGGL((k_lp_wave<512, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_W512));
    if (l.used[LPV_S64])
        hipLaunchKernelGGL((k_lp_small<64, 8, false, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_S64));
    if (l.used[LPV_S32])
        hipLaunchKernelGGL((k_lp_small<32, 16, false, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_S32));
    if (l.used[LPV_S16])
        hipLaunchKernelGGL((k_lp_small<16, 32, false, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_S16));
    if (l.used[LPV_S8])
        hipLaunchKernelGGL((k_lp_small<8, 64, false, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_S8));
    if (l.used[LPV_S4])
        hipLaunchKernelGGL((k_lp_small<4, 64, false, RK_BATCH, ItemSrc>), dim3(cu * 8), dim3(256), 0, stream, src(LPV_S4));
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

}  // namespace pcd
