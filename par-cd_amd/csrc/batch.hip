// Batched refinement: many independent graphs advance together, one
// refinement iteration of main.cpp:226-266 per round for every running
// graph, with the round's launches captured once in a hipGraph and
// replayed.
//
// Layout.  Each graph owns a region of one device arena: its CSR, an
// optional degree order, two label buffers, the compression scratch
// (first occurrence, new ids) and its community records.  The kernels see
// the batch through work-item lists:
//   vertex tiles     2048 consecutive vertex ids of one graph (compression,
//                    statistics clearing, metric sums)
//   small rows       up to CP positions of one graph's degree class <= G
//                    (statistics with G lanes per vertex)
//   long rows        up to ROWS_PER positions of one graph's rows of
//                    degree > 64 (statistics)
//   propagation      per kernel variant, position ranges of one graph's
//                    degree range (the bin map chooses which variant serves
//                    which degree range)
// Every kernel runs a fixed grid that strides over its list; the lists'
// lengths are read on the device, so a captured round stays valid while
// graphs are admitted and retired.  Per-graph state (iteration, score,
// community count, status) stays on the device; the stop rule of
// main.cpp:241-242 is evaluated there, so a round needs no host read-back.
//
// Identity with the plain path: a graph's labels, statistics and scores
// in a batch equal those of the same graph run alone (same seeds, same
// per-graph vertex and label numbering); the kernels only change who
// computes what.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "hostutil.hpp"
#include "work.h"

namespace pcd {

constexpr int VT = 2048;  // vertices per tile item
constexpr int ROWS_PER = 64;  // row positions per long-row item (one per lane of the update kernel)
static_assert(ROWS_PER == WAVE, "k_b_delta_rows checks one row per lane");
constexpr int STAT_G[5] = {64, 32, 16, 8, 4};
constexpr int STAT_CP[5] = {8, 16, 32, 64, 64};

// --------------------------------------------------------------- kernels
__global__ __launch_bounds__(256) void k_b_fill_first(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                      const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const int hi = min(w.p1, st[w.g].label_bound);
        int32_t *first = gs[w.g].first;
        for (int v = w.p0 + (int)threadIdx.x; v < hi; v += blockDim.x) first[v] = INT32_MAX;
    }
}

__global__ __launch_bounds__(256) void k_b_first_occurrence(const WorkItem *tiles, const int *count,
                                                            const BatchGraph *gs, const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const int32_t *next = gs[w.g].next;
        int32_t *first = gs[w.g].first;
        for (int v = w.p0 + (int)threadIdx.x; v < w.p1; v += blockDim.x) {
            int l = next[v];
            if (l >= 0 && first[l] > v) atomicMin(&first[l], v);  // read first: see k_first_occurrence
        }
    }
}

// framework.cpp:146-170: a vertex opens a new id if its label is NEW or it
// is its label's first occurrence
__device__ inline int first_flag(const BatchGraph &g, int v) {
    int l = g.next[v];
    return (l < 0 || g.first[l] == v) ? 1 : 0;
}

constexpr int VT_PER = VT / 256;

__global__ __launch_bounds__(256) void k_b_tile_count(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                      const BatchState *st, int32_t *tile_sum) {
    __shared__ int lds[256 / WAVE];
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        int c = 0;
        if (st[w.g].status == BG_RUNNING) {
            const BatchGraph &g = gs[w.g];
            for (int v = w.p0 + (int)threadIdx.x; v < w.p1; v += blockDim.x) c += first_flag(g, v);
        }
        c = wave_sum(c);
        if (__lane_id() == 0) lds[threadIdx.x / WAVE] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int i = 0; i < 256 / WAVE; i++) t += lds[i];
            tile_sum[u] = t;
        }
        __syncthreads();
    }
}

// one workgroup: exclusive scan of the tile sums (one graph's tiles are
// consecutive, so a graph's local offsets are differences of this scan)
__global__ __launch_bounds__(1024) void k_b_tile_scan(const int *count, const int32_t *tile_sum, int32_t *tile_pre) {
    __shared__ int lds[1024 / WAVE];
    const int nt = *count;
    int run = 0;
    for (int base = 0; base < nt; base += 1024) {
        int i = base + threadIdx.x;
        int x = i < nt ? tile_sum[i] : 0;
        int tot;
        int ex = block_excl_scan_n<1024>(x, &tot, lds);
        if (i < nt) tile_pre[i] = run + ex;
        run += tot;
    }
}

__global__ __launch_bounds__(256) void k_b_tile_apply(const WorkItem *tiles, const int *count, const int2 *gtiles,
                                                      const BatchGraph *gs, BatchState *st, const int32_t *tile_sum,
                                                      const int32_t *tile_pre) {
    __shared__ int lds[256 / WAVE];
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const BatchGraph &g = gs[w.g];
        const int2 gt = gtiles[w.g];
        const int gbase = tile_pre[u] - tile_pre[gt.x];
        // each thread owns VT_PER consecutive vertices
        const int v0 = w.p0 + (int)threadIdx.x * VT_PER;
        int flags = 0, c = 0;
#pragma unroll
        for (int e = 0; e < VT_PER; e++) {
            int v = v0 + e;
            int f = (v < w.p1) ? first_flag(g, v) : 0;
            flags |= f << e;
            c += f;
        }
        int tot;
        int ex = block_excl_scan_n<256>(c, &tot, lds);
        int id = gbase + ex;
        // graphs whose statistics follow the labels: the records move to
        // the new ids, a NEW_LABEL vertex opens its singleton record
        const bool ready = st[w.g].ready != 0;
#pragma unroll
        for (int e = 0; e < VT_PER; e++) {
            if ((flags >> e) & 1) {
                int v = v0 + e;
                int l = g.next[v];
                if (l < 0) g.cur[v] = id;
                else g.newid[l] = id;
                if (ready)
                    g.stats_alt[id] = l < 0 ? CommStats{1, 0, g.lp.indices[v + 1] - g.lp.indices[v], 0} : g.stats[l];
                id++;
            }
        }
        if (threadIdx.x == 0 && u == gt.x + gt.y - 1) st[w.g].k = gbase + tile_sum[u];
    }
}

__global__ __launch_bounds__(256) void k_b_relabel(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                   const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const BatchGraph &g = gs[w.g];
        for (int v = w.p0 + (int)threadIdx.x; v < w.p1; v += blockDim.x) {
            int l = g.next[v];
            if (l >= 0) g.cur[v] = g.newid[l];
        }
    }
}

// renumbered statistics back into place (graphs whose statistics follow
// the labels)
__global__ __launch_bounds__(256) void k_b_stats_copy(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                      const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING || !st[w.g].ready) continue;
        const int hi = min(w.p1, st[w.g].k);
        const BatchGraph &g = gs[w.g];
        for (int c = w.p0 + (int)threadIdx.x; c < hi; c += blockDim.x) g.stats[c] = g.stats_alt[c];
    }
}

// The full count runs in a graph's first round only
__global__ __launch_bounds__(256) void k_b_clear_stats(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                       const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING || st[w.g].ready) continue;
        const int hi = min(w.p1, st[w.g].k);
        CommStats *s = gs[w.g].stats;
        for (int c = w.p0 + (int)threadIdx.x; c < hi; c += blockDim.x) s[c] = CommStats{0, 0, 0, 0};
    }
}

// framework.cpp:69-143 over rows of degree <= G: G lanes per vertex
template <int G>
__global__ __launch_bounds__(256) void k_b_stats_small(const WorkItem *items, const int *count, const BatchGraph *gs,
                                                       const BatchState *st) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    constexpr int per_wave = WAVE / G;
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int nu = *count;
    for (int u = wave_g; u < nu; u += nwaves) {
        const WorkItem w = items[u];
        if (st[w.g].status != BG_RUNNING || st[w.g].ready) continue;
        const BatchGraph &g = gs[w.g];
        const LPArgs &a = g.lp;
        for (int base = w.p0; base < w.p1; base += per_wave) {
            int p = base + lane / G;
            bool has = p < w.p1;
            int v = 0, start = 0, deg = 0, l = 0;
            if (has) {
                v = a.order ? a.order[p] : p;
                start = a.indices[v];
                deg = a.indices[v + 1] - start;
                l = g.cur[v];
            }
            int same = 0;
            for (int j = glane; j < deg; j += G) same += (g.cur[a.adj[start + j]] == l) ? 1 : 0;
#pragma unroll
            for (int d = G / 2; d > 0; d >>= 1) same += __shfl_xor(same, d, WAVE);
            if (has && glane == 0) add_stats(g.stats, l, 1, same, deg - same);
        }
    }
}

// rows of degree > 64 (at most 4096 entries in a batch): one wave per item
// of up to ROWS_PER row positions, row after row
__global__ __launch_bounds__(256) void k_b_stats_rows(const WorkItem *items, const int *count, const BatchGraph *gs,
                                                      const BatchState *st) {
    constexpr int PF = 8;
    const int lane = __lane_id();
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int nu = *count;
    for (int u = wave_g; u < nu; u += nwaves) {
        const WorkItem w = items[u];
        if (st[w.g].status != BG_RUNNING || st[w.g].ready) continue;
        const BatchGraph &g = gs[w.g];
        const LPArgs &a = g.lp;
        for (int p = w.p0; p < w.p1; p++) {
            const int v = a.order ? a.order[p] : p;
            const int l = g.cur[v];
            const int b = a.indices[v], e = a.indices[v + 1];
            int same = 0;
            for (int j0 = b + lane; j0 < e; j0 += PF * WAVE) {
                int nb[PF];
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    const int j = j0 + k * WAVE;
                    nb[k] = j < e ? a.adj[j] : -1;
                }
#pragma unroll
                for (int k = 0; k < PF; k++) same += (nb[k] >= 0 && g.cur[nb[k]] == l) ? 1 : 0;
            }
            same = wave_sum(same);
            if (lane == 0) add_stats(g.stats, l, 1, same, (e - b) - same);
        }
    }
}

// Statistics after a propagation step, from the moved vertices only (the
// update of partition.hip's k_delta_small / k_delta_items): cur holds the
// labels the step read (L), next those it wrote (N), and the records of
// graphs whose statistics follow the labels are brought from L to N under
// L's ids; the compression that follows renumbers them.

// rows of degree <= G: G lanes per vertex; unmoved vertices read no row
template <int G>
__global__ __launch_bounds__(256) void k_b_delta_small(const WorkItem *items, const int *count, const BatchGraph *gs,
                                                       const BatchState *st) {
    const int lane = __lane_id(), glane = lane & (G - 1);
    constexpr int per_wave = WAVE / G;
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int nu = *count;
    for (int u = wave_g; u < nu; u += nwaves) {
        const WorkItem w = items[u];
        if (st[w.g].status != BG_RUNNING || !st[w.g].ready) continue;
        const BatchGraph &g = gs[w.g];
        const LPArgs &a = g.lp;
        for (int base = w.p0; base < w.p1; base += per_wave) {
            const int p = base + lane / G;
            int start = 0, deg = 0, lo = 0, ln = 0;
            bool moved = false;
            if (p < w.p1) {
                const int v = a.order ? a.order[p] : p;
                lo = g.cur[v];
                ln = g.next[v];
                moved = lo != ln;
                if (moved) {
                    start = a.indices[v];
                    deg = a.indices[v + 1] - start;
                }
            }
            if (!__ballot(moved)) continue;
            int sa = 0, sb = 0;
            for (int j = glane; j < deg; j += G) {
                const int x = a.adj[start + j];
                const int lx = g.cur[x], nx = g.next[x];
                const int f = lx != nx ? 1 : 2;
                sa += lx == lo ? f : 0;
                sb += nx == ln ? f : 0;
            }
#pragma unroll
            for (int d = G / 2; d > 0; d >>= 1) {
                sa += __shfl_xor(sa, d, WAVE);
                sb += __shfl_xor(sb, d, WAVE);
            }
            if (moved && glane == 0) {
                add_stats(g.stats, lo, -1, -sa, sa - deg);
                if (ln >= 0) add_stats(g.stats, ln, 1, sb, deg - sb);
            }
        }
    }
}

// rows of degree > 64: one wave per item of up to ROWS_PER (= 64) row
// positions; lane j checks row p0 + j, then the wave runs the moved rows
__global__ __launch_bounds__(256) void k_b_delta_rows(const WorkItem *items, const int *count, const BatchGraph *gs,
                                                      const BatchState *st) {
    constexpr int PF = 4;
    const int lane = __lane_id();
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    const int nu = *count;
    for (int u = wave_g; u < nu; u += nwaves) {
        const WorkItem w = items[u];
        if (st[w.g].status != BG_RUNNING || !st[w.g].ready) continue;
        const BatchGraph &g = gs[w.g];
        const LPArgs &a = g.lp;
        int v = 0, lo = 0, ln = 0;
        bool moved = false;
        if (w.p0 + lane < w.p1) {
            const int p = w.p0 + lane;
            v = a.order ? a.order[p] : p;
            lo = g.cur[v];
            ln = g.next[v];
            moved = lo != ln;
        }
        uint64_t mv = __ballot(moved);
        while (mv) {
            const int src = __builtin_ctzll(mv);
            mv &= mv - 1;
            const int xv = __builtin_amdgcn_readlane(v, src);
            const int xlo = __builtin_amdgcn_readlane(lo, src), xln = __builtin_amdgcn_readlane(ln, src);
            const int xb = a.indices[xv], xe = a.indices[xv + 1];
            int sa = 0, sb = 0;
            for (int j0 = xb + lane; j0 < xe; j0 += PF * WAVE) {
                int nb[PF];
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    const int j = j0 + k * WAVE;
                    nb[k] = j < xe ? a.adj[j] : -1;
                }
                int lx[PF], nx[PF];
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    lx[k] = nb[k] < 0 ? -2 : g.cur[nb[k]];
                    nx[k] = nb[k] < 0 ? -2 : g.next[nb[k]];
                }
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    if (nb[k] < 0) continue;
                    const int f = lx[k] != nx[k] ? 1 : 2;
                    sa += lx[k] == xlo ? f : 0;
                    sb += nx[k] == xln ? f : 0;
                }
            }
            sa = wave_sum(sa);
            sb = wave_sum(sb);
            if (lane == 0) {
                const int len = xe - xb;
                add_stats(g.stats, xlo, -1, -sa, sa - len);
                if (xln >= 0) add_stats(g.stats, xln, 1, sb, len - sb);
            }
        }
    }
}

// exact int64 sums of int_volume, volume^2 and size*(size-1) per graph
__global__ __launch_bounds__(256) void k_b_metric(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                  BatchState *st) {
    __shared__ long long lds[3][256 / WAVE];
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const int hi = min(w.p1, st[w.g].k);
        if (w.p0 >= hi) continue;
        const CommStats *s = gs[w.g].stats;
        long long a = 0, b = 0, c = 0;
        for (int i = w.p0 + (int)threadIdx.x; i < hi; i += blockDim.x) {
            CommStats x = s[i];
            long long vol = (long long)x.int_vol + x.ext_vol;
            a += x.int_vol;
            b += vol * vol;
            c += (long long)x.size * (x.size - 1);
        }
        a = wave_sum64(a);
        b = wave_sum64(b);
        c = wave_sum64(c);
        const int wv = threadIdx.x / WAVE;
        if (__lane_id() == 0) { lds[0][wv] = a; lds[1][wv] = b; lds[2][wv] = c; }
        __syncthreads();
        if (threadIdx.x < 3) {
            long long t = 0;
            for (int i = 0; i < 256 / WAVE; i++) t += lds[threadIdx.x][i];
            atomicAdd(&st[w.g].acc[threadIdx.x], (unsigned long long)t);
        }
        __syncthreads();
    }
}

// next = cur for the graphs still running: propagation then writes the
// active vertices only
__global__ __launch_bounds__(256) void k_b_copy_next(const WorkItem *tiles, const int *count, const BatchGraph *gs,
                                                     const BatchState *st) {
    const int nt = *count;
    for (int u = blockIdx.x; u < nt; u += gridDim.x) {
        const WorkItem w = tiles[u];
        if (st[w.g].status != BG_RUNNING) continue;
        const BatchGraph &g = gs[w.g];
        for (int v = w.p0 + (int)threadIdx.x; v < w.p1; v += blockDim.x) g.next[v] = g.cur[v];
    }
}

// Score (rules.cpp:9-22, 121-129), stop rule (main.cpp:241-242) and the
// iteration bookkeeping of main.cpp:262-264, one thread per graph.
__global__ void k_b_finalize(const BatchGraph *gs, BatchState *st, int nslots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    BatchState s = st[i];
    if (s.status != BG_RUNNING) return;
    const BatchGraph &g = gs[i];
    const long long a = (long long)s.acc[0], b = (long long)s.acc[1], c = (long long)s.acc[2];
    const int m = g.lp.rule.m;
    double score;
    if (g.lp.rule.kind == PCD_RULE_MODULARITY) {
        double two_m = 2.0 * m;
        score = a / two_m - (double)b / (two_m * two_m);
    } else {
        score = (a - g.lp.rule.lambda * (double)c) / (double)(2 * m);
    }
    const int it = s.iter;
    bool done = g.max_iter != -1 ? it >= g.max_iter
                                 : (s.prev_score > 0 && score > 0 && score / s.prev_score - 1 <= g.threshold);
    if (!done && it >= g.iter_cap) {
        done = true;
        s.capped = 1;
    }
    s.acc[0] = s.acc[1] = s.acc[2] = 0;
    s.score = score;
    s.prev_score = score;
    s.step = it;
    s.iter = it + 1;
    s.label_bound = s.k;
    s.ready = 1;  // counted in full this round; the propagation steps that follow update the counts
    if (done) s.status = BG_DONE;
    st[i] = s;
}


// Copies of page-locked host arrays into the arena, read by the device
// directly (one launch for many graphs: a stream of small copies reaches
// about 45 GB/s over PCIe, device reads of mapped host memory about 56).
struct PullPiece {
    const int32_t *src;
    int32_t *dst;
    int64_t words;
};

__global__ __launch_bounds__(256) void k_b_pull(const PullPiece *pieces, int count) {
    for (int i = blockIdx.x; i < count; i += gridDim.x) {
        const PullPiece pc = pieces[i];
        if ((((uintptr_t)pc.src | (uintptr_t)pc.dst) & 15) == 0 && (pc.words & 3) == 0) {
            const int4 *s4 = (const int4 *)pc.src;
            int4 *d4 = (int4 *)pc.dst;
            for (int64_t w = threadIdx.x; w < pc.words / 4; w += blockDim.x) d4[w] = s4[w];
        } else {
            // unaligned pieces: four loads in flight per lane before the stores
            const int32_t *__restrict__ src = pc.src;
            int32_t *__restrict__ dst = pc.dst;
            const int64_t st = blockDim.x;
            int64_t w = threadIdx.x;
            for (; w + 3 * st < pc.words; w += 4 * st) {
                const int32_t a0 = src[w], a1 = src[w + st], a2 = src[w + 2 * st], a3 = src[w + 3 * st];
                dst[w] = a0;
                dst[w + st] = a1;
                dst[w + 2 * st] = a2;
                dst[w + 3 * st] = a3;
            }
            for (; w < pc.words; w += st) dst[w] = src[w];
        }
    }
}

// Word copy between device memory and page-locked host memory by the
// device (the batch stream's record, state and item-list uploads and its
// state read-back): a copy-engine transfer would wait behind the copies an
// asynchronous admission queued.
__global__ __launch_bounds__(256) void k_b_copy(const int32_t *__restrict__ src, int32_t *__restrict__ dst,
                                                int64_t words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Format check of admitted graphs' rows in the arena (what prepare_graph
// checks on the host for a single graph): ascending, loop-free, ids in
// [0, n).  One wave per item of up to 256 rows; bad[slot] = 1 on failure.
__global__ __launch_bounds__(256) void k_b_check_rows(const WorkItem *items, int count, const BatchGraph *gs,
                                                      int32_t *bad) {
    const int lane = __lane_id();
    const int wave_g = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int nwaves = gridDim.x * blockDim.x / WAVE;
    for (int u = wave_g; u < count; u += nwaves) {
        const WorkItem w = items[u];
        const BatchGraph &g = gs[w.g];
        const int n = g.n;
        bool fail = false;
        for (int v = w.p0; v < w.p1; v++) {
            const int b = g.lp.indices[v], e = g.lp.indices[v + 1];
            for (int j = b + lane; j < e; j += WAVE) {
                const int x = g.lp.adj[j];
                fail |= x < 0 || x >= n || x == v || (j > b && g.lp.adj[j - 1] >= x);
            }
        }
        if (__ballot(fail) && lane == 0) bad[w.g] = 1;
    }
}

// Labels of several finished graphs, graph after graph: out[offs[i] + v] =
// cur[v] of the graph in slots[i].  One workgroup per graph at a time.
__global__ void __launch_bounds__(256) k_b_gather_labels(const BatchGraph *graphs, const int32_t *slots,
                                                         const int64_t *offs, int count, int32_t *out) {
    for (int i = blockIdx.x; i < count; i += gridDim.x) {
        const BatchGraph &g = graphs[slots[i]];
        const int32_t *src = g.cur;
        int32_t *dst = out + offs[i];
        for (int v = threadIdx.x; v < g.n; v += blockDim.x) dst[v] = src[v];
    }
}

}  // namespace pcd

// ===================================================================== host
using namespace pcd;

namespace {

// degree thresholds of the class boundaries (positions are degree-descending)
constexpr int NUM_TH = 11;
constexpr int TH[NUM_TH] = {4096, 1024, 512, 256, 128, 64, 32, 16, 8, 4, 0};
int th_index(int t) {
    for (int i = 0; i < NUM_TH; i++)
        if (TH[i] == t) return i;
    return -1;
}

// A bin map: variants with the largest degree each serves, descending.
struct BinMap {
    const char *name;
    int n;
    int var[10];
    int hi[10];
};
const BinMap MAPS[] = {
    {"s4-s64,w512<=256,b256<=1024,b8192", 8,
     {LPV_B512X8192, LPV_B256, LPV_W512, LPV_S64, LPV_S32, LPV_S16, LPV_S8, LPV_S4},
     {4096, 1024, 256, 64, 32, 16, 8, 4}},
    {"s4-s32,w512<=256,b256<=1024,b8192", 7,
     {LPV_B512X8192, LPV_B256, LPV_W512, LPV_S32, LPV_S16, LPV_S8, LPV_S4}, {4096, 1024, 256, 32, 16, 8, 4}},
    {"s4-s64,w1024<=512,b256<=1024,b8192", 8,
     {LPV_B512X8192, LPV_B256, LPV_W1024, LPV_S64, LPV_S32, LPV_S16, LPV_S8, LPV_S4},
     {4096, 1024, 512, 64, 32, 16, 8, 4}},
    {"s4-s64,w512<=256,b512<=1024,b8192", 8,
     {LPV_B512X8192, LPV_B512, LPV_W512, LPV_S64, LPV_S32, LPV_S16, LPV_S8, LPV_S4},
     {4096, 1024, 256, 64, 32, 16, 8, 4}},
    {"s4-s64,w512<=128,b256<=1024,b8192", 8,
     {LPV_B512X8192, LPV_B256, LPV_W512, LPV_S64, LPV_S32, LPV_S16, LPV_S8, LPV_S4},
     {4096, 1024, 128, 64, 32, 16, 8, 4}},
    {"s4-s64,w512<=256,w1024<=512,b256<=1024,b8192", 9,
     {LPV_B512X8192, LPV_B256, LPV_W1024, LPV_W512, LPV_S64, LPV_S32, LPV_S16, LPV_S8, LPV_S4},
     {4096, 1024, 512, 256, 64, 32, 16, 8, 4}},
};
constexpr int NUM_MAPS = sizeof(MAPS) / sizeof(MAPS[0]);

struct Slot {
    bool used = false;
    int status = BG_FREE;
    int n = 0, m = 0;
    size_t off = 0, bytes = 0;
    int64_t admitted = 0;      // rounds run before admission
    int pos_le[NUM_TH] = {0};  // first position with degree <= TH[i]
    BatchGraph rec;
};

// item list layout inside one device buffer (and its pinned mirror)
enum ListId { L_TILES = 0, L_SMALL0, L_ROWS = L_SMALL0 + 5, L_LP0, NUM_LISTS = L_LP0 + NUM_LPV };

}  // namespace

struct pcd_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    int cap = 0;  // graph slots
    char *arena = nullptr;
    size_t arena_bytes = 0;
    std::map<size_t, size_t> free_list;  // offset -> bytes
    std::vector<Slot> slots;
    BatchGraph *d_graphs = nullptr;
    BatchState *d_state = nullptr;
    BatchState *h_state = nullptr;  // pinned
    BatchGraph *h_graphs = nullptr; // pinned: graph records on their way to d_graphs
    int2 *d_gtiles = nullptr;
    int2 *h_gtiles = nullptr;       // pinned
    // item lists
    void *d_items = nullptr, *h_items = nullptr;
    size_t items_cap = 0;           // bytes of each
    int32_t *d_counts = nullptr, *h_counts = nullptr;
    size_t list_off[NUM_LISTS] = {0};
    int list_len[NUM_LISTS] = {0};
    int64_t list_cap[NUM_LISTS] = {0};  // items each list's region holds
    std::vector<int> pending_append;    // admitted since the lists were built
    std::vector<char> slot_listed;      // the slot has items in the lists
    int32_t *d_tile_sum = nullptr, *d_tile_pre = nullptr;
    size_t tile_cap = 0;
    char *d_gather = nullptr;  // label download staging (pcd_batch_labels_many)
    size_t gather_cap = 0;
    std::vector<PullPiece> pulls;  // admission copies the device reads from host memory
    PullPiece *d_pulls = nullptr;
    size_t pulls_cap = 0;
    WorkItem *d_check = nullptr;   // row-check items of pcd_batch_add_many
    size_t check_cap = 0;
    int32_t *d_bad = nullptr;      // per slot: rows failed the check
    // asynchronous admission (pcd_batch_add_many_async): uploads and row
    // checks on their own stream, one event per chunk of graphs; the
    // buffers belong to the last call and are reused once its chunks joined
    struct AChunk {
        int first = 0, count = 0;  // call-local graph indices
        hipEvent_t ev = nullptr;
    };
    hipStream_t astream = nullptr;
    std::vector<hipEvent_t> aevents;
    std::deque<AChunk> achunks;  // queued, not yet joined
    std::vector<int> aslots;     // call-local graph -> slot (-1: refused on the host)
    int apending = 0;            // graphs in BG_PENDING
    PullPiece *d_apulls = nullptr;
    size_t apulls_cap = 0;
    WorkItem *d_acheck = nullptr;
    size_t acheck_cap = 0;
    BatchGraph *d_arecs = nullptr;  // row-check records, call-local
    size_t arecs_cap = 0;
    char *d_adma = nullptr;  // device staging of one chunk's copy-engine transfers
    size_t adma_cap = 0;
    int32_t *d_abad = nullptr, *h_abad = nullptr;
    size_t abad_cap = 0;
    int32_t *h_astage = nullptr;  // page-locked staging of pageable labels and degree orders
    size_t astage_cap = 0;
    // buffers replaced while admission chunks were queued: hipFree and
    // hipHostFree wait for the whole device, the queued uploads included,
    // so they are freed once the queue has drained
    std::vector<void *> retired_dev, retired_host;
    bool items_dirty = true;
    int64_t listed_done_vertices = 0, listed_vertices = 0;
    // bin map and captured round
    int map = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t rounds = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double lp_ms_last = 0;
};

static int bfail(int code, const std::string &msg) { return fail(code, msg); }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t footprint(int n, int64_t E, bool sorted) {
    size_t nb = align256(4 * ((size_t)n + 1));
    return nb /*indices*/ + align256(4 * (size_t)std::max<int64_t>(E, 1)) /*adj*/ + (sorted ? 0 : nb) /*order*/ +
           4 * nb /*cur next first newid*/ + 2 * align256(16 * ((size_t)n + 1)) /*stats, stats_alt*/;
}

static bool arena_alloc(pcd_batch *b, size_t bytes, size_t *off) {
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            size_t rest = it->second - bytes;
            size_t roff = it->first + bytes;
            b->free_list.erase(it);
            if (rest) b->free_list[roff] = rest;
            return true;
        }
    }
    return false;
}

static void arena_free(pcd_batch *b, size_t off, size_t bytes) {
    auto it = b->free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != b->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_list.erase(nx);
    }
    if (it != b->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_list.erase(it);
        }
    }
}

static void drop_graph(pcd_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Frees a buffer the batch no longer uses, or holds it back while
// admission chunks are queued (see retired_dev).
static void free_device(pcd_batch *b, void *p) {
    if (!p) return;
    if (b->achunks.empty()) (void)hipFree(p);
    else b->retired_dev.push_back(p);
}
static void free_host(pcd_batch *b, void *p) {
    if (!p) return;
    if (b->achunks.empty()) (void)hipHostFree(p);
    else b->retired_host.push_back(p);
}
static void free_retired(pcd_batch *b) {
    if (!b->achunks.empty()) return;
    for (void *p : b->retired_dev) (void)hipFree(p);
    for (void *p : b->retired_host) (void)hipHostFree(p);
    b->retired_dev.clear();
    b->retired_host.clear();
}

// Queues k_b_copy on the batch stream: `bytes` (a multiple of 4) between
// page-locked host memory and the device, either way.  The host keeps a
// host source unchanged, and reads a host destination only, after the
// stream has passed the launch.
static int kcopy(pcd_batch *b, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return PCD_OK;
    if (bytes % 4) return bfail(PCD_E_STATE, "kcopy: size is not a multiple of 4");
    const int64_t words = (int64_t)(bytes / 4);
    const int blocks = (int)std::min<int64_t>(std::max<int64_t>(1, words / 2048), (int64_t)b->cu_count);
    hipLaunchKernelGGL(k_b_copy, dim3(blocks), dim3(256), 0, b->stream, (const int32_t *)src, (int32_t *)dst, words);
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

extern "C" int pcd_batch_map_count(void) { return NUM_MAPS; }
extern "C" const char *pcd_batch_map_name(int map) { return (map >= 0 && map < NUM_MAPS) ? MAPS[map].name : ""; }

extern "C" int pcd_batch_footprint(int n, int m, int64_t *bytes) {
    if (!bytes || n <= 0 || m < 0) return bfail(PCD_E_INVALID, "bad footprint arguments");
    *bytes = (int64_t)footprint(n, 2 * (int64_t)m, false);
    return PCD_OK;
}

extern "C" int pcd_batch_destroy(pcd_batch *b) {
    if (!b) return PCD_OK;
    (void)hipSetDevice(b->device);
    if (b->astream) (void)hipStreamSynchronize(b->astream);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->arena) (void)hipFree(b->arena);
    if (b->d_apulls) (void)hipFree(b->d_apulls);
    if (b->d_acheck) (void)hipFree(b->d_acheck);
    if (b->d_arecs) (void)hipFree(b->d_arecs);
    if (b->d_adma) (void)hipFree(b->d_adma);
    if (b->d_abad) (void)hipFree(b->d_abad);
    if (b->h_abad) (void)hipHostFree(b->h_abad);
    if (b->h_astage) (void)hipHostFree(b->h_astage);
    for (void *p : b->retired_dev) (void)hipFree(p);
    for (void *p : b->retired_host) (void)hipHostFree(p);
    for (hipEvent_t e : b->aevents) (void)hipEventDestroy(e);
    if (b->astream) (void)hipStreamDestroy(b->astream);
    if (b->d_graphs) (void)hipFree(b->d_graphs);
    if (b->d_state) (void)hipFree(b->d_state);
    if (b->d_gtiles) (void)hipFree(b->d_gtiles);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_counts) (void)hipFree(b->d_counts);
    if (b->d_tile_sum) (void)hipFree(b->d_tile_sum);
    if (b->d_tile_pre) (void)hipFree(b->d_tile_pre);
    if (b->d_gather) (void)hipFree(b->d_gather);
    if (b->d_pulls) (void)hipFree(b->d_pulls);
    if (b->d_check) (void)hipFree(b->d_check);
    if (b->d_bad) (void)hipFree(b->d_bad);
    if (b->h_state) (void)hipHostFree(b->h_state);
    if (b->h_graphs) (void)hipHostFree(b->h_graphs);
    if (b->h_gtiles) (void)hipHostFree(b->h_gtiles);
    if (b->h_items) (void)hipHostFree(b->h_items);
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PCD_OK;
}

extern "C" int pcd_batch_create(int device, int max_graphs, int64_t arena_bytes, pcd_batch **out) {
    if (!out) return bfail(PCD_E_INVALID, "null output pointer");
    *out = nullptr;
    if (max_graphs <= 0 || arena_bytes <= 0) return bfail(PCD_E_INVALID, "max_graphs and arena_bytes must be > 0");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return bfail(PCD_E_NODEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return bfail(PCD_E_INVALID, "device index out of range");
    hipDeviceProp_t prop;
    PCD_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return bfail(PCD_E_NODEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    PCD_HIP(hipSetDevice(device));
    pcd_batch *b = new pcd_batch();
    b->device = device;
    b->cu_count = prop.multiProcessorCount;
    b->cap = max_graphs;
    b->slots.resize(max_graphs);
    b->slot_listed.assign(max_graphs, 0);
    auto bad = [&](int code, const char *what) {
        pcd_batch_destroy(b);
        return bfail(code, what);
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bad(PCD_E_HIP, "stream");
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return bad(PCD_E_HIP, "event");
    b->arena_bytes = align256((size_t)arena_bytes);
    if (hipMalloc((void **)&b->arena, b->arena_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return bad(PCD_E_NOMEM, "arena allocation failed");
    }
    b->free_list[0] = b->arena_bytes;
    if (hipMalloc((void **)&b->d_graphs, sizeof(BatchGraph) * max_graphs) != hipSuccess ||
        hipMalloc((void **)&b->d_state, sizeof(BatchState) * max_graphs) != hipSuccess ||
        hipMalloc((void **)&b->d_gtiles, sizeof(int2) * max_graphs) != hipSuccess ||
        hipMalloc((void **)&b->d_counts, sizeof(int32_t) * NUM_LISTS) != hipSuccess) {
        (void)hipGetLastError();
        return bad(PCD_E_NOMEM, "batch tables");
    }
    // coherent: k_b_copy reads and writes these mirrors on the device
    if (hipHostMalloc((void **)&b->h_state, sizeof(BatchState) * max_graphs, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&b->h_graphs, sizeof(BatchGraph) * max_graphs, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&b->h_gtiles, sizeof(int2) * max_graphs, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&b->h_counts, sizeof(int32_t) * NUM_LISTS, hipHostMallocCoherent) != hipSuccess)
        return bad(PCD_E_NOMEM, "pinned batch tables");
    for (int i = 0; i < max_graphs; i++) {
        std::memset(&b->h_state[i], 0, sizeof(BatchState));
        b->h_state[i].status = BG_FREE;
    }
    std::memset(b->h_gtiles, 0, sizeof(int2) * max_graphs);
    std::memset(b->h_counts, 0, sizeof(int32_t) * NUM_LISTS);
    if (hipMemcpyAsync(b->d_state, b->h_state, sizeof(BatchState) * max_graphs, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->d_gtiles, b->h_gtiles, sizeof(int2) * max_graphs, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->d_counts, b->h_counts, sizeof(int32_t) * NUM_LISTS, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bad(PCD_E_HIP, "batch table upload");
    *out = b;
    return PCD_OK;
}

namespace {
// Host-side preparation of one graph: argument and format checks, the
// degree order (when the ids are not degree-sorted), the bin boundaries and
// the statistics row items.  Touches no batch or device state, so graphs
// can be prepared on several threads at once.
struct Prepared {
    int code = PCD_OK;
    const char *msg = "";
    bool sorted = true;
    int pos_le[NUM_TH] = {0};
    std::vector<int32_t> order;
};
}  // namespace

// check_rows false leaves the adjacency rows to the device check of
// pcd_batch_add_many (k_b_check_rows after the upload).
static void prepare_graph(int n, int m, const int32_t *indices, const int32_t *adj, const int32_t *labels,
                          const pcd_job *job, bool parallel, bool check_rows, Prepared &pr) {
    auto reject = [&](int code, const char *msg) {
        pr.code = code;
        pr.msg = msg;
    };
    if (!job) return reject(PCD_E_INVALID, "null job");
    if (n <= 0 || m <= 0 || !indices || !adj || !labels) return reject(PCD_E_INVALID, "bad graph arguments");
    if ((int64_t)2 * m > INT32_MAX) return reject(PCD_E_INVALID, "2m exceeds int32");
    if (job->rule.kind != PCD_RULE_MODULARITY && job->rule.kind != PCD_RULE_CPM)
        return reject(PCD_E_INVALID, "batches run modularity or CPM");
    if (!(job->frac > 0.0)) return reject(PCD_E_INVALID, "frac must be > 0");
    if (job->tie != PCD_TIE_RANDOM && job->tie != PCD_TIE_LOWEST) return reject(PCD_E_INVALID, "unknown tie order");
    if (indices[0] != 0 || indices[n] != 2 * m) return reject(PCD_E_FORMAT, "indices[0] != 0 or indices[n] != 2m");
    int maxd = 0;
    bool sorted = true;
    for (int v = 0; v < n; v++) {
        int d = indices[v + 1] - indices[v];
        if (d < 0) return reject(PCD_E_FORMAT, "indices must be monotone");
        if (v > 0 && d > indices[v] - indices[v - 1]) sorted = false;
        maxd = std::max(maxd, d);
        if (labels[v] < -1 || labels[v] >= n) return reject(PCD_E_INVALID, "labels must be in [0,n) or NEW_LABEL");
    }
    auto rows_ok = [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++)
            for (int j = indices[v]; j < indices[v + 1]; j++) {
                int u = adj[j];
                if (u < 0 || u >= n || u == v || (j > indices[v] && adj[j - 1] >= u)) return false;
            }
        return true;
    };
    bool ok = true;
    if (!check_rows) {
    } else if (parallel) {
        std::atomic<int> bad{0};
        parallel_for(n, [&](int64_t lo, int64_t hi) {
            if (!rows_ok(lo, hi)) bad = 1;
        }, 2048);
        ok = !bad;
    } else {
        ok = rows_ok(0, n);
    }
    if (!ok) return reject(PCD_E_FORMAT, "adjacency rows must be ascending, loop-free, with ids in [0,n)");
    if (maxd > 4096) return reject(PCD_E_INVALID, "degree above 4096: run this graph in a plain context");
    pr.sorted = sorted;
    // degree order (stable, descending) when the graph is not degree-sorted
    if (!sorted) {
        std::vector<int64_t> bucket((size_t)maxd + 2, 0);
        for (int v = 0; v < n; v++) bucket[maxd - (indices[v + 1] - indices[v]) + 1]++;
        for (int d = 0; d <= maxd; d++) bucket[d + 1] += bucket[d];
        pr.order.resize(n);
        for (int v = 0; v < n; v++) pr.order[bucket[maxd - (indices[v + 1] - indices[v])]++] = v;
    }
    auto deg_at = [&](int p) {
        int v = sorted ? p : pr.order[p];
        return indices[v + 1] - indices[v];
    };
    for (int i = 0; i < NUM_TH; i++) {
        int lo = 0, hi = n;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (deg_at(mid) <= TH[i]) hi = mid;
            else lo = mid + 1;
        }
        pr.pos_le[i] = lo;
    }
}

// True when the device may read p directly: page-locked host memory mapped
// at the same address (pcd_host_alloc, hipHostMalloc).  Pageable memory
// answers with an error, which is cleared.
static bool device_readable(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost && a.devicePointer == p;
}

// Queues a host-to-arena copy of `words` int32: a pull piece (flush_pulls
// launches them) for device-readable sources, else a copy on `stream`.
constexpr size_t PULL_WORDS = 1 << 18;  // 1 MiB pieces
static void pull_pieces(pcd_batch *b, int32_t *dst, const int32_t *src, size_t words) {
    for (size_t o = 0; o < words; o += PULL_WORDS)
        b->pulls.push_back(PullPiece{src + o, dst + o, (int64_t)std::min(PULL_WORDS, words - o)});
}
static int upload_words(pcd_batch *b, int32_t *dst, const int32_t *src, size_t words, hipStream_t stream) {
    if (words == 0) return PCD_OK;
    if (device_readable(src) && device_readable(src + words - 1)) {
        pull_pieces(b, dst, src, words);
        return PCD_OK;
    }
    PCD_HIP(hipMemcpyAsync(dst, src, 4 * words, hipMemcpyHostToDevice, stream));
    return PCD_OK;
}

// Copies `words` int32 into the page-locked staging buffer at *cursor and
// queues a pull piece from there (asynchronous admission of pageable
// labels and of the degree orders the host computes).
static void stage_words(pcd_batch *b, int32_t *dst, const int32_t *src, size_t words, int32_t **cursor) {
    if (words == 0) return;
    std::memcpy(*cursor, src, 4 * words);
    pull_pieces(b, dst, *cursor, words);
    *cursor += words;
}

// Launches the queued pull pieces on the batch stream.  The host sources
// must stay valid until the stream has passed the launch.
static int flush_pulls(pcd_batch *b) {
    if (b->pulls.empty()) return PCD_OK;
    if (b->pulls.size() > b->pulls_cap) {
        if (b->d_pulls) PCD_HIP(hipFree(b->d_pulls));
        b->d_pulls = nullptr;
        b->pulls_cap = 0;
        const size_t cap = b->pulls.size() + b->pulls.size() / 2 + 64;
        if (hipMalloc((void **)&b->d_pulls, sizeof(PullPiece) * cap) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(PCD_E_NOMEM, "admission copy list");
        }
        b->pulls_cap = cap;
    }
    // pageable source: staged before the call returns
    PCD_HIP(hipMemcpyAsync(b->d_pulls, b->pulls.data(), sizeof(PullPiece) * b->pulls.size(), hipMemcpyHostToDevice,
                           b->stream));
    const int count = (int)b->pulls.size();
    hipLaunchKernelGGL(k_b_pull, dim3(std::min(count, b->cu_count * 8)), dim3(256), 0, b->stream, b->d_pulls, count);
    PCD_HIP(hipGetLastError());
    b->pulls.clear();
    return PCD_OK;
}

// A slot's graph starts: its state record (host copy) is reset to running
// and it is listed for the next round.
static void start_graph(pcd_batch *b, int s) {
    BatchState &st = b->h_state[s];
    std::memset(&st, 0, sizeof(st));
    st.status = BG_RUNNING;
    st.label_bound = b->slots[s].n;
    // items of the slot's previous graph may still be listed: rebuild then
    if (b->items_cap == 0 || b->slot_listed[s]) b->items_dirty = true;
    if (!b->items_dirty) b->pending_append.push_back(s);
}

// Places a prepared graph in the first free slot at or after *hint and
// queues its uploads on the batch stream.  The slot's records (graph and
// state) are written on the host only; upload_records sends them.  The
// caller synchronises the stream before the host buffers may change.
// With `stage` (asynchronous admission) the copies go to the admission
// stream, the degree order and, with stage_labels, the labels through the
// staging buffer; the slot is left BG_PENDING, outside the item lists and
// with its state record untouched, until join_chunk takes it in.
static int admit_prepared(pcd_batch *b, int n, int m, const int32_t *indices, const int32_t *adj,
                          const int32_t *labels, const pcd_job *job, Prepared &pr, int *hint, int *slot_out,
                          int32_t **stage = nullptr, bool stage_labels = false) {
    const bool async = stage != nullptr;
    hipStream_t us = async ? b->astream : b->stream;
    int s = -1;
    for (int i = *hint; i < b->cap; i++)
        if (!b->slots[i].used) { s = i; break; }
    if (s < 0) return bfail(PCD_E_NOMEM, "no free graph slot");
    const int64_t E = 2 * (int64_t)m;
    const bool sorted = pr.sorted;
    const size_t bytes = footprint(n, E, sorted);
    size_t off = 0;
    if (!arena_alloc(b, bytes, &off)) return bfail(PCD_E_NOMEM, "batch arena full");
    *hint = s + 1;
    Slot &sl = b->slots[s];
    sl = Slot();
    sl.used = true;
    sl.status = async ? BG_PENDING : BG_RUNNING;
    sl.n = n;
    sl.m = m;
    sl.off = off;
    sl.bytes = bytes;
    sl.admitted = b->rounds;
    std::memcpy(sl.pos_le, pr.pos_le, sizeof(sl.pos_le));
    // device layout
    char *base = b->arena + off;
    size_t nb = align256(4 * ((size_t)n + 1));
    int32_t *d_ind = (int32_t *)base;
    int32_t *d_adj = (int32_t *)(base + nb);
    char *p = base + nb + align256(4 * (size_t)E);
    int32_t *d_order = nullptr;
    if (!sorted) {
        d_order = (int32_t *)p;
        p += nb;
    }
    int32_t *d_cur = (int32_t *)p;
    int32_t *d_next = (int32_t *)(p + nb);
    int32_t *d_first = (int32_t *)(p + 2 * nb);
    int32_t *d_newid = (int32_t *)(p + 3 * nb);
    CommStats *d_stats = (CommStats *)(p + 4 * nb);
    CommStats *d_stats_alt = (CommStats *)(p + 4 * nb + align256(16 * ((size_t)n + 1)));
    PCD_TRY(upload_words(b, d_ind, indices, (size_t)n + 1, us));
    PCD_TRY(upload_words(b, d_adj, adj, (size_t)E, us));
    if (!sorted) {
        if (async) stage_words(b, d_order, pr.order.data(), (size_t)n, stage);
        else PCD_HIP(hipMemcpyAsync(d_order, pr.order.data(), 4 * (size_t)n, hipMemcpyHostToDevice, b->stream));
    }
    if (async && stage_labels) stage_words(b, d_next, labels, (size_t)n, stage);
    else PCD_TRY(upload_words(b, d_next, labels, (size_t)n, us));
    BatchGraph &g = sl.rec;
    std::memset(&g, 0, sizeof(g));
    g.lp.indices = d_ind;
    g.lp.adj = d_adj;
    g.lp.labels = d_cur;
    g.lp.order = d_order;
    g.lp.lv = nullptr;
    g.lp.tot = nullptr;
    g.lp.stats = d_stats;
    g.lp.new_labels = d_next;
    g.lp.rule.kind = job->rule.kind;
    g.lp.rule.lambda = job->rule.lambda;
    set_edges(g.lp.rule, m);
    g.lp.rule.cc = 0.0;
    g.lp.seed = job->seed;
    g.lp.step = 0;
    g.lp.tie = job->tie;
    g.lp.f = (int)((job->frac < 1.0 ? job->frac : 1.0) * 2147483647.0);  // refinement.cpp:116
    g.lp.use_aux = 0;
    g.lp.counters = nullptr;
    g.next = d_next;
    g.cur = d_cur;
    g.stats = d_stats;
    g.stats_alt = d_stats_alt;
    g.first = d_first;
    g.newid = d_newid;
    g.n = n;
    g.max_iter = job->max_iter;
    g.threshold = job->threshold;
    g.iter_cap = job->iter_cap > 0 ? job->iter_cap : 1000;
    if (!async) start_graph(b, s);
    *slot_out = s;
    return PCD_OK;
}

// Uploads the graph and state records of `slots` (ascending), one launch
// per run of consecutive slots, through the page-locked record mirrors.
static int upload_records(pcd_batch *b, const std::vector<int> &slots) {
    for (int s : slots) b->h_graphs[s] = b->slots[s].rec;
    for (size_t i = 0; i < slots.size();) {
        size_t j = i + 1;
        while (j < slots.size() && slots[j] == slots[j - 1] + 1) j++;
        const int s0 = slots[i];
        const size_t cnt = j - i;
        PCD_TRY(kcopy(b, b->d_graphs + s0, b->h_graphs + s0, sizeof(BatchGraph) * cnt));
        PCD_TRY(kcopy(b, b->d_state + s0, b->h_state + s0, sizeof(BatchState) * cnt));
        i = j;
    }
    return PCD_OK;
}

extern "C" int pcd_batch_add(pcd_batch *b, int n, int m, const int32_t *indices, const int32_t *adj,
                             const int32_t *labels, const pcd_job *job, int *slot_out) {
    if (!b || !job || !slot_out) return bfail(PCD_E_INVALID, "null argument");
    *slot_out = -1;
    Prepared pr;
    prepare_graph(n, m, indices, adj, labels, job, true, true, pr);
    if (pr.code != PCD_OK) return bfail(pr.code, pr.msg);
    PCD_HIP(hipSetDevice(b->device));
    int hint = 0, s = -1;
    int rc = admit_prepared(b, n, m, indices, adj, labels, job, pr, &hint, &s);
    if (rc != PCD_OK) return rc;
    PCD_TRY(flush_pulls(b));
    rc = upload_records(b, std::vector<int>{s});
    if (rc != PCD_OK) return rc;
    PCD_HIP(hipStreamSynchronize(b->stream));
    *slot_out = s;
    return PCD_OK;
}

extern "C" int pcd_batch_add_many(pcd_batch *b, int count, const pcd_graph_ref *graphs, const pcd_job *jobs,
                                  int *slots_out, int *codes_out) {
    if (!b || count < 0 || (count > 0 && (!graphs || !jobs || !slots_out)))
        return bfail(PCD_E_INVALID, "bad arguments");
    for (int i = 0; i < count; i++) {
        slots_out[i] = -1;
        if (codes_out) codes_out[i] = PCD_OK;
    }
    if (count == 0) return PCD_OK;
    PCD_HIP(hipSetDevice(b->device));
    // Graphs are prepared CHUNK at a time on the host threads (one graph per
    // thread); the uploads of one chunk run while the next is prepared.
    constexpr int CHUNK = 256;
    std::vector<Prepared> pr;
    std::vector<int> admitted;
    int first_code = PCD_OK, first_index = -1;
    std::string first_msg;
    int hint = 0;
    for (int c0 = 0; c0 < count; c0 += CHUNK) {
        const int c1 = std::min(count, c0 + CHUNK);
        pr.assign(c1 - c0, Prepared());
        parallel_for(c1 - c0, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; i++) {
                const pcd_graph_ref &g = graphs[c0 + i];
                prepare_graph(g.n, g.m, g.indices, g.adj, g.labels, jobs + c0 + i, false, false, pr[i]);
            }
        }, 1);
        for (int i = c0; i < c1; i++) {
            const pcd_graph_ref &g = graphs[i];
            Prepared &p = pr[i - c0];
            int rc = p.code;
            std::string msg = p.msg;
            if (rc == PCD_OK) {
                int s = -1;
                rc = admit_prepared(b, g.n, g.m, g.indices, g.adj, g.labels, jobs + i, p, &hint, &s);
                if (rc == PCD_OK) {
                    slots_out[i] = s;
                    admitted.push_back(s);
                } else {
                    msg = pcd_last_error();
                }
            }
            if (rc == PCD_E_HIP) return bfail(rc, "graph " + std::to_string(i) + ": " + msg);
            if (codes_out) codes_out[i] = rc;
            if (rc != PCD_OK && first_code == PCD_OK) {
                first_code = rc;
                first_index = i;
                first_msg = msg;
            }
        }
        PCD_TRY(flush_pulls(b));  // this chunk's copies run while the next one is checked
    }
    // the rows of the admitted graphs, checked on the device after the upload
    if (!admitted.empty()) {
        std::vector<WorkItem> items;
        for (int s : admitted)
            for (int p = 0; p < b->slots[s].n; p += 256) items.push_back(WorkItem{s, p, std::min(b->slots[s].n, p + 256)});
        if (items.size() > b->check_cap) {
            if (b->d_check) PCD_HIP(hipFree(b->d_check));
            b->d_check = nullptr;
            b->check_cap = 0;
            if (hipMalloc((void **)&b->d_check, sizeof(WorkItem) * items.size()) != hipSuccess) {
                (void)hipGetLastError();
                return bfail(PCD_E_NOMEM, "row check list");
            }
            b->check_cap = items.size();
        }
        if (!b->d_bad && hipMalloc((void **)&b->d_bad, sizeof(int32_t) * b->cap) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(PCD_E_NOMEM, "row check flags");
        }
        PCD_TRY(upload_records(b, admitted));
        PCD_HIP(hipMemsetAsync(b->d_bad, 0, sizeof(int32_t) * b->cap, b->stream));
        PCD_HIP(hipMemcpyAsync(b->d_check, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice,
                               b->stream));
        hipLaunchKernelGGL(k_b_check_rows, dim3(std::min<size_t>(items.size(), (size_t)b->cu_count * 8)), dim3(256), 0,
                           b->stream, b->d_check, (int)items.size(), (const BatchGraph *)b->d_graphs, b->d_bad);
        PCD_HIP(hipGetLastError());
        std::vector<int32_t> bad((size_t)b->cap);
        PCD_HIP(hipMemcpyAsync(bad.data(), b->d_bad, sizeof(int32_t) * b->cap, hipMemcpyDeviceToHost, b->stream));
        PCD_HIP(hipStreamSynchronize(b->stream));
        std::vector<char> rejected((size_t)b->cap, 0);
        for (int s : admitted) rejected[s] = bad[s] != 0;
        // a rejected graph gives its slot and arena space back; its state
        // record goes back to free
        std::vector<int> freed;
        for (int i = 0; i < count; i++) {
            const int s = slots_out[i];
            if (s < 0 || !rejected[s]) continue;
            Slot &sl = b->slots[s];
            arena_free(b, sl.off, sl.bytes);
            sl = Slot();
            std::memset(&b->h_state[s], 0, sizeof(BatchState));
            b->h_state[s].status = BG_FREE;
            freed.push_back(s);
            slots_out[i] = -1;
            if (codes_out) codes_out[i] = PCD_E_FORMAT;
            if (first_code == PCD_OK || i < first_index) {
                first_code = PCD_E_FORMAT;
                first_index = i;
                first_msg = "adjacency rows must be ascending, loop-free, with ids in [0,n)";
            }
        }
        if (!freed.empty()) {
            auto &pa = b->pending_append;
            pa.erase(std::remove_if(pa.begin(), pa.end(), [&](int s) { return rejected[s] != 0; }), pa.end());
            for (int s : freed)
                PCD_HIP(hipMemcpyAsync(b->d_state + s, b->h_state + s, sizeof(BatchState), hipMemcpyHostToDevice,
                                       b->stream));
            PCD_HIP(hipStreamSynchronize(b->stream));
        }
    }
    if (first_code != PCD_OK) return bfail(first_code, "graph " + std::to_string(first_index) + ": " + first_msg);
    return PCD_OK;
}

// Takes in one chunk of an asynchronous admission once its uploads and row
// check are done: a graph whose rows failed is rejected (its arena space
// goes back, the slot stays taken until released), the others start and
// their records go to the device ahead of the next round.
static int join_chunk(pcd_batch *b, const pcd_batch::AChunk &c) {
    std::vector<int> joined;
    for (int i = c.first; i < c.first + c.count; i++) {
        const int s = b->aslots[i];
        if (s < 0) continue;
        Slot &sl = b->slots[s];
        if (!sl.used || sl.status != BG_PENDING) continue;
        b->apending--;
        if (b->h_abad[i]) {
            arena_free(b, sl.off, sl.bytes);
            sl.off = sl.bytes = 0;
            sl.status = BG_REJECTED;
            continue;
        }
        sl.status = BG_RUNNING;
        sl.admitted = b->rounds;
        start_graph(b, s);
        joined.push_back(s);
    }
    if (joined.empty()) return PCD_OK;
    std::sort(joined.begin(), joined.end());
    PCD_TRY(upload_records(b, joined));
    PCD_HIP(hipStreamSynchronize(b->stream));
    return PCD_OK;
}

// Joins the queued chunks in order: every chunk whose event has passed, all
// of them with wait_all, and at least the first with wait_one.
static int join_pending(pcd_batch *b, bool wait_all, bool wait_one) {
    bool waited = false;
    while (!b->achunks.empty()) {
        const pcd_batch::AChunk c = b->achunks.front();
        if (wait_all || (wait_one && !waited)) {
            PCD_HIP(hipEventSynchronize(c.ev));
            waited = true;
        } else {
            const hipError_t e = hipEventQuery(c.ev);
            if (e == hipErrorNotReady) {
                (void)hipGetLastError();
                break;
            }
            if (e != hipSuccess) return bfail(PCD_E_HIP, std::string("admission event: ") + hipGetErrorString(e));
        }
        b->achunks.pop_front();
        PCD_TRY(join_chunk(b, c));
    }
    free_retired(b);
    return PCD_OK;
}

// Grows a device buffer of `count` elements of T (contents dropped).
template <class T>
static int grow_device(T **p, size_t *cap, size_t count, const char *what) {
    if (count <= *cap) return PCD_OK;
    if (*p) PCD_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    const size_t c = count + count / 2 + 64;
    if (hipMalloc((void **)p, sizeof(T) * c) != hipSuccess) {
        (void)hipGetLastError();
        return bfail(PCD_E_NOMEM, what);
    }
    *cap = c;
    return PCD_OK;
}

// Gives back the slots and arena space of an asynchronous admission's
// graphs from call-local index `from` on, whose chunks were not queued
// because an error ended the call: they are reported refused with `code`.
static void abandon_async(pcd_batch *b, int from, int count, int code, int *slots_out, int *codes_out) {
    if (b->astream) (void)hipStreamSynchronize(b->astream);
    for (int i = from; i < count; i++) {
        const int s = b->aslots[i];
        if (s < 0) continue;
        Slot &sl = b->slots[s];
        if (sl.used && sl.status == BG_PENDING) {
            arena_free(b, sl.off, sl.bytes);
            sl = Slot();
            b->apending--;
        }
        b->aslots[i] = -1;
        slots_out[i] = -1;
        if (codes_out) codes_out[i] = code;
    }
}

static void slot_counts(const BinMap &mp, const Slot &sl, bool lp_only, int64_t *cnt);
static size_t list_esz(int);

// Sizes the item lists and tile-scan buffers for every running and pending
// graph under every bin map, as rebuild_items lays them out, so that no
// rebuild while admission copies are queued has to allocate: an
// allocation waits for the copy engines, which the queued copies hold.
static int reserve_lists(pcd_batch *b) {
    size_t need = 0;
    int64_t tiles = 0;
    for (int mp = 0; mp < NUM_MAPS; mp++) {
        int64_t cnt[NUM_LISTS] = {0};
        for (const Slot &sl : b->slots)
            if (sl.used && (sl.status == BG_RUNNING || sl.status == BG_PENDING)) slot_counts(MAPS[mp], sl, false, cnt);
        size_t off = 0;
        for (int l = 0; l < NUM_LISTS; l++)
            off = align256(off + (cnt[l] + std::max<int64_t>(cnt[l] / 2, 1024)) * list_esz(l));
        need = std::max(need, off);
        tiles = std::max(tiles, cnt[L_TILES] + std::max<int64_t>(cnt[L_TILES] / 2, 1024));
    }
    if (need > b->items_cap) {
        drop_graph(b);
        PCD_HIP(hipStreamSynchronize(b->stream));
        free_device(b, b->d_items);
        free_host(b, b->h_items);
        b->d_items = b->h_items = nullptr;
        b->items_cap = 0;
        b->items_dirty = true;
        const size_t cap = need + need / 4 + 4096;
        if (hipMalloc(&b->d_items, cap) != hipSuccess || hipHostMalloc(&b->h_items, cap, hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(PCD_E_NOMEM, "work item lists");
        }
        b->items_cap = cap;
    }
    if ((size_t)tiles > b->tile_cap) {
        drop_graph(b);
        PCD_HIP(hipStreamSynchronize(b->stream));
        free_device(b, b->d_tile_sum);
        free_device(b, b->d_tile_pre);
        b->d_tile_sum = b->d_tile_pre = nullptr;
        b->tile_cap = 0;
        b->items_dirty = true;
        const size_t cap = (size_t)tiles + 64;
        if (hipMalloc((void **)&b->d_tile_sum, 4 * cap) != hipSuccess ||
            hipMalloc((void **)&b->d_tile_pre, 4 * cap) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(PCD_E_NOMEM, "tile scan buffers");
        }
        b->tile_cap = cap;
    }
    return PCD_OK;
}

extern "C" int pcd_batch_add_many_async(pcd_batch *b, int count, const pcd_graph_ref *graphs, const pcd_job *jobs,
                                        int *slots_out, int *codes_out) {
    if (!b || count < 0 || (count > 0 && (!graphs || !jobs || !slots_out)))
        return bfail(PCD_E_INVALID, "bad arguments");
    for (int i = 0; i < count; i++) {
        slots_out[i] = -1;
        if (codes_out) codes_out[i] = PCD_OK;
    }
    if (count == 0) return PCD_OK;
    PCD_HIP(hipSetDevice(b->device));
    // the previous call's buffers are reused: its chunks join first
    PCD_TRY(join_pending(b, true, false));
    if (!b->astream && hipStreamCreateWithFlags(&b->astream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        b->astream = nullptr;
        return bfail(PCD_E_HIP, "admission stream");
    }
    // host checks of every graph (several threads), then the staging need
    std::vector<Prepared> pr((size_t)count);
    parallel_for(count, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            const pcd_graph_ref &g = graphs[i];
            prepare_graph(g.n, g.m, g.indices, g.adj, g.labels, jobs + i, false, false, pr[i]);
        }
    }, 1);
    std::vector<char> stage_lab((size_t)count, 0);
    size_t stage_words_needed = 0;
    for (int i = 0; i < count; i++) {
        if (pr[i].code != PCD_OK) continue;
        const pcd_graph_ref &g = graphs[i];
        stage_lab[i] = !(device_readable(g.labels) && device_readable(g.labels + g.n - 1));
        stage_words_needed += (stage_lab[i] ? (size_t)g.n : 0) + (pr[i].sorted ? 0 : (size_t)g.n);
    }
    if (stage_words_needed > b->astage_cap) {
        if (b->h_astage) PCD_HIP(hipHostFree(b->h_astage));
        b->h_astage = nullptr;
        b->astage_cap = 0;
        const size_t c = stage_words_needed + stage_words_needed / 4 + 1024;
        if (hipHostMalloc((void **)&b->h_astage, 4 * c, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            b->h_astage = nullptr;
            return bfail(PCD_E_NOMEM, "admission staging");
        }
        b->astage_cap = c;
    }
    // slots, arena space and copies, chunk by chunk
    constexpr int CHUNK = 256;
    struct Span {
        size_t p0, p1, i0, i1;
    };
    std::vector<Span> spans;
    std::vector<PullPiece> pieces;
    std::vector<WorkItem> items;
    std::vector<BatchGraph> recs((size_t)count);
    std::memset(recs.data(), 0, sizeof(BatchGraph) * recs.size());
    b->aslots.assign((size_t)count, -1);
    b->pulls.clear();
    int32_t *cursor = b->h_astage;
    int first_code = PCD_OK, first_index = -1;
    std::string first_msg;
    int hint = 0;
    for (int c0 = 0; c0 < count; c0 += CHUNK) {
        const int c1 = std::min(count, c0 + CHUNK);
        Span sp;
        sp.p0 = pieces.size();
        sp.i0 = items.size();
        for (int i = c0; i < c1; i++) {
            const pcd_graph_ref &g = graphs[i];
            int rc = pr[i].code;
            std::string msg = pr[i].msg;
            if (rc == PCD_OK) {
                int s = -1;
                rc = admit_prepared(b, g.n, g.m, g.indices, g.adj, g.labels, jobs + i, pr[i], &hint, &s, &cursor,
                                    stage_lab[i] != 0);
                if (rc == PCD_OK) {
                    slots_out[i] = s;
                    b->aslots[i] = s;
                    b->apending++;
                    recs[i] = b->slots[s].rec;
                    for (int p = 0; p < g.n; p += 256) items.push_back(WorkItem{i, p, std::min(g.n, p + 256)});
                } else {
                    msg = pcd_last_error();
                }
            }
            if (rc == PCD_E_HIP) {
                abandon_async(b, 0, count, rc, slots_out, codes_out);
                return bfail(rc, "graph " + std::to_string(i) + ": " + msg);
            }
            if (codes_out) codes_out[i] = rc;
            if (rc != PCD_OK && first_code == PCD_OK) {
                first_code = rc;
                first_index = i;
                first_msg = msg;
            }
            pr[i].order = std::vector<int32_t>();
        }
        pieces.insert(pieces.end(), b->pulls.begin(), b->pulls.end());
        b->pulls.clear();
        sp.p1 = pieces.size();
        sp.i1 = items.size();
        spans.push_back(sp);
    }
    // everything below queues work; on an error the graphs of the chunks
    // not yet queued give their slots back
    int queued = 0;  // graphs whose chunk has its event
    auto queue_all = [&]() -> int {
        PCD_TRY(reserve_lists(b));
        // Kernel reads of host memory slow the rounds running beside them (each
        // waiting read holds the memory pipeline); copy-engine transfers do not.
        // So per chunk the pieces are sorted by source address, pieces that
        // continue one another inside one page-locked allocation form one copy
        // into device staging, and the chunk's launch then moves each piece from
        // there to its place.  A piece whose allocation the runtime does not
        // report is read from host memory by that launch.
        struct Run {
            const char *src;
            size_t off, bytes;
        };
        std::vector<Run> runs;
        std::vector<size_t> run_first(spans.size() + 1, 0), soff(pieces.size(), SIZE_MAX);
        std::vector<const char *> pbase(pieces.size(), nullptr);
        {
            const char *cb = nullptr, *ce = nullptr;  // the last allocation found
            const char *sb = (const char *)b->h_astage, *se = sb + 4 * b->astage_cap;
            for (size_t k = 0; k < pieces.size(); k++) {
                const char *p = (const char *)pieces[k].src, *q = p + 4 * (size_t)pieces[k].words;
                if (sb && p >= sb && q <= se) {
                    pbase[k] = sb;
                    continue;
                }
                if (!(cb && p >= cb && q <= ce)) {
                    hipDeviceptr_t base = nullptr;
                    size_t sz = 0;
                    if (hipMemGetAddressRange(&base, &sz, (hipDeviceptr_t)p) == hipSuccess && base &&
                        (const char *)base <= p && q <= (const char *)base + sz) {
                        cb = (const char *)base;
                        ce = cb + sz;
                    } else {
                        (void)hipGetLastError();
                        cb = ce = nullptr;
                    }
                }
                pbase[k] = cb;
            }
        }
        size_t stage_bytes = 0;
        for (size_t c = 0; c < spans.size(); c++) {
            run_first[c] = runs.size();
            std::vector<size_t> ord;
            for (size_t k = spans[c].p0; k < spans[c].p1; k++)
                if (pbase[k]) ord.push_back(k);
            std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return pieces[x].src < pieces[y].src; });
            size_t off = 0;
            const char *run_base = nullptr;
            for (size_t k : ord) {
                const char *p = (const char *)pieces[k].src;
                const size_t nb = 4 * (size_t)pieces[k].words;
                if (runs.size() > run_first[c] && pbase[k] == run_base && p == runs.back().src + runs.back().bytes) {
                    soff[k] = runs.back().off + runs.back().bytes;
                    runs.back().bytes += nb;
                } else {
                    off = ((off + 15) & ~(size_t)15) + ((uintptr_t)p & 15);  // same alignment as the source
                    runs.push_back(Run{p, off, nb});
                    run_base = pbase[k];
                    soff[k] = off;
                }
                off = runs.back().off + runs.back().bytes;
            }
            stage_bytes = std::max(stage_bytes, off);
        }
        run_first[spans.size()] = runs.size();
        PCD_TRY(grow_device(&b->d_adma, &b->adma_cap, std::max<size_t>(stage_bytes, 16), "admission staging"));
        for (size_t k = 0; k < pieces.size(); k++)
            if (soff[k] != SIZE_MAX) pieces[k].src = (const int32_t *)(b->d_adma + soff[k]);
        // device copies of the piece, item and record lists, then per chunk:
        // copies into staging, the pieces to their places, row check, flags back
        // to the host, event
        PCD_TRY(grow_device(&b->d_apulls, &b->apulls_cap, std::max<size_t>(pieces.size(), 1), "admission copy list"));
        PCD_TRY(grow_device(&b->d_acheck, &b->acheck_cap, std::max<size_t>(items.size(), 1), "row check list"));
        PCD_TRY(grow_device(&b->d_arecs, &b->arecs_cap, (size_t)count, "row check records"));
        if ((size_t)count > b->abad_cap) {
            if (b->d_abad) PCD_HIP(hipFree(b->d_abad));
            if (b->h_abad) PCD_HIP(hipHostFree(b->h_abad));
            b->d_abad = b->h_abad = nullptr;
            b->abad_cap = 0;
            const size_t c = (size_t)count + count / 2 + 64;
            if (hipMalloc((void **)&b->d_abad, 4 * c) != hipSuccess ||
                hipHostMalloc((void **)&b->h_abad, 4 * c, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                if (b->d_abad) (void)hipFree(b->d_abad);
                b->d_abad = nullptr;
                b->h_abad = nullptr;
                return bfail(PCD_E_NOMEM, "row check flags");
            }
            b->abad_cap = c;
        }
        hipStream_t as = b->astream;
        if (!pieces.empty())
            PCD_HIP(hipMemcpyAsync(b->d_apulls, pieces.data(), sizeof(PullPiece) * pieces.size(), hipMemcpyHostToDevice, as));
        if (!items.empty())
            PCD_HIP(hipMemcpyAsync(b->d_acheck, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, as));
        PCD_HIP(hipMemcpyAsync(b->d_arecs, recs.data(), sizeof(BatchGraph) * count, hipMemcpyHostToDevice, as));
        PCD_HIP(hipMemsetAsync(b->d_abad, 0, 4 * (size_t)count, as));
        for (size_t c = 0; c < spans.size(); c++) {
            const Span &sp = spans[c];
            const int first = (int)c * CHUNK, cnt = std::min(count - first, CHUNK);
            // at most 16 MiB per transfer: a copy on the batch stream waits for
            // the transfer in flight on the engine, not for the queue behind it
            constexpr size_t DMA_BYTES = 16 << 20;
            for (size_t r = run_first[c]; r < run_first[c + 1]; r++)
                for (size_t o = 0; o < runs[r].bytes; o += DMA_BYTES)
                    PCD_HIP(hipMemcpyAsync(b->d_adma + runs[r].off + o, runs[r].src + o,
                                           std::min(DMA_BYTES, runs[r].bytes - o), hipMemcpyHostToDevice, as));
            if (sp.p1 > sp.p0)
                hipLaunchKernelGGL(k_b_pull, dim3(std::min<size_t>(sp.p1 - sp.p0, (size_t)b->cu_count * 2)), dim3(256), 0, as,
                                   (const PullPiece *)(b->d_apulls + sp.p0), (int)(sp.p1 - sp.p0));
            if (sp.i1 > sp.i0)
                hipLaunchKernelGGL(k_b_check_rows, dim3(std::min<size_t>(sp.i1 - sp.i0, (size_t)b->cu_count * 8)), dim3(256),
                                   0, as, (const WorkItem *)(b->d_acheck + sp.i0), (int)(sp.i1 - sp.i0),
                                   (const BatchGraph *)b->d_arecs, b->d_abad);
            PCD_HIP(hipGetLastError());
            PCD_HIP(hipMemcpyAsync(b->h_abad + first, b->d_abad + first, 4 * (size_t)cnt, hipMemcpyDeviceToHost, as));
            if (b->aevents.size() <= c) {
                hipEvent_t e = nullptr;
                PCD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                b->aevents.push_back(e);
            }
            PCD_HIP(hipEventRecord(b->aevents[c], as));
            b->achunks.push_back(pcd_batch::AChunk{first, cnt, b->aevents[c]});
            queued = first + cnt;
        }
        return PCD_OK;
    };
    const int qrc = queue_all();
    if (qrc != PCD_OK) {
        const std::string msg = pcd_last_error();
        abandon_async(b, queued, count, qrc, slots_out, codes_out);
        return bfail(qrc, msg);
    }
    if (first_code != PCD_OK) return bfail(first_code, "graph " + std::to_string(first_index) + ": " + first_msg);
    return PCD_OK;
}

// Items of one graph: vertex tiles, statistics classes (degree in (G/2, G],
// G4 takes 0..4), long rows, and the propagation ranges of the bin map.
static void lp_range(const BinMap &mp, const Slot &sl, int i, int &p0, int &p1) {
    p0 = sl.pos_le[th_index(mp.hi[i])];
    p1 = i + 1 < mp.n ? sl.pos_le[th_index(mp.hi[i + 1])] : sl.n;
}
static void small_range(const Slot &sl, int i, int &p0, int &p1) {
    int G = STAT_G[i];
    p0 = sl.pos_le[th_index(G)];
    p1 = G == 4 ? sl.n : sl.pos_le[th_index(G / 2)];
}
static int64_t nitems(int p0, int p1, int per) { return p1 > p0 ? (p1 - p0 + per - 1) / per : 0; }

static void slot_counts(const BinMap &mp, const Slot &sl, bool lp_only, int64_t *cnt) {
    int p0, p1;
    if (!lp_only) {
        cnt[L_TILES] += (sl.n + VT - 1) / VT;
        for (int i = 0; i < 5; i++) {
            small_range(sl, i, p0, p1);
            cnt[L_SMALL0 + i] += nitems(p0, p1, STAT_CP[i]);
        }
        cnt[L_ROWS] += nitems(0, sl.pos_le[th_index(64)], ROWS_PER);
    }
    for (int i = 0; i < mp.n; i++) {
        lp_range(mp, sl, i, p0, p1);
        cnt[L_LP0 + mp.var[i]] += nitems(p0, p1, LPV_ITEM[mp.var[i]]);
    }
}

static void slot_fill(const BinMap &mp, const Slot &sl, int s, bool lp_only, WorkItem **wp) {
    auto push_range = [&](int l, int p0, int p1, int per) {
        for (int p = p0; p < p1; p += per) *wp[l]++ = WorkItem{s, p, std::min(p1, p + per)};
    };
    int p0, p1;
    if (!lp_only) {
        push_range(L_TILES, 0, sl.n, VT);
        for (int i = 0; i < 5; i++) {
            small_range(sl, i, p0, p1);
            push_range(L_SMALL0 + i, p0, p1, STAT_CP[i]);
        }
        push_range(L_ROWS, 0, sl.pos_le[th_index(64)], ROWS_PER);
    }
    for (int i = 0; i < mp.n; i++) {
        lp_range(mp, sl, i, p0, p1);
        push_range(L_LP0 + mp.var[i], p0, p1, LPV_ITEM[mp.var[i]]);
    }
}

static size_t list_esz(int) { return sizeof(WorkItem); }

// Host-side check of the work-item list layout: every list's region lies
// inside the item buffer, after the previous list's, and holds its items;
// the host counts match the lengths; the tile list fits the tile-scan
// buffers.  Run after every rebuild and append, so that a kernel of the
// captured round can only read items that were written.
static int check_lists(const pcd_batch *b) {
    size_t end = 0;
    for (int l = 0; l < NUM_LISTS; l++) {
        const size_t off = b->list_off[l], bytes = (size_t)b->list_cap[l] * list_esz(l);
        if (b->list_len[l] < 0 || b->list_len[l] > b->list_cap[l] || b->h_counts[l] != b->list_len[l] ||
            off < end || off + bytes > b->items_cap)
            return bfail(PCD_E_STATE, "work item list " + std::to_string(l) + " out of its region (len " +
                                          std::to_string(b->list_len[l]) + ", cap " + std::to_string(b->list_cap[l]) +
                                          ", offset " + std::to_string(off) + ", buffer " + std::to_string(b->items_cap) + ")");
        end = off + bytes;
    }
    if ((size_t)b->list_cap[L_TILES] > b->tile_cap) return bfail(PCD_E_STATE, "tile list exceeds the tile-scan buffers");
    return PCD_OK;
}

// Rebuilds the work-item lists from the running graphs and uploads them:
// one counting pass fixes the layout (every list with room to grow, so
// later admissions append in place), a second writes the items straight
// into the pinned staging buffer.  With settled_only, graphs admitted after
// the last round are left out (their labels are not compressed yet); with
// lp_only, only the propagation lists (which depend on the bin map) are
// rewritten, behind unchanged base lists; stride > 1 lists every
// stride-th slot only.
static int rebuild_items(pcd_batch *b, bool settled_only = false, bool lp_only = false, int stride = 1) {
    const BinMap &mp = MAPS[b->map];
    auto listed_slot = [&](const Slot &sl) {
        return sl.used && sl.status == BG_RUNNING && !(settled_only && sl.admitted >= b->rounds) &&
               (stride == 1 || (&sl - b->slots.data()) % stride == 0);
    };
    // pass 1: counts (room also for the graphs of queued admission chunks,
    // so that they join by appending, without a new capture)
    int64_t cnt[NUM_LISTS] = {0}, room[NUM_LISTS] = {0};
    int64_t listed = 0;
    for (int s = 0; s < b->cap; s++) {
        const Slot &sl = b->slots[s];
        if (sl.used && sl.status == BG_PENDING && stride == 1) slot_counts(mp, sl, lp_only, room);
        if (!listed_slot(sl)) continue;
        listed += sl.n;
        slot_counts(mp, sl, lp_only, cnt);
    }
    // layout with growth room
    const int first_list = lp_only ? L_LP0 : 0;
    size_t off = lp_only ? align256(b->list_off[L_ROWS] + std::max<int64_t>(b->list_cap[L_ROWS], 1) * sizeof(WorkItem))
                         : 0;
    size_t list_off[NUM_LISTS];
    int64_t list_cap[NUM_LISTS];
    for (int l = first_list; l < NUM_LISTS; l++) {
        if (cnt[l] + room[l] > INT32_MAX / 2) return bfail(PCD_E_NOMEM, "work item list too long");
        list_cap[l] = cnt[l] + room[l] + std::max<int64_t>(cnt[l] / 2, 1024);
        list_off[l] = off;
        off = align256(off + list_cap[l] * list_esz(l));
    }
    const size_t ntiles = lp_only ? b->tile_cap : list_cap[L_TILES];
    PCD_HIP(hipStreamSynchronize(b->stream));  // the staging buffer may still feed a copy
    if (off > b->items_cap) {
        if (lp_only) return rebuild_items(b, settled_only, false, stride);
        drop_graph(b);
        free_device(b, b->d_items);
        free_host(b, b->h_items);
        b->d_items = b->h_items = nullptr;
        size_t cap = off + off / 4 + 4096;
        if (hipMalloc(&b->d_items, cap) != hipSuccess || hipHostMalloc(&b->h_items, cap, hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            b->items_cap = 0;
            return bfail(PCD_E_NOMEM, "work item lists");
        }
        b->items_cap = cap;
    }
    if (ntiles > b->tile_cap) {
        drop_graph(b);
        free_device(b, b->d_tile_sum);
        free_device(b, b->d_tile_pre);
        size_t cap = ntiles + 64;
        if (hipMalloc((void **)&b->d_tile_sum, 4 * cap) != hipSuccess ||
            hipMalloc((void **)&b->d_tile_pre, 4 * cap) != hipSuccess) {
            (void)hipGetLastError();
            b->tile_cap = 0;
            return bfail(PCD_E_NOMEM, "tile scan buffers");
        }
        b->tile_cap = cap;
    }
    for (int l = first_list; l < NUM_LISTS; l++) {
        b->list_off[l] = list_off[l];
        b->list_cap[l] = l == L_TILES ? std::min<int64_t>(list_cap[l], (int64_t)b->tile_cap) : list_cap[l];
        b->list_len[l] = (int)cnt[l];
        b->h_counts[l] = (int)cnt[l];
    }
    // pass 2: fill
    char *hb = (char *)b->h_items;
    WorkItem *wp[NUM_LISTS];
    for (int l = first_list; l < NUM_LISTS; l++) wp[l] = (WorkItem *)(hb + b->list_off[l]);
    int ntile = 0;
    for (int s = 0; s < b->cap; s++) {
        const Slot &sl = b->slots[s];
        if (!lp_only) b->slot_listed[s] = listed_slot(sl) ? 1 : 0;
        if (!listed_slot(sl)) {
            if (!lp_only) b->h_gtiles[s] = make_int2(0, 0);
            continue;
        }
        if (!lp_only) {
            const int nt = (sl.n + VT - 1) / VT;
            b->h_gtiles[s] = make_int2(ntile, nt);
            ntile += nt;
        }
        slot_fill(mp, sl, s, lp_only, wp);
    }
    const size_t up0 = lp_only ? b->list_off[L_LP0] : 0;
    PCD_TRY(kcopy(b, (char *)b->d_items + up0, hb + up0, off - up0));
    PCD_TRY(kcopy(b, b->d_counts, b->h_counts, sizeof(int32_t) * NUM_LISTS));
    if (!lp_only) PCD_TRY(kcopy(b, b->d_gtiles, b->h_gtiles, sizeof(int2) * b->cap));
    PCD_HIP(hipStreamSynchronize(b->stream));
    if (!lp_only) {
        b->items_dirty = false;
        b->listed_vertices = listed;
        b->listed_done_vertices = 0;
        b->pending_append.clear();
    }
    // offsets inside the item buffer moved: the captured round is stale
    drop_graph(b);
    return check_lists(b);
}

// Appends the items of the graphs admitted since the last rebuild behind
// the current lists (offsets unchanged, so the captured round stays valid);
// falls back to a rebuild when a list is out of room.
static int append_items(pcd_batch *b) {
    const BinMap &mp = MAPS[b->map];
    int64_t add[NUM_LISTS] = {0};
    for (int s : b->pending_append) slot_counts(mp, b->slots[s], false, add);
    for (int l = 0; l < NUM_LISTS; l++)
        if (b->list_len[l] + add[l] > b->list_cap[l]) return rebuild_items(b);
    PCD_HIP(hipStreamSynchronize(b->stream));  // the staging buffer may still feed a copy
    char *hb = (char *)b->h_items;
    WorkItem *wp[NUM_LISTS];
    for (int l = 0; l < NUM_LISTS; l++) wp[l] = (WorkItem *)(hb + b->list_off[l] + b->list_len[l] * list_esz(l));
    int ntile = b->list_len[L_TILES];
    int s_lo = b->cap, s_hi = -1;
    for (int s : b->pending_append) {
        const Slot &sl = b->slots[s];
        const int nt = (sl.n + VT - 1) / VT;
        b->h_gtiles[s] = make_int2(ntile, nt);
        ntile += nt;
        slot_fill(mp, sl, s, false, wp);
        b->slot_listed[s] = 1;
        b->listed_vertices += sl.n;
        s_lo = std::min(s_lo, s);
        s_hi = std::max(s_hi, s);
    }
    // the tile ranges of the slots in between are unchanged on both sides
    if (s_hi >= s_lo) PCD_TRY(kcopy(b, b->d_gtiles + s_lo, b->h_gtiles + s_lo, sizeof(int2) * (size_t)(s_hi - s_lo + 1)));
    for (int l = 0; l < NUM_LISTS; l++) {
        if (!add[l]) continue;
        size_t o = b->list_off[l] + (size_t)b->list_len[l] * list_esz(l);
        PCD_TRY(kcopy(b, (char *)b->d_items + o, hb + o, add[l] * list_esz(l)));
        b->list_len[l] += (int)add[l];
        b->h_counts[l] = b->list_len[l];
    }
    PCD_TRY(kcopy(b, b->d_counts, b->h_counts, sizeof(int32_t) * NUM_LISTS));
    PCD_HIP(hipStreamSynchronize(b->stream));
    b->pending_append.clear();
    return check_lists(b);
}

template <class T>
static T *list_ptr(pcd_batch *b, int l) {
    return (T *)((char *)b->d_items + b->list_off[l]);
}

static BatchLPLaunch lp_launch(pcd_batch *b) {
    BatchLPLaunch l;
    std::memset(&l, 0, sizeof(l));
    const BinMap &mp = MAPS[b->map];
    for (int v = 0; v < NUM_LPV; v++) {
        l.items[v] = list_ptr<WorkItem>(b, L_LP0 + v);
        l.count[v] = b->d_counts + L_LP0 + v;
        l.used[v] = false;
    }
    for (int i = 0; i < mp.n; i++) l.used[mp.var[i]] = true;
    l.graphs = b->d_graphs;
    l.state = b->d_state;
    l.cu_count = b->cu_count;
    return l;
}

// The launches of one round (main.cpp:226-266 for every running graph).
static int launch_round(pcd_batch *b) {
    hipStream_t s = b->stream;
    const int cu = b->cu_count;
    const WorkItem *tiles = list_ptr<WorkItem>(b, L_TILES);
    const int *ntile = b->d_counts + L_TILES;
    const BatchGraph *gs = b->d_graphs;
    BatchState *st = b->d_state;
    // statistics of the last round's propagation step (graphs past their
    // first round)
    hipLaunchKernelGGL(k_b_delta_rows, dim3(cu * 8), dim3(256), 0, s, list_ptr<WorkItem>(b, L_ROWS),
                       b->d_counts + L_ROWS, gs, st);
#define PCD_BDELTA(I, G)                                                                                    \
    hipLaunchKernelGGL(k_b_delta_small<G>, dim3(cu * 8), dim3(256), 0, s, list_ptr<WorkItem>(b, L_SMALL0 + I), \
                       b->d_counts + L_SMALL0 + I, gs, st);
    PCD_BDELTA(0, 64)
    PCD_BDELTA(1, 32)
    PCD_BDELTA(2, 16)
    PCD_BDELTA(3, 8)
    PCD_BDELTA(4, 4)
#undef PCD_BDELTA
    // compress_labels (framework.cpp:146-170)
    hipLaunchKernelGGL(k_b_fill_first, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    hipLaunchKernelGGL(k_b_first_occurrence, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    hipLaunchKernelGGL(k_b_tile_count, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st, b->d_tile_sum);
    hipLaunchKernelGGL(k_b_tile_scan, dim3(1), dim3(1024), 0, s, ntile, b->d_tile_sum, b->d_tile_pre);
    hipLaunchKernelGGL(k_b_tile_apply, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, b->d_gtiles, gs, st,
                       b->d_tile_sum, b->d_tile_pre);
    hipLaunchKernelGGL(k_b_relabel, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    hipLaunchKernelGGL(k_b_stats_copy, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    // collect_statistics (framework.cpp:69-143) in a graph's first round
    hipLaunchKernelGGL(k_b_clear_stats, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    hipLaunchKernelGGL(k_b_stats_rows, dim3(cu * 8), dim3(256), 0, s, list_ptr<WorkItem>(b, L_ROWS),
                       b->d_counts + L_ROWS, gs, st);
#define PCD_BSTATS(I, G)                                                                                    \
    hipLaunchKernelGGL(k_b_stats_small<G>, dim3(cu * 8), dim3(256), 0, s, list_ptr<WorkItem>(b, L_SMALL0 + I), \
                       b->d_counts + L_SMALL0 + I, gs, st);
    PCD_BSTATS(0, 64)
    PCD_BSTATS(1, 32)
    PCD_BSTATS(2, 16)
    PCD_BSTATS(3, 8)
    PCD_BSTATS(4, 4)
#undef PCD_BSTATS
    // Rule::calculate and the stop rule
    hipLaunchKernelGGL(k_b_metric, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    hipLaunchKernelGGL(k_b_finalize, dim3((b->cap + 255) / 256), dim3(256), 0, s, gs, st, b->cap);
    // PropagationRefinement::refine for the graphs still running
    hipLaunchKernelGGL(k_b_copy_next, dim3(cu * 4), dim3(256), 0, s, tiles, ntile, gs, st);
    PCD_TRY(batch_propagation_launch(s, lp_launch(b)));
    PCD_HIP(hipGetLastError());
    return PCD_OK;
}

// Host-side check of a captured round before it is instantiated: only
// kernel nodes (no memset or copy nodes), each with a kernel, a grid of at
// most 64 workgroups per CU, 64-1024 threads per workgroup and no dynamic
// LDS (every batch kernel declares its LDS statically).
static int check_graph(pcd_batch *b, hipGraph_t g) {
    size_t nn = 0;
    PCD_HIP(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    if (nn) PCD_HIP(hipGraphGetNodes(g, nodes.data(), &nn));
    for (size_t i = 0; i < nn; i++) {
        hipGraphNodeType t;
        PCD_HIP(hipGraphNodeGetType(nodes[i], &t));
        if (t != hipGraphNodeTypeKernel)
            return bfail(PCD_E_STATE, "captured round holds a node of type " + std::to_string((int)t));
        hipKernelNodeParams kp;
        std::memset(&kp, 0, sizeof(kp));
        PCD_HIP(hipGraphKernelNodeGetParams(nodes[i], &kp));
        const uint64_t blocks = (uint64_t)kp.gridDim.x * kp.gridDim.y * kp.gridDim.z;
        const uint64_t threads = (uint64_t)kp.blockDim.x * kp.blockDim.y * kp.blockDim.z;
        if (!kp.func || blocks == 0 || blocks > (uint64_t)b->cu_count * 64 || threads < 64 || threads > 1024 ||
            kp.sharedMemBytes != 0)
            return bfail(PCD_E_STATE, "captured kernel node " + std::to_string(i) + " has grid " + std::to_string(blocks) +
                                          ", block " + std::to_string(threads) + ", dynamic LDS " +
                                          std::to_string(kp.sharedMemBytes));
    }
    if (nn == 0) return bfail(PCD_E_STATE, "captured round is empty");
    return PCD_OK;
}

static int ensure_graph(pcd_batch *b) {
    if (b->items_dirty) PCD_TRY(rebuild_items(b));
    else if (!b->pending_append.empty()) PCD_TRY(append_items(b));
    if (b->exec) return PCD_OK;
    // relaxed: worker threads of a server keep allocating and copying on
    // their own streams while this stream captures
    PCD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int rc = launch_round(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != PCD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(PCD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    PCD_TRY(check_graph(b, g));
    PCD_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return PCD_OK;
}

static int poll_state(pcd_batch *b, int *running) {
    PCD_TRY(kcopy(b, b->h_state, b->d_state, sizeof(BatchState) * b->cap));
    PCD_HIP(hipStreamSynchronize(b->stream));
    int r = 0;
    for (int s = 0; s < b->cap; s++) {
        Slot &sl = b->slots[s];
        if (!sl.used || sl.status == BG_PENDING || sl.status == BG_REJECTED) continue;
        int now = b->h_state[s].status;
        if (sl.status == BG_RUNNING && now != BG_RUNNING) b->listed_done_vertices += sl.n;
        sl.status = now;
        if (now == BG_RUNNING) r++;
    }
    // retire finished graphs from the lists once they are a quarter of the work
    if (b->listed_done_vertices * 4 > b->listed_vertices) b->items_dirty = true;
    if (running) *running = r;
    return PCD_OK;
}

extern "C" int pcd_batch_step(pcd_batch *b, int rounds, int *running) {
    if (!b) return bfail(PCD_E_INVALID, "null batch");
    if (rounds <= 0) return bfail(PCD_E_INVALID, "rounds must be positive");
    PCD_HIP(hipSetDevice(b->device));
    if (!b->achunks.empty()) {
        bool any = false;
        for (const Slot &sl : b->slots) any |= sl.used && sl.status == BG_RUNNING;
        PCD_TRY(join_pending(b, false, !any));
    }
    PCD_TRY(ensure_graph(b));
    for (int r = 0; r < rounds; r++) PCD_HIP(hipGraphLaunch(b->exec, b->stream));
    b->rounds += rounds;
    PCD_TRY(poll_state(b, running));
    if (running) *running += b->apending;
    return PCD_OK;
}

extern "C" int pcd_batch_status(pcd_batch *b, int slot, int *status, int *iterations, int *k, double *score) {
    if (!b || slot < 0 || slot >= b->cap) return bfail(PCD_E_INVALID, "bad slot");
    const Slot &sl = b->slots[slot];
    const BatchState &st = b->h_state[slot];
    if (status) *status = sl.used ? sl.status : BG_FREE;
    if (iterations) *iterations = st.iter;
    if (k) *k = st.k;
    if (score) *score = st.score;
    return PCD_OK;
}

extern "C" int pcd_batch_finished(pcd_batch *b, int *slots, int cap, int *count) {
    if (!b || !count || (cap > 0 && !slots)) return bfail(PCD_E_INVALID, "bad arguments");
    int c = 0;
    for (int s = 0; s < b->cap; s++)
        if (b->slots[s].used && b->slots[s].status == BG_DONE) {
            if (c < cap) slots[c] = s;
            c++;
        }
    *count = c;
    return PCD_OK;
}

extern "C" int pcd_batch_labels(pcd_batch *b, int slot, int32_t *labels) {
    if (!b || slot < 0 || slot >= b->cap || !labels) return bfail(PCD_E_INVALID, "bad arguments");
    const Slot &sl = b->slots[slot];
    if (!sl.used) return bfail(PCD_E_STATE, "slot is free");
    if (sl.status != BG_DONE) return bfail(PCD_E_STATE, "graph is still running");
    PCD_HIP(hipSetDevice(b->device));
    PCD_HIP(hipMemcpyAsync(labels, sl.rec.cur, 4 * (size_t)sl.n, hipMemcpyDeviceToHost, b->stream));
    PCD_HIP(hipStreamSynchronize(b->stream));
    return PCD_OK;
}

extern "C" int pcd_batch_labels_many(pcd_batch *b, int count, const int *slots, int32_t *labels) {
    if (!b || count < 0 || (count > 0 && (!slots || !labels))) return bfail(PCD_E_INVALID, "bad arguments");
    if (count == 0) return PCD_OK;
    std::vector<int64_t> head((size_t)count + (count + 1) / 2);  // offsets, then the slots as int32
    int32_t *hslots = (int32_t *)(head.data() + count);
    int64_t total = 0;
    for (int i = 0; i < count; i++) {
        const int s = slots[i];
        if (s < 0 || s >= b->cap) return bfail(PCD_E_INVALID, "bad slot " + std::to_string(s));
        const Slot &sl = b->slots[s];
        if (!sl.used) return bfail(PCD_E_STATE, "slot " + std::to_string(s) + " is free");
        if (sl.status != BG_DONE) return bfail(PCD_E_STATE, "graph in slot " + std::to_string(s) + " is still running");
        head[i] = total;
        hslots[i] = s;
        total += sl.n;
    }
    const size_t hbytes = align256(sizeof(int64_t) * head.size());
    const size_t need = hbytes + 4 * (size_t)total;
    PCD_HIP(hipSetDevice(b->device));
    if (need > b->gather_cap) {
        free_device(b, b->d_gather);
        b->d_gather = nullptr;
        b->gather_cap = 0;
        if (hipMalloc((void **)&b->d_gather, need) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(PCD_E_NOMEM, "label staging allocation failed");
        }
        b->gather_cap = need;
    }
    const int64_t *d_offs = (const int64_t *)b->d_gather;
    const int32_t *d_slots = (const int32_t *)(d_offs + count);
    int32_t *d_out = (int32_t *)(b->d_gather + hbytes);
    PCD_HIP(hipMemcpyAsync(b->d_gather, head.data(), sizeof(int64_t) * head.size(), hipMemcpyHostToDevice, b->stream));
    const int grid = std::min(count, b->cu_count * 8);
    hipLaunchKernelGGL(k_b_gather_labels, dim3(grid), dim3(256), 0, b->stream, (const BatchGraph *)b->d_graphs,
                       d_slots, d_offs, count, d_out);
    PCD_HIP(hipGetLastError());
    PCD_HIP(hipMemcpyAsync(labels, d_out, 4 * (size_t)total, hipMemcpyDeviceToHost, b->stream));
    PCD_HIP(hipStreamSynchronize(b->stream));
    return PCD_OK;
}

extern "C" int pcd_batch_stats(pcd_batch *b, int slot, int32_t *sizes, int32_t *int_volume, int32_t *ext_volume) {
    if (!b || slot < 0 || slot >= b->cap) return bfail(PCD_E_INVALID, "bad arguments");
    const Slot &sl = b->slots[slot];
    if (!sl.used || sl.status != BG_DONE) return bfail(PCD_E_STATE, "graph is not finished");
    const int k = b->h_state[slot].k;
    std::vector<CommStats> st(std::max(k, 1));
    PCD_HIP(hipSetDevice(b->device));
    PCD_HIP(hipMemcpyAsync(st.data(), sl.rec.stats, sizeof(CommStats) * k, hipMemcpyDeviceToHost, b->stream));
    PCD_HIP(hipStreamSynchronize(b->stream));
    for (int c = 0; c < k; c++) {
        if (sizes) sizes[c] = st[c].size;
        if (int_volume) int_volume[c] = st[c].int_vol;
        if (ext_volume) ext_volume[c] = st[c].ext_vol;
    }
    return PCD_OK;
}

extern "C" int pcd_batch_release(pcd_batch *b, int slot) {
    if (!b || slot < 0 || slot >= b->cap) return bfail(PCD_E_INVALID, "bad slot");
    if (!b->slots[slot].used) return bfail(PCD_E_STATE, "slot is free");
    PCD_HIP(hipSetDevice(b->device));
    if (b->slots[slot].status == BG_PENDING) PCD_TRY(join_pending(b, true, false));
    Slot &sl = b->slots[slot];
    BatchState &st = b->h_state[slot];
    std::memset(&st, 0, sizeof(st));
    st.status = BG_FREE;
    // stream order: earlier rounds finish before the slot is reused
    PCD_TRY(kcopy(b, b->d_state + slot, &st, sizeof(BatchState)));
    PCD_HIP(hipStreamSynchronize(b->stream));
    if (sl.status == BG_RUNNING) b->items_dirty = true;
    if (sl.bytes) arena_free(b, sl.off, sl.bytes);
    sl = Slot();
    return PCD_OK;
}

extern "C" int pcd_batch_info(pcd_batch *b, int64_t *free_bytes, int *graphs, int *running, int64_t *rounds) {
    if (!b) return bfail(PCD_E_INVALID, "null batch");
    size_t fb = 0;
    for (auto &kv : b->free_list) fb += kv.second;
    int g = 0, r = 0;
    for (auto &sl : b->slots) {
        if (!sl.used) continue;
        g++;
        if (sl.status == BG_RUNNING || sl.status == BG_PENDING) r++;
    }
    if (free_bytes) *free_bytes = (int64_t)fb;
    if (graphs) *graphs = g;
    if (running) *running = r;
    if (rounds) *rounds = b->rounds;
    return PCD_OK;
}

extern "C" int pcd_batch_set_map(pcd_batch *b, int map) {
    if (!b || map < 0 || map >= NUM_MAPS) return bfail(PCD_E_INVALID, "bad bin map");
    if (map != b->map) {
        b->map = map;
        b->items_dirty = true;
        drop_graph(b);
    }
    return PCD_OK;
}

// Times the propagation phase of the current round under every bin map
// and keeps the fastest.  Propagation reads the compressed labels and
// statistics of the last completed round and writes only the next labels,
// which it fully determines, so repeating it changes no state.  Needs one
// completed round.  Large batches are timed on a sample of about 512 of
// their running graphs (every stride-th slot).
extern "C" int pcd_batch_autotune(pcd_batch *b, int reps, int *chosen, double *ms_per_map) {
    if (!b || reps <= 0) return bfail(PCD_E_INVALID, "bad arguments");
    if (b->rounds == 0) return bfail(PCD_E_STATE, "autotune needs one completed round");
    PCD_HIP(hipSetDevice(b->device));
    int running = 0;
    for (auto &sl : b->slots) running += sl.used && sl.status == BG_RUNNING;
    const int stride = std::max(1, running / 512);
    int best = b->map;
    double best_ms = 1e300;
    for (int mp = 0; mp < NUM_MAPS; mp++) {
        b->map = mp;
        PCD_TRY(rebuild_items(b, true, mp > 0, stride));
        double t = 1e300;
        for (int r = 0; r < reps; r++) {
            PCD_HIP(hipEventRecord(b->ev0, b->stream));
            PCD_TRY(batch_propagation_launch(b->stream, lp_launch(b)));
            PCD_HIP(hipEventRecord(b->ev1, b->stream));
            PCD_HIP(hipEventSynchronize(b->ev1));
            float ms = 0;
            PCD_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            t = std::min(t, (double)ms);
        }
        if (ms_per_map) ms_per_map[mp] = t;
        if (t < best_ms) {
            best_ms = t;
            best = mp;
        }
    }
    b->map = best;
    b->items_dirty = true;
    drop_graph(b);
    if (chosen) *chosen = best;
    return PCD_OK;
}
