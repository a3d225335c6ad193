// Internal declarations of the HIP side: the context and the launchers.
#ifndef PARCD_INTERNAL_H
#define PARCD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "parcd.h"
#include "rule_formulas.h"

namespace pcd {

int fail(int code, const std::string &msg);

#define PCD_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ::pcd::fail(PCD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define PCD_TRY(expr)               \
    do {                            \
        int r_ = (expr);            \
        if (r_ != PCD_OK) return r_; \
    } while (0)

constexpr int WAVE = 64;
constexpr int EMPTY_KEY = -1;

// Degree bins over vertex positions (positions are vertex ids when the
// graph is degree-sorted descending, as `convert` writes it, else the
// positions of a degree-descending permutation `order`).
enum Bin { BIN_HUB = 0, BIN_B4096, BIN_B1024, BIN_B512, BIN_W256, BIN_G64, BIN_G32, BIN_G16, BIN_G8, BIN_G4, NUM_BINS };
// degree upper bound of each bin (HUB: unbounded)
constexpr int BIN_MAX_DEG[NUM_BINS] = {0x7fffffff, 4096, 1024, 512, 256, 64, 32, 16, 8, 4};

struct BinLayout {
    int start[NUM_BINS + 1];  // position ranges, bin b = [start[b], start[b+1])
    int split2048;            // first position of degree <= 2048 (inside B4096)
};

// The rows one propagation launch counts: their positions in bin order (a
// bin is a position range), each bin's entries, and the layout of the rows
// in the HUB bin -- fixed-size chunks, label-hash partitions, the
// partitioned-entry buffer, the partial results and the fallback global
// tables.  Either the graph's vertices (fields of the context) or, in a
// merging step, its communities.
struct RowSet {
    BinLayout bins{};
    int64_t bin_entries[NUM_BINS]{};
    int num_hubs = 0, num_hub_items = 0, num_hub_parts = 0;
    const int4 *hub_items = nullptr;       // chunk: {hub, row, begin, end}
    const int32_t *hub_item_off = nullptr; // chunk -> offset of its (P+1) partition offsets
    const int4 *hub_part_items = nullptr;  // {hub, part, first chunk, number of chunks}
    const int4 *hub_info = nullptr;        // per hub: {row, P, first part item, first chunk}
    const int4 *hub_part_where = nullptr;  // per partition item: {row, P, poff of the first chunk, buffer base}
    const int64_t *hub_buf_base = nullptr; // per hub: offset into hub_buf
    int32_t *hub_poff = nullptr;
    int32_t *hub_buf = nullptr;
    void *hub_partial = nullptr;
    int32_t *hub_over = nullptr;
    const int64_t *hub_table_off = nullptr;  // per hub: fallback global table
    const int32_t *hub_table_log = nullptr;
    int32_t *hub_keys = nullptr, *hub_cnts = nullptr;
    // Laid out on the device (no host read-back): `layout` holds the bin
    // starts (NUM_BINS + 1 positions) and the fused small-row launch's
    // workgroup split (7 entries), `hub_counts` {chunks, partition items,
    // hubs}; the launches then use fixed grids and read these.
    const int *layout = nullptr;
    const int *hub_counts = nullptr;
};
constexpr int ROW_LAYOUT_INTS = NUM_BINS + 1 + 7;

// Tunables (the autotuner rewrites these per graph shape).
struct Tuning {
    int hub_chunk = 4096;        // entries per hub partition work item
    int hub_part_entries = 2048; // target entries per hub label partition
    int hub_table_slots = 4096;  // LDS table slots of a hub partition (multiple of 64, <= HUB_TB)
    int stat_chunk = 2048;       // entries per statistics work item for deg > 64
    int small_aux = 0;           // small bins gather {label, aux} records (1) or labels (0)
    int block_aux = 0;           // wave/workgroup-per-vertex bins gather {label, aux} records
    int hub_aux = 0;             // hub partitions count {label, aux} records (1) or labels (0)
    int hub_nt = 512;            // hub partition workgroups: 512 threads, 4096 slots or 256, 2048
    int small_fused = 1;         // the G64..G4 bins in one launch (1) or one launch each (0)
    int bin_streams = 1;         // a propagation step runs the hub rows and degrees 1025-4096 on
                                 // two side streams, beside the other bins (0: one stream)
    int bin_streams_min_edges = 1 << 19;  // ... for graphs with at least this many edges
    int small_occ8 = 1;          // the fused small-bin launch compiled for 8 waves per SIMD
                                 // (modularity and CPM builds; registers beyond spill to lanes)
    int stats_cache = 1;         // statistics over long rows read the labels of the first
                                 // STATS_CACHE vertices (the highest degrees) from LDS
    int inc_stats = 1;           // a propagation step updates the statistics from the rows
                                 // of the vertices it moved (else a full recount)
    int b1024_wave = 1;          // degrees 257-512 of B1024 run one wave per vertex with a
                                 // 1024-slot table (twice the vertices in flight per CU)
    int b1024_u16 = 1;           // those wave tables keep 16-bit counts (five workgroups per
                                 // CU instead of four)
    int b4096_u16 = 1;           // degrees 1025-4096: tables with 16-bit counts (56 instead of
                                 // 72 KB of LDS: beside the other bins, 1.012 -> 1.005 ms)
    int b4096_split = 0;         // degrees 1025-2048 of B4096 run 256-thread workgroups with a
                                 // 4096-slot table (measured slower on R-MAT-22: off)
    int cluster_rounds = 128;    // device rounds of the clustering partition before the
                                 // remaining dependency chains are finished on the host
    int onesweep = 0;            // radix sorts: one launch per pass with decoupled look-back (1)
                                 // or tile histograms, a scan and a scatter per pass (0).  Off:
                                 // its scatter costs what the three-launch pass's does, and its
                                 // all-digit histogram is slower on keys with long equal runs
    int os_spin_limit = 1 << 20; // polls a one-sweep look-back waits before the sort falls back
    int scan_tiles_per_cu = 8;   // scans: fewer rounds per tile until there are this many tiles per CU
                                 // (R-MAT-22 iteration 1.120 ms at 2, 1.102 at 8, 1.131 at 32)
    int graph_loop = 1;          // modularity / CPM with -s 0: passes replayed from one captured
                                 // pass with the stop rule on the device (pcd_iterate_many)
    int graph_loop_min_entries = 0;  // graphs with fewer CSR entries run pass by pass
    int merge_sort = 0;          // merging picks partners from the sorted unique community pairs
                                 // (1) or counts over community rows, no sort of the entries (0)
};
constexpr int STATS_CACHE = 16384;
constexpr int HUB_MAX_PARTS = 1024;
constexpr int HUB_TB = 4096;          // LDS table capacity of a hub partition
static_assert(Tuning{}.hub_table_slots <= HUB_TB && Tuning{}.hub_table_slots % 64 == 0,
              "default hub table must fit the LDS table");
constexpr int HUB_CHUNK = 4096;      // entries per hub chunk item (pass 1 holds 512 x 8)
constexpr int HUB_PARTIAL_BYTES = 32; // storage per (hub, partition) partial result

// Kernel timer slots (HIP events around each launch when timing is on).
enum KSlot {
    KS_COMPRESS = 0, KS_STATS_ITEMS, KS_STATS_SMALL, KS_METRIC,
    KS_LP_SELECT, KS_LP_HUB_INSERT, KS_LP_HUB_DECIDE, KS_LP_B4096, KS_LP_B1024, KS_LP_W256,
    KS_LP_G64, KS_LP_G32, KS_LP_G16, KS_LP_G8, KS_LP_G4,
    KS_MERGE_KEYS, KS_MERGE_SORT, KS_MERGE_RLE, KS_MERGE_ROUNDS, KS_TRIANGLES, KS_STATS_DELTA, KS_LP_SMALL,
    KS_LP_B512, KS_LP_HUB_FINAL, KS_MERGE_STATS,
    NUM_KSLOTS
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Host-built layout of HUB-bin rows (see RowSet), from each row's {id,
// first entry, end entry}.
struct HubLayoutHost {
    std::vector<int4> chunks, parts, where, info;
    std::vector<int32_t> chunk_off, tlog;
    std::vector<int64_t> buf_base, toff;  // per row, plus the total at the end
    int64_t poff_total = 0;
};
int build_hub_layout(const std::vector<int4> &rows, int part_entries, HubLayoutHost &L);

struct pcd_ctx_impl;

}  // namespace pcd

struct pcd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    // streams a propagation step forks independent degree bins onto (tuning
    // bin_streams), and the events that fork them from `stream` and join them
    // back; all at the default priority
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    // stream pcd_iterate runs the metric on, beside the propagation (which
    // does not read it), and the event that forks it from `stream`
    hipStream_t mstream = nullptr;
    hipEvent_t ev_mfork = nullptr;

    // graph
    int n = 0, m = 0, max_degree = 0;
    bool has_graph = false;
    int32_t *d_indices = nullptr, *d_adj = nullptr;
    int32_t *d_order = nullptr;  // nullptr when degree-sorted
    std::vector<int32_t> h_degree_sorted_pos;
    pcd::BinLayout bins{};
    int64_t bin_entries[pcd::NUM_BINS]{};
    pcd::Tuning tune{};
    // static work items: hub chunks (hub index, begin, end) and stats chunks
    int num_hubs = 0;
    int num_hub_items = 0;               // chunk items
    int num_hub_parts = 0;               // (hub, partition) items
    int4 *d_hub_items = nullptr;         // chunk: {hub, v, begin, end}
    int32_t *d_hub_item_off = nullptr;   // chunk -> offset of its (P+1) partition offsets
    int4 *d_hub_part_items = nullptr;    // {hub, part, first chunk, number of chunks}
    int4 *d_hub_info = nullptr;          // per hub: {v, P, first part item, first chunk}
    int4 *d_hub_part_where = nullptr;    // per partition item: {v, P, poff of the first chunk, buffer base}
    int32_t *d_hub_poff = nullptr;       // partition offsets of all chunks
    uint64_t *d_moved_bits = nullptr;    // per vertex: label changed in the last step (statistics update)
    int32_t *d_tot = nullptr;            // per community the total the modularity / CPM gains read
    int32_t *d_hub_buf = nullptr;        // partitioned {label, aux} records (sum of hub degrees)
    int64_t *d_hub_buf_base = nullptr;   // per hub: offset into d_hub_buf
    void *d_hub_partial = nullptr;       // per (hub, part) partial argmax
    int32_t *d_hub_over = nullptr;       // per hub: partition overflow flag
    int64_t *d_hub_table_off = nullptr;  // per hub (fallback global tables)
    int32_t *d_hub_table_log = nullptr;  // log2 table size per hub
    int32_t *d_hub_keys = nullptr, *d_hub_cnts = nullptr;
    int64_t hub_table_slots = 0;
    int num_stat_items = 0;
    int4 *d_stat_items = nullptr;  // {v, begin, end, first}

    // partition
    bool has_labels = false;
    bool labels_compressed = false;
    bool labels_have_new = false;  // labels may hold NEW_LABEL
    int64_t first_cap = 0;         // capacity of d_first / d_newid
    bool has_stats = false;
    // d_stats holds the statistics of d_labels under the label ids of the
    // last compression, vertices with NEW_LABEL not counted (after a
    // propagation step); the next compression renumbers them
    bool stats_pending = false;
    int label_bound = 0;  // labels in [-1, label_bound)
    int k = 0;
    int32_t *d_labels = nullptr, *d_labels_alt = nullptr;
    int2 *d_lv = nullptr;  // per vertex {label, aux} for the propagation gathers
    pcd::CommStats *d_stats = nullptr;
    pcd::CommStats *d_stats_alt = nullptr;  // compression writes the renumbered statistics here
    int32_t *d_moved = nullptr;  // 8 list bounds, then the moved small-bin vertices and stats items
    int32_t *d_first = nullptr, *d_newid = nullptr;
    // per compressed label, the vertex the last compression found first
    // (n + 1 entries, each a vertex id or -1): the next compression's
    // first-occurrence table starts from it where that vertex kept the label
    int32_t *d_firstv = nullptr;
    int32_t *d_active = nullptr;   // per-bin active lists (bin ranges)
    int32_t *d_counts = nullptr;   // per-bin active counts + misc device counters
    int32_t *d_hub_active = nullptr;
    int32_t *d_scan_part = nullptr;  // scan partials
    int scan_part_cap = 0;
    int32_t *d_k = nullptr;          // device copy of k
    int64_t *d_acc = nullptr;        // exact integer accumulators
    double *d_dacc = nullptr;        // FP64 partial sums
    int dacc_cap = 0;
    double *d_result = nullptr;
    double *h_result = nullptr;      // pinned
    bool k_deferred = false;         // k is on the device; h_small[K_SLOT] receives it
    hipEvent_t ev_metric = nullptr;  // recorded after the score's read-back
    int32_t *h_small = nullptr;      // pinned scratch ints
    int64_t *h_small64 = nullptr;    // pinned scratch int64

    // triangles
    bool has_triangles = false;
    int32_t *d_tri = nullptr, *d_tri_deg = nullptr, *d_et = nullptr;
    int32_t *d_lowpos = nullptr, *d_it = nullptr, *d_itd = nullptr;
    double cc = 0.0;

    // merge / contraction scratch
    uint64_t *d_keys0 = nullptr, *d_keys1 = nullptr;
    size_t keys_cap = 0;
    int32_t *d_pair_links = nullptr, *d_pair_head = nullptr;
    int32_t *d_rs_hist = nullptr;
    size_t rs_hist_cap = 0;
    // device-resident loop (loop.hip): one captured pass, replayed
    bool capturing = false;          // a pass is being captured: no host reads, no pointer swaps of stats
    const int *graph_step = nullptr; // the captured pass reads its step number here
    void *loop = nullptr;            // the captured pass and its device state (loop.hip)
    int64_t alloc_gen = 0;           // counts device allocations and frees (a capture's buffers)
    int64_t tune_gen = 0;            // counts tuning changes (a capture's kernel forms)
    uint32_t *d_os = nullptr;  // one-sweep sort: look-back status words, histograms, tickets
    size_t os_cap = 0;
    int64_t os_fallbacks = 0;  // passes redone after a look-back gave up waiting
    int32_t *d_bp = nullptr, *d_found = nullptr;
    uint64_t *d_best_ord = nullptr, *d_best_key = nullptr;
    void *d_mcand = nullptr;         // merge candidates of runs crossing waves
    size_t mcand_cap = 0;
    void *mrows = nullptr;           // community rows of the merging step (merge.hip)

    // measurement
    bool timing = false;
    uint32_t timer_mask = 0xFFFFFFFFu;  // kernel timer slots that record events
    bool counters = false;
    hipEvent_t ev0[PCD_PHASE_COUNT]{}, ev1[PCD_PHASE_COUNT]{};
    bool ev_valid[PCD_PHASE_COUNT]{};
    int64_t cnt_active = 0, cnt_entries = 0, cnt_distinct = 0;
    unsigned long long *d_bin_counters = nullptr;  // NUM_BINS x {active, entries, distinct}
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ktimes[pcd::NUM_KSLOTS];
    std::vector<hipEvent_t> event_pool;
    hipEvent_t kt_open = nullptr;

    std::vector<void *> allocations;
};

namespace pcd {

// allocation helpers (all device memory of a context is tracked)
int dev_alloc(pcd_ctx *ctx, void **p, size_t bytes);
void dev_free(pcd_ctx *ctx, void *p);
template <typename T>
int dev_alloc_t(pcd_ctx *ctx, T **p, size_t count) {
    return dev_alloc(ctx, (void **)p, sizeof(T) * (count ? count : 1));
}

// launch-size helper: grid for a grid-stride kernel
inline int grid_for(pcd_ctx *ctx, int64_t items, int per_block, int max_per_cu = 8) {
    int64_t g = (items + per_block - 1) / per_block;
    int64_t cap = (int64_t)ctx->cu_count * max_per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// primitives (prims.hip)
int fill_i32(pcd_ctx *ctx, int32_t *p, int32_t v, int64_t count);
int fill_u64(pcd_ctx *ctx, uint64_t *p, uint64_t v, int64_t count);
int ensure_scan_parts(pcd_ctx *ctx, int64_t n);
int radix_sort_u64(pcd_ctx *ctx, uint64_t *keys, uint64_t *tmp, int64_t n, int bits, uint64_t **out);
// the tile histograms of a radix sort of n keys (a captured sort may not allocate)
int ensure_sort_space(pcd_ctx *ctx, int64_t n);

// phases (partition.hip, propagate.hip, merge.hip, triangles.hip)
// defer_k: k stays on the device (d_k) and its read-back lands in pinned
// memory; resolve_k takes it once the stream has passed the copy.
int compress_labels_dev(pcd_ctx *ctx, bool defer_k = false);
void resolve_k(pcd_ctx *ctx);
// drops the captured pass of the device-resident loop (loop.hip)
void loop_release(pcd_ctx *ctx);
// drops the merging step's row state (merge.hip); its device buffers belong
// to the context's allocation list
void merge_release(pcd_ctx *ctx);
int collect_statistics_dev(pcd_ctx *ctx);
int statistics_delta_dev(pcd_ctx *ctx, const int32_t *before, const int32_t *after);
int clustering_labels_dev(pcd_ctx *ctx, int *num_communities);
int metric_dev(pcd_ctx *ctx, pcd_rule rule, double *out);
// modularity, CPM, conductance: kernels and the score's copy into
// ctx->h_result, no wait
int metric_enqueue(pcd_ctx *ctx, pcd_rule rule);
int wcc_metric_enqueue(pcd_ctx *ctx);
int propagation_dev(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie);
// propagation_dev in two parts: the kernels that write the next labels into
// d_labels_alt, then the statistics update and the swap that make them the
// current labels (a launch that is not committed changes nothing else)
int propagation_launch(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie);
int propagation_commit(pcd_ctx *ctx);
int merging_dev(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie);
// reserves the buffers of a merging step over community rows (a captured
// step may not allocate)
int merge_rows_reserve(pcd_ctx *ctx);
// fails when a merging step since the last check left hub rows out
int merge_rows_check(pcd_ctx *ctx);
int contract_dev(pcd_ctx *ctx, int64_t *num_pairs);
int contract_download(pcd_ctx *ctx, int64_t U, int32_t *l, int32_t *k, int32_t *links);
int triangles_dev(pcd_ctx *ctx);
int wcc_metric_dev(pcd_ctx *ctx, double *out);

RuleParams make_rule(pcd_ctx *ctx, pcd_rule rule);

// phase timing
void phase_begin(pcd_ctx *ctx, int phase);
void phase_end(pcd_ctx *ctx, int phase);
// kernel timing: bracket one or more launches of a slot
void kt_begin(pcd_ctx *ctx, int slot);
void kt_end(pcd_ctx *ctx, int slot);

}  // namespace pcd

#endif
