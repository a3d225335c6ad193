/*
 * parcd.h -- C ABI of the MI355X (gfx950) community-detection hot path.
 *
 * Plain C types only: int32 ids and offsets, host pointers that are borrowed
 * for the duration of a call, an opaque context that owns all device memory
 * and its HIP streams: one that orders every call, plus side streams that a
 * propagation step and pcd_iterate's metric fork onto and join back into it
 * with events, so results are ordered as on one stream.  A context is used
 * by one thread at a time.  Every function returns PCD_OK (0) or a negative
 * PCD_E* status; pcd_last_error() gives the message of the last failure on
 * the calling thread.  Nothing here exits the process or throws.
 *
 * Each entry point names the reference interface it replaces (stijnh/Par-CD,
 * src/).  Semantics follow that code; differences are stated per function.
 */
#ifndef PARCD_H
#define PARCD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCD_OK 0
#define PCD_E_INVALID (-1)   /* bad argument */
#define PCD_E_HIP (-2)       /* HIP runtime error */
#define PCD_E_NOMEM (-3)     /* device or host allocation failed / over budget */
#define PCD_E_IO (-4)        /* file error */
#define PCD_E_FORMAT (-5)    /* malformed PCD file or edge list */
#define PCD_E_STATE (-6)     /* call out of order (no graph, no labels, ...) */
#define PCD_E_NODEVICE (-7)  /* no usable gfx950 device */

#define PCD_NEW_LABEL (-1)   /* common.hpp:20 NEW_LABEL (~0) */

/* Metric rules (rules.hpp:9-75). */
#define PCD_RULE_MODULARITY 0
#define PCD_RULE_CPM 1
#define PCD_RULE_WCC 2
#define PCD_RULE_CONDUCTANCE 3
/* a rule with no device form (the default of parcd::Rule::descriptor in the
 * C++ mirror): refused with PCD_E_INVALID by every call that takes a rule */
#define PCD_RULE_HOST (-1)

/* Tie order among equal best gains.  RANDOM is a uniform choice keyed by a
 * counter hash of (seed, step, vertex|community, label) -- the counterpart of
 * the reservoir sampling at refinement.cpp:158-162,249-252; LOWEST takes the
 * lowest label (the deterministic synchronous mode). */
#define PCD_TIE_RANDOM 0
#define PCD_TIE_LOWEST 1

typedef struct pcd_rule {
    int kind;      /* PCD_RULE_* */
    double lambda; /* CPM resolution (CPMRule(l), rules.hpp:52) */
} pcd_rule;

typedef struct pcd_ctx pcd_ctx;

/* ---------------------------------------------------------------- context */
/* Replaces framework-cuda.cuh:14-15 cuda_init/cuda_deinit. */
int pcd_ctx_create(int device, pcd_ctx **out);
int pcd_ctx_destroy(pcd_ctx *ctx);
const char *pcd_last_error(void);
int pcd_device_count(int *count);
int pcd_synchronize(pcd_ctx *ctx);

/* ------------------------------------------------------------------ graph */
/* Uploads a CSR graph (graph.hpp:6-40 Graph{n, m, indices, adj}).  indices
 * has n+1 entries, adj 2m entries, rows sorted ascending without loops or
 * duplicates (what `convert` writes).  Replaces framework-cuda.cuh:16
 * cuda_transfer_graph.  Previous graph, labels and statistics are dropped. */
int pcd_graph_upload(pcd_ctx *ctx, int n, int m, const int32_t *indices, const int32_t *adj);
int pcd_graph_info(pcd_ctx *ctx, int *n, int *m, int *max_degree);

/* ----------------------------------------------------------------- labels */
/* Replaces framework-cuda.cuh:17-18 cuda_labels_to_device/_from_device.
 * Labels are >= 0 or PCD_NEW_LABEL. */
int pcd_labels_upload(pcd_ctx *ctx, const int32_t *labels);
int pcd_labels_download(pcd_ctx *ctx, int32_t *labels);

/* framework.cpp:146-170 compress_labels: dense relabel in order of first
 * occurrence over vertex ids, each NEW_LABEL becomes a fresh id.  Bit-exact
 * with the reference.  *num_labels (may be NULL) receives k. */
int pcd_compress_labels(pcd_ctx *ctx, int *num_labels);

/* No counterpart in the reference: an allocation-free compress_labels for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * pcd_compress_labels, bit for bit, on caller-owned DEVICE arrays: labels[n]
 * (values in [0,n) or PCD_NEW_LABEL) is read, out_labels[n] receives the
 * compressed labels and *num_labels (one device int32) receives k, or -1
 * when a label lies outside [-1, n).  `stream` is a hipStream_t (NULL: the
 * null stream); the call only enqueues kernels and returns.  labels and
 * out_labels must be different arrays.  Scratch (first-occurrence and
 * new-id tables, scan partials) is a file-scope device array of about
 * 32 MiB that each device holds until the process ends: n is limited to
 * PCD_NOALLOC_MAX_VERTICES (PCD_E_INVALID above).  Every call on one device
 * shares that array, so calls that overlap in time on one device race.
 * No hipMalloc, hipFree, hipMemset or allocating library routine is called. */
#define PCD_NOALLOC_MAX_VERTICES 4194304
int pcd_compress_labels_noalloc(int n, const int32_t *labels, int32_t *out_labels, int32_t *num_labels,
                                void *stream);

/* framework.cpp:103-143 collect_statistics: sizes, int_volume, ext_volume
 * for the current (compressed) labels.  k = max label + 1.  Bit-exact. */
int pcd_collect_statistics(pcd_ctx *ctx, int *num_labels);
/* Copies the statistics to host arrays of k entries each (any may be NULL). */
int pcd_stats_download(pcd_ctx *ctx, int32_t *sizes, int32_t *int_volume, int32_t *ext_volume);

/* Rule::calculate (rules.hpp:11; rules.cpp:9-22, 58-70, 121-129, 150-235)
 * on the current labels and statistics.  Modularity and CPM use exact
 * integer sums of the community totals; WCC and conductance sum in FP64. */
int pcd_metric(pcd_ctx *ctx, pcd_rule rule, double *score);

/* ------------------------------------------------------------- refinement */
/* One label-propagation step, PropagationRefinement::refine
 * (refinement.cpp:114-182): each vertex is active with probability frac
 * (counter hash of (seed, step, v) against frac*RAND_MAX); an active vertex
 * joins the neighbour community of best score_join if best + score_leave > 0
 * and > score_leave, else leaves to a fresh singleton if score_leave > 0, else
 * stays.  All reads see the labels and statistics before the step.  The new
 * labels replace the current ones (NEW_LABEL for leavers); call
 * pcd_compress_labels before the next statistics.  Needs current statistics. */
int pcd_propagation_step(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie);

/* One parallel-merging step, MergingRefinement::refine
 * (refinement.cpp:185-288): per round every unmatched community picks the
 * neighbour community of best score_merge (>= 0), matched pairs are found by
 * handshake, and matched communities take min(l, partner).  Community-pair
 * link counts come from the contraction below.  Needs current statistics. */
int pcd_merging_step(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie);

/* Contraction: the community graph of the current labels as the list of
 * (l, k, links) for l != k, sorted by (l, k) -- the table MergingRefinement
 * builds at refinement.cpp:228-238 and framework-cuda.cu:744-807 computes by
 * sort + reduce-by-key.  *num_pairs receives U; arrays (may be NULL) need U
 * entries; call first with NULL arrays to get U. */
int pcd_contract(pcd_ctx *ctx, int64_t *num_pairs, int32_t *l, int32_t *k, int32_t *links);

/* Graph::count_triangles (graph.cpp:107-175): per-vertex tri (twice the
 * triangles at v), tri_deg (incident edges on a triangle) and the mean local
 * clustering coefficient cc.  Results stay on the device for WCC; host
 * arrays (n entries) and cc may be NULL. */
int pcd_count_triangles(pcd_ctx *ctx, int32_t *tri, int32_t *tri_deg, double *cc);
/* Overrides the clustering coefficient the WCC gains use (default: the one
 * computed by pcd_count_triangles). */
int pcd_set_clustering_coef(pcd_ctx *ctx, double cc);

/* ------------------------------------------------------------ measurement */
/* Per-phase device time (HIP events on the context stream) of the most
 * recent call of each phase, in milliseconds. */
#define PCD_PHASE_COMPRESS 0
#define PCD_PHASE_STATS 1
#define PCD_PHASE_METRIC 2
#define PCD_PHASE_PROPAGATE 3
#define PCD_PHASE_MERGE 4
#define PCD_PHASE_TRIANGLES 5
#define PCD_PHASE_COUNT 6
int pcd_enable_timing(pcd_ctx *ctx, int enable);
int pcd_phase_time(pcd_ctx *ctx, int phase, double *ms);
/* Work counters of the propagation steps since the last
 * pcd_bin_counters_reset, summed over degree bins: active vertices, their
 * edge entries E_A, and distinct neighbour labels D (for the
 * algorithmic-bytes model).  Counting costs extra atomics; enable first. */
int pcd_enable_counters(pcd_ctx *ctx, int enable);
int pcd_propagation_counters(pcd_ctx *ctx, int64_t *active, int64_t *entries, int64_t *distinct);

/* Per-kernel timers: with timing enabled, every launch of the hot kernels
 * is bracketed by HIP events on the context stream.  pcd_kernel_timer sums
 * the durations (ms) and counts the launches of one kernel slot since the
 * last reset.  Slots are numbered 0 .. pcd_kernel_slot_count()-1. */
int pcd_kernel_slot_count(void);
const char *pcd_kernel_slot_name(int slot);
int pcd_kernel_timer_reset(pcd_ctx *ctx);
int pcd_kernel_timer(pcd_ctx *ctx, int slot, double *total_ms, int64_t *launches);
/* Restricts the timers to the slots whose bit is set in mask (all by
 * default): each timed launch adds two events to the stream. */
int pcd_kernel_timer_select(pcd_ctx *ctx, uint32_t mask);
/* Work counters of the propagation step per degree bin (summed since the
 * last reset): active vertices, their edge entries and distinct neighbour
 * labels.  Bins are numbered 0 .. pcd_bin_count()-1 (hub first). */
int pcd_bin_count(void);
const char *pcd_bin_name(int bin);
int pcd_bin_counters_reset(pcd_ctx *ctx);
int pcd_bin_counters(pcd_ctx *ctx, int bin, int64_t *active, int64_t *entries, int64_t *distinct);
/* Tunables of the kernel schedule: "hub_table_slots" (LDS slots of a hub
 * partition table, multiple of 64 in [64, 4096]; takes effect at once),
 * "hub_part_entries" (target entries per hub label partition) and
 * "stat_chunk" (entries per statistics work item), which take effect at the
 * next pcd_graph_upload, "small_aux", "block_aux" and "hub_aux" (0/1: the
 * small-degree, per-vertex table and hub kernels gather {label, community
 * total} records instead of labels, for modularity and CPM) and "hub_nt"
 * (512 or 256: hub partitions are counted by 512-thread workgroups with
 * 4096-slot tables or 256-thread ones with 2048), "small_fused" (0/1: the
 * small-degree bins in one launch), "small_occ8" (0/1: that launch built
 * for 8 waves per SIMD, modularity and CPM), "b1024_u16" and "b4096_u16"
 * (0/1: the degree 257-512 and 1025-4096 tables keep 16-bit counts),
 * "bin_streams" (1, the
 * default: a propagation step runs the hub rows and degrees 1025-4096 on two
 * side streams of the context, beside the other bins; 0: one stream) and
 * "bin_streams_min_edges" (graphs with fewer edges keep one stream; default
 * 2^19), "onesweep" (0/1: radix sorts by one launch per pass with decoupled
 * look-back instead of three) and "os_spin_limit" (polls a look-back waits
 * before it gives up and the sort is finished by three-launch passes;
 * default 2^20; 0 makes the second pass of every sort of more than one tile
 * give up, for tests of that fallback) and "scan_tiles_per_cu" (scans use
 * shorter tiles until there are this many per CU; default 8, in [1, 64]).
 * Results do not depend on them. */
int pcd_set_tuning(pcd_ctx *ctx, const char *key, int value);
/* Hub label partitions whose distinct labels overflowed their LDS table
 * since the last call (those hubs were recounted in global memory); the
 * counter is reset by the call. */
int pcd_hub_overflows(pcd_ctx *ctx, int64_t *count);
/* One-sweep radix sorts (tunable "onesweep") whose look-back gave up
 * waiting, so that three-launch passes finished them, since the last call;
 * the counter is reset by the call. */
int pcd_sort_fallbacks(pcd_ctx *ctx, int64_t *count);
/* Streaming device-to-device copy of `bytes` (16-byte accesses), `reps`
 * times: *gbs receives the attained read+write bandwidth in GB/s. */
int pcd_copy_probe(pcd_ctx *ctx, int64_t bytes, int reps, double *gbs);
/* Random 4-byte gathers over the uploaded graph's own adjacency, `reps`
 * times: every entry's neighbour id (coalesced) and its label, with
 * levels = 2 that label's label as well (a dependent gather by community id,
 * as the gains read community totals).  *gathers_per_s receives the rate
 * (E x levels gathers per rep): the gather roof the propagation kernels run
 * against, with no hashing or scoring between the gathers.  Needs labels;
 * labels outside [0, n) are not followed.  No counterpart in the reference:
 * a measurement probe, like pcd_copy_probe. */
int pcd_gather_probe(pcd_ctx *ctx, int levels, int reps, double *gathers_per_s);
/* Vertices per degree bin of the uploaded graph (bin layout). */
int pcd_bin_size(pcd_ctx *ctx, int bin, int64_t *vertices, int64_t *entries);

/* ---------------------------------------------------------------- batches */
/* Many independent graphs refined together: one round runs one iteration of
 * the main.cpp:226-266 loop (compress, statistics, metric, stop rule,
 * propagation with -s 0) for every running graph, as one captured hipGraph
 * of persistent kernels over (graph, position-range) work items.  Graphs
 * are admitted into a device arena between rounds and keep their state
 * there until released.  Each graph's labels, statistics and scores equal
 * those of the same graph run alone through a context (same job
 * parameters), whatever else is in the batch.  No reference counterpart:
 * the reference runs one graph per process. */
typedef struct pcd_batch pcd_batch;

/* Per-graph options: the ./main flags of one run (main.cpp:55-120). */
typedef struct pcd_job {
    pcd_rule rule;    /* -m: MODULARITY or CPM */
    double frac;      /* -f (0.5) */
    double threshold; /* -t (0) */
    int max_iter;     /* -n (-1: stop on the threshold rule) */
    int tie;          /* PCD_TIE_* */
    uint64_t seed;    /* counter-hash seed of selection and ties */
    int iter_cap;     /* safety bound on iterations (<= 0: 1000) */
} pcd_job;

#define PCD_BATCH_RUNNING 0
#define PCD_BATCH_DONE 1
#define PCD_BATCH_FREE 2
/* admitted by pcd_batch_add_many_async, uploads not yet joined */
#define PCD_BATCH_PENDING 3
/* admitted by pcd_batch_add_many_async, rows failed the device check
 * (PCD_E_FORMAT); the slot holds no arena space until released */
#define PCD_BATCH_REJECTED 4

/* One pass of the main.cpp:226-266 loop body (main.cpp:229-260) on the
 * context's graph and labels: compress, statistics, metric (*score), the stop
 * rule of -n / -t against prev_score (*done), and, unless done, one
 * propagation (-s 0; -s 2 on even it) or merging step numbered `it`.  *k
 * receives the communities after the compression.  The propagation kernels
 * are queued behind the metric's before the score is read back, and their
 * result is dropped when the stop rule ends the loop, so the labels and
 * statistics on return are those of a loop that never ran the step. */
int pcd_iterate(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, double prev_score, double *score,
                int *k, int *done);

/* Passes it, it+1, ..., it+count-1 of the main.cpp:226-266 loop, each as
 * pcd_iterate would run it, stopping after the pass at which the stop rule
 * (main.cpp:241-242, -n, the iteration cap) fires.  prev_score is the score
 * of pass it-1 (0 for it == 0).  *passes receives the number of passes run,
 * *score and *num_labels those of the last pass run, *done whether the rule
 * fired.  Once the statistics follow the labels (after the first pass on
 * labels from outside), the passes run device-resident: one propagation pass
 * and one merging pass are captured as hipGraphs (stop rule, pass number and
 * step number on the device) and replayed in the order of the schedule
 * (-s 0 propagates, -s 1 merges, -s 2 alternates from propagation), and the
 * host reads a small state record every few passes instead of a score per
 * pass.  Every rule (WCC once the graph's triangles are counted).  The labels
 * and statistics of the last pass are then left in the compressed form
 * pcd_iterate's next compression would give (the same partition).  With the
 * tunables graph_loop 0, inc_stats 0 or (for -s 1/2) merge_sort 1, every
 * pass runs as pcd_iterate. */
int pcd_iterate_many(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, int count,
                     double prev_score, int *passes, double *score, int *num_labels, int *done);

/* The whole main.cpp:226-266 loop on the context's graph and current labels
 * (-s strategy 0/1/2, -r rounds): on return the context holds the final
 * compressed labels and statistics; *score receives the last score (what
 * main.cpp:273 prints), *iterations the final `it`, *k the communities. */
int pcd_refine(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, double *score, int *iterations, int *k);

/* `max_graphs` slots and an arena of `arena_bytes` device bytes. */
int pcd_batch_create(int device, int max_graphs, int64_t arena_bytes, pcd_batch **out);
int pcd_batch_destroy(pcd_batch *b);
/* Arena bytes a graph of n vertices and m edges occupies (upper bound). */
int pcd_batch_footprint(int n, int m, int64_t *bytes);
/* Admits a graph (CSR as pcd_graph_upload, degrees <= 4096) with its
 * initial labels (in [0,n) or PCD_NEW_LABEL).  PCD_E_NOMEM when no slot or
 * arena space is free: retry after releasing finished graphs. */
int pcd_batch_add(pcd_batch *b, int n, int m, const int32_t *indices, const int32_t *adj, const int32_t *labels,
                  const pcd_job *job, int *slot);
/* One graph of pcd_batch_add_many: the arguments of pcd_batch_add. */
typedef struct pcd_graph_ref {
    int n, m;
    const int32_t *indices, *adj, *labels;
} pcd_graph_ref;
/* Admits `count` graphs (graphs[i] with jobs[i]) in one call: the host
 * checks run on several threads, arrays in page-locked memory
 * (pcd_host_alloc) are read by the device directly, the adjacency rows are
 * checked on the device after the upload (a graph that fails is released
 * and reports PCD_E_FORMAT), and the stream is synchronised once.
 * slots[i] receives graph i's slot or -1; codes[i] (optional) its
 * pcd_batch_add return code.  Returns PCD_OK when every graph was admitted,
 * else the code of the first one that was not (pcd_last_error names it);
 * the others stay admitted.  Stands for one pcd_batch_add per graph, that
 * is, one reference main.cpp run per graph (main.cpp:170-225). */
int pcd_batch_add_many(pcd_batch *b, int count, const pcd_graph_ref *graphs, const pcd_job *jobs, int *slots,
                       int *codes);
/* pcd_batch_add_many without the wait, so that admission overlaps the
 * rounds of graphs already running.  The host checks run and the slots are
 * given out before the call returns (slots[i], codes[i] and the return code
 * as pcd_batch_add_many's for everything checked on the host); the uploads
 * and the device row check are queued on the batch's admission stream, 256
 * graphs to a chunk, and the graphs are PCD_BATCH_PENDING.  A later
 * pcd_batch_step joins every chunk whose uploads are done (waiting for the
 * first when no graph is running): a joined graph runs from the next round
 * on (PCD_BATCH_RUNNING), or is PCD_BATCH_REJECTED when its rows failed the
 * device check.  The arrays of every pending graph must stay valid and
 * unchanged until it has left PCD_BATCH_PENDING.  Labels (and the degree
 * order the host computes) from pageable memory are staged in page-locked
 * memory by the call; indices and adj are read by the device directly when
 * they are page-locked (pcd_host_alloc), else copied on the admission
 * stream.  A second call first joins the chunks of the previous one. */
int pcd_batch_add_many_async(pcd_batch *b, int count, const pcd_graph_ref *graphs, const pcd_job *jobs, int *slots,
                             int *codes);
/* Runs `rounds` rounds; *running receives the number of graphs still
 * running afterwards, pending ones (pcd_batch_add_many_async) included. */
int pcd_batch_step(pcd_batch *b, int rounds, int *running);
/* State after the last step: PCD_BATCH_*, iterations (main.cpp's final
 * `it`), communities and the last score (the value main.cpp:273 prints). */
int pcd_batch_status(pcd_batch *b, int slot, int *status, int *iterations, int *k, double *score);
/* Slots that finished and are not yet released (*count may exceed cap). */
int pcd_batch_finished(pcd_batch *b, int *slots, int cap, int *count);
/* Final compressed labels / community statistics of a finished graph. */
int pcd_batch_labels(pcd_batch *b, int slot, int32_t *labels);
/* Labels of `count` finished graphs in one copy, concatenated in the order
 * of `slots` (graph i's n labels follow graph i-1's). */
int pcd_batch_labels_many(pcd_batch *b, int count, const int *slots, int32_t *labels);
int pcd_batch_stats(pcd_batch *b, int slot, int32_t *sizes, int32_t *int_volume, int32_t *ext_volume);
/* Frees a slot and its arena space (a running graph is abandoned; a
 * pending one is waited for first). */
int pcd_batch_release(pcd_batch *b, int slot);
int pcd_batch_info(pcd_batch *b, int64_t *free_bytes, int *graphs, int *running, int64_t *rounds);
/* Bin maps: which propagation kernel serves which degree range. */
int pcd_batch_map_count(void);
const char *pcd_batch_map_name(int map);
int pcd_batch_set_map(pcd_batch *b, int map);
/* Times the propagation phase under every bin map on the current state
 * (`reps` runs each; needs one completed round; changes no results) and
 * keeps the fastest; ms_per_map (pcd_batch_map_count entries) may be NULL. */
int pcd_batch_autotune(pcd_batch *b, int reps, int *chosen, double *ms_per_map);

/* ---------------------------------------------------------------- serving */
/* A request loop with continuous batching: a scheduler thread admits
 * requests into a batch (pcd_batch_*) as slots and arena bytes free up, runs
 * one round at a time for everything resident and retires finished graphs;
 * requests the batch does not take (WCC, initial partition 1, degrees above
 * 4096, more than batch_max_edges edges, graphs larger than the arena) run on
 * worker threads with their own
 * contexts and streams, admitted against plain_bytes.  Each result equals
 * pcd_refine (-s 0) of the same request alone.  No reference counterpart. */
typedef struct pcd_server pcd_server;
typedef struct pcd_server_config {
    int device;
    int64_t batch_arena_bytes; /* device arena of the batch: its admission budget */
    int batch_slots;           /* graphs resident in the batch at once */
    int plain_workers;         /* worker threads, one context and stream each */
    int64_t plain_bytes;       /* device-byte budget of the workers together */
    int64_t batch_max_edges;   /* graphs with more edges go to the workers (0: no limit) */
} pcd_server_config;

int pcd_server_create(const pcd_server_config *cfg, pcd_server **out);
/* Stops the server after the admitted work; unfinished requests are dropped. */
int pcd_server_destroy(pcd_server *s);
/* Queues a request.  indices/adj are borrowed until its result is taken;
 * labels (n entries, copied) are the initial partition when initial == 0;
 * initial == 1 builds create_clustering_partition (framework.cpp:27-59). */
int pcd_server_submit(pcd_server *s, int n, int m, const int32_t *indices, const int32_t *adj, const int32_t *labels,
                      const pcd_job *job, int initial, int64_t *id);
/* Ids of finished requests (blocks up to timeout_ms for the first one). */
int pcd_server_wait(pcd_server *s, int timeout_ms, int64_t *ids, int cap, int *count);
/* Result of a finished request (returns its error code if it failed) and
 * forgets it: labels (n entries), last score, iterations, communities,
 * latency from submission to completion, route (0 batch, 1 worker). */
int pcd_server_result(pcd_server *s, int64_t id, int32_t *labels, double *score, int *iterations, int *k,
                      double *latency_ms, int *route);
int pcd_server_info(pcd_server *s, int *queued, int *running_batch, int *running_plain, int64_t *rounds);

/* --------------------------------------------------------- host utilities */
/* Page-locked host memory (hipHostMalloc): graph arrays placed here upload
 * at full link rate.  Free with pcd_host_free. */
int pcd_host_alloc(int64_t bytes, void **p);
int pcd_host_free(void *p);
/* graph.cpp:40-105 Graph::read: reads a PCD file into malloc'ed arrays
 * (free with pcd_free).  Same validation: magic, sizes, no trailing bytes. */
int pcd_read_pcd(const char *path, int *n, int *m, int32_t **indices, int32_t **adj, int32_t **map);
/* convert.cpp:244-278 write_edges (file bytes only). */
int pcd_write_pcd(const char *path, int n, int m, const int32_t *indices, const int32_t *adj,
                  const int32_t *map);
/* convert.cpp:14-74 read_edges: text edge list -> arrays (free with pcd_free). */
int pcd_read_edge_text(const char *path, int64_t *num_edges, int32_t **a, int32_t **b, int *num_nodes);
/* convert.cpp:76-196: edge list -> degree-sorted, loop- and duplicate-free
 * CSR identical to what `convert` writes.  Outputs malloc'ed.  Self-loop
 * input edges are dropped (convert.cpp:94-103 leaves an uninitialised slot
 * for them, so the reference output is undefined there). */
int pcd_convert_edges(int num_nodes, int64_t num_edges, const int32_t *a, const int32_t *b, int *n, int *m,
                      int32_t **indices, int32_t **adj, int32_t **map);
void pcd_free(void *p);

/* ------------------------------------------------------------ graph sets */
/* Many PCD graphs in one file, for batching and serving.  Each member holds
 * the exact bytes of one PCD file (graph.cpp:40-105 reads them,
 * convert.cpp:244-278 writes them).  Layout: "PCDSET1\0", int64 count, count
 * x {int64 offset, int64 bytes}, then the members in table order, each
 * starting 11 bytes before a 64-byte boundary (its arrays 4-byte aligned, its
 * indices 64-byte aligned in a mapping), zero bytes between them, nothing
 * after the last.  The layout is canonical: the same graphs always give the
 * same bytes.  No reference counterpart (the reference reads one graph per
 * run, main.cpp:170-225). */
typedef struct pcd_graphset pcd_graphset;
typedef struct pcd_graphset_writer pcd_graphset_writer;
/* Maps the file read-only and checks its table and each member's header and
 * length (PCD_E_FORMAT otherwise). */
int pcd_graphset_open(const char *path, pcd_graphset **out);
/* Unmaps the file: pointers from pcd_graphset_graph become invalid. */
int pcd_graphset_close(pcd_graphset *s);
int pcd_graphset_count(const pcd_graphset *s, int *count);
/* Graph i as pointers into the mapping (valid until pcd_graphset_close; any
 * output may be NULL).  The first call for a member makes pcd_read_pcd's
 * structural checks on its arrays; a member that fails them reports
 * PCD_E_FORMAT on every call. */
int pcd_graphset_graph(pcd_graphset *s, int i, int *n, int *m, const int32_t **indices, const int32_t **adj,
                       const int32_t **map);
/* Writes member i as a PCD file: the same bytes the set holds for it. */
int pcd_graphset_extract(pcd_graphset *s, int i, const char *path);
/* Streaming writer: open with the number of graphs, add them in order (map
 * may be NULL: vertex v keeps id v; a graph that fails the structural checks
 * is refused with PCD_E_INVALID), close.  Closing after fewer than `count`
 * graphs removes the file and returns PCD_E_STATE. */
int pcd_graphset_writer_open(const char *path, int count, pcd_graphset_writer **out);
int pcd_graphset_writer_add(pcd_graphset_writer *w, int n, int m, const int32_t *indices, const int32_t *adj,
                            const int32_t *map);
int pcd_graphset_writer_close(pcd_graphset_writer *w);
/* Packs PCD files (read with pcd_read_pcd's checks) into a set, in order;
 * on an error no file is left behind. */
int pcd_graphset_pack(const char *path, int count, const char *const *pcd_paths);

/* framework.cpp:14-24 create_singleton_partition: identity permutation
 * shuffled with glibc rand() (the caller's generator state, as there). */
int pcd_singleton_partition(int n, int32_t *labels);

/* The serving-form workload: `count` requests, each an SBM or R-MAT graph
 * (1/2 each) with a log-uniform target edge count in [m_lo, m_hi] and a
 * metric (modularity 0.6, CPM 0.01 0.3, WCC with initial partition 1 0.1),
 * built in parallel as pcd_gen_sbm_batch builds its graphs.  params holds
 * {family, target m, kind, lambda, initial, 0} per request. */
int pcd_gen_requests(uint64_t seed, int count, int64_t m_lo, int64_t m_hi, int32_t **n, int32_t **m,
                     int64_t **ind_off, int64_t **adj_off, int32_t **indices, int32_t **adj, double **params);

/* framework.cpp:27-59 create_clustering_partition on the device: the
 * context's labels become the triangle-based initial partition (`-i 1`),
 * equal to the reference's serial greedy; triangles are counted first if
 * needed.  *num_communities receives the number of communities. */
int pcd_clustering_labels(pcd_ctx *ctx, int *num_communities);

/* framework.cpp:27-59 create_clustering_partition from per-vertex triangle
 * counts tri (as pcd_count_triangles returns them), on the host. */
int pcd_clustering_partition(int n, const int32_t *indices, const int32_t *adj, const int32_t *tri,
                             int32_t *labels);

/* Normalized mutual information of two labelings of the n vertices
 * (common.cpp:165-173 compare_communities(g, a, b) with
 * common.cpp:75-127 norm_mutual_info; labels >= 0).  As in the reference
 * the count product n_i * n_j is int32 arithmetic: past 2^31 it wraps, and
 * the result can exceed 1 or be NaN; both labelings with one community
 * give 0/0 = NaN. */
int pcd_nmi(int n, const int32_t *a, const int32_t *b, double *nmi);
/* The same against a ground-truth file (common.cpp:129-163
 * compare_communities(g, labels, filename)): each line that is not empty and
 * does not start with '#' lists the vertices of one community; N stays n. */
int pcd_nmi_file(int n, const int32_t *labels, const char *path, double *nmi);

/* Seeded synthetic graphs (edge lists, loop-free; duplicates possible).
 * SplitMix64 streams; same seed -> same edges on any host. */
int pcd_gen_sbm(uint64_t seed, int num_blocks, const int32_t *block_sizes, double p_in, double p_out,
                int64_t *num_edges, int32_t **a, int32_t **b, int *num_nodes);
int pcd_gen_rmat(uint64_t seed, int scale, int64_t num_draws, double pa, double pb, double pc,
                 int64_t *num_edges, int32_t **a, int32_t **b, int *num_nodes);
int pcd_gen_chung_lu(uint64_t seed, int n, double exponent, int dmin, int dmax, int num_comms, int cmin,
                     int cmax, double boost, int64_t *num_edges, int32_t **a, int32_t **b, int *num_nodes);

/* The batched-form workload: `count` SBM graphs built in parallel (graph i
 * from streams of (seed, i): log-uniform n, uniform block count, p_in,
 * p_out, metric and CPM resolution), each converted like pcd_convert_edges.
 * Concatenated CSR arrays with per-graph offsets; params holds 6 doubles per
 * graph {nominal n, blocks, p_in, p_out, kind, lambda}.  Free with pcd_free. */
int pcd_gen_sbm_batch(uint64_t seed, int count, int nmin, int nmax, int bmin, int bmax, double pin_lo, double pin_hi,
                      double pout_lo, double pout_hi, double lam_lo, double lam_hi, int32_t **n, int32_t **m,
                      int64_t **ind_off, int64_t **adj_off, int32_t **indices, int32_t **adj, double **params);

#ifdef __cplusplus
}
#endif
#endif /* PARCD_H */
