// Work sources of the propagation kernels.
//
// A kernel asks its source for units of work: each unit is a range of
// vertex positions of one graph together with that graph's propagation
// arguments.  RangeSrc serves one degree bin of a single graph (the plain
// path); ItemSrc serves the work items of a batch of graphs, each a
// position range inside one graph's bin, and skips graphs that are no
// longer running.  Kernels loop over units with a fixed grid (persistent
// form), so a batch's item lists can change without relaunch geometry.
#ifndef PARCD_WORK_H
#define PARCD_WORK_H

#include "prims.h"

namespace pcd {

struct LPArgs {
    const int32_t *indices, *adj, *labels, *order;
    const int2 *lv;  // per vertex {label, aux}
    const CommStats *stats;
    const int32_t *tot;  // per community the one total the gains read (aux), or nullptr: read stats
    int32_t *new_labels;
    RuleParams rule;
    uint64_t seed;
    int step;
    int tie;
    int f;        // selection threshold (refinement.cpp:116)
    int use_aux;  // the rule's gains read only aux
    unsigned long long *counters;  // {active, entries, distinct} of the bin, or nullptr
    const int *step_ptr;  // the step number in device memory (replayed graphs), or nullptr: `step`
    const int32_t *found;  // merging rows: per community "matched" (refinement.cpp:206), else unused
};

__device__ inline int vertex_at(const LPArgs &a, int p) { return a.order ? a.order[p] : p; }

// the step number of the selection hash and the tie keys
__device__ inline int step_of(const LPArgs &a) { return a.step_ptr ? *a.step_ptr : a.step; }

// refinement.cpp:129: the vertex takes part in this step.  A replayed graph
// whose loop has stopped sets its device step number to -1: no vertex does.
__device__ inline bool lp_selected(const LPArgs &a, int v) {
    if (a.step_ptr) {
        const int st = *a.step_ptr;
        return st >= 0 && selected(a.seed, st, v, a.f);
    }
    return selected(a.seed, a.step, v, a.f);
}

// positions [p0, p1) of one bin in units of `per` positions
struct RangeSrc {
    int p0, p1, per;
    LPArgs a;
    __device__ int units() const { return (p1 - p0 + per - 1) / per; }
    __device__ bool unit(int u, LPArgs &out, int &q0, int &q1) const {
        q0 = p0 + u * per;
        q1 = min(p1, q0 + per);
        out = a;
        return true;
    }
};

// the same with the range [bounds[0], bounds[1]) in device memory (bins laid
// out by kernels, for launches that are captured or not waited on)
struct DevRangeSrc {
    const int *bounds;
    int per;
    LPArgs a;
    __device__ int units() const { return (bounds[1] - bounds[0] + per - 1) / per; }
    __device__ bool unit(int u, LPArgs &out, int &q0, int &q1) const {
        q0 = bounds[0] + u * per;
        q1 = min(bounds[1], q0 + per);
        out = a;
        return true;
    }
};

// ------------------------------------------------------------ batches
// One work item: positions [p0, p1) of graph g (bin ranges, vertex tiles
// or statistics rows, depending on the list it is in).
struct WorkItem {
    int g, p0, p1;
};

// BG_PENDING and BG_REJECTED are host-side slot states of asynchronous
// admission; a device state record never holds them (it stays BG_FREE
// until the graph joins).
enum BatchStatus { BG_RUNNING = 0, BG_DONE = 1, BG_FREE = 2, BG_PENDING = 3, BG_REJECTED = 4 };

// Static description of one graph of a batch (written at admission).
// Labels live in two buffers with fixed roles: propagation reads `cur`
// (lp.labels) and writes `next` (lp.new_labels); compression reads `next`
// and writes `cur`.
struct BatchGraph {
    LPArgs lp;
    int32_t *next;        // == lp.new_labels
    int32_t *cur;         // == lp.labels
    CommStats *stats;     // == lp.stats
    CommStats *stats_alt; // compression renumbers the statistics into it
    int32_t *first, *newid;
    int n, max_iter;      // max_iter -1: stop on the score threshold
    double threshold;
    int iter_cap, pad;
};

// Dynamic state of one graph (device resident, polled by the host).
struct BatchState {
    int status;       // BatchStatus
    int iter;         // completed iterations (main.cpp's it)
    int step;         // step number of this round's propagation
    int k;            // communities after this round's compression
    int label_bound;  // labels in `next` are < label_bound or NEW_LABEL
    int capped;       // stopped by iter_cap rather than the stop rule
    int ready;        // stats follow the labels: each round updates them from the moved vertices
    int pad;
    double score, prev_score;
    unsigned long long acc[3];
};
static_assert(sizeof(BatchState) == 72, "batch state record");

struct ItemSrc {
    const WorkItem *items;
    const int *count;
    const BatchGraph *graphs;
    const BatchState *state;
    __device__ int units() const { return *count; }
    __device__ bool unit(int u, LPArgs &out, int &q0, int &q1) const {
        const WorkItem w = items[u];
        if (state[w.g].status != BG_RUNNING) return false;
        out = graphs[w.g].lp;
        out.step = state[w.g].step;
        q0 = w.p0;
        q1 = w.p1;
        return true;
    }
};

// Propagation kernel variants a batch's bin map can route a degree range to.
enum LPVariant {
    LPV_S4 = 0, LPV_S8, LPV_S16, LPV_S32, LPV_S64,  // G lanes per vertex
    LPV_W512, LPV_W1024,                            // one wave per vertex, table slots
    LPV_B256, LPV_B512,                             // workgroup per vertex, 2048 slots
    LPV_B512X8192,                                  // workgroup per vertex, 8192 slots
    NUM_LPV
};
// largest degree each variant handles (its table holds 2 x degree slots)
constexpr int LPV_MAX_DEG[NUM_LPV] = {4, 8, 16, 32, 64, 256, 512, 1024, 1024, 4096};
// positions per work item of each variant
constexpr int LPV_ITEM[NUM_LPV] = {64, 64, 32, 16, 8, 16, 8, 4, 4, 2};

struct BatchLPLaunch {
    const WorkItem *items[NUM_LPV];
    const int *count[NUM_LPV];
    bool used[NUM_LPV];
    const BatchGraph *graphs;
    const BatchState *state;
    int cu_count;
};

// launches the propagation kernels of a batch round (propagate.hip)
int batch_propagation_launch(hipStream_t stream, const BatchLPLaunch &l);

// the graph's vertices as rows of the propagation kernels (propagate.hip)
RowSet graph_rows(pcd_ctx *ctx);
// One merging round over community rows (propagate.hip): a.indices are the
// rows' offsets into a.adj (labels), a.order the bin order of `rows`,
// a.new_labels the partner array, a.found the matched flags.
int merge_rows_launch(pcd_ctx *ctx, const RowSet &rows, const LPArgs &a);

}  // namespace pcd

#endif
