// The refinement loop of main.cpp:226-266 kept on the device.
//
// One pass (compress, metric, stop rule, then a propagation step or a
// merging step, with the statistics update) is captured once as a hipGraph
// per step kind and replayed in the order of the schedule: -s 0 propagates
// every pass, -s 1 merges every pass, -s 2 alternates, starting with
// propagation (main.cpp:247-258); every rule, WCC once the graph's
// triangles are counted.  The pass number, the previous score and the
// stop rule (main.cpp:241-242, -n, the iteration cap) live in a device
// record; the captured step reads its step number from it, so every
// replay is the next pass.  When the rule fires, or the requested number of
// passes is reached, the pass freezes the record and keeps a copy of the
// compressed labels and the renumbered statistics; from then on the step
// number is -1, so further replays select no vertex and change nothing.
// The host launches a few replays at a time and reads the record once per
// batch of replays, never once per pass; at the end it puts the kept copy
// back.  The result equals pcd_iterate called pass by pass.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "../host/hostutil.hpp"
#include "prims.h"

namespace pcd {

struct LoopState {
    int it;        // pass the next replay runs (main.cpp's it)
    int step;      // step number the captured propagation reads: it, or -1 once frozen
    int it_end;    // first pass not requested
    int frozen;    // stopped: done, or it_end reached
    int done;      // the stop rule fired
    int just;      // frozen in the current pass: the snapshot copies
    int k_freeze;  // communities after the frozen pass's compression
    int cur_k;     // communities of the current pass
    int prev_k;    // communities of the last completed pass
    int pad;
    double prev;        // score of the last completed pass
    double cur;         // score of the current pass
    double score_done;  // score of the pass at which the rule fired
};

struct LoopParams {
    int max_iter, cap;
    double threshold;
};

__global__ void k_loop_rule(LoopState *s, const double *score, const int32_t *dk, LoopParams p) {
    if (s->frozen) return;
    const int it = s->it;
    const int k = *dk;
    if (it >= s->it_end) {
        s->frozen = 1;
        s->just = 1;
        s->k_freeze = k;
        s->step = -1;
        return;
    }
    const double sc = *score;
    const bool known = (p.max_iter != -1 && it >= p.max_iter) || it >= p.cap;
    const bool d = known || (p.max_iter == -1 && s->prev > 0 && sc > 0 && sc / s->prev - 1 <= p.threshold);
    s->cur = sc;
    s->cur_k = k;
    if (d) {
        s->frozen = 1;
        s->done = 1;
        s->just = 1;
        s->k_freeze = k;
        s->score_done = sc;
        s->step = -1;
    } else {
        s->step = it;
    }
}

__global__ void k_loop_snapshot(const LoopState *s, int n, const int32_t *labels, const CommStats *stats,
                                int32_t *snap_labels, CommStats *snap_stats) {
    if (!s->just) return;
    const int k = s->k_freeze;
    const int stride = gridDim.x * blockDim.x;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) snap_labels[v] = labels[v];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) snap_stats[c] = stats[c];
}

__global__ void k_loop_next(LoopState *s) {
    if (!s->frozen) {
        s->prev = s->cur;
        s->prev_k = s->cur_k;
        s->it++;
    }
    s->just = 0;
}

enum PassKind { PASS_PROPAGATE = 0, PASS_MERGE = 1 };

// step kind of pass `it` under schedule `strategy` (main.cpp:247-258)
static int pass_kind(int strategy, int it) {
    if (strategy == 0) return PASS_PROPAGATE;
    if (strategy == 1) return PASS_MERGE;
    return it % 2 == 0 ? PASS_PROPAGATE : PASS_MERGE;
}

// The captured passes, their state and the buffers they were captured with.
struct Loop {
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // by PassKind
    LoopState *d_state = nullptr, *h_state = nullptr;
    int32_t *snap_labels = nullptr;
    CommStats *snap_stats = nullptr;
    // what the capture depends on
    const void *ptrs[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t alloc_gen = -1;
    int64_t tune_gen = -1;
    pcd_rule rule{};
    double frac = 0, threshold = 0;
    uint64_t seed = 0;
    int tie = 0, max_iter = 0, cap = 0, rounds = 0;
    bool broken = false;  // capture failed once: this context runs pass by pass
};

static Loop *loop_of(pcd_ctx *ctx) { return (Loop *)ctx->loop; }

void loop_release(pcd_ctx *ctx) {
    Loop *L = loop_of(ctx);
    if (!L) return;
    for (hipGraphExec_t &e : L->exec)
        if (e) (void)hipGraphExecDestroy(e);
    if (L->h_state) (void)hipHostFree(L->h_state);
    // device buffers belong to the context's allocation list
    delete L;
    ctx->loop = nullptr;
}

// whether the captured passes of `kinds` (bit per PassKind) were captured
// for these buffers, tuning, job and merging rounds
static bool same_key(const Loop *L, pcd_ctx *ctx, const pcd_job *job, int cap, int rounds, unsigned kinds) {
    const void *p[4] = {ctx->d_labels, ctx->d_labels_alt, ctx->d_stats, ctx->d_stats_alt};
    for (int i = 0; i < 4; i++)
        if (p[i] != L->ptrs[i]) return false;
    for (int k = 0; k < 2; k++)
        if (((kinds >> k) & 1) && !L->exec[k]) return false;
    return L->alloc_gen == ctx->alloc_gen && L->tune_gen == ctx->tune_gen && L->rule.kind == job->rule.kind &&
           L->rule.lambda == job->rule.lambda && L->frac == job->frac && L->threshold == job->threshold &&
           L->seed == job->seed && L->tie == job->tie && L->max_iter == job->max_iter && L->cap == cap &&
           (!((kinds >> PASS_MERGE) & 1) || L->rounds == rounds);
}

// Captures one pass of step kind `kind`.  Nothing runs: the host
// bookkeeping is put back as it was after the capture.  Any failure leaves
// no graph behind.
static int capture_pass(pcd_ctx *ctx, Loop *L, const pcd_job *job, int cap, int kind, int rounds) {
    hipGraphExec_t &exec = L->exec[kind];
    if (exec) {
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    // buffers the pass may grow lazily are sized first (no allocation may
    // happen while the stream captures)
    PCD_TRY(ensure_scan_parts(ctx, (int64_t)ctx->n + ctx->num_stat_items + 64));
    if (kind == PASS_MERGE) PCD_TRY(merge_rows_reserve(ctx));
    if (!L->d_state) PCD_TRY(dev_alloc_t(ctx, &L->d_state, 1));
    if (!L->h_state && hipHostMalloc((void **)&L->h_state, sizeof(LoopState), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        L->h_state = nullptr;
        return fail(PCD_E_NOMEM, "pinned loop state");
    }
    if (!L->snap_labels) PCD_TRY(dev_alloc_t(ctx, &L->snap_labels, ctx->n + 1));
    if (!L->snap_stats) PCD_TRY(dev_alloc_t(ctx, &L->snap_stats, ctx->n + 1));
    // host bookkeeping as it stands (a pass changes it while capturing)
    const bool s_timing = ctx->timing, s_counters = ctx->counters, s_comp = ctx->labels_compressed,
               s_has = ctx->has_stats, s_pend = ctx->stats_pending, s_new = ctx->labels_have_new,
               s_kdef = ctx->k_deferred;
    const int s_bound = ctx->label_bound, s_k = ctx->k;
    const void *p0[4] = {ctx->d_labels, ctx->d_labels_alt, ctx->d_stats, ctx->d_stats_alt};
    ctx->timing = false;
    ctx->counters = false;
    ctx->label_bound = ctx->n;
    const pcd_rule rule = job->rule;
    LoopParams lp{job->max_iter, cap, job->threshold};
    PCD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    ctx->graph_step = &L->d_state->step;
    int rc = compress_labels_dev(ctx, true);
    if (rc == PCD_OK) rc = metric_enqueue(ctx, rule);
    if (rc == PCD_OK) {
        hipLaunchKernelGGL(k_loop_rule, dim3(1), dim3(1), 0, ctx->stream, L->d_state, ctx->d_result, ctx->d_k, lp);
        hipLaunchKernelGGL(k_loop_snapshot, dim3(grid_for(ctx, ctx->n + 1, 256, 8)), dim3(256), 0, ctx->stream,
                           L->d_state, ctx->n, ctx->d_labels, ctx->d_stats, L->snap_labels, L->snap_stats);
        if (kind == PASS_PROPAGATE) {
            rc = propagation_launch(ctx, rule, job->frac, job->seed, 0, job->tie);
            if (rc == PCD_OK) rc = propagation_commit(ctx);
        } else {
            rc = merging_dev(ctx, rule, rounds, job->seed, 0, job->tie);
        }
    }
    if (rc == PCD_OK) hipLaunchKernelGGL(k_loop_next, dim3(1), dim3(1), 0, ctx->stream, L->d_state);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    ctx->graph_step = nullptr;
    const bool same_ptrs = p0[0] == ctx->d_labels && p0[1] == ctx->d_labels_alt && p0[2] == ctx->d_stats &&
                           p0[3] == ctx->d_stats_alt;
    ctx->timing = s_timing;
    ctx->counters = s_counters;
    ctx->labels_compressed = s_comp;
    ctx->has_stats = s_has;
    ctx->stats_pending = s_pend;
    ctx->labels_have_new = s_new;
    ctx->k_deferred = s_kdef;
    ctx->label_bound = s_bound;
    ctx->k = s_k;
    (void)hipGetLastError();
    if (rc == PCD_OK && (e != hipSuccess || !g)) rc = fail(PCD_E_HIP, std::string("loop capture: ") + hipGetErrorString(e));
    if (rc == PCD_OK && !same_ptrs) rc = fail(PCD_E_STATE, "loop capture: the pass does not return its buffers");
    if (rc == PCD_OK && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        exec = nullptr;
        rc = fail(PCD_E_HIP, "loop graph instantiation");
    }
    if (g) (void)hipGraphDestroy(g);
    if (rc != PCD_OK) return rc;
    for (int i = 0; i < 4; i++) L->ptrs[i] = p0[i];
    L->alloc_gen = ctx->alloc_gen;
    L->tune_gen = ctx->tune_gen;
    L->rule = rule;
    L->frac = job->frac;
    L->threshold = job->threshold;
    L->seed = job->seed;
    L->tie = job->tie;
    L->max_iter = job->max_iter;
    L->cap = cap;
    if (kind == PASS_MERGE) L->rounds = rounds;
    return PCD_OK;
}

// Passes [it, it + count) on the device; see pcd_iterate_many.
static int device_passes(pcd_ctx *ctx, Loop *L, const pcd_job *job, int strategy, int cap, int it, int count,
                         double prev_score, int *passes, double *score, int *num_labels, int *done) {
    LoopState &h = *L->h_state;
    std::memset(&h, 0, sizeof(h));
    h.it = it;
    h.step = it;
    h.it_end = it + count;
    h.prev = prev_score;
    h.prev_k = ctx->k;
    PCD_HIP(hipMemcpyAsync(L->d_state, &h, sizeof(LoopState), hipMemcpyHostToDevice, ctx->stream));
    // replays per read of the record: about half a millisecond of passes
    // (one pass at a time for large graphs), never beyond the first pass
    // that can freeze
    int per = 1;
    double pass_ms = -1.0;
    for (;;) {
        const int left = h.it_end - h.it + 1;  // the pass at it_end freezes
        const int c = std::max(1, std::min(per, left));
        const auto t0 = std::chrono::steady_clock::now();
        // replay r of this batch is pass h.it + r (once frozen, a replay
        // of either kind changes nothing)
        for (int r = 0; r < c; r++) PCD_HIP(hipGraphLaunch(L->exec[pass_kind(strategy, h.it + r)], ctx->stream));
        PCD_HIP(hipMemcpyAsync(&h, L->d_state, sizeof(LoopState), hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        if (h.frozen) break;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / c;
        pass_ms = pass_ms < 0 ? ms : std::min(pass_ms, ms);
        per = std::max(1, std::min(32, (int)(0.5 / std::max(pass_ms, 1e-3))));
    }
    // the labels and statistics of the frozen pass back in place
    PCD_HIP(hipMemcpyAsync(ctx->d_labels, L->snap_labels, sizeof(int32_t) * (size_t)ctx->n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    if (h.k_freeze > 0)
        PCD_HIP(hipMemcpyAsync(ctx->d_stats, L->snap_stats, sizeof(CommStats) * (size_t)h.k_freeze,
                               hipMemcpyDeviceToDevice, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->k = h.k_freeze;
    ctx->label_bound = h.k_freeze;
    ctx->k_deferred = false;
    ctx->labels_compressed = true;
    ctx->has_stats = true;
    ctx->stats_pending = false;
    ctx->labels_have_new = false;
    if (h.done) {
        *passes = h.it - it + 1;
        *score = h.score_done;
        *num_labels = h.k_freeze;
        *done = 1;
    } else {
        *passes = h.it - it;
        *score = h.prev;
        *num_labels = h.prev_k;
        *done = 0;
    }
    return PCD_OK;
}

}  // namespace pcd

using namespace pcd;

extern "C" int pcd_iterate_many(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, int count,
                                double prev_score, int *passes, double *score, int *num_labels, int *done) {
    if (!ctx || !job || !passes || !score || !num_labels || !done) return fail(PCD_E_INVALID, "null argument");
    if (count <= 0) return fail(PCD_E_INVALID, "count must be positive");
    if (strategy < 0 || strategy > 2) return fail(PCD_E_INVALID, "strategy must be 0, 1 or 2");
    if (job->rule.kind < PCD_RULE_MODULARITY || job->rule.kind > PCD_RULE_CONDUCTANCE)
        return fail(PCD_E_INVALID, "unknown rule kind " + std::to_string(job->rule.kind));
    if (strategy != 1 && !(job->frac > 0.0)) return fail(PCD_E_INVALID, "frac must be > 0");
    if (strategy != 0 && rounds <= 0) return fail(PCD_E_INVALID, "rounds must be positive");
    if (job->tie != PCD_TIE_RANDOM && job->tie != PCD_TIE_LOWEST) return fail(PCD_E_INVALID, "unknown tie order");
    if (!ctx->has_graph) return fail(PCD_E_STATE, "no graph uploaded");
    if (!ctx->has_labels) return fail(PCD_E_STATE, "no labels uploaded");
    if (it < 0 || (int64_t)it + count > INT32_MAX / 2) return fail(PCD_E_INVALID, "pass range out of range");
    *passes = 0;
    *done = 0;
    *score = prev_score;
    *num_labels = 0;
    const int cap = job->iter_cap > 0 ? job->iter_cap : 1000;
    double prev = prev_score;
    int run = 0;
    // pass by pass until the statistics follow the labels (the first pass
    // of labels from outside counts them in full), or for the rules and
    // strategies the captured pass does not cover
    // captured step kinds the schedule replays (bit per PassKind); merging
    // is captured in its community-row form
    const unsigned kinds = strategy == 0 ? 1u : strategy == 1 ? 2u : 3u;
    for (;;) {
        const bool device_ok = (strategy == 0 || !ctx->tune.merge_sort) &&
                               (job->rule.kind != PCD_RULE_WCC || ctx->has_triangles) &&
                               ctx->tune.graph_loop && ctx->tune.inc_stats && ctx->ev_metric && ctx->has_labels &&
                               (ctx->stats_pending || ctx->has_stats) &&
                               2 * (int64_t)ctx->m >= ctx->tune.graph_loop_min_entries &&
                               !(loop_of(ctx) && loop_of(ctx)->broken) && count - run >= 2;
        if (run >= count || device_ok) break;
        int d = 0, k = 0;
        double s = 0.0;
        PCD_TRY(pcd_iterate(ctx, job, strategy, rounds, it + run, prev, &s, &k, &d));
        run++;
        prev = s;
        *passes = run;
        *score = s;
        *num_labels = k;
        if (d) {
            *done = 1;
            return PCD_OK;
        }
    }
    if (run >= count) return PCD_OK;
    PCD_HIP(hipSetDevice(ctx->device));
    if (!ctx->loop) ctx->loop = new Loop();
    Loop *L = loop_of(ctx);
    if (!same_key(L, ctx, job, cap, rounds, kinds)) {
        // everything a captured merging step uses is allocated first, so
        // that neither capture allocates
        int rc = (kinds & 2u) ? merge_rows_reserve(ctx) : PCD_OK;
        for (int kd = 0; kd < 2 && rc == PCD_OK; kd++)
            if ((kinds >> kd) & 1u) rc = capture_pass(ctx, L, job, cap, kd, rounds);
        if (rc != PCD_OK) {
            // this context runs pass by pass from now on
            L->broken = true;
            int p2 = 0, d2 = 0, k2 = 0;
            double s2 = 0.0;
            PCD_TRY(pcd_iterate_many(ctx, job, strategy, rounds, it + run, count - run, prev, &p2, &s2, &k2, &d2));
            *passes = run + p2;
            *score = p2 ? s2 : *score;
            *num_labels = p2 ? k2 : *num_labels;
            *done = d2;
            return PCD_OK;
        }
    }
    int p2 = 0, d2 = 0, k2 = 0;
    double s2 = 0.0;
    PCD_TRY(device_passes(ctx, L, job, strategy, cap, it + run, count - run, prev, &p2, &s2, &k2, &d2));
    if (kinds & 2u) PCD_TRY(merge_rows_check(ctx));
    *passes = run + p2;
    if (p2) {
        *score = s2;
        *num_labels = k2;
    }
    *done = d2;
    return PCD_OK;
}
