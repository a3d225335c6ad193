// Serving loop: refinement requests on one GPU with continuous batching.
//
// A scheduler thread owns one batch (pcd_batch_*).  Every loop it admits
// queued requests into the batch while slots and arena bytes allow (the
// HBM-footprint admission), runs one round for everything resident, and
// retires the graphs that stopped: a request that arrives while others run
// joins at the next round boundary, and every graph keeps its labels and
// statistics on the device until it finishes.  Requests the batch does not
// take go to worker threads that each own a plain context and stream,
// admitted against their own device-byte budget; their kernels run
// concurrently with the batch rounds.  Those are WCC, the triangle-based
// initial partition (built on the device), degrees above 4096, graphs above the edge limit (their
// iterations would set the length of every round) and graphs larger than
// the whole arena.
//
// A request's result equals the result of running it alone (pcd_refine on a
// context with the same job): the batch gives every graph the numbering and
// counter-hash streams it has alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hostutil.hpp"
#include "parcd.h"

using namespace pcd;

namespace {

using clk = std::chrono::steady_clock;

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING, RQ_DONE, RQ_FAILED };

struct Req {
    int64_t id = 0;
    int n = 0, m = 0, maxd = 0;
    const int32_t *ind = nullptr, *adj = nullptr;
    std::vector<int32_t> labels;  // initial labels, then the result
    pcd_job job;
    int initial = 0;
    int route = 0;  // 0 batch, 1 plain context
    int64_t footprint = 0;
    clk::time_point t_submit, t_start, t_done;
    int state = RQ_QUEUED;
    int err = PCD_OK;
    std::string msg;
    double score = 0.0;
    int iters = 0, k = 0;
};

// device bytes a plain context holds for a graph: CSR, labels, statistics,
// scan and hub scratch, triangle arrays (an upper bound, not exact)
int64_t plain_footprint(int n, int m) { return 96 * ((int64_t)n + 1) + 24 * (int64_t)m * 2 + (16 << 20); }

}  // namespace

struct pcd_server {
    pcd_server_config cfg;
    pcd_batch *batch = nullptr;
    std::mutex mu;
    std::condition_variable cv_sched, cv_plain, cv_done;
    std::deque<Req *> incoming;   // submitted, not yet seen by the scheduler
    std::deque<Req *> plain_q;    // admitted to the plain workers
    std::deque<int64_t> completed;
    std::unordered_map<int64_t, Req *> reqs;
    int64_t next_id = 1;
    bool stop = false;
    int64_t plain_bytes_in_use = 0;
    int running_batch = 0, running_plain = 0, queued = 0;
    int64_t rounds = 0;
    int fatal = PCD_OK;
    std::string fatal_msg;
    std::thread sched;
    std::vector<std::thread> workers;
};

static void finish(pcd_server *s, Req *r, int err, const std::string &msg) {
    // caller holds s->mu
    r->t_done = clk::now();
    r->err = err;
    r->msg = msg;
    r->state = err == PCD_OK ? RQ_DONE : RQ_FAILED;
    s->completed.push_back(r->id);
    s->cv_done.notify_all();
}

static void plain_worker(pcd_server *s) {
    pcd_ctx *ctx = nullptr;
    int rc0 = pcd_ctx_create(s->cfg.device, &ctx);
    for (;;) {
        Req *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_plain.wait(lk, [&] { return s->stop || !s->plain_q.empty(); });
            if (s->plain_q.empty()) break;
            r = s->plain_q.front();
            s->plain_q.pop_front();
            r->state = RQ_RUNNING;
            r->t_start = clk::now();
            s->running_plain++;
        }
        int rc = rc0;
        std::string msg = rc0 != PCD_OK ? std::string(pcd_last_error()) : std::string();
        if (rc == PCD_OK) {
            rc = pcd_graph_upload(ctx, r->n, r->m, r->ind, r->adj);
            // the triangle-based initial partition (-i 1) is built on the
            // device; WCC counts its triangles inside the first metric call
            int nc = 0;
            if (rc == PCD_OK) rc = r->initial == 1 ? pcd_clustering_labels(ctx, &nc)
                                                   : pcd_labels_upload(ctx, r->labels.data());
            if (rc == PCD_OK) rc = pcd_refine(ctx, &r->job, 0, 1, &r->score, &r->iters, &r->k);
            if (rc == PCD_OK) {
                r->labels.resize(r->n);
                rc = pcd_labels_download(ctx, r->labels.data());
            }
            if (rc != PCD_OK) msg = pcd_last_error();
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->running_plain--;
        s->plain_bytes_in_use -= r->footprint;
        finish(s, r, rc, msg);
        s->cv_sched.notify_all();
    }
    if (ctx) pcd_ctx_destroy(ctx);
}

static void scheduler(pcd_server *s) {
    std::vector<Req *> pend_batch, pend_plain;
    std::unordered_map<int, Req *> resident;
    std::vector<int> done_slots(std::max(1, s->cfg.batch_slots));
    auto fatal = [&](int rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = rc;
        s->fatal_msg = pcd_last_error();
        for (auto &kv : resident) finish(s, kv.second, rc, s->fatal_msg);
        for (Req *r : pend_batch) finish(s, r, rc, s->fatal_msg);
        resident.clear();
        pend_batch.clear();
        s->running_batch = 0;
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_sched.wait(lk, [&] {
                return s->stop || !s->incoming.empty() || !resident.empty() ||
                       (!pend_plain.empty() && s->plain_bytes_in_use == 0) || !pend_batch.empty();
            });
            if (s->stop && s->incoming.empty() && resident.empty() && pend_batch.empty() && pend_plain.empty()) break;
            while (!s->incoming.empty()) {
                Req *r = s->incoming.front();
                s->incoming.pop_front();
                (r->route == 0 ? pend_batch : pend_plain).push_back(r);
            }
            // plain admission against the device-byte budget (one at a time always fits)
            for (auto it = pend_plain.begin(); it != pend_plain.end();) {
                Req *r = *it;
                if (s->plain_bytes_in_use == 0 || s->plain_bytes_in_use + r->footprint <= s->cfg.plain_bytes) {
                    s->plain_bytes_in_use += r->footprint;
                    s->plain_q.push_back(r);
                    s->queued--;
                    it = pend_plain.erase(it);
                } else {
                    ++it;
                }
            }
            if (!s->plain_q.empty()) s->cv_plain.notify_all();
        }
        if (s->fatal != PCD_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Req *r : pend_batch) finish(s, r, s->fatal, s->fatal_msg);
            pend_batch.clear();
            continue;
        }
        // batch admission by footprint: whatever fits, in arrival order
        int admitted = 0;
        for (auto it = pend_batch.begin(); it != pend_batch.end();) {
            Req *r = *it;
            int slot = -1;
            int rc = pcd_batch_add(s->batch, r->n, r->m, r->ind, r->adj, r->labels.data(), &r->job, &slot);
            if (rc == PCD_E_NOMEM) {
                ++it;
                continue;
            }
            std::lock_guard<std::mutex> lk(s->mu);
            s->queued--;
            if (rc != PCD_OK) {
                finish(s, r, rc, pcd_last_error());
            } else {
                r->state = RQ_RUNNING;
                r->t_start = clk::now();
                resident[slot] = r;
                s->running_batch++;
                admitted++;
            }
            it = pend_batch.erase(it);
        }
        if (resident.empty()) {
            if (!pend_batch.empty()) {
                // nothing resident and still no room: the graph can never fit
                std::lock_guard<std::mutex> lk(s->mu);
                for (Req *r : pend_batch) {
                    s->queued--;
                    finish(s, r, PCD_E_NOMEM, "graph exceeds the batch arena");
                }
                pend_batch.clear();
            }
            continue;
        }
        int running = 0;
        int rc = pcd_batch_step(s->batch, 1, &running);
        if (rc != PCD_OK) {
            fatal(rc);
            continue;
        }
        int nd = 0;
        rc = pcd_batch_finished(s->batch, done_slots.data(), (int)done_slots.size(), &nd);
        if (rc != PCD_OK) {
            fatal(rc);
            continue;
        }
        for (int i = 0; i < std::min(nd, (int)done_slots.size()); i++) {
            int slot = done_slots[i];
            Req *r = resident[slot];
            int st = 0;
            rc = pcd_batch_status(s->batch, slot, &st, &r->iters, &r->k, &r->score);
            r->labels.resize(r->n);
            if (rc == PCD_OK) rc = pcd_batch_labels(s->batch, slot, r->labels.data());
            std::string msg = rc == PCD_OK ? std::string() : std::string(pcd_last_error());
            int rc2 = pcd_batch_release(s->batch, slot);
            if (rc == PCD_OK && rc2 != PCD_OK) {
                rc = rc2;
                msg = pcd_last_error();
            }
            resident.erase(slot);
            std::lock_guard<std::mutex> lk(s->mu);
            s->running_batch--;
            finish(s, r, rc, msg);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->rounds++;
    }
}

extern "C" int pcd_server_create(const pcd_server_config *cfg, pcd_server **out) {
    if (!cfg || !out) return fail(PCD_E_INVALID, "null argument");
    *out = nullptr;
    if (cfg->batch_slots <= 0 || cfg->batch_arena_bytes <= 0 || cfg->plain_workers < 0 || cfg->plain_bytes < 0)
        return fail(PCD_E_INVALID, "bad server configuration");
    pcd_server *s = new pcd_server();
    s->cfg = *cfg;
    int rc = pcd_batch_create(cfg->device, cfg->batch_slots, cfg->batch_arena_bytes, &s->batch);
    if (rc != PCD_OK) {
        delete s;
        return rc;
    }
    s->sched = std::thread(scheduler, s);
    for (int i = 0; i < cfg->plain_workers; i++) s->workers.emplace_back(plain_worker, s);
    *out = s;
    return PCD_OK;
}

extern "C" int pcd_server_destroy(pcd_server *s) {
    if (!s) return PCD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->cv_plain.notify_all();
    if (s->sched.joinable()) s->sched.join();
    for (auto &t : s->workers)
        if (t.joinable()) t.join();
    pcd_batch_destroy(s->batch);
    for (auto &kv : s->reqs) delete kv.second;
    for (Req *r : s->incoming) delete r;
    delete s;
    return PCD_OK;
}

extern "C" int pcd_server_submit(pcd_server *s, int n, int m, const int32_t *indices, const int32_t *adj,
                                 const int32_t *labels, const pcd_job *job, int initial, int64_t *id) {
    if (!s || !job || !id || n <= 0 || m < 0 || !indices || (m > 0 && !adj))
        return fail(PCD_E_INVALID, "bad request arguments");
    if (initial != 0 && initial != 1) return fail(PCD_E_INVALID, "initial partition must be 0 or 1");
    if (initial == 0 && !labels) return fail(PCD_E_INVALID, "initial labels required");
    Req *r = new Req();
    r->n = n;
    r->m = m;
    r->ind = indices;
    r->adj = adj;
    r->job = *job;
    r->initial = initial;
    if (labels) r->labels.assign(labels, labels + n);
    for (int v = 0; v < n; v++) r->maxd = std::max(r->maxd, indices[v + 1] - indices[v]);
    int64_t bfp = 0;
    pcd_batch_footprint(n, std::max(m, 1), &bfp);
    const bool batchable = (job->rule.kind == PCD_RULE_MODULARITY || job->rule.kind == PCD_RULE_CPM) &&
                           initial == 0 && r->maxd <= 4096 && m > 0 && bfp <= s->cfg.batch_arena_bytes &&
                           (s->cfg.batch_max_edges <= 0 || m <= s->cfg.batch_max_edges);
    if (!batchable && s->cfg.plain_workers == 0) {
        delete r;
        return fail(PCD_E_INVALID, "request needs a plain worker and the server has none");
    }
    r->route = batchable ? 0 : 1;
    r->footprint = batchable ? bfp : plain_footprint(n, m);
    r->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->incoming.push_back(r);
        s->queued++;
    }
    s->cv_sched.notify_all();
    *id = r->id;
    return PCD_OK;
}

extern "C" int pcd_server_wait(pcd_server *s, int timeout_ms, int64_t *ids, int cap, int *count) {
    if (!s || !count || (cap > 0 && !ids)) return fail(PCD_E_INVALID, "bad arguments");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completed.empty() && timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->completed.empty(); });
    int c = 0;
    while (c < cap && !s->completed.empty()) {
        ids[c++] = s->completed.front();
        s->completed.pop_front();
    }
    *count = c;
    return PCD_OK;
}

extern "C" int pcd_server_result(pcd_server *s, int64_t id, int32_t *labels, double *score, int *iterations, int *k,
                                 double *latency_ms, int *route) {
    if (!s) return fail(PCD_E_INVALID, "null server");
    Req *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(PCD_E_INVALID, "unknown request id");
        r = it->second;
        if (r->state != RQ_DONE && r->state != RQ_FAILED) return fail(PCD_E_STATE, "request not finished");
        s->reqs.erase(it);
    }
    int rc = r->err;
    if (rc == PCD_OK) {
        if (labels) std::memcpy(labels, r->labels.data(), sizeof(int32_t) * r->n);
        if (score) *score = r->score;
        if (iterations) *iterations = r->iters;
        if (k) *k = r->k;
    } else {
        fail(rc, r->msg);
    }
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (route) *route = r->route;
    delete r;
    return rc;
}

extern "C" int pcd_server_info(pcd_server *s, int *queued, int *running_batch, int *running_plain, int64_t *rounds) {
    if (!s) return fail(PCD_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = s->queued;
    if (running_batch) *running_batch = s->running_batch;
    if (running_plain) *running_plain = s->running_plain;
    if (rounds) *rounds = s->rounds;
    return PCD_OK;
}
