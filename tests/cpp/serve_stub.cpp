// The serving loop (par-cd_amd/host/serve.cpp) on the CPU, against a stub of
// the device API it calls (pcd_ctx_*, pcd_graph_upload, pcd_labels_*,
// pcd_clustering_labels, pcd_refine, pcd_batch_*).  Built with
// -fsanitize=thread, or address+undefined, it checks the server's own host
// logic: locking, request lifetimes, the scheduler / worker hand-offs and
// shutdown.  The stub enforces the threading contract of the C ABI (a
// context and a batch are used by one thread at a time) and deep-copies
// nothing the caller lends: it reads the borrowed arrays when it computes.
//
// Each stub "refinement" is a deterministic function of the request, so a
// request's result is the same in the batch, on a worker and alone; the
// driver checks that for every request of a mixed, staggered stream.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "parcd.h"

int pcd_stub_fail(int code, const char *msg);

namespace {

std::atomic<int> g_violations{0};

void violation(const char *what) {
    fprintf(stderr, "contract violation: %s\n", what);
    g_violations++;
}

// exclusive use by one thread at a time
struct Owner {
    std::atomic<int> busy{0};
    void enter(const char *what) {
        if (busy.exchange(1)) violation(what);
    }
    void leave() { busy = 0; }
};

uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// the stub's refinement: labels, score, iterations and k from the inputs
struct Result {
    std::vector<int32_t> labels;
    double score = 0;
    int iters = 0, k = 0;
};

Result stub_refine(int n, int m, const int32_t *ind, const int32_t *adj, const int32_t *lab0, const pcd_job &job) {
    uint64_t h = mix(job.seed * 31 + (uint64_t)job.rule.kind * 7 + (uint64_t)n);
    for (int v = 0; v <= n; v++) h = mix(h ^ (uint64_t)ind[v]);
    for (int64_t j = 0; j < 2 * (int64_t)m; j += 1 + (2 * (int64_t)m) / 64) h = mix(h ^ (uint64_t)adj[j]);
    for (int v = 0; v < n; v++) h = mix(h ^ (uint64_t)(uint32_t)lab0[v]);
    Result r;
    r.iters = 1 + (int)(h % 7);
    const int groups = std::max(1, n / 5);
    r.labels.resize(n);
    std::vector<char> seen(groups, 0);
    for (int v = 0; v < n; v++) {
        const int g = (int)(mix(h + (uint64_t)v) % (uint64_t)groups);
        r.labels[v] = g;
        if (!seen[g]) {
            seen[g] = 1;
            r.k++;
        }
    }
    r.score = (double)(h % 100000) / 100000.0;
    return r;
}

}  // namespace

// ------------------------------------------------------------ context stub
struct pcd_ctx {
    Owner own;
    int n = 0, m = 0;
    const int32_t *ind = nullptr, *adj = nullptr;  // borrowed for the duration of a call only
    std::vector<int32_t> ind_copy, adj_copy;       // "device" copies
    std::vector<int32_t> labels;
    bool has_graph = false, has_labels = false;
};

extern "C" int pcd_ctx_create(int device, pcd_ctx **out) {
    if (device != 0) return pcd_stub_fail(PCD_E_INVALID, "device");
    *out = new pcd_ctx();
    return PCD_OK;
}

extern "C" int pcd_ctx_destroy(pcd_ctx *ctx) {
    if (!ctx) return PCD_OK;
    ctx->own.enter("context destroyed while in use");
    ctx->own.leave();
    delete ctx;
    return PCD_OK;
}

extern "C" int pcd_graph_upload(pcd_ctx *ctx, int n, int m, const int32_t *indices, const int32_t *adj) {
    ctx->own.enter("context used by two threads");
    ctx->n = n;
    ctx->m = m;
    ctx->ind_copy.assign(indices, indices + n + 1);
    ctx->adj_copy.assign(adj, adj + 2 * (int64_t)m);
    ctx->has_graph = true;
    ctx->has_labels = false;
    std::this_thread::sleep_for(std::chrono::microseconds(50));
    ctx->own.leave();
    return PCD_OK;
}

extern "C" int pcd_labels_upload(pcd_ctx *ctx, const int32_t *labels) {
    ctx->own.enter("context used by two threads");
    if (!ctx->has_graph) violation("labels before graph");
    ctx->labels.assign(labels, labels + ctx->n);
    ctx->has_labels = true;
    ctx->own.leave();
    return PCD_OK;
}

extern "C" int pcd_clustering_labels(pcd_ctx *ctx, int *num) {
    ctx->own.enter("context used by two threads");
    ctx->labels.resize(ctx->n);
    for (int v = 0; v < ctx->n; v++) ctx->labels[v] = v / 3;
    ctx->has_labels = true;
    if (num) *num = (ctx->n + 2) / 3;
    ctx->own.leave();
    return PCD_OK;
}

extern "C" int pcd_refine(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, double *score, int *iters,
                          int *k) {
    ctx->own.enter("context used by two threads");
    if (!ctx->has_labels) violation("refine without labels");
    if (strategy != 0 || rounds != 1) violation("serving runs -s 0");
    Result r = stub_refine(ctx->n, ctx->m, ctx->ind_copy.data(), ctx->adj_copy.data(), ctx->labels.data(), *job);
    std::this_thread::sleep_for(std::chrono::microseconds(200 * r.iters));
    ctx->labels = r.labels;
    *score = r.score;
    *iters = r.iters;
    *k = r.k;
    ctx->own.leave();
    return PCD_OK;
}

extern "C" int pcd_labels_download(pcd_ctx *ctx, int32_t *labels) {
    ctx->own.enter("context used by two threads");
    std::memcpy(labels, ctx->labels.data(), sizeof(int32_t) * ctx->n);
    ctx->own.leave();
    return PCD_OK;
}

// -------------------------------------------------------------- batch stub
struct StubSlot {
    bool used = false;
    int status = 0;  // 1 running, 2 done
    int64_t bytes = 0;
    int rounds_left = 0;
    Result res;
};

struct pcd_batch {
    Owner own;
    int64_t arena = 0, used_bytes = 0;
    std::vector<StubSlot> slots;
};

extern "C" int pcd_batch_footprint(int n, int m, int64_t *bytes) {
    *bytes = 64 * ((int64_t)n + 1) + 8 * (int64_t)m + 4096;
    return PCD_OK;
}

extern "C" int pcd_batch_create(int device, int max_graphs, int64_t arena_bytes, pcd_batch **out) {
    if (device != 0 || max_graphs <= 0) return pcd_stub_fail(PCD_E_INVALID, "batch arguments");
    pcd_batch *b = new pcd_batch();
    b->arena = arena_bytes;
    b->slots.resize(max_graphs);
    *out = b;
    return PCD_OK;
}

extern "C" int pcd_batch_destroy(pcd_batch *b) {
    if (!b) return PCD_OK;
    b->own.enter("batch destroyed while in use");
    b->own.leave();
    delete b;
    return PCD_OK;
}

extern "C" int pcd_batch_add(pcd_batch *b, int n, int m, const int32_t *indices, const int32_t *adj,
                             const int32_t *labels, const pcd_job *job, int *slot) {
    b->own.enter("batch used by two threads");
    int rc = PCD_OK;
    for (int v = 0; v < n && rc == PCD_OK; v++)
        if (labels[v] < -1 || labels[v] >= n) rc = pcd_stub_fail(PCD_E_INVALID, "labels must be in [0,n)");
    int64_t fp = 0;
    pcd_batch_footprint(n, m, &fp);
    int s = -1;
    if (rc == PCD_OK) {
        for (int i = 0; i < (int)b->slots.size(); i++)
            if (!b->slots[i].used) {
                s = i;
                break;
            }
        if (s < 0 || b->used_bytes + fp > b->arena) rc = pcd_stub_fail(PCD_E_NOMEM, "batch full");
    }
    if (rc == PCD_OK) {
        StubSlot &sl = b->slots[s];
        sl = StubSlot();
        sl.used = true;
        sl.status = 1;
        sl.bytes = fp;
        sl.res = stub_refine(n, m, indices, adj, labels, *job);  // reads the borrowed arrays now
        sl.rounds_left = sl.res.iters;
        b->used_bytes += fp;
        *slot = s;
    }
    b->own.leave();
    return rc;
}

extern "C" int pcd_batch_step(pcd_batch *b, int rounds, int *running) {
    b->own.enter("batch used by two threads");
    int r = 0;
    std::this_thread::sleep_for(std::chrono::microseconds(150));
    for (auto &sl : b->slots) {
        if (!sl.used || sl.status != 1) continue;
        sl.rounds_left -= rounds;
        if (sl.rounds_left <= 0) sl.status = 2;
        else r++;
    }
    if (running) *running = r;
    b->own.leave();
    return PCD_OK;
}

extern "C" int pcd_batch_finished(pcd_batch *b, int *slots, int cap, int *count) {
    b->own.enter("batch used by two threads");
    int c = 0;
    for (int s = 0; s < (int)b->slots.size(); s++)
        if (b->slots[s].used && b->slots[s].status == 2) {
            if (c < cap) slots[c] = s;
            c++;
        }
    *count = c;
    b->own.leave();
    return PCD_OK;
}

extern "C" int pcd_batch_status(pcd_batch *b, int slot, int *status, int *iters, int *k, double *score) {
    b->own.enter("batch used by two threads");
    const StubSlot &sl = b->slots.at(slot);
    if (status) *status = sl.status;
    if (iters) *iters = sl.res.iters;
    if (k) *k = sl.res.k;
    if (score) *score = sl.res.score;
    b->own.leave();
    return PCD_OK;
}

extern "C" int pcd_batch_labels(pcd_batch *b, int slot, int32_t *labels) {
    b->own.enter("batch used by two threads");
    const StubSlot &sl = b->slots.at(slot);
    std::memcpy(labels, sl.res.labels.data(), sizeof(int32_t) * sl.res.labels.size());
    b->own.leave();
    return PCD_OK;
}

extern "C" int pcd_batch_release(pcd_batch *b, int slot) {
    b->own.enter("batch used by two threads");
    StubSlot &sl = b->slots.at(slot);
    if (!sl.used) violation("release of a free slot");
    b->used_bytes -= sl.bytes;
    sl = StubSlot();
    b->own.leave();
    return PCD_OK;
}

// ------------------------------------------------------------------ driver
namespace {

struct Graph {
    int n, m;
    std::vector<int32_t> ind, adj, lab;
};

// a random simple graph with sorted rows
Graph make_graph(std::mt19937_64 &rng, int n, int extra) {
    std::vector<std::vector<int32_t>> rows(n);
    for (int v = 1; v < n; v++) {
        int u = (int)(rng() % v);
        rows[v].push_back(u);
        rows[u].push_back(v);
    }
    for (int e = 0; e < extra; e++) {
        int a = (int)(rng() % n), b = (int)(rng() % n);
        if (a == b) continue;
        rows[a].push_back(b);
        rows[b].push_back(a);
    }
    Graph g;
    g.n = n;
    g.ind.assign(n + 1, 0);
    for (int v = 0; v < n; v++) {
        std::sort(rows[v].begin(), rows[v].end());
        rows[v].erase(std::unique(rows[v].begin(), rows[v].end()), rows[v].end());
        g.ind[v + 1] = g.ind[v] + (int)rows[v].size();
        g.adj.insert(g.adj.end(), rows[v].begin(), rows[v].end());
    }
    g.m = g.ind[n] / 2;
    g.lab.resize(n);
    for (int v = 0; v < n; v++) g.lab[v] = (int)(rng() % n);
    return g;
}

}  // namespace

int main(int argc, char **argv) {
    const int count = argc > 1 ? atoi(argv[1]) : 400;
    std::mt19937_64 rng(5);
    std::vector<Graph> gs;
    std::vector<pcd_job> jobs;
    std::vector<int> initial;
    for (int i = 0; i < count; i++) {
        const int n = 20 + (int)(rng() % (i % 9 == 0 ? 3000 : 300));
        gs.push_back(make_graph(rng, n, n * (1 + (int)(rng() % 4))));
        pcd_job j;
        std::memset(&j, 0, sizeof(j));
        const int r = (int)(rng() % 10);
        j.rule.kind = r < 6 ? PCD_RULE_MODULARITY : r < 9 ? PCD_RULE_CPM : PCD_RULE_WCC;
        j.rule.lambda = j.rule.kind == PCD_RULE_CPM ? 0.01 : 0.0;
        j.frac = 0.5;
        j.max_iter = -1;
        j.seed = (uint64_t)i;
        jobs.push_back(j);
        initial.push_back(j.rule.kind == PCD_RULE_WCC && i % 2 == 0 ? 1 : 0);
    }
    // expected: every request alone
    std::vector<Result> want(count);
    for (int i = 0; i < count; i++) {
        std::vector<int32_t> l0 = gs[i].lab;
        if (initial[i]) {
            l0.resize(gs[i].n);
            for (int v = 0; v < gs[i].n; v++) l0[v] = v / 3;
        }
        want[i] = stub_refine(gs[i].n, gs[i].m, gs[i].ind.data(), gs[i].adj.data(), l0.data(), jobs[i]);
    }
    pcd_server_config cfg;
    cfg.device = 0;
    cfg.batch_arena_bytes = 256 << 10;  // small: admission waits for room
    cfg.batch_slots = 6;
    cfg.plain_workers = 3;
    cfg.plain_bytes = 64 << 20;
    cfg.batch_max_edges = 1500;         // larger graphs go to the workers
    pcd_server *srv = nullptr;
    if (pcd_server_create(&cfg, &srv) != PCD_OK) {
        fprintf(stderr, "create failed: %s\n", pcd_last_error());
        return 1;
    }
    std::vector<int64_t> ids(count, -1);
    std::mutex ids_mu;
    // two submitting threads, staggered
    auto submit = [&](int parity) {
        for (int i = parity; i < count; i += 2) {
            int64_t id = -1;
            int rc = pcd_server_submit(srv, gs[i].n, gs[i].m, gs[i].ind.data(), gs[i].adj.data(),
                                       initial[i] ? nullptr : gs[i].lab.data(), &jobs[i], initial[i], &id);
            if (rc != PCD_OK) {
                fprintf(stderr, "submit %d failed: %s\n", i, pcd_last_error());
                violation("submit");
            }
            {
                std::lock_guard<std::mutex> lk(ids_mu);
                ids[i] = id;
            }
            if (i % 7 == 0) std::this_thread::sleep_for(std::chrono::microseconds(300));
        }
    };
    std::thread t0(submit, 0), t1(submit, 1);
    t0.join();
    t1.join();
    std::vector<int> index_of_id(count + 1, -1);
    for (int i = 0; i < count; i++) index_of_id[ids[i]] = i;
    int got = 0, bad = 0, routed[2] = {0, 0};
    auto t_start = std::chrono::steady_clock::now();
    while (got < count) {
        int64_t done[64];
        int c = 0;
        pcd_server_wait(srv, 50, done, 64, &c);
        for (int j = 0; j < c; j++) {
            const int i = index_of_id[done[j]];
            std::vector<int32_t> lab(gs[i].n);
            double score = 0, lat = 0;
            int it = 0, k = 0, route = -1;
            int rc = pcd_server_result(srv, done[j], lab.data(), &score, &it, &k, &lat, &route);
            got++;
            if (rc != PCD_OK || lab != want[i].labels || score != want[i].score || it != want[i].iters ||
                k != want[i].k) {
                fprintf(stderr, "request %d differs from its run alone (rc %d)\n", i, rc);
                bad++;
            }
            if (route == 0 || route == 1) routed[route]++;
        }
        if (std::chrono::steady_clock::now() - t_start > std::chrono::seconds(120)) {
            fprintf(stderr, "timeout with %d of %d results\n", got, count);
            return 1;
        }
    }
    // an invalid request reports its error through the result
    {
        Graph &g = gs[1];
        std::vector<int32_t> badlab = g.lab;
        badlab[0] = g.n + 5;
        int64_t id = -1;
        pcd_job j = jobs[1];
        j.rule.kind = PCD_RULE_MODULARITY;
        if (pcd_server_submit(srv, g.n, g.m, g.ind.data(), g.adj.data(), badlab.data(), &j, 0, &id) != PCD_OK)
            violation("submit of the invalid request");
        int64_t d = -1;
        int c = 0;
        for (int tries = 0; c == 0 && tries < 100; tries++) pcd_server_wait(srv, 50, &d, 1, &c);
        if (c != 1 || pcd_server_result(srv, d, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == PCD_OK)
            violation("an invalid request must fail");
    }
    // shutdown with work still queued and running: the server drops it
    for (int i = 0; i < 40; i++) {
        int64_t id = -1;
        pcd_server_submit(srv, gs[i].n, gs[i].m, gs[i].ind.data(), gs[i].adj.data(),
                          initial[i] ? nullptr : gs[i].lab.data(), &jobs[i], initial[i], &id);
    }
    pcd_server_destroy(srv);
    printf("results %d, differing %d, batch %d, workers %d, contract violations %d\n", got, bad, routed[0],
           routed[1], g_violations.load());
    return (bad == 0 && g_violations == 0 && routed[0] > 0 && routed[1] > 0) ? 0 : 1;
}

#include "hostutil.hpp"
int pcd_stub_fail(int code, const char *msg) { return pcd::fail(code, msg); }
