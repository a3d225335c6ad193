// C ABI entry points: context, graph and label transfer, phase dispatch.
#include <algorithm>
#include <cstring>

#include "../host/hostutil.hpp"
#include "prims.h"

using namespace pcd;

namespace pcd {

int dev_alloc(pcd_ctx *ctx, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(PCD_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
    }
    ctx->allocations.push_back(*p);
    ctx->alloc_gen++;
    return PCD_OK;
}

void dev_free(pcd_ctx *ctx, void *p) {
    if (!p) return;
    auto it = std::find(ctx->allocations.begin(), ctx->allocations.end(), p);
    if (it != ctx->allocations.end()) ctx->allocations.erase(it);
    (void)hipFree(p);
    ctx->alloc_gen++;
}

RuleParams make_rule(pcd_ctx *ctx, pcd_rule rule) {
    RuleParams r;
    r.kind = rule.kind;
    r.lambda = rule.lambda;
    set_edges(r, ctx->m);
    r.cc = ctx->cc;
    return r;
}

void phase_begin(pcd_ctx *ctx, int phase) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev0[phase], ctx->stream);
}

void phase_end(pcd_ctx *ctx, int phase) {
    if (ctx->timing) {
        (void)hipEventRecord(ctx->ev1[phase], ctx->stream);
        ctx->ev_valid[phase] = true;
    }
}

static hipEvent_t pool_get(pcd_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Events around the launches of one timer slot (only the selected slots:
// each event costs a few microseconds of stream time)
void kt_begin(pcd_ctx *ctx, int slot) {
    if (!ctx->timing || !((ctx->timer_mask >> slot) & 1u)) return;
    ctx->kt_open = pool_get(ctx);
    (void)hipEventRecord(ctx->kt_open, ctx->stream);
}

void kt_end(pcd_ctx *ctx, int slot) {
    if (!ctx->timing || !ctx->kt_open || !((ctx->timer_mask >> slot) & 1u)) return;
    hipEvent_t e1 = pool_get(ctx);
    (void)hipEventRecord(e1, ctx->stream);
    ctx->ktimes[slot].push_back(std::make_pair(ctx->kt_open, e1));
    ctx->kt_open = nullptr;
}

static const char *kslot_names[NUM_KSLOTS] = {
    "compress", "stats_items", "stats_small", "metric", "lp_records", "lp_hub_partition", "lp_hub_count",
    "lp_b4096", "lp_b1024", "lp_w256", "lp_g64", "lp_g32", "lp_g16", "lp_g8", "lp_g4",
    "merge_keys", "merge_sort", "merge_rle", "merge_rounds", "triangles", "stats_delta", "lp_small",
    "lp_b512", "lp_hub_final", "merge_stats"};

static const char *bin_names[NUM_BINS] = {"hub", "b4096", "b1024", "b512", "w256", "g64", "g32", "g16", "g8", "g4"};

}  // namespace pcd

static int check_ctx(pcd_ctx *ctx) {
    if (!ctx) return fail(PCD_E_INVALID, "null context");
    PCD_HIP(hipSetDevice(ctx->device));
    return PCD_OK;
}

static int check_graph(pcd_ctx *ctx) {
    PCD_TRY(check_ctx(ctx));
    if (!ctx->has_graph) return fail(PCD_E_STATE, "no graph uploaded");
    return PCD_OK;
}

static int check_labels(pcd_ctx *ctx) {
    PCD_TRY(check_graph(ctx));
    if (!ctx->has_labels) return fail(PCD_E_STATE, "no labels uploaded");
    return PCD_OK;
}

static int check_rule(pcd_rule rule) {
    if (rule.kind < PCD_RULE_MODULARITY || rule.kind > PCD_RULE_CONDUCTANCE)
        return fail(PCD_E_INVALID, "unknown rule kind " + std::to_string(rule.kind));
    return PCD_OK;
}

extern "C" int pcd_device_count(int *count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return PCD_OK;
}

extern "C" int pcd_ctx_create(int device, pcd_ctx **out) {
    if (!out) return fail(PCD_E_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(PCD_E_NODEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return fail(PCD_E_INVALID, "device index out of range");
    hipDeviceProp_t prop;
    PCD_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PCD_E_NODEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    PCD_HIP(hipSetDevice(device));
    pcd_ctx *ctx = new pcd_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(PCD_E_HIP, "hipStreamCreate failed");
    }
    for (int i = 0; i < PCD_PHASE_COUNT; i++) {
        (void)hipEventCreate(&ctx->ev0[i]);
        (void)hipEventCreate(&ctx->ev1[i]);
    }
    if (hipEventCreateWithFlags(&ctx->ev_metric, hipEventDisableTiming) != hipSuccess) ctx->ev_metric = nullptr;
    {
        bool ok = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 2 && ok; i++)
            ok = hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming) == hipSuccess;
        if (!ok) (void)hipGetLastError();  // without them, bins run on `stream` alone
        if (hipStreamCreateWithFlags(&ctx->mstream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->ev_mfork, hipEventDisableTiming) != hipSuccess)
            (void)hipGetLastError();  // the metric then runs on `stream`
    }
    if (hipHostMalloc((void **)&ctx->h_result, 64, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&ctx->h_small, 256, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&ctx->h_small64, 256, hipHostMallocDefault) != hipSuccess) {
        pcd_ctx_destroy(ctx);
        return fail(PCD_E_NOMEM, "pinned host allocation failed");
    }
    int rc = dev_alloc_t(ctx, &ctx->d_k, 4);
    if (rc == PCD_OK) rc = dev_alloc_t(ctx, &ctx->d_acc, 16);
    if (rc == PCD_OK) rc = dev_alloc_t(ctx, &ctx->d_dacc, 4096);
    if (rc == PCD_OK) rc = dev_alloc_t(ctx, &ctx->d_result, 8);
    if (rc == PCD_OK) rc = dev_alloc_t(ctx, &ctx->d_counts, NUM_BINS + 16);
    if (rc == PCD_OK) rc = fill_i32(ctx, ctx->d_counts, 0, NUM_BINS + 16);
    if (rc == PCD_OK) rc = dev_alloc_t(ctx, &ctx->d_bin_counters, 3 * NUM_BINS);
    if (rc == PCD_OK) rc = fill_u64(ctx, (uint64_t *)ctx->d_bin_counters, 0, 3 * NUM_BINS);
    if (rc != PCD_OK) {
        pcd_ctx_destroy(ctx);
        return rc;
    }
    ctx->dacc_cap = 4096;
    *out = ctx;
    return PCD_OK;
}

static void free_graph_state(pcd_ctx *ctx) {
    loop_release(ctx);  // its device buffers go with the others below
    merge_release(ctx);
    // everything except the per-context scalars allocated at creation
    void *keep[] = {ctx->d_k, ctx->d_acc, ctx->d_dacc, ctx->d_result, ctx->d_counts, ctx->d_bin_counters};
    std::vector<void *> all = ctx->allocations;
    for (void *p : all) {
        bool k = false;
        for (void *q : keep) k = k || (p == q);
        if (!k) dev_free(ctx, p);
    }
    ctx->d_indices = ctx->d_adj = ctx->d_order = nullptr;
    ctx->d_hub_items = nullptr;
    ctx->d_hub_item_off = nullptr;
    ctx->d_hub_part_items = ctx->d_hub_info = ctx->d_hub_part_where = nullptr;
    ctx->d_hub_poff = ctx->d_hub_buf = ctx->d_hub_over = nullptr;
    ctx->d_hub_buf_base = nullptr;
    ctx->d_hub_partial = nullptr;
    ctx->num_hub_parts = 0;
    ctx->d_hub_table_off = nullptr;
    ctx->d_hub_table_log = ctx->d_hub_keys = ctx->d_hub_cnts = nullptr;
    ctx->d_stat_items = nullptr;
    ctx->d_labels = ctx->d_labels_alt = nullptr;
    ctx->d_lv = nullptr;
    ctx->d_tot = nullptr;
    ctx->d_moved_bits = nullptr;
    ctx->d_stats = ctx->d_stats_alt = nullptr;
    ctx->d_moved = nullptr;
    ctx->d_first = ctx->d_newid = ctx->d_active = ctx->d_hub_active = nullptr;
    ctx->d_firstv = nullptr;
    ctx->d_scan_part = nullptr;
    ctx->scan_part_cap = 0;
    ctx->d_tri = ctx->d_tri_deg = ctx->d_et = ctx->d_lowpos = ctx->d_it = ctx->d_itd = nullptr;
    ctx->d_keys0 = ctx->d_keys1 = nullptr;
    ctx->keys_cap = 0;
    ctx->d_pair_links = ctx->d_pair_head = nullptr;
    ctx->d_rs_hist = nullptr;
    ctx->rs_hist_cap = 0;
    ctx->d_os = nullptr;
    ctx->os_cap = 0;
    ctx->d_bp = ctx->d_found = nullptr;
    ctx->d_best_ord = ctx->d_best_key = nullptr;
    ctx->d_mcand = nullptr;
    ctx->mcand_cap = 0;
    ctx->first_cap = 0;
    ctx->has_graph = ctx->has_labels = ctx->has_stats = ctx->has_triangles = false;
    ctx->stats_pending = false;
    ctx->labels_compressed = false;
    ctx->num_hubs = ctx->num_hub_items = ctx->num_stat_items = 0;
    ctx->cc = 0.0;
}

extern "C" int pcd_ctx_destroy(pcd_ctx *ctx) {
    if (!ctx) return PCD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2; i++)
        if (ctx->side[i]) (void)hipStreamSynchronize(ctx->side[i]);
    if (ctx->mstream) (void)hipStreamSynchronize(ctx->mstream);
    loop_release(ctx);
    merge_release(ctx);
    for (void *p : ctx->allocations) (void)hipFree(p);
    ctx->allocations.clear();
    for (int i = 0; i < PCD_PHASE_COUNT; i++) {
        if (ctx->ev0[i]) (void)hipEventDestroy(ctx->ev0[i]);
        if (ctx->ev1[i]) (void)hipEventDestroy(ctx->ev1[i]);
    }
    for (int s = 0; s < NUM_KSLOTS; s++)
        for (auto &pr : ctx->ktimes[s]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->kt_open) (void)hipEventDestroy(ctx->kt_open);
    if (ctx->ev_metric) (void)hipEventDestroy(ctx->ev_metric);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_mfork) (void)hipEventDestroy(ctx->ev_mfork);
    if (ctx->mstream) (void)hipStreamDestroy(ctx->mstream);
    for (int i = 0; i < 2; i++) {
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
    }
    if (ctx->h_result) (void)hipHostFree(ctx->h_result);
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    if (ctx->h_small64) (void)hipHostFree(ctx->h_small64);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PCD_OK;
}

extern "C" int pcd_synchronize(pcd_ctx *ctx) {
    PCD_TRY(check_ctx(ctx));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    return PCD_OK;
}

static int ceil_log2_host(int64_t x) {
    int b = 0;
    while (((int64_t)1 << b) < x) b++;
    return b;
}

namespace pcd {
// Chunks of HUB_CHUNK entries and P label partitions per HUB-bin row; the
// partitioned-entry buffer holds the rows one after another, and each row
// has a fallback global table of nextpow2(2 entries) slots.
int build_hub_layout(const std::vector<int4> &rows, int part_entries, HubLayoutHost &L) {
    const size_t nh = rows.size();
    L.chunks.clear();
    L.parts.clear();
    L.where.clear();
    L.chunk_off.clear();
    L.info.assign(nh, make_int4(0, 0, 0, 0));
    L.buf_base.assign(nh + 1, 0);
    L.toff.assign(nh + 1, 0);
    L.tlog.assign(nh + 1, 0);
    int64_t poff_total = 0;
    for (size_t h = 0; h < nh; h++) {
        const int row = rows[h].x, b = rows[h].y, e = rows[h].z;
        const int64_t d = (int64_t)e - b;
        const int P = (int)std::min<int64_t>(HUB_MAX_PARTS,
                                             std::max<int64_t>(1, (d + part_entries - 1) / part_entries));
        const int first_chunk = (int)L.chunks.size();
        for (int b0 = b; b0 < e; b0 += HUB_CHUNK) {
            L.chunks.push_back(make_int4((int)h, row, b0, (int)std::min<int64_t>(e, (int64_t)b0 + HUB_CHUNK)));
            L.chunk_off.push_back((int32_t)poff_total);
            poff_total += P + 1;
        }
        const int nch = (int)L.chunks.size() - first_chunk;
        L.info[h] = make_int4(row, P, (int)L.parts.size(), first_chunk);
        for (int q = 0; q < P; q++) {
            // a partition's chunk offsets sit P + 1 apart from the first
            // chunk's; the buffer base fits int32 (hub entries < 2^31)
            L.parts.push_back(make_int4((int)h, q, first_chunk, nch));
            L.where.push_back(make_int4(row, P, nch > 0 ? L.chunk_off[first_chunk] : 0, (int)L.buf_base[h]));
        }
        L.buf_base[h + 1] = L.buf_base[h] + d;
        L.tlog[h] = std::max(6, ceil_log2_host(2 * d));
        L.toff[h + 1] = L.toff[h] + ((int64_t)1 << L.tlog[h]);
        if (poff_total > INT32_MAX || L.buf_base[h + 1] > INT32_MAX) return fail(PCD_E_INVALID, "hub layout too large");
    }
    L.poff_total = poff_total;
    return PCD_OK;
}
}  // namespace pcd

// graph.hpp:6-40 layout; builds degree bins and static work items.
extern "C" int pcd_graph_upload(pcd_ctx *ctx, int n, int m, const int32_t *indices, const int32_t *adj) {
    PCD_TRY(check_ctx(ctx));
    if (n <= 0 || m < 0 || !indices || (m > 0 && !adj)) return fail(PCD_E_INVALID, "bad graph arguments");
    if ((int64_t)2 * m > INT32_MAX) return fail(PCD_E_INVALID, "2m exceeds int32");
    if (indices[0] != 0 || indices[n] != 2 * m) return fail(PCD_E_FORMAT, "indices[0] != 0 or indices[n] != 2m");
    // validation: monotone offsets, ids in range, rows strictly ascending
    std::vector<int> bad(host_threads() + 1, 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        int b = 0;
        for (int64_t v = lo; v < hi && !b; v++) {
            if (indices[v + 1] < indices[v]) { b = 1; break; }
            for (int32_t j = indices[v]; j < indices[v + 1]; j++) {
                int32_t u = adj[j];
                if (u < 0 || u >= n || u == v || (j > indices[v] && adj[j - 1] >= u)) { b = 1; break; }
            }
        }
        if (b) bad[0] = 1;
    }, 4096);
    if (bad[0]) return fail(PCD_E_FORMAT, "adjacency rows must be ascending, loop-free, with ids in [0,n)");
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    free_graph_state(ctx);
    ctx->n = n;
    ctx->m = m;
    const int64_t E = (int64_t)2 * m;
    // degree order
    std::vector<int32_t> deg(n);
    int maxd = 0;
    bool sorted = true;
    for (int v = 0; v < n; v++) {
        deg[v] = indices[v + 1] - indices[v];
        maxd = std::max(maxd, deg[v]);
        if (v > 0 && deg[v] > deg[v - 1]) sorted = false;
    }
    ctx->max_degree = maxd;
    std::vector<int32_t> order;
    if (!sorted) {
        // stable counting sort by degree, descending
        std::vector<int64_t> bucket((size_t)maxd + 2, 0);
        for (int v = 0; v < n; v++) bucket[maxd - deg[v] + 1]++;
        for (int d = 0; d <= maxd; d++) bucket[d + 1] += bucket[d];
        order.resize(n);
        for (int v = 0; v < n; v++) order[bucket[maxd - deg[v]]++] = v;
    }
    auto deg_at = [&](int p) { return sorted ? deg[p] : deg[order[p]]; };
    // bin boundaries by degree thresholds (positions are degree-descending);
    // bin b holds degrees in (BIN_MAX_DEG[b+1], BIN_MAX_DEG[b]]
    auto first_pos_le = [&](int t) {  // first position with degree <= t
        int lo = 0, hi = n;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (deg_at(mid) <= t) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    };
    ctx->bins.start[0] = 0;
    for (int b = 1; b < NUM_BINS; b++) ctx->bins.start[b] = first_pos_le(BIN_MAX_DEG[b]);
    ctx->bins.start[NUM_BINS] = n;
    ctx->bins.split2048 = first_pos_le(2048);
    for (int b = 0; b < NUM_BINS; b++) {
        int64_t e = 0;
        for (int p = ctx->bins.start[b]; p < ctx->bins.start[b + 1]; p++) e += deg_at(p);
        ctx->bin_entries[b] = e;
    }
    // hub layout: chunks of HUB_CHUNK entries, P label partitions per hub
    const int hub_end = ctx->bins.start[BIN_HUB + 1];
    std::vector<int4> hub_rows(hub_end);
    for (int h = 0; h < hub_end; h++) {
        const int v = sorted ? h : order[h];
        hub_rows[h] = make_int4(v, indices[v], indices[v + 1], 0);
    }
    HubLayoutHost hl;
    PCD_TRY(build_hub_layout(hub_rows, ctx->tune.hub_part_entries, hl));
    const std::vector<int4> &hub_chunks = hl.chunks, &hub_parts = hl.parts, &hub_where = hl.where, &hub_info = hl.info;
    const std::vector<int32_t> &chunk_off = hl.chunk_off, &tlog = hl.tlog;
    const std::vector<int64_t> &buf_base = hl.buf_base, &toff = hl.toff;
    const int64_t poff_total = hl.poff_total;
    ctx->num_hubs = hub_end;
    ctx->num_hub_items = (int)hub_chunks.size();
    ctx->num_hub_parts = (int)hub_parts.size();
    ctx->hub_table_slots = toff[hub_end];
    // statistics items: rows longer than 64, in chunks
    std::vector<int4> stat_items;
    const int big_end = ctx->bins.start[BIN_G64];
    for (int p = 0; p < big_end; p++) {
        int v = sorted ? p : order[p];
        for (int b = indices[v]; b < indices[v + 1]; b += ctx->tune.stat_chunk)
            stat_items.push_back(make_int4(v, b, std::min(indices[v + 1], b + ctx->tune.stat_chunk), b == indices[v]));
    }
    ctx->num_stat_items = (int)stat_items.size();

    PCD_TRY(dev_alloc_t(ctx, &ctx->d_indices, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_adj, E + 1));
    PCD_HIP(hipMemcpyAsync(ctx->d_indices, indices, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    if (E > 0) PCD_HIP(hipMemcpyAsync(ctx->d_adj, adj, sizeof(int32_t) * E, hipMemcpyHostToDevice, ctx->stream));
    if (!sorted) {
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_order, n));
        PCD_HIP(hipMemcpyAsync(ctx->d_order, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    }
    if (hub_end > 0) {
        auto up = [&](auto **dst, const auto &vec) -> int {
            PCD_TRY(dev_alloc_t(ctx, dst, vec.size()));
            PCD_HIP(hipMemcpyAsync(*dst, vec.data(), sizeof(vec[0]) * vec.size(), hipMemcpyHostToDevice, ctx->stream));
            return PCD_OK;
        };
        PCD_TRY(up(&ctx->d_hub_items, hub_chunks));
        PCD_TRY(up(&ctx->d_hub_item_off, chunk_off));
        PCD_TRY(up(&ctx->d_hub_part_items, hub_parts));
        PCD_TRY(up(&ctx->d_hub_info, hub_info));
        PCD_TRY(up(&ctx->d_hub_part_where, hub_where));
        PCD_TRY(up(&ctx->d_hub_buf_base, buf_base));
        PCD_TRY(up(&ctx->d_hub_table_off, toff));
        PCD_TRY(up(&ctx->d_hub_table_log, tlog));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_hub_poff, poff_total));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_hub_buf, 2 * buf_base[hub_end]));
        PCD_TRY(dev_alloc(ctx, &ctx->d_hub_partial, (size_t)HUB_PARTIAL_BYTES * hub_parts.size()));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_hub_over, hub_end));
        PCD_TRY(fill_i32(ctx, ctx->d_hub_over, 0, hub_end));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_hub_keys, ctx->hub_table_slots));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_hub_cnts, ctx->hub_table_slots));
        PCD_TRY(fill_i32(ctx, ctx->d_hub_keys, EMPTY_KEY, ctx->hub_table_slots));
        PCD_TRY(fill_i32(ctx, ctx->d_hub_cnts, 0, ctx->hub_table_slots));
    }
    if (!stat_items.empty()) {
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_stat_items, stat_items.size()));
        PCD_HIP(hipMemcpyAsync(ctx->d_stat_items, stat_items.data(), sizeof(int4) * stat_items.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_labels, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_labels_alt, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_lv, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_tot, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_moved_bits, (size_t)(n + WAVE - 1) / WAVE + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_stats, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_stats_alt, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_moved, 8 + (size_t)n + ctx->num_stat_items + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_active, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_first, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_newid, n + 1));
    PCD_TRY(dev_alloc_t(ctx, &ctx->d_firstv, n + 1));
    PCD_TRY(fill_i32(ctx, ctx->d_firstv, -1, n + 1));
    ctx->first_cap = n + 1;
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->has_graph = true;
    return PCD_OK;
}

extern "C" int pcd_graph_info(pcd_ctx *ctx, int *n, int *m, int *max_degree) {
    PCD_TRY(check_graph(ctx));
    if (n) *n = ctx->n;
    if (m) *m = ctx->m;
    if (max_degree) *max_degree = ctx->max_degree;
    return PCD_OK;
}

extern "C" int pcd_labels_upload(pcd_ctx *ctx, const int32_t *labels) {
    PCD_TRY(check_graph(ctx));
    if (!labels) return fail(PCD_E_INVALID, "null labels");
    int mx = -1;
    bool has_new = false;
    for (int v = 0; v < ctx->n; v++) {
        if (labels[v] < -1) return fail(PCD_E_INVALID, "labels must be >= 0 or NEW_LABEL");
        if (labels[v] == -1) has_new = true;
        mx = std::max(mx, labels[v]);
    }
    int64_t bound = (int64_t)mx + 1;
    if (bound + 1 > ctx->first_cap) {
        dev_free(ctx, ctx->d_first);
        dev_free(ctx, ctx->d_newid);
        ctx->d_first = ctx->d_newid = nullptr;
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_first, bound + 1));
        PCD_TRY(dev_alloc_t(ctx, &ctx->d_newid, bound + 1));
        ctx->first_cap = bound + 1;
    }
    PCD_HIP(hipMemcpyAsync(ctx->d_labels, labels, sizeof(int32_t) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->label_bound = (int)bound;
    ctx->has_labels = true;
    ctx->labels_compressed = false;
    ctx->labels_have_new = has_new;
    ctx->has_stats = false;
    ctx->stats_pending = false;
    return PCD_OK;
}

// framework.cpp:27-59 create_clustering_partition, on the device
extern "C" int pcd_clustering_labels(pcd_ctx *ctx, int *num_communities) {
    PCD_TRY(check_graph(ctx));
    if (!ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    int nc = 0;
    PCD_TRY(clustering_labels_dev(ctx, &nc));
    ctx->label_bound = nc;
    ctx->has_labels = true;
    ctx->labels_compressed = false;
    ctx->labels_have_new = false;
    ctx->has_stats = false;
    ctx->stats_pending = false;
    if (num_communities) *num_communities = nc;
    return PCD_OK;
}

extern "C" int pcd_labels_download(pcd_ctx *ctx, int32_t *labels) {
    PCD_TRY(check_labels(ctx));
    PCD_HIP(hipMemcpyAsync(labels, ctx->d_labels, sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    return PCD_OK;
}

extern "C" int pcd_compress_labels(pcd_ctx *ctx, int *num_labels) {
    PCD_TRY(check_labels(ctx));
    phase_begin(ctx, PCD_PHASE_COMPRESS);
    PCD_TRY(compress_labels_dev(ctx));
    phase_end(ctx, PCD_PHASE_COMPRESS);
    ctx->labels_have_new = false;
    if (num_labels) *num_labels = ctx->k;
    return PCD_OK;
}

__global__ void k_max_label(int n, const int32_t *labels, int32_t *out) {
    int mx = -1;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) mx = max(mx, labels[v]);
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d, WAVE));
    if (__lane_id() == 0) atomicMax(out, mx + 1);
}

extern "C" int pcd_collect_statistics(pcd_ctx *ctx, int *num_labels) {
    PCD_TRY(check_labels(ctx));
    if (ctx->labels_have_new) return fail(PCD_E_STATE, "labels contain NEW_LABEL; compress them first");
    phase_begin(ctx, PCD_PHASE_STATS);
    if (ctx->has_stats && ctx->labels_compressed) {
        // the last compression renumbered statistics that a propagation step
        // had updated: they are those of the current labels already
        phase_end(ctx, PCD_PHASE_STATS);
        if (num_labels) *num_labels = ctx->k;
        return PCD_OK;
    }
    if (!ctx->labels_compressed) {
        // framework.cpp:110: k = max label + 1
        PCD_TRY(fill_i32(ctx, ctx->d_k, 0, 1));
        hipLaunchKernelGGL(k_max_label, dim3(grid_for(ctx, ctx->n, 256, 4)), dim3(256), 0, ctx->stream, ctx->n,
                           ctx->d_labels, ctx->d_k);
        PCD_HIP(hipMemcpyAsync(ctx->h_small, ctx->d_k, 4, hipMemcpyDeviceToHost, ctx->stream));
        PCD_HIP(hipStreamSynchronize(ctx->stream));
        ctx->k = ctx->h_small[0];
        if (ctx->k > ctx->n + 1) {
            // stats arrays hold n+1 communities; larger label ranges need compression
            return fail(PCD_E_STATE, "labels exceed n; compress them first");
        }
    }
    PCD_TRY(collect_statistics_dev(ctx));
    phase_end(ctx, PCD_PHASE_STATS);
    if (num_labels) *num_labels = ctx->k;
    return PCD_OK;
}

__global__ void k_split_stats(int k, const CommStats *st, int32_t *s, int32_t *iv, int32_t *ev) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x) {
        CommStats x = st[c];
        s[c] = x.size;
        iv[c] = x.int_vol;
        ev[c] = x.ext_vol;
    }
}

extern "C" int pcd_stats_download(pcd_ctx *ctx, int32_t *sizes, int32_t *int_volume, int32_t *ext_volume) {
    PCD_TRY(check_labels(ctx));
    if (!ctx->has_stats) return fail(PCD_E_STATE, "no statistics; call pcd_collect_statistics");
    const int k = ctx->k;
    if (k == 0) return PCD_OK;
    std::vector<CommStats> h(k);
    PCD_HIP(hipMemcpyAsync(h.data(), ctx->d_stats, sizeof(CommStats) * k, hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < k; c++) {
        if (sizes) sizes[c] = h[c].size;
        if (int_volume) int_volume[c] = h[c].int_vol;
        if (ext_volume) ext_volume[c] = h[c].ext_vol;
    }
    return PCD_OK;
}

extern "C" int pcd_metric(pcd_ctx *ctx, pcd_rule rule, double *score) {
    PCD_TRY(check_labels(ctx));
    PCD_TRY(check_rule(rule));
    if (!ctx->has_stats) return fail(PCD_E_STATE, "no statistics; call pcd_collect_statistics");
    if (rule.kind == PCD_RULE_WCC && !ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    phase_begin(ctx, PCD_PHASE_METRIC);
    PCD_TRY(metric_dev(ctx, rule, score));
    phase_end(ctx, PCD_PHASE_METRIC);
    return PCD_OK;
}

extern "C" int pcd_propagation_step(pcd_ctx *ctx, pcd_rule rule, double frac, uint64_t seed, int step, int tie) {
    PCD_TRY(check_labels(ctx));
    PCD_TRY(check_rule(rule));
    if (!ctx->has_stats) return fail(PCD_E_STATE, "no statistics; call pcd_collect_statistics");
    if (!(frac > 0.0)) return fail(PCD_E_INVALID, "frac must be > 0");
    if (tie != PCD_TIE_RANDOM && tie != PCD_TIE_LOWEST) return fail(PCD_E_INVALID, "unknown tie order");
    if (rule.kind == PCD_RULE_WCC && !ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    phase_begin(ctx, PCD_PHASE_PROPAGATE);
    PCD_TRY(propagation_dev(ctx, rule, frac, seed, step, tie));
    phase_end(ctx, PCD_PHASE_PROPAGATE);
    ctx->labels_have_new = true;
    return PCD_OK;
}

extern "C" int pcd_merging_step(pcd_ctx *ctx, pcd_rule rule, int rounds, uint64_t seed, int step, int tie) {
    PCD_TRY(check_labels(ctx));
    PCD_TRY(check_rule(rule));
    if (!ctx->has_stats) return fail(PCD_E_STATE, "no statistics; call pcd_collect_statistics");
    if (rounds <= 0) return fail(PCD_E_INVALID, "rounds must be positive");
    if (tie != PCD_TIE_RANDOM && tie != PCD_TIE_LOWEST) return fail(PCD_E_INVALID, "unknown tie order");
    if (rule.kind == PCD_RULE_WCC && !ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    phase_begin(ctx, PCD_PHASE_MERGE);
    PCD_TRY(merging_dev(ctx, rule, rounds, seed, step, tie));
    phase_end(ctx, PCD_PHASE_MERGE);
    ctx->labels_have_new = false;
    return PCD_OK;
}

// One pass of the main.cpp:226-266 loop body.  With a propagation step due,
// its kernels are launched right after the metric's, before the score is
// back on the host; the host then waits on an event behind the metric only,
// applies the stop rule, and either commits the step (statistics update and
// label swap, queued behind the kernels) or drops it (the kernels wrote only
// the spare label buffer and scratch).  When the statistics follow the
// labels, k also stays on the device until that wait, so the GPU runs the
// iterations back to back.
static int iterate_pass(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, double prev_score,
                        double *score, int *num_labels, int *done);

extern "C" int pcd_iterate(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, double prev_score,
                           double *score, int *num_labels, int *done) {
    const int rc = iterate_pass(ctx, job, strategy, rounds, it, prev_score, score, num_labels, done);
    if (ctx && ctx->k_deferred) {  // a failure after the compression: take k before returning
        (void)hipStreamSynchronize(ctx->stream);
        resolve_k(ctx);
    }
    return rc;
}

static int iterate_pass(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, int it, double prev_score,
                        double *score, int *num_labels, int *done) {
    if (!ctx || !job) return fail(PCD_E_INVALID, "null argument");
    if (strategy < 0 || strategy > 2) return fail(PCD_E_INVALID, "strategy must be 0, 1 or 2");
    PCD_TRY(check_labels(ctx));
    PCD_TRY(check_rule(job->rule));
    const pcd_rule rule = job->rule;
    const bool prop = strategy == 0 || (strategy == 2 && it % 2 == 0);
    if (prop && !(job->frac > 0.0)) return fail(PCD_E_INVALID, "frac must be > 0");
    if (!prop && rounds <= 0) return fail(PCD_E_INVALID, "rounds must be positive");
    if (job->tie != PCD_TIE_RANDOM && job->tie != PCD_TIE_LOWEST) return fail(PCD_E_INVALID, "unknown tie order");
    const int cap = job->iter_cap > 0 ? job->iter_cap : 1000;
    if (rule.kind == PCD_RULE_WCC && !ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    // k can stay on the device when the compression renumbers statistics
    // that follow the labels (no recount) and the metric reads k there
    const bool defer = rule.kind != PCD_RULE_WCC && ctx->ev_metric && (ctx->stats_pending || ctx->has_stats);
    phase_begin(ctx, PCD_PHASE_COMPRESS);
    PCD_TRY(compress_labels_dev(ctx, defer));
    phase_end(ctx, PCD_PHASE_COMPRESS);
    ctx->labels_have_new = false;
    phase_begin(ctx, PCD_PHASE_STATS);
    if (!ctx->has_stats) PCD_TRY(collect_statistics_dev(ctx));
    phase_end(ctx, PCD_PHASE_STATS);
    // -n and the iteration cap decide without the score
    const bool known_done = (job->max_iter != -1 && it >= job->max_iter) || it >= cap;
    double s = 0.0;
    bool launched = false;
    if (!ctx->ev_metric) {
        phase_begin(ctx, PCD_PHASE_METRIC);
        PCD_TRY(metric_dev(ctx, rule, &s));
        phase_end(ctx, PCD_PHASE_METRIC);
    } else {
        // The metric on its own stream, forked after the compression: the
        // propagation does not read it and starts at once (under WCC the
        // metric's internal-triangle pass, its own scratch arrays, overlaps the
        // whole propagation); phase timing and kernel timers keep it on
        // `stream` (their events are recorded there).  The statistics update
        // waits for it on the device below.
        const bool mfork = ctx->mstream && ctx->ev_mfork && !ctx->timing;
        hipStream_t main_stream = ctx->stream;
        if (mfork) {
            PCD_HIP(hipEventRecord(ctx->ev_mfork, ctx->stream));
            PCD_HIP(hipStreamWaitEvent(ctx->mstream, ctx->ev_mfork, 0));
            ctx->stream = ctx->mstream;  // the metric's launches and copy go there
        }
        phase_begin(ctx, PCD_PHASE_METRIC);
        const int mrc = metric_enqueue(ctx, rule);
        phase_end(ctx, PCD_PHASE_METRIC);
        const hipError_t erc = hipEventRecord(ctx->ev_metric, ctx->stream);
        ctx->stream = main_stream;
        PCD_TRY(mrc);
        PCD_HIP(erc);
        if (prop && !known_done) {
            phase_begin(ctx, PCD_PHASE_PROPAGATE);
            PCD_TRY(propagation_launch(ctx, rule, job->frac, job->seed, it, job->tie));
            launched = true;
        }
        // what follows on `stream` (the statistics update, a merging step, the
        // next compression) waits for the metric's reads of the statistics
        if (mfork) PCD_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_metric, 0));
        PCD_HIP(hipEventSynchronize(ctx->ev_metric));
        s = ctx->h_result[0];
    }
    resolve_k(ctx);  // the event is behind k's read-back too
    const bool d = known_done ||
                   (job->max_iter == -1 && prev_score > 0 && s > 0 && s / prev_score - 1 <= job->threshold);
    if (!d) {
        if (prop) {
            if (!launched) {
                phase_begin(ctx, PCD_PHASE_PROPAGATE);
                PCD_TRY(propagation_launch(ctx, rule, job->frac, job->seed, it, job->tie));
            }
            PCD_TRY(propagation_commit(ctx));
            phase_end(ctx, PCD_PHASE_PROPAGATE);
            ctx->labels_have_new = true;
        } else {
            phase_begin(ctx, PCD_PHASE_MERGE);
            PCD_TRY(merging_dev(ctx, rule, rounds, job->seed, it, job->tie));
            phase_end(ctx, PCD_PHASE_MERGE);
            ctx->labels_have_new = false;
        }
    } else if (launched) {
        phase_end(ctx, PCD_PHASE_PROPAGATE);  // dropped: labels and statistics unchanged
    }
    if (score) *score = s;
    if (num_labels) *num_labels = ctx->k;
    if (done) *done = d ? 1 : 0;
    return PCD_OK;
}

extern "C" int pcd_contract(pcd_ctx *ctx, int64_t *num_pairs, int32_t *l, int32_t *k, int32_t *links) {
    PCD_TRY(check_labels(ctx));
    if (!ctx->has_stats) return fail(PCD_E_STATE, "no statistics; call pcd_collect_statistics");
    if (!num_pairs) return fail(PCD_E_INVALID, "null num_pairs");
    int64_t U = 0;
    PCD_TRY(contract_dev(ctx, &U));
    *num_pairs = U;
    if (l || k || links) PCD_TRY(contract_download(ctx, U, l, k, links));
    return PCD_OK;
}

extern "C" int pcd_count_triangles(pcd_ctx *ctx, int32_t *tri, int32_t *tri_deg, double *cc) {
    PCD_TRY(check_graph(ctx));
    phase_begin(ctx, PCD_PHASE_TRIANGLES);
    PCD_TRY(triangles_dev(ctx));
    phase_end(ctx, PCD_PHASE_TRIANGLES);
    if (tri) PCD_HIP(hipMemcpyAsync(tri, ctx->d_tri, 4 * (size_t)ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    if (tri_deg)
        PCD_HIP(hipMemcpyAsync(tri_deg, ctx->d_tri_deg, 4 * (size_t)ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    if (cc) *cc = ctx->cc;
    return PCD_OK;
}

extern "C" int pcd_set_clustering_coef(pcd_ctx *ctx, double cc) {
    PCD_TRY(check_graph(ctx));
    if (!ctx->has_triangles) {
        phase_begin(ctx, PCD_PHASE_TRIANGLES);
        PCD_TRY(triangles_dev(ctx));
        phase_end(ctx, PCD_PHASE_TRIANGLES);
    }
    ctx->cc = cc;
    return PCD_OK;
}

extern "C" int pcd_enable_timing(pcd_ctx *ctx, int enable) {
    PCD_TRY(check_ctx(ctx));
    ctx->timing = enable != 0;
    return PCD_OK;
}

extern "C" int pcd_phase_time(pcd_ctx *ctx, int phase, double *ms) {
    PCD_TRY(check_ctx(ctx));
    if (phase < 0 || phase >= PCD_PHASE_COUNT || !ms) return fail(PCD_E_INVALID, "bad phase");
    if (!ctx->ev_valid[phase]) {
        *ms = 0.0;
        return PCD_OK;
    }
    PCD_HIP(hipEventSynchronize(ctx->ev1[phase]));
    float f = 0;
    PCD_HIP(hipEventElapsedTime(&f, ctx->ev0[phase], ctx->ev1[phase]));
    *ms = f;
    return PCD_OK;
}

extern "C" int pcd_set_tuning(pcd_ctx *ctx, const char *key, int value) {
    PCD_TRY(check_ctx(ctx));
    if (!key) return fail(PCD_E_INVALID, "null key");
    std::string k(key);
    if (k == "hub_table_slots") {
        if (value < 64 || value > HUB_TB || (value % 64)) return fail(PCD_E_INVALID, "hub_table_slots: multiple of 64 in [64, 4096]");
        ctx->tune.hub_table_slots = value;
    } else if (k == "hub_part_entries") {
        if (value < 64) return fail(PCD_E_INVALID, "hub_part_entries must be >= 64");
        ctx->tune.hub_part_entries = value;  // applies to the next graph upload
    } else if (k == "small_aux") {
        ctx->tune.small_aux = value != 0;
    } else if (k == "block_aux") {
        ctx->tune.block_aux = value != 0;
    } else if (k == "hub_aux") {
        ctx->tune.hub_aux = value != 0;
    } else if (k == "small_fused") {
        ctx->tune.small_fused = value != 0;
    } else if (k == "bin_streams") {
        ctx->tune.bin_streams = value != 0;
    } else if (k == "bin_streams_min_edges") {
        if (value < 0) return fail(PCD_E_INVALID, "bin_streams_min_edges must be >= 0");
        ctx->tune.bin_streams_min_edges = value;
    } else if (k == "small_occ8") {
        ctx->tune.small_occ8 = value != 0;
    } else if (k == "b1024_u16") {
        ctx->tune.b1024_u16 = value != 0;
    } else if (k == "hub_nt") {
        if (value != 256 && value != 512) return fail(PCD_E_INVALID, "hub_nt: 256 or 512");
        ctx->tune.hub_nt = value;
    } else if (k == "stats_cache") {
        ctx->tune.stats_cache = value != 0;
    } else if (k == "inc_stats") {
        ctx->tune.inc_stats = value != 0;
    } else if (k == "b1024_wave") {
        ctx->tune.b1024_wave = value != 0;
    } else if (k == "b4096_u16") {
        ctx->tune.b4096_u16 = value != 0;
    } else if (k == "b4096_split") {
        ctx->tune.b4096_split = value != 0;
    } else if (k == "cluster_rounds") {
        if (value < 0) return fail(PCD_E_INVALID, "cluster_rounds must be >= 0");
        ctx->tune.cluster_rounds = value;
    } else if (k == "onesweep") {
        ctx->tune.onesweep = value != 0;
    } else if (k == "scan_tiles_per_cu") {
        if (value < 1 || value > 64) return fail(PCD_E_INVALID, "scan_tiles_per_cu must be in [1, 64]");
        ctx->tune.scan_tiles_per_cu = value;
    } else if (k == "os_spin_limit") {
        if (value < 0) return fail(PCD_E_INVALID, "os_spin_limit must be >= 0");
        ctx->tune.os_spin_limit = value;
    } else if (k == "merge_sort") {
        ctx->tune.merge_sort = value != 0;
    } else if (k == "graph_loop") {
        ctx->tune.graph_loop = value != 0;
    } else if (k == "graph_loop_min_entries") {
        if (value < 0) return fail(PCD_E_INVALID, "graph_loop_min_entries must be >= 0");
        ctx->tune.graph_loop_min_entries = value;
    } else if (k == "stat_chunk") {
        if (value < 64) return fail(PCD_E_INVALID, "stat_chunk must be >= 64");
        ctx->tune.stat_chunk = value;  // applies to the next graph upload
    } else {
        return fail(PCD_E_INVALID, "unknown tuning key " + k);
    }
    ctx->tune_gen++;  // a captured pass holds the old kernel forms
    return PCD_OK;
}

extern "C" int pcd_enable_counters(pcd_ctx *ctx, int enable) {
    PCD_TRY(check_ctx(ctx));
    ctx->counters = enable != 0;
    return PCD_OK;
}

extern "C" int pcd_propagation_counters(pcd_ctx *ctx, int64_t *active, int64_t *entries, int64_t *distinct) {
    int64_t a = 0, e = 0, d = 0;
    for (int b = 0; b < NUM_BINS; b++) {
        int64_t x = 0, y = 0, z = 0;
        PCD_TRY(pcd_bin_counters(ctx, b, &x, &y, &z));
        a += x; e += y; d += z;
    }
    if (active) *active = a;
    if (entries) *entries = e;
    if (distinct) *distinct = d;
    return PCD_OK;
}

extern "C" int pcd_kernel_timer_select(pcd_ctx *ctx, uint32_t mask) {
    PCD_TRY(check_ctx(ctx));
    ctx->timer_mask = mask;
    return PCD_OK;
}

extern "C" int pcd_kernel_slot_count(void) { return NUM_KSLOTS; }

extern "C" const char *pcd_kernel_slot_name(int slot) {
    return (slot >= 0 && slot < NUM_KSLOTS) ? kslot_names[slot] : "";
}

extern "C" int pcd_kernel_timer_reset(pcd_ctx *ctx) {
    PCD_TRY(check_ctx(ctx));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < NUM_KSLOTS; s++) {
        for (auto &pr : ctx->ktimes[s]) {
            ctx->event_pool.push_back(pr.first);
            ctx->event_pool.push_back(pr.second);
        }
        ctx->ktimes[s].clear();
    }
    return PCD_OK;
}

extern "C" int pcd_kernel_timer(pcd_ctx *ctx, int slot, double *total_ms, int64_t *launches) {
    PCD_TRY(check_ctx(ctx));
    if (slot < 0 || slot >= NUM_KSLOTS) return fail(PCD_E_INVALID, "bad kernel slot");
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    double t = 0;
    for (auto &pr : ctx->ktimes[slot]) {
        float f = 0;
        PCD_HIP(hipEventElapsedTime(&f, pr.first, pr.second));
        t += f;
    }
    if (total_ms) *total_ms = t;
    if (launches) *launches = (int64_t)ctx->ktimes[slot].size();
    return PCD_OK;
}

extern "C" int pcd_sort_fallbacks(pcd_ctx *ctx, int64_t *count) {
    PCD_TRY(check_ctx(ctx));
    if (count) *count = ctx->os_fallbacks;
    ctx->os_fallbacks = 0;
    return PCD_OK;
}

extern "C" int pcd_hub_overflows(pcd_ctx *ctx, int64_t *count) {
    PCD_TRY(check_ctx(ctx));
    int32_t h = 0;
    PCD_HIP(hipMemcpyAsync(&h, ctx->d_counts + NUM_BINS + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    if (count) *count = h;
    PCD_TRY(fill_i32(ctx, ctx->d_counts + NUM_BINS + 2, 0, 1));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    return PCD_OK;
}

namespace pcd {
__global__ void k_stream_copy(const int4 *__restrict__ src, int4 *__restrict__ dst, int64_t n);
}

extern "C" int pcd_copy_probe(pcd_ctx *ctx, int64_t bytes, int reps, double *gbs) {
    PCD_TRY(check_ctx(ctx));
    if (bytes < 1024 || reps < 1 || !gbs) return fail(PCD_E_INVALID, "bad copy probe arguments");
    const int64_t n = bytes / 16;
    int4 *a = nullptr, *b = nullptr;
    PCD_TRY(dev_alloc_t(ctx, &a, n));
    int rc = dev_alloc_t(ctx, &b, n);
    if (rc != PCD_OK) { dev_free(ctx, a); return rc; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = ctx->cu_count * 8;
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, a, b, n);  // warm-up
    (void)hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, (r & 1) ? b : a, (r & 1) ? a : b, n);
    (void)hipEventRecord(e1, ctx->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(ctx, a);
    dev_free(ctx, b);
    if (e != hipSuccess) return fail(PCD_E_HIP, std::string("copy probe: ") + hipGetErrorString(e));
    *gbs = 2.0 * 16.0 * (double)n * reps / (ms / 1000.0) / 1e9;
    return PCD_OK;
}

namespace pcd {
__global__ void k_gather_probe(int64_t E, const int32_t *adj, const int32_t *labels, int n, int levels,
                               unsigned long long *sink);
}

extern "C" int pcd_gather_probe(pcd_ctx *ctx, int levels, int reps, double *gathers_per_s) {
    PCD_TRY(check_labels(ctx));
    if ((levels != 1 && levels != 2) || reps < 1 || !gathers_per_s)
        return fail(PCD_E_INVALID, "bad gather probe arguments");
    const int64_t E = (int64_t)2 * ctx->m;
    unsigned long long *sink = nullptr;
    PCD_TRY(dev_alloc_t(ctx, &sink, 1));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = ctx->cu_count * 8;
    hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, ctx->stream, E, ctx->d_adj, ctx->d_labels, ctx->n,
                       levels, sink);  // warm-up
    (void)hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_gather_probe, dim3(grid), dim3(256), 0, ctx->stream, E, ctx->d_adj, ctx->d_labels,
                           ctx->n, levels, sink);
    (void)hipEventRecord(e1, ctx->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(ctx, sink);
    if (e != hipSuccess) return fail(PCD_E_HIP, std::string("gather probe: ") + hipGetErrorString(e));
    *gathers_per_s = (double)E * levels * reps / (ms / 1000.0);
    return PCD_OK;
}

extern "C" int pcd_host_alloc(int64_t bytes, void **p) {
    if (!p || bytes <= 0) return fail(PCD_E_INVALID, "bad arguments");
    *p = nullptr;
    if (hipHostMalloc(p, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(PCD_E_NOMEM, "pinned host allocation failed");
    }
    return PCD_OK;
}

extern "C" int pcd_host_free(void *p) {
    if (p) PCD_HIP(hipHostFree(p));
    return PCD_OK;
}

extern "C" int pcd_bin_count(void) { return NUM_BINS; }

extern "C" const char *pcd_bin_name(int bin) { return (bin >= 0 && bin < NUM_BINS) ? bin_names[bin] : ""; }

extern "C" int pcd_bin_counters_reset(pcd_ctx *ctx) {
    PCD_TRY(check_ctx(ctx));
    PCD_TRY(fill_u64(ctx, (uint64_t *)ctx->d_bin_counters, 0, 3 * NUM_BINS));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    return PCD_OK;
}

extern "C" int pcd_bin_counters(pcd_ctx *ctx, int bin, int64_t *active, int64_t *entries, int64_t *distinct) {
    PCD_TRY(check_ctx(ctx));
    if (bin < 0 || bin >= NUM_BINS) return fail(PCD_E_INVALID, "bad bin");
    unsigned long long h[3];
    PCD_HIP(hipMemcpyAsync(h, ctx->d_bin_counters + 3 * bin, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PCD_HIP(hipStreamSynchronize(ctx->stream));
    if (active) *active = (int64_t)h[0];
    if (entries) *entries = (int64_t)h[1];
    if (distinct) *distinct = (int64_t)h[2];
    return PCD_OK;
}

extern "C" int pcd_bin_size(pcd_ctx *ctx, int bin, int64_t *vertices, int64_t *entries) {
    PCD_TRY(check_graph(ctx));
    if (bin < 0 || bin >= NUM_BINS) return fail(PCD_E_INVALID, "bad bin");
    if (vertices) *vertices = ctx->bins.start[bin + 1] - ctx->bins.start[bin];
    if (entries) *entries = ctx->bin_entries[bin];
    return PCD_OK;
}
