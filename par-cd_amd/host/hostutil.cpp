// Host-side utilities of the C ABI: PCD file I/O, the text->CSR converter
// and seeded synthetic graph generators.
//
// The converter produces the same bytes as the reference `convert`
// (convert.cpp:76-278) but runs its per-row work on all host cores.
#include <algorithm>
#include <atomic>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "parcd.h"
#include "hostutil.hpp"

namespace pcd {

thread_local std::string g_last_error;
// set in workers that already run one task per thread: nested loops run serially
static thread_local bool g_serial = false;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int host_threads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    if (const char *e = getenv("PCD_HOST_THREADS")) {
        int v = atoi(e);
        if (v > 0) return v;
    }
    return (int)std::min(h, 16u);
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn, int64_t grain) {
    int nt = g_serial ? 1 : host_threads();
    if (n <= grain || nt <= 1) {
        if (n > 0) fn(0, n);
        return;
    }
    int64_t chunks = std::min<int64_t>(nt, (n + grain - 1) / grain);
    std::vector<std::thread> th;
    for (int64_t c = 0; c < chunks; c++) {
        int64_t lo = n * c / chunks, hi = n * (c + 1) / chunks;
        th.emplace_back([=, &fn] { fn(lo, hi); });
    }
    for (auto &t : th) t.join();
}

}  // namespace pcd

using namespace pcd;

extern "C" const char *pcd_last_error(void) { return g_last_error.c_str(); }

extern "C" void pcd_free(void *p) { free(p); }

// framework.cpp:14-24
extern "C" int pcd_singleton_partition(int n, int32_t *labels) {
    if (n <= 0 || !labels) return fail(PCD_E_INVALID, "bad arguments");
    for (int v = 0; v < n; v++) labels[v] = v;
    for (int v = 0; v < n; v++) {
        int j = rand() % n;
        int32_t t = labels[v];
        labels[v] = labels[j];
        labels[j] = t;
    }
    return PCD_OK;
}

template <typename T>
static T *xmalloc(size_t count) {
    return (T *)malloc(sizeof(T) * (count ? count : 1));
}

// graph.cpp:40-105 Graph::read
extern "C" int pcd_read_pcd(const char *path, int *n_out, int *m_out, int32_t **indices, int32_t **adj,
                            int32_t **map) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(PCD_E_IO, std::string("failed to open ") + path);
    char w[3];
    int n = 0, m = 0;
    if (fread(w, 1, 3, fp) != 3) { fclose(fp); return fail(PCD_E_FORMAT, "error while reading header"); }
    if (memcmp(w, "PCD", 3) != 0) { fclose(fp); return fail(PCD_E_FORMAT, "invalid file type"); }
    if (fread(&n, 4, 1, fp) != 1 || fread(&m, 4, 1, fp) != 1 || n < 0 || m < 0) {
        fclose(fp);
        return fail(PCD_E_FORMAT, "error while reading sizes");
    }
    int32_t *ind = xmalloc<int32_t>((size_t)n + 1), *a = xmalloc<int32_t>((size_t)2 * m),
            *mp = xmalloc<int32_t>((size_t)n);
    if (!ind || !a || !mp) { fclose(fp); free(ind); free(a); free(mp); return fail(PCD_E_NOMEM, "out of host memory"); }
    bool err = fread(ind, 4, (size_t)n + 1, fp) != (size_t)n + 1;
    err = err || fread(a, 4, (size_t)2 * m, fp) != (size_t)2 * m;
    err = err || fread(mp, 4, (size_t)n, fp) != (size_t)n;
    if (!err && fgetc(fp) != EOF) err = true;  // graph.cpp:90-93: no trailing bytes
    fclose(fp);
    if (!err) {
        // structural checks the kernels rely on (the reference trusts the file)
        if (ind[0] != 0 || ind[n] != 2 * m) err = true;
        for (int i = 0; !err && i < n; i++) if (ind[i + 1] < ind[i]) err = true;
        for (int64_t j = 0; !err && j < (int64_t)2 * m; j++) if (a[j] < 0 || a[j] >= n) err = true;
    }
    if (err) {
        free(ind); free(a); free(mp);
        return fail(PCD_E_FORMAT, std::string("error while reading ") + path);
    }
    *n_out = n; *m_out = m; *indices = ind; *adj = a; *map = mp;
    return PCD_OK;
}

// convert.cpp:244-278 write_edges (file bytes)
extern "C" int pcd_write_pcd(const char *path, int n, int m, const int32_t *indices, const int32_t *adj,
                             const int32_t *map) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(PCD_E_IO, std::string("failed to open ") + path);
    bool ok = fwrite("PCD", 1, 3, fp) == 3;
    ok = ok && fwrite(&n, 4, 1, fp) == 1 && fwrite(&m, 4, 1, fp) == 1;
    ok = ok && fwrite(indices, 4, (size_t)n + 1, fp) == (size_t)n + 1;
    ok = ok && fwrite(adj, 4, (size_t)2 * m, fp) == (size_t)2 * m;
    ok = ok && fwrite(map, 4, (size_t)n, fp) == (size_t)n;
    ok = (fclose(fp) == 0) && ok;
    return ok ? PCD_OK : fail(PCD_E_IO, "an error occured while writing output file");
}

// convert.cpp:14-74 read_edges: same acceptance rules (blank and '#' lines
// skipped, two non-negative decimal integers per line, lines < 1024 chars).
extern "C" int pcd_read_edge_text(const char *path, int64_t *num_edges, int32_t **a_out, int32_t **b_out,
                                  int *num_nodes) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(PCD_E_IO, std::string("failed to open ") + path);
    std::vector<int32_t> a, b;
    std::vector<char> buf(1 << 20);
    std::string line;
    int nn = 0;
    long lineno = 0;
    size_t got;
    bool eof = false;
    size_t pos = 0, len = 0;
    auto getc_ = [&]() -> int {
        if (pos == len) {
            if (eof) return EOF;
            len = fread(buf.data(), 1, buf.size(), fp);
            pos = 0;
            if (len == 0) { eof = true; return EOF; }
        }
        return (unsigned char)buf[pos++];
    };
    (void)got;
    auto isspace_ = [](int c) { return c == ' ' || (c >= '\t' && c <= '\r'); };
    auto isdigit_ = [](int c) { return c >= '0' && c <= '9'; };
    for (;;) {
        lineno++;
        line.clear();
        int ch;
        while ((ch = getc_()) != EOF && ch != '\n') {
            if (line.size() >= 1023) { fclose(fp); return fail(PCD_E_FORMAT, "an error occured while reading input file"); }
            line.push_back((char)ch);
        }
        bool at_eof = ch == EOF;
        if (at_eof && line.empty()) break;
        const char *s = line.c_str();
        int idx = 0;
        while (isspace_(s[idx])) idx++;
        if (s[idx] == 0 || s[idx] == '#') { if (at_eof) break; continue; }
        if (!isdigit_(s[idx])) { fclose(fp); return fail(PCD_E_FORMAT, "line " + std::to_string(lineno) + " is corrupt"); }
        uint32_t x = 0, y = 0;
        while (isdigit_(s[idx])) x = 10u * x + (uint32_t)(s[idx++] - '0');
        while (isspace_(s[idx])) idx++;
        if (!isdigit_(s[idx])) { fclose(fp); return fail(PCD_E_FORMAT, "line " + std::to_string(lineno) + " is corrupt"); }
        while (isdigit_(s[idx])) y = 10u * y + (uint32_t)(s[idx++] - '0');
        a.push_back((int32_t)x); b.push_back((int32_t)y);
        int mx = std::max((int32_t)x, (int32_t)y) + 1;
        if (mx > nn) nn = mx;
        if (at_eof) break;
    }
    fclose(fp);
    int32_t *pa = xmalloc<int32_t>(a.size()), *pb = xmalloc<int32_t>(b.size());
    if (!pa || !pb) { free(pa); free(pb); return fail(PCD_E_NOMEM, "out of host memory"); }
    if (!a.empty()) { memcpy(pa, a.data(), 4 * a.size()); memcpy(pb, b.data(), 4 * b.size()); }
    *a_out = pa; *b_out = pb; *num_edges = (int64_t)a.size(); *num_nodes = nn;
    return PCD_OK;
}

// convert.cpp:76-196: CSR build, per-row dedup, degree sort (degree desc,
// id desc: convert.cpp:150 greater<pair<int,int>>), row re-sort, isolated
// vertex truncation.  Rows are independent, so they are sorted in parallel.
extern "C" int pcd_convert_edges(int num_nodes, int64_t ne, const int32_t *ea, const int32_t *eb, int *n_out,
                                 int *m_out, int32_t **indices_out, int32_t **adj_out, int32_t **map_out) {
    if (num_nodes < 0 || ne < 0) return fail(PCD_E_INVALID, "negative size");
    for (int64_t i = 0; i < ne; i++)
        if (ea[i] < 0 || eb[i] < 0 || ea[i] >= num_nodes || eb[i] >= num_nodes)
            return fail(PCD_E_INVALID, "edge endpoint out of range");
    const int N = num_nodes;
    std::vector<int64_t> ind((size_t)N + 1, 0);
    for (int64_t i = 0; i < ne; i++) {
        if (ea[i] == eb[i]) continue;  // self-loops: see parcd.h
        ind[ea[i] + 1]++; ind[eb[i] + 1]++;
    }
    for (int i = 0; i < N; i++) ind[i + 1] += ind[i];
    if (ind[N] > (int64_t)INT32_MAX) return fail(PCD_E_INVALID, "more than 2^31-1 endpoints");
    std::vector<int32_t> adj((size_t)std::max<int64_t>(ind[N], 1));
    {
        std::vector<int64_t> cur(ind.begin(), ind.end() - 1);
        for (int64_t i = 0; i < ne; i++) {
            int32_t a = ea[i], b = eb[i];
            if (a == b) continue;
            adj[cur[a]++] = b;
            adj[cur[b]++] = a;
        }
    }
    // per-row sort + unique (convert.cpp:109-133)
    std::vector<int32_t> ndeg((size_t)N + 1, 0);
    parallel_for(N, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++) {
            int32_t *s = adj.data() + ind[v], *e = adj.data() + ind[v + 1];
            std::sort(s, e);
            ndeg[v] = (int32_t)(std::unique(s, e) - s);
        }
    }, 4096);
    // degree sort (convert.cpp:136-156): counting sort on degree, ids
    // descending inside a degree
    int32_t maxd = 0;
    for (int v = 0; v < N; v++) maxd = std::max(maxd, ndeg[v]);
    std::vector<int64_t> bucket((size_t)maxd + 2, 0);
    for (int v = 0; v < N; v++) bucket[maxd - ndeg[v] + 1]++;
    for (int d = 0; d <= maxd; d++) bucket[d + 1] += bucket[d];
    std::vector<int32_t> rmap((size_t)std::max(N, 1)), fmap((size_t)std::max(N, 1));
    for (int v = N - 1; v >= 0; v--) {
        int32_t i = (int32_t)bucket[maxd - ndeg[v]]++;
        rmap[i] = v;
        fmap[v] = i;
    }
    int n = N;
    for (int i = 0; i < N; i++) if (ndeg[rmap[i]] == 0) { n = i; break; }
    std::vector<int64_t> nind((size_t)N + 1, 0);
    for (int i = 0; i < N; i++) nind[i + 1] = nind[i] + ndeg[rmap[i]];
    int32_t *oind = xmalloc<int32_t>((size_t)n + 1), *oadj = xmalloc<int32_t>((size_t)nind[n]),
            *omap = xmalloc<int32_t>((size_t)n);
    if (!oind || !oadj || !omap) { free(oind); free(oadj); free(omap); return fail(PCD_E_NOMEM, "out of host memory"); }
    for (int i = 0; i <= n; i++) oind[i] = (int32_t)nind[i];
    for (int i = 0; i < n; i++) omap[i] = rmap[i];
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            int32_t v = rmap[i];
            int32_t *dst = oadj + nind[i];
            const int32_t *src = adj.data() + ind[v];
            for (int32_t j = 0; j < ndeg[v]; j++) dst[j] = fmap[src[j]];
            std::sort(dst, dst + ndeg[v]);
        }
    }, 4096);
    *n_out = n;
    *m_out = (int)(nind[n] / 2);
    *indices_out = oind; *adj_out = oadj; *map_out = omap;
    return PCD_OK;
}

// ------------------------------------------------------------- generators
// Streams: SplitMix64 (Steele, Lea, Flood 2014); uniform doubles take the
// top 53 bits.
namespace {
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};
uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t stream_seed(uint64_t seed, uint64_t a, uint64_t b) {
    return mix(seed ^ mix(a * 0xD6E8FEB86659FD93ULL + mix(b + 0x632BE59BD9B4E019ULL)));
}
// Number of failures before the next success of Bernoulli(p) trials.
int64_t geometric_skip(SplitMix &r, double logq) {
    double u = r.uniform();
    if (u <= 0.0) u = 1e-300;
    return (int64_t)std::floor(std::log(u) / logq);
}
int emit(std::vector<int32_t> &a, std::vector<int32_t> &b, int64_t *ne, int32_t **pa, int32_t **pb) {
    int32_t *oa = xmalloc<int32_t>(a.size()), *ob = xmalloc<int32_t>(b.size());
    if (!oa || !ob) { free(oa); free(ob); return fail(PCD_E_NOMEM, "out of host memory"); }
    if (!a.empty()) { memcpy(oa, a.data(), 4 * a.size()); memcpy(ob, b.data(), 4 * b.size()); }
    *ne = (int64_t)a.size(); *pa = oa; *pb = ob;
    return PCD_OK;
}
}  // namespace

// Planted-partition SBM: every unordered pair inside a block is an edge with
// probability p_in, across blocks p_out.  Pairs of each block pair (bi<=bj)
// are enumerated row-major and sampled by geometric skipping on the stream
// stream_seed(seed, bi, bj), so the cost is O(n + m).
extern "C" int pcd_gen_sbm(uint64_t seed, int nb, const int32_t *bs, double p_in, double p_out, int64_t *ne,
                           int32_t **pa, int32_t **pb, int *num_nodes) {
    if (nb <= 0 || p_in < 0 || p_in > 1 || p_out < 0 || p_out > 1) return fail(PCD_E_INVALID, "bad SBM parameters");
    std::vector<int64_t> off((size_t)nb + 1, 0);
    for (int i = 0; i < nb; i++) {
        if (bs[i] < 0) return fail(PCD_E_INVALID, "negative block size");
        off[i + 1] = off[i] + bs[i];
    }
    if (off[nb] > INT32_MAX) return fail(PCD_E_INVALID, "too many vertices");
    std::vector<int32_t> a, b;
    for (int bi = 0; bi < nb; bi++) {
        for (int bj = bi; bj < nb; bj++) {
            double p = bi == bj ? p_in : p_out;
            if (p <= 0) continue;
            int64_t si = bs[bi], sj = bs[bj];
            int64_t total = bi == bj ? si * (si - 1) / 2 : si * sj;
            if (total == 0) continue;
            SplitMix r(stream_seed(seed, (uint64_t)bi, (uint64_t)bj));
            double logq = p < 1 ? std::log1p(-p) : 0;
            int64_t t = p < 1 ? geometric_skip(r, logq) : 0;
            // row bookkeeping for the upper triangle of a diagonal block
            int64_t row = 0, row_start = 0;
            while (t < total) {
                int64_t x, y;
                if (bi == bj) {
                    while (t >= row_start + (si - 1 - row)) { row_start += si - 1 - row; row++; }
                    x = row; y = row + 1 + (t - row_start);
                } else {
                    x = t / sj; y = t % sj;
                }
                a.push_back((int32_t)(off[bi] + x));
                b.push_back((int32_t)(off[bj] + y));
                t += 1 + (p < 1 ? geometric_skip(r, logq) : 0);
            }
        }
    }
    *num_nodes = (int)off[nb];
    return emit(a, b, ne, pa, pb);
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): draw i descends `scale` levels
// with quadrant probabilities a, b, c, d = 1-a-b-c using the stream
// stream_seed(seed, i, 0); quadrant b sets the destination bit, c the source
// bit, d both.  Self-loops are dropped; duplicates are left to the converter.
extern "C" int pcd_gen_rmat(uint64_t seed, int scale, int64_t draws, double pa_, double pb_, double pc_,
                            int64_t *ne, int32_t **oa, int32_t **ob, int *num_nodes) {
    if (scale < 1 || scale > 30 || draws < 0) return fail(PCD_E_INVALID, "bad R-MAT parameters");
    std::vector<int32_t> src((size_t)draws), dst((size_t)draws);
    const double ab = pa_ + pb_, abc = pa_ + pb_ + pc_;
    parallel_for(draws, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            SplitMix r(stream_seed(seed, (uint64_t)i, 0));
            uint32_t s = 0, d = 0;
            for (int l = 0; l < scale; l++) {
                double u = r.uniform();
                uint32_t sb = 0, db = 0;
                if (u < pa_) {
                } else if (u < ab) {
                    db = 1;
                } else if (u < abc) {
                    sb = 1;
                } else {
                    sb = 1; db = 1;
                }
                s = 2 * s + sb; d = 2 * d + db;
            }
            src[i] = (int32_t)s; dst[i] = (int32_t)d;
        }
    }, 1 << 16);
    std::vector<int32_t> a, b;
    a.reserve((size_t)draws); b.reserve((size_t)draws);
    int32_t mx = -1;
    for (int64_t i = 0; i < draws; i++) {
        if (src[i] == dst[i]) continue;
        a.push_back(src[i]); b.push_back(dst[i]);
        mx = std::max(mx, std::max(src[i], dst[i]));
    }
    // vertex count = max id + 1, as convert.cpp:61 computes it from the text
    *num_nodes = mx + 1;
    return emit(a, b, ne, oa, ob);
}

// Chung-Lu with planted communities.  Expected degrees w_i follow a discrete
// power law P(d) ~ d^-exponent on [dmin, dmax] (inverse-CDF draws).  The
// background graph is Chung-Lu, P(i,j) = min(1, w_i w_j / S), sampled in
// O(n + m) by the skipping method of Miller & Hagberg (2011) over weights
// sorted descending.  num_comms planted communities of log-uniform sizes in
// [cmin, cmax] (disjoint random vertex sets) add, inside each community, the
// pairs again with probability (boost-1)*min(1, w_i w_j / S), so intra pairs
// are about `boost` times as likely.
extern "C" int pcd_gen_chung_lu(uint64_t seed, int n, double expo, int dmin, int dmax, int num_comms, int cmin,
                                int cmax, double boost, int64_t *ne, int32_t **oa, int32_t **ob, int *num_nodes) {
    if (n <= 1 || dmin < 1 || dmax < dmin || expo <= 1) return fail(PCD_E_INVALID, "bad Chung-Lu parameters");
    SplitMix r(stream_seed(seed, 1, 1));
    // discrete power-law CDF
    std::vector<double> cdf((size_t)(dmax - dmin + 1));
    double acc = 0;
    for (int d = dmin; d <= dmax; d++) { acc += std::pow((double)d, -expo); cdf[d - dmin] = acc; }
    std::vector<double> w((size_t)n);
    for (int i = 0; i < n; i++) {
        double u = r.uniform() * acc;
        w[i] = dmin + (double)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    }
    double S = 0;
    for (int i = 0; i < n; i++) S += w[i];
    std::vector<int32_t> order((size_t)n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return w[x] > w[y]; });
    std::vector<int32_t> a, b;
    auto skip_pass = [&](const std::vector<int32_t> &vs, double scale, SplitMix &rr) {
        // vs sorted by weight descending
        int64_t k = (int64_t)vs.size();
        for (int64_t u = 0; u + 1 < k; u++) {
            int64_t v = u + 1;
            double p = std::min(scale * w[vs[u]] * w[vs[v]] / S, 1.0);
            while (v < k && p > 0) {
                if (p < 1) v += (int64_t)std::floor(std::log(std::max(rr.uniform(), 1e-300)) / std::log1p(-p));
                if (v < k) {
                    double q = std::min(scale * w[vs[u]] * w[vs[v]] / S, 1.0);
                    if (rr.uniform() < q / p) { a.push_back(vs[u]); b.push_back(vs[v]); }
                    p = q;
                    v++;
                }
            }
        }
    };
    SplitMix rb(stream_seed(seed, 2, 0));
    skip_pass(order, 1.0, rb);
    // planted communities
    std::vector<int32_t> perm((size_t)n);
    for (int i = 0; i < n; i++) perm[i] = i;
    SplitMix rp(stream_seed(seed, 3, 0));
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int64_t)(rp.next() % (uint64_t)(i + 1))]);
    int64_t used = 0;
    for (int c = 0; c < num_comms && boost > 1; c++) {
        double lo = std::log((double)cmin), hi = std::log((double)cmax);
        int64_t sz = (int64_t)std::floor(std::exp(lo + (hi - lo) * rp.uniform()));
        sz = std::min<int64_t>(sz, n - used);
        if (sz < 2) break;
        std::vector<int32_t> mem(perm.begin() + used, perm.begin() + used + sz);
        used += sz;
        std::stable_sort(mem.begin(), mem.end(), [&](int32_t x, int32_t y) { return w[x] > w[y]; });
        SplitMix rc(stream_seed(seed, 4, (uint64_t)c));
        skip_pass(mem, boost - 1.0, rc);
    }
    *num_nodes = n;
    return emit(a, b, ne, oa, ob);
}

// Batched-form workload: `count` planted-partition SBM graphs.  Graph i
// draws its shape from stream_seed(seed, i, 0): n log-uniform in
// [nmin, nmax], a block count uniform in [bmin, bmax] (equal blocks, the
// first n mod B one larger), p_in uniform in [pin_lo, pin_hi], p_out in
// [pout_lo, pout_hi], then the metric (modularity or CPM, 1/2 each) and a
// CPM resolution uniform in [lam_lo, lam_hi]; its edges come from
// pcd_gen_sbm with seed stream_seed(seed, i, 1), converted as pcd_convert_edges
// does.  Graphs are built in parallel, one per worker.  Outputs (malloc'ed,
// free with pcd_free): per graph n, m, the offsets of its rows and entries in
// the concatenated indices / adj arrays, and params[6 i ..] = {nominal n, B,
// p_in, p_out, kind, lambda}.
namespace {
struct GenGraph {
    int n = 0, m = 0;
    int32_t *ind = nullptr, *adj = nullptr;
    double prm[6] = {0, 0, 0, 0, 0, 0};
    int rc = PCD_OK;
};

// Builds `count` graphs on all host threads (one graph per worker at a
// time; nested loops run serially) and concatenates them.
int gen_many(int count, const std::function<void(int, GenGraph &)> &make, int32_t **n_out, int32_t **m_out,
             int64_t **ind_off, int64_t **adj_off, int32_t **indices, int32_t **adj, double **params) {
    std::vector<GenGraph> gs(count);
    std::atomic<int> next{0};
    auto worker = [&]() {
        g_serial = true;
        for (int i; (i = next.fetch_add(1)) < count;) make(i, gs[i]);
        g_serial = false;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < host_threads(); t++) th.emplace_back(worker);
    for (auto &t : th) t.join();
    int64_t tn = 0, te = 0;
    int rc = PCD_OK;
    for (auto &g : gs) {
        if (g.rc != PCD_OK) rc = g.rc;
        tn += g.n + 1;
        te += 2 * (int64_t)g.m;
    }
    int32_t *on = xmalloc<int32_t>(count), *om = xmalloc<int32_t>(count), *oi = xmalloc<int32_t>((size_t)tn),
            *oa = xmalloc<int32_t>((size_t)std::max<int64_t>(te, 1));
    int64_t *oio = xmalloc<int64_t>(count), *oao = xmalloc<int64_t>(count);
    double *op = xmalloc<double>((size_t)count * 6);
    if (rc == PCD_OK && (!on || !om || !oi || !oa || !oio || !oao || !op)) rc = fail(PCD_E_NOMEM, "out of host memory");
    if (rc == PCD_OK) {
        int64_t pi = 0, pa = 0;
        for (int i = 0; i < count; i++) {
            GenGraph &g = gs[i];
            on[i] = g.n;
            om[i] = g.m;
            oio[i] = pi;
            oao[i] = pa;
            if (g.ind) memcpy(oi + pi, g.ind, 4 * ((size_t)g.n + 1));
            else oi[pi] = 0;
            if (g.m) memcpy(oa + pa, g.adj, 8 * (size_t)g.m);
            memcpy(op + 6 * (size_t)i, g.prm, sizeof(g.prm));
            pi += g.n + 1;
            pa += 2 * (int64_t)g.m;
        }
    }
    for (auto &g : gs) { free(g.ind); free(g.adj); }
    if (rc != PCD_OK) {
        free(on); free(om); free(oi); free(oa); free(oio); free(oao); free(op);
        return rc;
    }
    *n_out = on; *m_out = om; *ind_off = oio; *adj_off = oao; *indices = oi; *adj = oa; *params = op;
    return PCD_OK;
}

void convert_into(GenGraph &g, int nodes, int64_t ne, int32_t *ea, int32_t *eb) {
    int32_t *mp = nullptr;
    g.rc = pcd_convert_edges(nodes, ne, ea, eb, &g.n, &g.m, &g.ind, &g.adj, &mp);
    free(mp);
}
}  // namespace

// Batched-form workload: `count` planted-partition SBM graphs.  Graph i
// draws its shape from stream_seed(seed, i, 0): n log-uniform in
// [nmin, nmax], a block count uniform in [bmin, bmax] (equal blocks, the
// first n mod B one larger), p_in uniform in [pin_lo, pin_hi], p_out in
// [pout_lo, pout_hi], then the metric (modularity or CPM, 1/2 each) and a
// CPM resolution uniform in [lam_lo, lam_hi]; its edges come from
// pcd_gen_sbm with seed stream_seed(seed, i, 1), converted as pcd_convert_edges
// does.  Outputs (malloc'ed, free with pcd_free): per graph n, m, the offsets
// of its rows and entries in the concatenated indices / adj arrays, and
// params[6 i ..] = {nominal n, B, p_in, p_out, kind, lambda}.
extern "C" int pcd_gen_sbm_batch(uint64_t seed, int count, int nmin, int nmax, int bmin, int bmax, double pin_lo,
                                 double pin_hi, double pout_lo, double pout_hi, double lam_lo, double lam_hi,
                                 int32_t **n_out, int32_t **m_out, int64_t **ind_off, int64_t **adj_off,
                                 int32_t **indices, int32_t **adj, double **params) {
    if (count <= 0 || nmin < 1 || nmax < nmin || bmin < 1 || bmax < bmin) return fail(PCD_E_INVALID, "bad batch shape");
    auto make = [&](int i, GenGraph &g) {
        SplitMix r(stream_seed(seed, (uint64_t)i, 0));
        double ln = std::log((double)nmin) + r.uniform() * (std::log((double)nmax) - std::log((double)nmin));
        int nn = std::min(nmax, std::max(nmin, (int)std::lround(std::exp(ln))));
        int B = bmin + (int)(r.next() % (uint64_t)(bmax - bmin + 1));
        double pin = pin_lo + r.uniform() * (pin_hi - pin_lo);
        double pout = pout_lo + r.uniform() * (pout_hi - pout_lo);
        int kind = (r.next() >> 63) ? PCD_RULE_CPM : PCD_RULE_MODULARITY;
        double lam = lam_lo + r.uniform() * (lam_hi - lam_lo);
        g.prm[0] = nn; g.prm[1] = B; g.prm[2] = pin; g.prm[3] = pout; g.prm[4] = kind; g.prm[5] = lam;
        std::vector<int32_t> bs(B, nn / B);
        for (int b = 0; b < nn % B; b++) bs[b]++;
        int64_t ne = 0;
        int32_t *ea = nullptr, *eb = nullptr;
        int nodes = 0;
        g.rc = pcd_gen_sbm(stream_seed(seed, (uint64_t)i, 1), B, bs.data(), pin, pout, &ne, &ea, &eb, &nodes);
        if (g.rc == PCD_OK) convert_into(g, nodes, ne, ea, eb);
        free(ea);
        free(eb);
    };
    return gen_many(count, make, n_out, m_out, ind_off, adj_off, indices, adj, params);
}

// Serving-form workload: request i draws from stream_seed(seed, i, 0) its
// graph family (SBM or R-MAT, 1/2 each), a target edge count log-uniform in
// [m_lo, m_hi] and its metric (modularity 0.6, CPM with resolution 0.01
// 0.3, WCC with the triangle-based initial partition 0.1).  SBM: n = m/8
// (average degree 16), blocks of 200, 80 % of the edges inside blocks.
// R-MAT: scale ceil(log2(m/8)), m draws, a/b/c = .57/.19/.19.  Edges from
// stream_seed(seed, i, 1), converted as pcd_convert_edges does.
// params[6 i ..] = {family (0 SBM, 1 R-MAT), target m, kind, lambda,
// initial partition, 0}.
extern "C" int pcd_gen_requests(uint64_t seed, int count, int64_t m_lo, int64_t m_hi, int32_t **n_out,
                                int32_t **m_out, int64_t **ind_off, int64_t **adj_off, int32_t **indices,
                                int32_t **adj, double **params) {
    if (count <= 0 || m_lo < 16 || m_hi < m_lo) return fail(PCD_E_INVALID, "bad request shape");
    auto make = [&](int i, GenGraph &g) {
        SplitMix r(stream_seed(seed, (uint64_t)i, 0));
        const int fam = (int)(r.next() >> 63);
        double lm = std::log((double)m_lo) + r.uniform() * (std::log((double)m_hi) - std::log((double)m_lo));
        const int64_t mt = std::max<int64_t>(m_lo, std::min<int64_t>(m_hi, std::llround(std::exp(lm))));
        const double u = r.uniform();
        const int kind = u < 0.6 ? PCD_RULE_MODULARITY : (u < 0.9 ? PCD_RULE_CPM : PCD_RULE_WCC);
        g.prm[0] = fam; g.prm[1] = (double)mt; g.prm[2] = kind; g.prm[3] = kind == PCD_RULE_CPM ? 0.01 : 0.0;
        g.prm[4] = kind == PCD_RULE_WCC ? 1 : 0;
        int64_t ne = 0;
        int32_t *ea = nullptr, *eb = nullptr;
        int nodes = 0;
        const uint64_t es = stream_seed(seed, (uint64_t)i, 1);
        if (fam == 0) {
            int64_t nn = std::max<int64_t>(32, mt / 8);
            int B = (int)std::max<int64_t>(2, nn / 200);
            std::vector<int32_t> bs(B, (int32_t)(nn / B));
            for (int b = 0; b < nn % B; b++) bs[b]++;
            double pairs_in = 0;
            for (int b = 0; b < B; b++) pairs_in += 0.5 * (double)bs[b] * (bs[b] - 1);
            double pairs_out = 0.5 * (double)nn * (nn - 1) - pairs_in;
            double pin = std::min(1.0, 0.8 * mt / std::max(1.0, pairs_in));
            double pout = std::min(1.0, 0.2 * mt / std::max(1.0, pairs_out));
            g.rc = pcd_gen_sbm(es, B, bs.data(), pin, pout, &ne, &ea, &eb, &nodes);
        } else {
            int scale = std::max(5, (int)std::ceil(std::log2(std::max(2.0, mt / 8.0))));
            g.rc = pcd_gen_rmat(es, scale, mt, 0.57, 0.19, 0.19, &ne, &ea, &eb, &nodes);
        }
        if (g.rc == PCD_OK) convert_into(g, nodes, ne, ea, eb);
        free(ea);
        free(eb);
    };
    return gen_many(count, make, n_out, m_out, ind_off, adj_off, indices, adj, params);
}

// framework.cpp:27-59 create_clustering_partition: vertices in order of
// descending local clustering coefficient tri/(d(d-1)) (int32 product as
// there), then descending degree, then id; each unlabelled vertex opens a
// community and takes its unlabelled neighbours.
extern "C" int pcd_clustering_partition(int n, const int32_t *indices, const int32_t *adj, const int32_t *tri,
                                        int32_t *labels) {
    if (n <= 0 || !indices || !adj || !tri || !labels) return fail(PCD_E_INVALID, "bad arguments");
    struct Key {
        double ncc;
        int ndeg, v;
    };
    std::vector<Key> keys(n);
    for (int v = 0; v < n; v++) {
        int deg = indices[v + 1] - indices[v];
        double cc = deg > 1 ? tri[v] / (double)(int)((unsigned)deg * (unsigned)(deg - 1)) : 0.0;
        keys[v] = Key{-cc, -deg, v};
    }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        if (a.ncc < b.ncc) return true;
        if (b.ncc < a.ncc) return false;
        if (a.ndeg != b.ndeg) return a.ndeg < b.ndeg;
        return a.v < b.v;
    });
    int nc = 0;
    for (int v = 0; v < n; v++) labels[v] = PCD_NEW_LABEL;
    for (const Key &k : keys) {
        if (labels[k.v] != PCD_NEW_LABEL) continue;
        labels[k.v] = nc;
        for (int j = indices[k.v]; j < indices[k.v + 1]; j++)
            if (labels[adj[j]] == PCD_NEW_LABEL) labels[adj[j]] = nc;
        nc++;
    }
    return PCD_OK;
}

// ----------------------------------------------------------------- NMI
// norm_mutual_info (common.cpp:75-127) over the confusion counts n_ij of two
// labelings: -2 sum n_ij log(n_ij N / (n_i n_j)) / (sum n_i log(n_i / N) +
// sum n_j log(n_j / N)), N = the graph's vertex count.  The reference keeps
// n_ij in a std::map, so the first sum runs over (i, j) ascending and the
// second over i, then j ascending; here the (i, j) keys are sorted once.
// n_i * n_j is an int32 product there and wraps from 2^31 on; it is kept so
// (a result above 1, or NaN, flags a run beyond that range).
static double nmi_sorted_pairs(int n, std::vector<uint64_t> &keys) {
    std::sort(keys.begin(), keys.end());
    struct Count {
        int id, count;
    };
    std::vector<Count> rows, cols;
    std::vector<int32_t> js(keys.size());
    for (size_t t = 0; t < keys.size(); t++) {
        const int i = (int)(keys[t] >> 32);
        if (rows.empty() || rows.back().id != i) rows.push_back(Count{i, 0});
        rows.back().count++;
        js[t] = (int32_t)(uint32_t)keys[t];
    }
    std::sort(js.begin(), js.end());
    for (int32_t j : js) {
        if (cols.empty() || cols.back().id != j) cols.push_back(Count{j, 0});
        cols.back().count++;
    }
    double nom = 0.0, den = 0.0;
    size_t row = 0;
    for (size_t a = 0; a < keys.size();) {
        size_t b = a;
        while (b < keys.size() && keys[b] == keys[a]) b++;
        const int i = (int)(keys[a] >> 32), j = (int)(uint32_t)keys[a], val = (int)(b - a);
        while (rows[row].id != i) row++;
        const Count &col = *std::lower_bound(cols.begin(), cols.end(), j,
                                             [](const Count &c, int x) { return c.id < x; });
        const int prod = (int)((unsigned)rows[row].count * (unsigned)col.count);
        nom += val * std::log(((double)val * n) / prod);
        a = b;
    }
    for (const Count &r : rows) den += r.count * std::log(r.count / ((double)n));
    for (const Count &c : cols) den += c.count * std::log(c.count / ((double)n));
    return -2 * nom / den;
}

// compare_communities(g, a, b) (common.cpp:165-173)
extern "C" int pcd_nmi(int n, const int32_t *a, const int32_t *b, double *nmi) {
    if (n <= 0 || !a || !b || !nmi) return fail(PCD_E_INVALID, "bad arguments");
    std::vector<uint64_t> keys((size_t)n);
    for (int v = 0; v < n; v++) {
        if (a[v] < 0 || b[v] < 0) return fail(PCD_E_INVALID, "labels must be >= 0");
        keys[v] = ((uint64_t)(uint32_t)a[v] << 32) | (uint32_t)b[v];
    }
    *nmi = nmi_sorted_pairs(n, keys);
    return PCD_OK;
}

// compare_communities(g, labels, filename) (common.cpp:129-163): line c
// (from 1, empty and '#' lines not counted) lists the vertices of ground
// truth community c, read as integers up to the first token that is not one.
extern "C" int pcd_nmi_file(int n, const int32_t *labels, const char *path, double *nmi) {
    if (n <= 0 || !labels || !path || !nmi) return fail(PCD_E_INVALID, "bad arguments");
    std::ifstream f(path);
    if (!f.is_open()) return fail(PCD_E_IO, std::string("failed to open ") + path);
    std::vector<uint64_t> keys;
    std::string line;
    int comm = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        comm++;
        std::istringstream ss(line);
        int v;
        while (ss >> v) {
            if (v < 0 || v >= n) return fail(PCD_E_FORMAT, "vertex " + std::to_string(v) + " outside [0, n)");
            if (labels[v] < 0) return fail(PCD_E_INVALID, "labels must be >= 0");
            keys.push_back(((uint64_t)(uint32_t)comm << 32) | (uint32_t)labels[v]);
        }
    }
    if (f.fail() && !f.eof()) return fail(PCD_E_IO, std::string("error while reading ") + path);
    *nmi = nmi_sorted_pairs(n, keys);
    return PCD_OK;
}
