/*
 * oracle.c -- CPU restatement of the Par-CD reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py.  No product code links, loads or calls it.
 *
 * Every function restates the reference C++/OpenMP path (the src/ C++ files of
 * stijnh/Par-CD) at one thread, citing the file:line it follows.  Two RNG
 * modes exist for the refinement steps:
 *
 *   mode 0 ("libc")    glibc rand()/rand_r() in the reference's call order,
 *                      first-insertion tie order of FastHashTable and
 *                      reservoir tie-breaking.  Reproduces `main -p 1`.
 *   mode 1 ("counter") vertex selection by a counter-based hash and a fixed
 *                      tie order (lowest label, or a hashed key).  This is the
 *                      serial replay the GPU path must match label for label.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   convert pipeline      golden bytes from the compiled reference `convert`
 *   count_triangles       golden tri/tri_deg/cc + brute force + networkx
 *   compress/statistics   golden harness dumps + brute force
 *   calculate (mod/cpm)   golden harness dumps + networkx modularity +
 *                         exact integer closed forms
 *   calculate (wcc)       golden harness dumps + brute-force definition
 *   score_* (all rules)   golden harness dumps (incl. int32 overflow cases)
 *                         + join+leave == delta(calculate) invariant
 *   refine (libc mode)    golden `main -p 1` label dumps + harness refine dumps
 *   refine (counter mode) shares all code with libc mode except the RNG and
 *                         the tie order; its RNG is pinned by known-answer
 *                         values of the hash (tests/test_oracle_rng.py).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NEW_LABEL (-1) /* common.hpp:20  NEW_LABEL (~0) */

enum { ORC_MOD = 0, ORC_CPM = 1, ORC_WCC = 2, ORC_COND = 3 };
enum { ORC_TIE_RANDOM = 0, ORC_TIE_LOWEST = 1 };

/* ------------------------------------------------------------------------ */
/* counter-based RNG (specification shared, by value, with the GPU path)    */
/* ------------------------------------------------------------------------ */

uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_hash3(uint64_t seed, uint32_t a, uint32_t b) {
    uint64_t x = (((uint64_t)a << 32) | (uint64_t)b) + 1ULL;
    return orc_mix64(seed + 0x9E3779B97F4A7C15ULL * x);
}

/* reference: refinement.cpp:116,129 `f = min(frac,1)*RAND_MAX; rand_r() > f -> skip` */
int orc_frac_threshold(double frac) {
    double f = frac < 1.0 ? frac : 1.0;
    return (int)(f * 2147483647.0);
}

int orc_select(uint64_t seed, int step, int v, int f) {
    uint32_t r = (uint32_t)(orc_hash3(seed, (uint32_t)step, (uint32_t)v) >> 33);
    return (int)r <= f;
}

uint32_t orc_tie_key(int tie, uint64_t seed, int step, int v, int label) {
    if (tie == ORC_TIE_LOWEST) return (uint32_t)label;
    uint64_t base = orc_hash3(seed ^ 0x5851F42D4C957F2DULL, (uint32_t)step, (uint32_t)v);
    return (uint32_t)(orc_mix64(base ^ ((uint64_t)(uint32_t)label * 0xD6E8FEB86659FD93ULL)) >> 32);
}

uint64_t orc_merge_seed(uint64_t seed, int round) {
    return orc_mix64(seed ^ (0xA0761D6478BD642FULL + (uint64_t)round));
}

/* true if candidate (s, key, label) beats (bs, bkey, blabel) */
static int better(double s, uint32_t key, int label, double bs, uint32_t bkey, int blabel) {
    if (s > bs) return 1;
    if (s < bs) return 0;
    if (key != bkey) return key < bkey;
    return label < blabel;
}

/* ------------------------------------------------------------------------ */
/* insertion-ordered counting table (refinement.cpp:24-110 FastHashTable:   */
/* the only observable property is first-insertion iteration order)         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int cap, mask, size;
    int *slot_key, *slot_val, *order; /* order[i] = slot of i-th inserted key */
} otable;

static void ot_init(otable *t, int cap) {
    int c = 16;
    while (c < 2 * cap) c <<= 1;
    t->cap = c; t->mask = c - 1; t->size = 0;
    t->slot_key = (int *)malloc(sizeof(int) * c);
    t->slot_val = (int *)malloc(sizeof(int) * c);
    t->order = (int *)malloc(sizeof(int) * c);
    for (int i = 0; i < c; i++) t->slot_key[i] = -1;
}
static void ot_free(otable *t) { free(t->slot_key); free(t->slot_val); free(t->order); }
static void ot_grow(otable *t) {
    otable n2; ot_init(&n2, t->cap);
    for (int i = 0; i < t->size; i++) {
        int k = t->slot_key[t->order[i]], v = t->slot_val[t->order[i]];
        int h = (int)(((uint32_t)k * 2654435761u) & (uint32_t)n2.mask);
        while (n2.slot_key[h] != -1) h = (h + 1) & n2.mask;
        n2.slot_key[h] = k; n2.slot_val[h] = v; n2.order[n2.size++] = h;
    }
    ot_free(t); *t = n2;
}
static void ot_insert(otable *t, int key) {
    int h = (int)(((uint32_t)key * 2654435761u) & (uint32_t)t->mask);
    while (t->slot_key[h] != -1 && t->slot_key[h] != key) h = (h + 1) & t->mask;
    if (t->slot_key[h] == key) { t->slot_val[h]++; return; }
    t->slot_key[h] = key; t->slot_val[h] = 1; t->order[t->size++] = h;
    if (2 * t->size > t->cap) ot_grow(t);
}
static void ot_clear(otable *t) {
    for (int i = 0; i < t->size; i++) t->slot_key[t->order[i]] = -1;
    t->size = 0;
}
#define OT_KEY(t, i) ((t)->slot_key[(t)->order[i]])
#define OT_VAL(t, i) ((t)->slot_val[(t)->order[i]])

/* ------------------------------------------------------------------------ */
/* convert pipeline (convert.cpp)                                           */
/* ------------------------------------------------------------------------ */

/* convert.cpp:14-74 read_edges. Returns 0 ok, -1 corrupt line, -2 io error.
 * Edge arrays grow by realloc; caller frees with orc_free. */
int orc_read_edges_text(const char *path, int **ea, int **eb, long *ne, int *num_nodes, long *bad_line) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -2;
    long cap = 1024, cnt = 0;
    int *a = (int *)malloc(sizeof(int) * cap), *b = (int *)malloc(sizeof(int) * cap);
    char line[1024];
    int nn = 0;
    long lineno = 0;
    int rc = 0;
    /* std::istream::getline(line,1024): reads up to 1023 chars; a longer line
       sets failbit, which ends the loop (convert.cpp:20) and is an error
       (convert.cpp:68-71). */
    for (;;) {
        lineno++;
        int len = 0, ch = 0, overflow = 0;
        while ((ch = fgetc(fp)) != EOF && ch != '\n') {
            if (len < 1023) line[len++] = (char)ch; else { overflow = 1; break; }
        }
        line[len] = 0;
        if (overflow) { rc = -2; break; }
        int at_eof = (ch == EOF);
        if (at_eof && len == 0) break;
        int idx = 0;
        while (line[idx] == ' ' || (line[idx] >= '\t' && line[idx] <= '\r')) idx++;
        if (line[idx] == 0 || line[idx] == '#') { if (at_eof) break; continue; }
        if (!(line[idx] >= '0' && line[idx] <= '9')) { rc = -1; *bad_line = lineno; break; }
        int x = 0;
        while (line[idx] >= '0' && line[idx] <= '9') { x = (int)(10u * (unsigned)x + (unsigned)(line[idx] - '0')); idx++; }
        while (line[idx] == ' ' || (line[idx] >= '\t' && line[idx] <= '\r')) idx++;
        if (!(line[idx] >= '0' && line[idx] <= '9')) { rc = -1; *bad_line = lineno; break; }
        int y = 0;
        while (line[idx] >= '0' && line[idx] <= '9') { y = (int)(10u * (unsigned)y + (unsigned)(line[idx] - '0')); idx++; }
        if (cnt == cap) { cap *= 2; a = (int *)realloc(a, sizeof(int) * cap); b = (int *)realloc(b, sizeof(int) * cap); }
        a[cnt] = x; b[cnt] = y; cnt++;
        int mx = (x > y ? x : y) + 1;
        if (mx > nn) nn = mx;
        if (at_eof) break;
    }
    fclose(fp);
    *ea = a; *eb = b; *ne = cnt; *num_nodes = nn;
    return rc;
}

void orc_free(void *p) { free(p); }

static int cmp_int(const void *x, const void *y) {
    int a = *(const int *)x, b = *(const int *)y;
    return (a > b) - (a < b);
}

typedef struct { int deg, id; } degid;
/* convert.cpp:150 sort(..., greater<pair<int,int>>()): degree desc, id desc */
static int cmp_degid_desc(const void *x, const void *y) {
    const degid *a = (const degid *)x, *b = (const degid *)y;
    if (a->deg != b->deg) return a->deg > b->deg ? -1 : 1;
    if (a->id != b->id) return a->id > b->id ? -1 : 1;
    return 0;
}

/* convert.cpp:76-278 (convert_to_csr, remove_duplicates_and_loops,
 * sort_on_degree, remove_isolated_vertices, write_edges layout).
 * Inputs must not contain self-loop lines: convert.cpp:94-103 leaves an
 * uninitialised adjacency slot for them (undefined output), so they are
 * rejected with -3.  Outputs: indices[n+1], adj[2m], map[n] (caller buffers
 * of num_nodes+1, 2*ne, num_nodes entries). */
int orc_convert_edges(int num_nodes, long ne, const int *ea, const int *eb,
                      int *indices_out, int *adj_out, int *map_out, int *n_out, int *m_out) {
    for (long i = 0; i < ne; i++) if (ea[i] == eb[i]) return -3;
    int *deg = (int *)calloc((size_t)num_nodes + 1, sizeof(int));
    int *ind = (int *)malloc(sizeof(int) * ((size_t)num_nodes + 1));
    int *adj = (int *)malloc(sizeof(int) * (size_t)(2 * ne + 1));
    /* convert.cpp:83-103 */
    for (long i = 0; i < ne; i++) { deg[ea[i]]++; deg[eb[i]]++; }
    ind[0] = 0;
    for (int i = 0; i < num_nodes; i++) { ind[i + 1] = ind[i] + deg[i]; deg[i] = 0; }
    for (long i = 0; i < ne; i++) {
        int a = ea[i], b = eb[i];
        adj[ind[a] + deg[a]] = b;
        adj[ind[b] + deg[b]] = a;
        deg[a]++; deg[b]++;
    }
    free(deg);
    /* convert.cpp:109-133 */
    int index = 0;
    for (int i = 0; i < num_nodes; i++) {
        int start = ind[i], end = ind[i + 1];
        ind[i] = index;
        qsort(adj + start, (size_t)(end - start), sizeof(int), cmp_int);
        int prev = -1;
        for (int j = start; j < end; j++) {
            int v = adj[j];
            if (v != i && v != prev) adj[index++] = v;
            prev = v;
        }
    }
    ind[num_nodes] = index;
    /* convert.cpp:136-182 */
    degid *di = (degid *)malloc(sizeof(degid) * ((size_t)num_nodes + 1));
    for (int i = 0; i < num_nodes; i++) { di[i].deg = ind[i + 1] - ind[i]; di[i].id = i; }
    qsort(di, (size_t)num_nodes, sizeof(degid), cmp_degid_desc);
    int *fmap = (int *)malloc(sizeof(int) * ((size_t)num_nodes + 1));
    for (int i = 0; i < num_nodes; i++) { fmap[di[i].id] = i; map_out[i] = di[i].id; }
    int ni = 0;
    for (int i = 0; i < num_nodes; i++) {
        int v = map_out[i];
        indices_out[i] = ni;
        for (int j = ind[v]; j < ind[v + 1]; j++) adj_out[ni++] = fmap[adj[j]];
    }
    indices_out[num_nodes] = ni;
    for (int i = 0; i < num_nodes; i++)
        qsort(adj_out + indices_out[i], (size_t)(indices_out[i + 1] - indices_out[i]), sizeof(int), cmp_int);
    /* convert.cpp:184-196 */
    int n = num_nodes;
    for (int i = 0; i < num_nodes; i++) if (indices_out[i + 1] - indices_out[i] == 0) { n = i; break; }
    free(di); free(fmap); free(ind); free(adj);
    *n_out = n;
    *m_out = indices_out[n] / 2; /* convert.cpp:245 */
    return 0;
}

/* convert.cpp:244-278 write_edges (file part only; the stdout dump is omitted) */
int orc_write_pcd(const char *path, int n, int m, const int *indices, const int *adj, const int *map) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return -2;
    fwrite("PCD", 1, 3, fp);
    fwrite(&n, 4, 1, fp);
    fwrite(&m, 4, 1, fp);
    fwrite(indices, 4, (size_t)n + 1, fp);
    fwrite(adj, 4, (size_t)2 * m, fp);
    fwrite(map, 4, (size_t)n, fp);
    int bad = ferror(fp);
    fclose(fp);
    return bad ? -2 : 0;
}

/* graph.cpp:40-105 Graph::read. Returns 0 ok, <0 error; arrays via malloc. */
int orc_read_pcd(const char *path, int *n_out, int *m_out, int **indices, int **adj, int **map) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return -2;
    char w[3];
    int n = 0, m = 0;
    if (fread(w, 1, 3, fp) != 3) { fclose(fp); return -4; }
    if (memcmp(w, "PCD", 3) != 0) { fclose(fp); return -5; }
    if (fread(&n, 4, 1, fp) != 1 || fread(&m, 4, 1, fp) != 1 || n < 0 || m < 0) { fclose(fp); return -4; }
    int *ind = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *a = (int *)malloc(sizeof(int) * ((size_t)2 * m + 1));
    int *mp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int err = fread(ind, 4, (size_t)n + 1, fp) != (size_t)n + 1;
    err |= fread(a, 4, (size_t)2 * m, fp) != (size_t)2 * m;
    err |= fread(mp, 4, (size_t)n, fp) != (size_t)n;
    if (!err && fgetc(fp) != EOF) err = 1; /* graph.cpp:90-93 trailing bytes */
    fclose(fp);
    if (err) { free(ind); free(a); free(mp); return -4; }
    *n_out = n; *m_out = m; *indices = ind; *adj = a; *map = mp;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* triangles (graph.cpp:107-175 Graph::count_triangles)                     */
/* ------------------------------------------------------------------------ */

void orc_count_triangles(int n, const int *indices, const int *adj, int *et, int *tri, int *tri_deg, double *cc) {
    long E = indices[n];
    for (long i = 0; i < E; i++) et[i] = 0;
    for (int i = 0; i < n; i++) { tri[i] = 0; tri_deg[i] = 0; }
    /* graph.cpp:117-144 */
    for (int v = 0; v < n; v++) {
        for (int r = indices[v]; r < indices[v + 1]; r++) {
            int u = adj[r];
            if (u >= v) break;
            int x = indices[v], xe = indices[v + 1], y = indices[u], ye = indices[u + 1];
            while (x != xe && y != ye && adj[x] < u && adj[y] < u) {
                int d = adj[x] - adj[y];
                if (d == 0) { et[x]++; et[y]++; et[r]++; }
                if (d <= 0) x++;
                if (d >= 0) y++;
            }
        }
    }
    /* graph.cpp:146-159 */
    for (int v = 0; v < n; v++) {
        for (int r = indices[v]; r < indices[v + 1]; r++) {
            int u = adj[r];
            int t = et[r];
            if (t > 0) { tri[v] += t; tri[u] += t; tri_deg[v]++; tri_deg[u]++; }
        }
    }
    /* graph.cpp:161-169, num_threads = 1 (graph.hpp:23 default): sequential */
    double c = 0;
    for (int v = 0; v < n; v++) {
        int t = tri[v];
        int deg = indices[v + 1] - indices[v];
        if (deg > 1) c += (t / (deg * (deg - 1.0))) / n;
    }
    *cc = c;
}

/* ------------------------------------------------------------------------ */
/* partitions (framework.cpp)                                               */
/* ------------------------------------------------------------------------ */

void orc_srand(unsigned s) { srand(s); }
int orc_rand(void) { return rand(); }

/* framework.cpp:14-24 */
void orc_singleton_partition(int n, int *labels) {
    for (int v = 0; v < n; v++) labels[v] = v;
    for (int v = 0; v < n; v++) {
        int j = rand() % n;
        int t = labels[v]; labels[v] = labels[j]; labels[j] = t;
    }
}

typedef struct { double ncc; int ndeg; int v; } ccpair;
static int cmp_ccpair(const void *x, const void *y) {
    const ccpair *a = (const ccpair *)x, *b = (const ccpair *)y;
    if (a->ncc < b->ncc) return -1;
    if (b->ncc < a->ncc) return 1;
    if (a->ndeg != b->ndeg) return a->ndeg < b->ndeg ? -1 : 1;
    return (a->v > b->v) - (a->v < b->v);
}

static int wrap_mul(int a, int b) { return (int)((unsigned)a * (unsigned)b); }

/* framework.cpp:27-59 (deg*(deg-1) is an int32 product in the reference) */
void orc_clustering_partition(int n, const int *indices, const int *adj, const int *tri, int *labels) {
    ccpair *p = (ccpair *)malloc(sizeof(ccpair) * ((size_t)n + 1));
    for (int v = 0; v < n; v++) {
        int deg = indices[v + 1] - indices[v];
        double cc = deg > 1 ? tri[v] / (double)wrap_mul(deg, deg - 1) : 0.0;
        p[v].ncc = -cc; p[v].ndeg = -deg; p[v].v = v;
    }
    qsort(p, (size_t)n, sizeof(ccpair), cmp_ccpair);
    int nc = 0;
    for (int i = 0; i < n; i++) labels[i] = ORC_NEW_LABEL;
    for (int i = 0; i < n; i++) {
        int v = p[i].v;
        if (labels[v] == ORC_NEW_LABEL) {
            labels[v] = nc;
            for (int j = indices[v]; j < indices[v + 1]; j++)
                if (labels[adj[j]] == ORC_NEW_LABEL) labels[adj[j]] = nc;
            nc++;
        }
    }
    free(p);
}

/* framework.cpp:146-170 */
int orc_compress_labels(int n, int *labels) {
    int mx = -1;
    for (int v = 0; v < n; v++) if (labels[v] > mx) mx = labels[v];
    int nl = mx + 1, next = 0;
    int *re = (int *)malloc(sizeof(int) * ((size_t)nl + 1));
    for (int i = 0; i < nl; i++) re[i] = ORC_NEW_LABEL;
    for (int v = 0; v < n; v++) {
        int o = labels[v], nw;
        if (o == ORC_NEW_LABEL) nw = next++;
        else if (re[o] != ORC_NEW_LABEL) nw = re[o];
        else nw = re[o] = next++;
        labels[v] = nw;
    }
    free(re);
    return next;
}

/* framework.cpp:69-100 (non-atomic path; identical results to :103-143).
 * Returns k = max label + 1. Output arrays must hold k entries. */
int orc_collect_statistics(int n, const int *indices, const int *adj, const int *labels,
                           int *sizes, int *intv, int *extv) {
    int k = -1;
    for (int v = 0; v < n; v++) if (labels[v] > k) k = labels[v];
    k += 1;
    for (int c = 0; c < k; c++) { sizes[c] = 0; intv[c] = 0; extv[c] = 0; }
    for (int i = 0; i < n; i++) {
        int l = labels[i], di = 0, de = 0;
        for (int j = indices[i]; j < indices[i + 1]; j++) {
            if (l == labels[adj[j]]) di++; else de++;
        }
        sizes[l]++; intv[l] += di; extv[l] += de;
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* rules (rules.cpp)                                                        */
/* ------------------------------------------------------------------------ */

/* rules.cpp:240-389 calculate_wcc_improvement; int32 sub-products wrap as
 * in the reference (rules.cpp:243, 265, 267-268, 297, 326, 328-329, 358). */
double orc_wcc_improvement(int n_x, int n_y, int c_x, int c_y, int m_x, int m_y, int m_xy,
                           int b_x, int b_y, double cc) {
    double p_x = n_x > 1 ? m_x / (n_x * (n_x - 1.0) / 2.0) : 0.0;
    double p_y = n_y > 1 ? m_y / (n_y * (n_y - 1.0) / 2.0) : 0.0;
    double p_xy = c_x > 0 && c_y > 0 ? m_xy / (double)wrap_mul(c_x, c_y) : 0.0;
    double q_x = n_x > 0 ? (b_x - m_xy) / (double)n_x : 0;
    double q_y = n_y > 0 ? (b_y - m_xy) / (double)n_y : 0;
    double score = 0.0;
    if (c_x > 0) {
        double t_x = (wrap_mul(n_x - 1, n_x - 2) / 2) * p_x * p_x * p_x;
        double t_xy = t_x + wrap_mul(c_x - 1, c_y) * p_x * p_y * p_xy
                      + (wrap_mul(c_y, c_y - 1) / 2) * p_xy * p_xy * p_y;
        double t_v = t_xy + q_x * (q_x - 1) / 2 * cc + q_x * (n_x - 1) * p_x * cc + c_y * q_x * p_xy * cc;
        double vt_xy = q_x;
        double vt_x = vt_xy + p_xy * c_y;
        double vt_v = vt_x + p_x * (n_x - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_x + n_y - 1 + vt_xy)) - (vt_v / t_v) * (t_x / (n_x - 1 + vt_x));
        if (t_v > 0) score += c_x * tr;
    }
    if (c_x < n_x) {
        double t_x = (wrap_mul(n_x - 1, n_x - 2) / 2) * p_x * p_x * p_x;
        double t_xy = t_x;
        double t_v = t_xy + q_x * (q_x - 1) / 2 * cc + q_x * (n_x - 1) * p_x * cc;
        double vt_xy = q_x;
        double vt_x = vt_xy;
        double vt_v = vt_x + p_x * (n_x - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_x + n_y - 1 + vt_xy)) - (vt_v / t_v) * (t_x / (n_x - 1 + vt_x));
        if (t_v > 0) score += (n_x - c_x) * tr;
    }
    if (c_y > 0) {
        double t_y = (wrap_mul(n_y - 1, n_y - 2) / 2) * p_y * p_y * p_y;
        double t_xy = t_y + wrap_mul(c_y - 1, c_x) * p_y * p_x * p_xy
                      + (wrap_mul(c_x, c_x - 1) / 2) * p_xy * p_xy * p_x;
        double t_v = t_xy + q_y * (q_y - 1) / 2 * cc + q_y * (n_y - 1) * p_y * cc + c_x * q_y * p_xy * cc;
        double vt_xy = q_y;
        double vt_y = vt_xy + p_xy * c_x;
        double vt_v = vt_y + p_y * (n_y - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_y + n_x - 1 + vt_xy)) - (vt_v / t_v) * (t_y / (n_y - 1 + vt_y));
        if (t_v > 0) score += c_y * tr;
    }
    if (c_y < n_y) {
        double t_y = (wrap_mul(n_y - 1, n_y - 2) / 2) * p_y * p_y * p_y;
        double t_xy = t_y;
        double t_v = t_xy + q_y * (q_y - 1) / 2 * cc + q_y * (n_y - 1) * p_y * cc;
        double vt_xy = q_y;
        double vt_y = vt_xy;
        double vt_v = vt_y + p_y * (n_y - 1);
        double tr = (vt_v / t_v) * (t_xy / (n_y + n_x - 1 + vt_xy)) - (vt_v / t_v) * (t_y / (n_y - 1 + vt_y));
        if (t_v > 0) score += (n_y - c_y) * tr;
    }
    return score;
}

/* Scalar gain functions.  Arguments carry exactly the values the reference
 * reads from Graph/Partition. */

/* rules.cpp:34-43 (mod), 136-139 (cpm), 403-417 (wcc), 93-102 (cond) */
double orc_score_join(int kind, double lambda, int m, double cc, int deg,
                      int size_c, int int_c, int ext_c, int num) {
    switch (kind) {
    case ORC_MOD: {
        double e_ij = num / (double)m;
        double a_i = (int_c + ext_c) / (2.0 * m);
        double a_j = deg / (2.0 * m);
        return e_ij - 2 * a_i * a_j;
    }
    case ORC_CPM: return num - lambda * size_c;
    case ORC_WCC: return orc_wcc_improvement(1, size_c, 1, num, 0, int_c / 2, num, deg, ext_c, cc);
    default: {
        int s = size_c, i = int_c, v = i + ext_c;
        return wrap_mul(s + 1, i + 2 * num) / (double)(v + deg) - s * (i / (double)v);
    }
    }
}

/* rules.cpp:45-56 (mod), 141-144 (cpm), 419-436 (wcc), 104-114 (cond) */
double orc_score_leave(int kind, double lambda, int m, double cc, int deg,
                       int size_own, int int_own, int ext_own, int num) {
    switch (kind) {
    case ORC_MOD: {
        double e_ij = num / (double)m;
        double a_i = (int_own + ext_own - deg) / (2.0 * m);
        double a_j = deg / (2.0 * m);
        return -(e_ij - 2 * a_i * a_j);
    }
    case ORC_CPM: return -(num - lambda * (size_own - 1));
    case ORC_WCC:
        return -orc_wcc_improvement(1, size_own - 1, 1, num, 0, int_own / 2 - num, num, deg,
                                    ext_own - deg + 2 * num, cc);
    default: {
        int s = size_own, i = int_own, v = i + ext_own;
        return wrap_mul(s - 1, i - 2 * num) / (double)(v - deg) - s * (i / (double)v);
    }
    }
}

/* rules.cpp:24-32 (mod), 131-134 (cpm), 391-401 (wcc), 76-91 (cond) */
double orc_score_merge(int kind, double lambda, int m, double cc,
                       int size_i, int int_i, int ext_i, int size_j, int int_j, int ext_j, int links) {
    switch (kind) {
    case ORC_MOD: {
        double e_ij = links / (double)m;
        double a_i = (int_i + ext_i) / (2.0 * m);
        double a_j = (int_j + ext_j) / (2.0 * m);
        return e_ij - 2 * a_i * a_j;
    }
    case ORC_CPM: return links - lambda * size_i * size_j;
    case ORC_WCC: {
        int cx = links < size_i ? links : size_i;
        int cy = links < size_j ? links : size_j;
        return orc_wcc_improvement(size_i, size_j, cx, cy, int_i / 2, int_j / 2, links, ext_i, ext_j, cc);
    }
    default: {
        int si = size_i, sj = size_j, ii = int_i, ij = int_j, vi = ii + ext_i, vj = ij + ext_j;
        return wrap_mul(si + sj, ii + ij + links) / (double)(vi + vj + links)
               - wrap_mul(si, ii) / (double)vi - wrap_mul(sj, ij) / (double)vj;
    }
    }
}

/* rules.cpp:150-235 WCCRule::calculate (sequential; the reference sums with
 * an OpenMP reduction, so its last bits depend on the thread count) */
double orc_wcc_calculate(int n, const int *indices, const int *adj, const int *labels,
                         const int *sizes, const int *tri, const int *tri_deg, int *it_out, int *itd_out) {
    long E = indices[n];
    int *eit = (int *)calloc((size_t)E + 1, sizeof(int));
    int *it = it_out ? it_out : (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *itd = itd_out ? itd_out : (int *)malloc(sizeof(int) * ((size_t)n + 1));
    for (int i = 0; i < n; i++) { it[i] = 0; itd[i] = 0; }
    for (int v = 0; v < n; v++) {
        int l = labels[v];
        for (int r = indices[v]; r < indices[v + 1]; r++) {
            int u = adj[r];
            if (u >= v) break;
            if (l != labels[u]) continue;
            int x = indices[v], xe = indices[v + 1], y = indices[u], ye = indices[u + 1];
            while (x != xe && adj[x] < u && y != ye && adj[y] < u) {
                int d = adj[x] - adj[y];
                if (d == 0 && labels[adj[x]] == l) { eit[x]++; eit[y]++; eit[r]++; }
                if (d <= 0) x++;
                if (d >= 0) y++;
            }
        }
    }
    for (int v = 0; v < n; v++) {
        for (int r = indices[v]; r < indices[v + 1]; r++) {
            int u = adj[r], t = eit[r];
            if (t > 0) { it[v] += t; itd[v]++; it[u] += t; itd[u]++; }
        }
    }
    double score = 0.0;
    for (int v = 0; v < n; v++) {
        int size = sizes[labels[v]], t = tri[v], dt = tri_deg[v], itr = it[v], ditr = itd[v];
        if (t > 0) {
            double a = itr / (double)t;
            double b = dt / (double)(size - 1 + dt - ditr);
            score += (a * b) / n;
        }
    }
    free(eit);
    if (!it_out) free(it);
    if (!itd_out) free(itd);
    return score;
}

/* rules.cpp:9-22 (mod), 121-129 (cpm), 58-70 (cond) plus WCC above */
double orc_calculate(int kind, double lambda, int n, int m, const int *indices, const int *adj,
                     const int *labels, int k, const int *sizes, const int *intv, const int *extv,
                     const int *tri, const int *tri_deg) {
    double s = 0.0;
    switch (kind) {
    case ORC_MOD:
        for (int i = 0; i < k; i++) {
            double actual = intv[i] / (2.0 * m);
            double e = (intv[i] + extv[i]) / (2.0 * m);
            s += actual - e * e;
        }
        return s;
    case ORC_CPM:
        for (int i = 0; i < k; i++) s += intv[i] - lambda * sizes[i] * (sizes[i] - 1);
        return s / (double)(2 * m);
    case ORC_WCC:
        return orc_wcc_calculate(n, indices, adj, labels, sizes, tri, tri_deg, NULL, NULL);
    default:
        for (int c = 0; c < k; c++) s += sizes[c] * (intv[c] / (double)(intv[c] + extv[c]));
        return s / n;
    }
}

/* ------------------------------------------------------------------------ */
/* refinement (refinement.cpp)                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    int kind; double lambda; int m; double cc;
    const int *indices, *adj, *labels, *sizes, *intv, *extv;
} orc_view;

static double join_v(const orc_view *g, int v, int c, int num) {
    return orc_score_join(g->kind, g->lambda, g->m, g->cc, g->indices[v + 1] - g->indices[v],
                          g->sizes[c], g->intv[c], g->extv[c], num);
}
static double leave_v(const orc_view *g, int v, int num) {
    int c = g->labels[v];
    return orc_score_leave(g->kind, g->lambda, g->m, g->cc, g->indices[v + 1] - g->indices[v],
                           g->sizes[c], g->intv[c], g->extv[c], num);
}
static double merge_v(const orc_view *g, int i, int j, int links) {
    return orc_score_merge(g->kind, g->lambda, g->m, g->cc, g->sizes[i], g->intv[i], g->extv[i],
                           g->sizes[j], g->intv[j], g->extv[j], links);
}

/* refinement.cpp:114-182 PropagationRefinement::refine at one thread.
 * mode 0: libc RNG (one rand() per call for the thread seed, rand_r per
 *         vertex and per tie); mode 1: counter RNG + fixed tie order. */
void orc_propagation_refine(int kind, double lambda, int n, int m, const int *indices, const int *adj,
                            const int *labels, const int *sizes, const int *intv, const int *extv,
                            double cc, double frac, int mode, uint64_t seed, int step, int tie,
                            int *new_labels) {
    orc_view g = {kind, lambda, m, cc, indices, adj, labels, sizes, intv, extv};
    int f = orc_frac_threshold(frac);
    unsigned int rseed = 0;
    if (mode == 0) rseed = (unsigned)rand();
    otable t;
    ot_init(&t, 4096);
    for (int v = 0; v < n; v++) {
        int active = mode == 0 ? !(rand_r(&rseed) > f) : orc_select(seed, step, v, f);
        if (!active) { new_labels[v] = labels[v]; continue; }
        for (int j = indices[v]; j < indices[v + 1]; j++) ot_insert(&t, labels[adj[j]]);
        int curr = labels[v], best_label = -1, best_count = 0, curr_freq = 0;
        double best = 0.0;
        uint32_t best_key = 0xFFFFFFFFu;
        for (int i = 0; i < t.size; i++) {
            int label = OT_KEY(&t, i), freq = OT_VAL(&t, i);
            if (label != curr) {
                double s = join_v(&g, v, label, freq);
                if (mode == 0) {
                    if (s > best) { best_label = label; best = s; best_count = 1; }
                    else if (s == best) { if (rand_r(&rseed) % (unsigned)(++best_count) == 0) best_label = label; }
                } else if (s > 0.0) {
                    uint32_t key = orc_tie_key(tie, seed, step, v, label);
                    if (best_label == -1 || better(s, key, label, best, best_key, best_label)) {
                        best = s; best_key = key; best_label = label;
                    }
                }
            } else {
                curr_freq = freq;
            }
        }
        ot_clear(&t);
        double leave = leave_v(&g, v, curr_freq);
        best += leave;
        if (best > 0 && best > leave) new_labels[v] = best_label;
        else if (leave > 0) new_labels[v] = ORC_NEW_LABEL;
        else new_labels[v] = curr;
    }
    ot_free(&t);
}

/* refinement.cpp:185-288 MergingRefinement::refine at one thread.
 * k = number of labels (labels in [0,k)).  Optional outputs: best_out[k]
 * (partner, -1 for none) and found_out[k] after the last round. */
void orc_merging_refine(int kind, double lambda, int n, int m, int k, const int *indices, const int *adj,
                        const int *labels, const int *sizes, const int *intv, const int *extv,
                        double cc, int rounds, int mode, uint64_t seed, int step, int tie,
                        int *new_labels, int *best_out, int *found_out) {
    orc_view g = {kind, lambda, m, cc, indices, adj, labels, sizes, intv, extv};
    int *ci = (int *)malloc(sizeof(int) * ((size_t)k + 2));
    int *cm = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    ci[0] = 0;
    if (k >= 1) ci[1] = 0;
    for (int i = 2; i < k + 1; i++) ci[i] = ci[i - 1] + sizes[i - 2];
    for (int v = 0; v < n; v++) cm[ci[labels[v] + 1]++] = v;
    int *bp = (int *)malloc(sizeof(int) * ((size_t)k + 1));
    char *fp = (char *)calloc((size_t)k + 1, 1);
    for (int l = 0; l < k; l++) bp[l] = -1;
    unsigned int rseed = 0;
    if (mode == 0) rseed = (unsigned)rand();
    otable t;
    ot_init(&t, 4096);
    int all_done = 0;
    for (int rnd = 0; rnd < rounds && !all_done; rnd++) {
        uint64_t rs = orc_merge_seed(seed, rnd);
        for (int l = 0; l < k; l++) {
            if (fp[l] || (bp[l] != -1 && !fp[bp[l]])) continue;
            for (int i = ci[l]; i != ci[l + 1]; i++) {
                int v = cm[i];
                for (int j = indices[v]; j < indices[v + 1]; j++) {
                    int kk = labels[adj[j]];
                    if (kk != l && !fp[kk]) ot_insert(&t, kk);
                }
            }
            int best_label = -1, count = 1;
            double best = 0;
            uint32_t best_key = 0xFFFFFFFFu;
            for (int i = 0; i < t.size; i++) {
                int label = OT_KEY(&t, i), freq = OT_VAL(&t, i);
                double s = merge_v(&g, l, label, freq);
                if (mode == 0) {
                    if (s == best) { if (rand_r(&rseed) % (unsigned)(count++) == 0) best_label = label; }
                    else if (s > best) { best_label = label; best = s; count = 1; }
                } else if (s >= 0.0) {
                    uint32_t key = orc_tie_key(tie, rs, step, l, label);
                    if (best_label == -1 || better(s, key, label, best, best_key, best_label)) {
                        best = s; best_key = key; best_label = label;
                    }
                }
            }
            bp[l] = best_label;
            ot_clear(&t);
        }
        all_done = 1;
        for (int l = 0; l < k; l++) {
            if (!fp[l] && (bp[l] == -1 || bp[bp[l]] == l)) fp[l] = 1;
            all_done &= fp[l];
        }
    }
    for (int v = 0; v < n; v++) {
        int l = labels[v];
        if (fp[l] && bp[l] != -1) new_labels[v] = l < bp[l] ? l : bp[l];
        else new_labels[v] = l;
    }
    if (best_out) for (int l = 0; l < k; l++) best_out[l] = bp[l];
    if (found_out) for (int l = 0; l < k; l++) found_out[l] = fp[l];
    ot_free(&t);
    free(ci); free(cm); free(bp); free(fp);
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Community-pair link table (the integer content of contraction):
 * for each community l, the distinct neighbour communities k != l of its
 * members and their link counts (refinement.cpp:228-238 with no community
 * matched yet).  Output is sorted by (l, k).  Returns U, the number of
 * pairs; arrays must hold up to indices[n] entries. */
long orc_community_pairs(int n, const int *indices, const int *adj, const int *labels, int k,
                         int *out_l, int *out_k, int *out_links) {
    long E = indices[n];
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)E + 1));
    long P = 0;
    for (int v = 0; v < n; v++) {
        int l = labels[v];
        for (int j = indices[v]; j < indices[v + 1]; j++) {
            int kk = labels[adj[j]];
            if (kk != l) keys[P++] = ((uint64_t)(uint32_t)l << 32) | (uint32_t)kk;
        }
    }
    qsort(keys, (size_t)P, sizeof(uint64_t), cmp_u64);
    long U = 0;
    for (long i = 0; i < P;) {
        long j = i;
        while (j < P && keys[j] == keys[i]) j++;
        out_l[U] = (int)(keys[i] >> 32);
        out_k[U] = (int)(uint32_t)keys[i];
        out_links[U] = (int)(j - i);
        U++;
        i = j;
    }
    (void)k;
    free(keys);
    return U;
}

/* ------------------------------------------------------------------------ */
/* outer loop (main.cpp:187-287)                                            */
/* ------------------------------------------------------------------------ */

/* strategy: 0 = propagation, 1 = merging, 2 = alternate (main.cpp:247).
 * initial: 0 singleton (rand()), 1 triangle clustering (main.cpp:193-212).
 * The glibc generator is reset to its default state first, as in a fresh
 * `main` process (srand(1)); in libc mode a non-zero seed s resets it with
 * srand(s) instead, which samples the reference's run-to-run spread as if
 * it had a seed option.  `init_labels` (if non-NULL) gives the initial labels
 * instead of the rand()-based singleton partition.
 * Outputs: labels_out[n] (final compressed labels), score, iterations, k. */
int orc_run(int n, int m, const int *indices, const int *adj, int kind, double lambda,
            double frac, int rounds, int max_iter, double threshold, int initial, int strategy,
            int mode, uint64_t seed, int tie, const int *init_labels, double cc_override,
            int *labels_out, double *score_out, int *iters_out, int *k_out,
            double *scores_trace, int trace_cap) {
    int *labels = labels_out;
    int *tri = NULL, *tri_deg = NULL;
    double cc = 0.0;
    srand(mode == 0 && seed != 0 ? (unsigned)seed : 1u);
    if (initial == 1 || kind == ORC_WCC) {
        long E = indices[n];
        int *et = (int *)malloc(sizeof(int) * ((size_t)E + 1));
        tri = (int *)malloc(sizeof(int) * ((size_t)n + 1));
        tri_deg = (int *)malloc(sizeof(int) * ((size_t)n + 1));
        orc_count_triangles(n, indices, adj, et, tri, tri_deg, &cc);
        free(et);
        if (cc_override == cc_override && cc_override >= 0) cc = cc_override;
    }
    if (init_labels) { for (int v = 0; v < n; v++) labels[v] = init_labels[v]; }
    else if (initial == 0) orc_singleton_partition(n, labels);
    else orc_clustering_partition(n, indices, adj, tri, labels);

    int *sizes = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *intv = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *extv = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *nl = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    double prev = 0.0;
    int done = 0, it = 0, k = 0;
    while (!done) {
        orc_compress_labels(n, labels);
        k = orc_collect_statistics(n, indices, adj, labels, sizes, intv, extv);
        double score = orc_calculate(kind, lambda, n, m, indices, adj, labels, k, sizes, intv, extv, tri, tri_deg);
        if (scores_trace && it < trace_cap) scores_trace[it] = score;
        done = max_iter != -1 ? it >= max_iter : (prev > 0 && score > 0 && score / prev - 1 <= threshold);
        if (!done) {
            if (strategy == 0 || (strategy == 2 && it % 2 == 0))
                orc_propagation_refine(kind, lambda, n, m, indices, adj, labels, sizes, intv, extv, cc,
                                       frac, mode, seed, it, tie, nl);
            else
                orc_merging_refine(kind, lambda, n, m, k, indices, adj, labels, sizes, intv, extv, cc,
                                   rounds, mode, seed, it, tie, nl, NULL, NULL);
            for (int v = 0; v < n; v++) labels[v] = nl[v];
        }
        it++;
        prev = score;
    }
    *score_out = prev; *iters_out = it; *k_out = k;
    free(sizes); free(intv); free(extv); free(nl);
    free(tri); free(tri_deg);
    return 0;
}

/* ------------------------------------------------------------------ NMI
 * norm_mutual_info (common.cpp:75-127) over the confusion counts of
 * compare_communities (common.cpp:129-173).  The reference keeps the counts
 * in a std::map keyed by (i, j): nom sums over pairs in ascending (i, j)
 * order, den over i ascending then j ascending; ni[i] * nj[j] is an int32
 * product (it wraps for products of 2^31 and more) converted to double
 * after the multiplication.  Labels and community indices must be >= 0. */
static int cmp_pair64(const void *x, const void *y) {
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : a > b;
}

static double nmi_of_pairs(int n, uint64_t *keys, long cnt) {
    qsort(keys, (size_t)cnt, sizeof(uint64_t), cmp_pair64);
    /* counts per row i and column j: sorted unique values with their sums */
    long ui = 0, uj = 0;
    int *ri = malloc(sizeof(int) * (cnt + 1)), *ci = malloc(sizeof(int) * (cnt + 1));
    int *rj = malloc(sizeof(int) * (cnt + 1)), *cj = malloc(sizeof(int) * (cnt + 1));
    uint32_t *js = malloc(sizeof(uint32_t) * (cnt + 1));
    for (long t = 0; t < cnt; t++) {
        int i = (int)(keys[t] >> 32);
        if (ui == 0 || ri[ui - 1] != i) { ri[ui] = i; ci[ui] = 0; ui++; }
        ci[ui - 1]++;
        js[t] = (uint32_t)keys[t];
    }
    uint32_t *jsorted = malloc(sizeof(uint32_t) * (cnt + 1));
    memcpy(jsorted, js, sizeof(uint32_t) * cnt);
    qsort(jsorted, (size_t)cnt, sizeof(uint32_t), cmp_int);  /* values < 2^31 */
    for (long t = 0; t < cnt; t++) {
        int j = (int)jsorted[t];
        if (uj == 0 || rj[uj - 1] != j) { rj[uj] = j; cj[uj] = 0; uj++; }
        cj[uj - 1]++;
    }
    double nom = 0.0, den = 0.0;
    long a = 0, ib = 0;
    while (a < cnt) {
        long b = a;
        while (b < cnt && keys[b] == keys[a]) b++;
        int i = (int)(keys[a] >> 32), j = (int)(uint32_t)keys[a], val = (int)(b - a);
        while (ri[ib] != i) ib++;
        long lo = 0, hi = uj - 1;
        while (lo < hi) { long mid = (lo + hi) / 2; if (rj[mid] < j) lo = mid + 1; else hi = mid; }
        nom += val * log(((double)val * n) / wrap_mul(ci[ib], cj[lo]));
        a = b;
    }
    for (long t = 0; t < ui; t++) den += ci[t] * log(ci[t] / ((double)n));
    for (long t = 0; t < uj; t++) den += cj[t] * log(cj[t] / ((double)n));
    free(ri); free(ci); free(rj); free(cj); free(js); free(jsorted);
    return -2 * nom / den;
}

/* compare_communities(g, a, b) (common.cpp:165-173); -1 on negative labels */
int orc_nmi(int n, const int *a, const int *b, double *out) {
    uint64_t *keys = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    for (int v = 0; v < n; v++) {
        if (a[v] < 0 || b[v] < 0) { free(keys); return -1; }
        keys[v] = ((uint64_t)(uint32_t)a[v] << 32) | (uint32_t)b[v];
    }
    *out = nmi_of_pairs(n, keys, n);
    free(keys);
    return 0;
}

/* compare_communities(g, labels, filename) (common.cpp:129-163): community
 * c (from 1, in line order; empty and '#' lines skipped) holds the integers
 * of its line up to the first token that is not one.  -1: unreadable file,
 * a vertex outside [0, n) or a negative label. */
int orc_nmi_file(int n, const int *labels, const char *path, double *out) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    long cap = 1024, cnt = 0;
    uint64_t *keys = malloc(sizeof(uint64_t) * cap);
    char *line = NULL;
    size_t lcap = 0;
    ssize_t len;
    int comm = 0, bad = 0;
    while (!bad && (len = getline(&line, &lcap, fp)) >= 0) {
        if (len > 0 && line[len - 1] == '\n') line[--len] = 0;
        if (len == 0 || line[0] == '#') continue;
        comm++;
        char *p = line;
        for (;;) {
            char *e;
            long v = strtol(p, &e, 10);
            if (e == p) break;
            p = e;
            if (v < 0 || v >= n || labels[v] < 0) { bad = 1; break; }
            if (cnt == cap) { cap *= 2; keys = realloc(keys, sizeof(uint64_t) * cap); }
            keys[cnt++] = ((uint64_t)(uint32_t)comm << 32) | (uint32_t)labels[v];
        }
    }
    free(line);
    fclose(fp);
    if (bad) { free(keys); return -1; }
    *out = nmi_of_pairs(n, keys, cnt);
    free(keys);
    return 0;
}
