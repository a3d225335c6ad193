// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A small driver linked against the reference's own CPU sources (graph.cpp,
// common.cpp, framework.cpp, rules.cpp, refinement.cpp, compiled where they
// lie under the reference tree by oracle/Makefile).  It calls the reference
// functions on inputs given in small binary files and writes what they
// return, so that tests/golden/make_golden.py can record golden fixtures.
// Nothing here is product code; nothing in the product links it.
//
// File format of every input/output: a sequence of records
//   int32 tag_len, char tag[tag_len], int32 kind (0=int32, 1=float64),
//   int64 count, payload.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"
#include "framework.hpp"
#include "graph.hpp"
#include "refinement.hpp"
#include "rules.hpp"

using namespace std;

struct Rec {
    int kind;
    vector<int> i;
    vector<double> d;
};

static map<string, Rec> read_recs(const char *path) {
    map<string, Rec> out;
    FILE *fp = fopen(path, "rb");
    if (!fp) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    int tl;
    while (fread(&tl, 4, 1, fp) == 1) {
        string tag(tl, ' ');
        if (fread(&tag[0], 1, tl, fp) != (size_t)tl) exit(3);
        Rec r;
        long long cnt;
        if (fread(&r.kind, 4, 1, fp) != 1 || fread(&cnt, 8, 1, fp) != 1) exit(3);
        if (r.kind == 0) { r.i.resize(cnt); if (cnt && fread(&r.i[0], 4, cnt, fp) != (size_t)cnt) exit(3); }
        else { r.d.resize(cnt); if (cnt && fread(&r.d[0], 8, cnt, fp) != (size_t)cnt) exit(3); }
        out[tag] = r;
    }
    fclose(fp);
    return out;
}

static FILE *g_out;
static void put_i(const char *tag, const int *p, long long cnt) {
    int tl = strlen(tag), kind = 0;
    fwrite(&tl, 4, 1, g_out); fwrite(tag, 1, tl, g_out); fwrite(&kind, 4, 1, g_out);
    fwrite(&cnt, 8, 1, g_out); if (cnt) fwrite(p, 4, cnt, g_out);
}
static void put_d(const char *tag, const double *p, long long cnt) {
    int tl = strlen(tag), kind = 1;
    fwrite(&tl, 4, 1, g_out); fwrite(tag, 1, tl, g_out); fwrite(&kind, 4, 1, g_out);
    fwrite(&cnt, 8, 1, g_out); if (cnt) fwrite(p, 8, cnt, g_out);
}

static const Graph *load(const char *path) {
    const Graph *g = Graph::read(path);
    if (!g) { fprintf(stderr, "cannot read graph %s\n", path); exit(2); }
    return g;
}

// graph <pcd> <out>: Graph::count_triangles (graph.cpp:107-175)
static int cmd_graph(int argc, char **argv) {
    const Graph *g = load(argv[2]);
    g->count_triangles();
    g_out = fopen(argv[3], "wb");
    put_i("tri", g->tri, g->n);
    put_i("tri_deg", g->tri_deg, g->n);
    double cc = g->clustering_coef;
    put_d("cc", &cc, 1);
    fclose(g_out);
    return 0;
}

// part <pcd> <in> <out>: compress_labels, collect_statistics, every rule's
// calculate and every score_join/score_leave/score_merge the refinement
// steps can ask for on this partition.
static int cmd_part(int argc, char **argv) {
    const Graph *g = load(argv[2]);
    g->count_triangles();
    map<string, Rec> in = read_recs(argv[3]);
    const int n = g->n;
    int *labels = new int[n];
    for (int v = 0; v < n; v++) labels[v] = in["labels"].i[v];
    double lambda = in["lambda"].d[0];
    int k = compress_labels(*g, labels);
    Partition p(labels);
    collect_statistics(*g, p, 1);
    g_out = fopen(argv[4], "wb");
    put_i("k", &k, 1);
    put_i("num_labels", &p.num_labels, 1);
    put_i("labels", p.labels, n);
    put_i("sizes", p.sizes, p.num_labels);
    put_i("int_volume", p.int_volume, p.num_labels);
    put_i("ext_volume", p.ext_volume, p.num_labels);

    ModularityRule mod;
    CPMRule cpm(lambda);
    WCCRule wcc;
    ConductanceRule cond;
    const Rule *rules[4] = {&mod, &cpm, &wcc, &cond};
    double calc[4];
    for (int r = 0; r < 4; r++) calc[r] = rules[r]->calculate(*g, p);
    put_d("calculate", calc, 4);

    // join/leave: for each v, each distinct neighbour label (count)
    vector<int> jv, jc, jn;
    vector<int> lv, ln;
    for (int v = 0; v < n; v++) {
        map<int, int> cnt;
        for (const int *u = g->begin_neighbors(v); u != g->end_neighbors(v); u++) cnt[p.labels[*u]]++;
        int own = 0;
        for (map<int, int>::iterator it = cnt.begin(); it != cnt.end(); ++it) {
            if (it->first == p.labels[v]) own = it->second;
            jv.push_back(v); jc.push_back(it->first); jn.push_back(it->second);
        }
        lv.push_back(v); ln.push_back(own);
    }
    put_i("join_v", &jv[0], jv.size()); put_i("join_c", &jc[0], jc.size()); put_i("join_n", &jn[0], jn.size());
    put_i("leave_v", &lv[0], lv.size()); put_i("leave_n", &ln[0], ln.size());
    for (int r = 0; r < 4; r++) {
        vector<double> js(jv.size()), ls(lv.size());
        for (size_t i = 0; i < jv.size(); i++) js[i] = rules[r]->score_join(*g, p, jv[i], jc[i], jn[i]);
        for (size_t i = 0; i < lv.size(); i++) ls[i] = rules[r]->score_leave(*g, p, lv[i], ln[i]);
        char t1[32], t2[32];
        sprintf(t1, "join_s%d", r); sprintf(t2, "leave_s%d", r);
        put_d(t1, js.empty() ? NULL : &js[0], js.size());
        put_d(t2, ls.empty() ? NULL : &ls[0], ls.size());
    }
    // merge: community pair link counts (the integer content of contraction)
    map<pair<int, int>, int> links;
    for (int v = 0; v < n; v++)
        for (const int *u = g->begin_neighbors(v); u != g->end_neighbors(v); u++)
            if (p.labels[*u] != p.labels[v]) links[make_pair(p.labels[v], p.labels[*u])]++;
    vector<int> mi, mj, ml;
    for (map<pair<int, int>, int>::iterator it = links.begin(); it != links.end(); ++it) {
        mi.push_back(it->first.first); mj.push_back(it->first.second); ml.push_back(it->second);
    }
    put_i("merge_i", mi.empty() ? NULL : &mi[0], mi.size());
    put_i("merge_j", mj.empty() ? NULL : &mj[0], mj.size());
    put_i("merge_links", ml.empty() ? NULL : &ml[0], ml.size());
    for (int r = 0; r < 4; r++) {
        vector<double> ms(mi.size());
        for (size_t i = 0; i < mi.size(); i++) ms[i] = rules[r]->score_merge(*g, p, mi[i], mj[i], ml[i]);
        char t[32];
        sprintf(t, "merge_s%d", r);
        put_d(t, ms.empty() ? NULL : &ms[0], ms.size());
    }
    fclose(g_out);
    return 0;
}

// synth <in> <out>: score functions on fabricated community statistics
// (reaches the int32-overflow sub-expressions of rules.cpp:265-358).
static int cmd_synth(int argc, char **argv) {
    map<string, Rec> in = read_recs(argv[2]);
    vector<int> &deg = in["deg"].i;
    int nv = deg.size();
    int *indices = new int[nv + 1];
    indices[0] = 0;
    for (int i = 0; i < nv; i++) indices[i + 1] = indices[i] + deg[i];
    int *adj = new int[indices[nv] + 1];
    memset(adj, 0, sizeof(int) * (indices[nv] + 1));
    int m = in["m"].i[0];
    Graph *g = new Graph(nv, m, adj, indices, NULL, NULL, NULL);
    *const_cast<double *>(&g->clustering_coef) = in["cc"].d[0];
    int K = in["sizes"].i.size();
    int *labels = new int[nv];
    for (int i = 0; i < nv; i++) labels[i] = in["own"].i[i];
    Partition p(labels);
    p.num_labels = K;
    p.sizes = new int[K]; p.int_volume = new int[K]; p.ext_volume = new int[K];
    for (int c = 0; c < K; c++) {
        p.sizes[c] = in["sizes"].i[c]; p.int_volume[c] = in["int_volume"].i[c]; p.ext_volume[c] = in["ext_volume"].i[c];
    }
    double lambda = in["lambda"].d[0];
    ModularityRule mod;
    CPMRule cpm(lambda);
    WCCRule wcc;
    ConductanceRule cond;
    const Rule *rules[4] = {&mod, &cpm, &wcc, &cond};
    vector<int> &qv = in["q_v"].i, &qc = in["q_c"].i, &qn = in["q_n"].i;
    vector<int> &mi = in["m_i"].i, &mj = in["m_j"].i, &ml = in["m_l"].i;
    g_out = fopen(argv[3], "wb");
    for (int r = 0; r < 4; r++) {
        vector<double> js(qv.size()), ls(qv.size()), ms(mi.size());
        for (size_t i = 0; i < qv.size(); i++) {
            js[i] = rules[r]->score_join(*g, p, qv[i], qc[i], qn[i]);
            ls[i] = rules[r]->score_leave(*g, p, qv[i], qn[i]);
        }
        for (size_t i = 0; i < mi.size(); i++) ms[i] = rules[r]->score_merge(*g, p, mi[i], mj[i], ml[i]);
        char t1[32], t2[32], t3[32];
        sprintf(t1, "join_s%d", r); sprintf(t2, "leave_s%d", r); sprintf(t3, "merge_s%d", r);
        put_d(t1, js.empty() ? NULL : &js[0], js.size());
        put_d(t2, ls.empty() ? NULL : &ls[0], ls.size());
        put_d(t3, ms.empty() ? NULL : &ms[0], ms.size());
    }
    fclose(g_out);
    return 0;
}

// refine <pcd> <in> <out>: one PropagationRefinement / MergingRefinement
// step at one thread after srand(seed) (refinement.cpp:114-288).
static int cmd_refine(int argc, char **argv) {
    const Graph *g = load(argv[2]);
    map<string, Rec> in = read_recs(argv[3]);
    int kind = in["kind"].i[0], strategy = in["strategy"].i[0], rounds = in["rounds"].i[0];
    unsigned seed = (unsigned)in["srand"].i[0];
    double lambda = in["lambda"].d[0], frac = in["frac"].d[0];
    if (kind == 2) g->count_triangles();
    const int n = g->n;
    int *labels = new int[n];
    for (int v = 0; v < n; v++) labels[v] = in["labels"].i[v];
    compress_labels(*g, labels);
    Partition p(labels);
    collect_statistics(*g, p, 1);
    Rule *r;
    if (kind == 0) r = new ModularityRule();
    else if (kind == 1) r = new CPMRule(lambda);
    else if (kind == 2) r = new WCCRule();
    else r = new ConductanceRule();
    srand(seed);
    int *nl = new int[n];
    if (strategy == 0) { PropagationRefinement pr(frac, 1); pr.refine(*g, p, *r, nl); }
    else { MergingRefinement mr(rounds, 1); mr.refine(*g, p, *r, nl); }
    g_out = fopen(argv[4], "wb");
    put_i("labels", p.labels, n);
    put_i("new_labels", nl, n);
    fclose(g_out);
    return 0;
}

// nmi <pcd> <in> <out>: compare_communities(g, a, b) (common.cpp:165-173)
static int cmd_nmi(int argc, char **argv) {
    const Graph *g = load(argv[2]);
    map<string, Rec> in = read_recs(argv[3]);
    const int n = g->n;
    int *a = new int[n], *b = new int[n];
    for (int v = 0; v < n; v++) { a[v] = in["a"].i[v]; b[v] = in["b"].i[v]; }
    double nmi = compare_communities(*g, a, b);
    g_out = fopen(argv[4], "wb");
    put_d("nmi", &nmi, 1);
    fclose(g_out);
    return 0;
}

// nmifile <pcd> <in> <truth> <out>: compare_communities(g, labels, file)
// (common.cpp:129-163)
static int cmd_nmifile(int argc, char **argv) {
    if (argc < 6) return 2;
    const Graph *g = load(argv[2]);
    map<string, Rec> in = read_recs(argv[3]);
    const int n = g->n;
    int *labels = new int[n];
    for (int v = 0; v < n; v++) labels[v] = in["labels"].i[v];
    double nmi = compare_communities(*g, labels, string(argv[4]));
    g_out = fopen(argv[5], "wb");
    put_d("nmi", &nmi, 1);
    fclose(g_out);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) { fprintf(stderr, "usage: ref_harness graph|part|synth|refine|nmi|nmifile ...\n"); return 2; }
    string c = argv[1];
    if (c == "graph") return cmd_graph(argc, argv);
    if (c == "part") return cmd_part(argc, argv);
    if (c == "synth") return cmd_synth(argc, argv);
    if (c == "refine") return cmd_refine(argc, argv);
    if (c == "nmi") return cmd_nmi(argc, argv);
    if (c == "nmifile") return cmd_nmifile(argc, argv);
    fprintf(stderr, "unknown command %s\n", argv[1]);
    return 2;
}
