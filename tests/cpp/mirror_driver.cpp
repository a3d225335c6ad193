// A driver written the way a user of the reference writes one: the loop of
// main.cpp:226-266 over the C++ mirror classes (Graph, Partition,
// ModularityRule / CPMRule, PropagationRefinement / MergingRefinement, the
// framework functions), plus a rule of its own declared exactly as
// rules.hpp:9-19 declares one.  It prints the final labels, one per line.
//
//   mirror_driver <graph.pcd> <iterations> <mod|cpm-G> <s: 0|1> [custom]
//
// With "custom", the loop runs the user rule instead, and the program prints
// "refused <code>" when the device refuses it (exit status 3).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "parcd.hpp"

// A rule of the caller's own, in the shape rules.hpp:9-19 gives (all five
// methods overridden, no descriptor): it must compile against the mirror.
class HalfModularity : public parcd::Rule {
  public:
    double calculate(const parcd::Graph &g, const parcd::Partition &p) const override {
        double s = 0;
        for (int c = 0; c < (int)p.sizes.size(); c++) {
            double v = p.int_volume[c] + p.ext_volume[c];
            s += p.int_volume[c] / (2.0 * g.m) - (v / (2.0 * g.m)) * (v / (2.0 * g.m));
        }
        return 0.5 * s;
    }
    double score_merge(const parcd::Graph &g, const parcd::Partition &p, int c1, int c2, int num_links) const override {
        return 0.5 * num_links;
    }
    double score_join(const parcd::Graph &g, const parcd::Partition &p, int v, int c, int num_neighbors) const override {
        return 0.5 * num_neighbors;
    }
    double score_leave(const parcd::Graph &g, const parcd::Partition &p, int v, int num_neighbors) const override {
        return -0.5 * num_neighbors;
    }
    std::string get_name() const override { return "half modularity"; }
};

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s graph.pcd iterations mod|cpm-G 0|1 [custom]\n", argv[0]);
        return 2;
    }
    const int iters = atoi(argv[2]);
    const std::string metric = argv[3];
    const int strategy = atoi(argv[4]);
    const bool custom = argc > 5 && std::strcmp(argv[5], "custom") == 0;
    try {
        parcd::Device dev(0);
        parcd::Graph *g = parcd::Graph::read(argv[1], dev);
        if (!g) return 2;
        parcd::ModularityRule mod;
        parcd::CPMRule cpm(metric.rfind("cpm-", 0) == 0 ? atof(metric.c_str() + 4) : 0.0);
        HalfModularity half;
        const parcd::Rule &r = custom ? (const parcd::Rule &)half
                                      : metric == "mod" ? (const parcd::Rule &)mod : (const parcd::Rule &)cpm;
        parcd::PropagationRefinement prop(0.5);
        parcd::MergingRefinement merge(1);
        parcd::Refinement &ref = strategy == 1 ? (parcd::Refinement &)merge : (parcd::Refinement &)prop;
        ref.tie = PCD_TIE_LOWEST;  // the deterministic synchronous mode

        std::vector<int32_t> labels(g->n);
        parcd::create_singleton_partition(*g, labels.data());  // framework.cpp:14-24
        parcd::Partition p(labels);
        std::vector<int32_t> next(g->n);
        for (int it = 0;; it++) {
            p.num_labels = parcd::compress_labels(*g, p.labels.data());
            if (it >= iters) break;
            parcd::collect_statistics(*g, p);
            double q = r.calculate(*g, p);
            fprintf(stderr, "iteration %d: %s %.9f, %d communities\n", it, r.get_name().c_str(), q, p.num_labels);
            ref.refine(*g, p, r, next.data());
            p.labels = next;
        }
        for (int v = 0; v < g->n; v++) printf("%d\n", p.labels[v]);
        delete g;
    } catch (const parcd::Error &e) {
        if (custom && e.code == PCD_E_INVALID) {
            printf("refused %d\n", e.code);
            return 3;
        }
        fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
