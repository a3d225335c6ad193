// C++ host API of the MI355X path, mirroring the reference interfaces:
//   Graph (graph.hpp:6-40), Partition (common.hpp:26-37),
//   Rule + ModularityRule / CPMRule / WCCRule / ConductanceRule
//   (rules.hpp:9-75), Refinement + PropagationRefinement /
//   MergingRefinement (refinement.hpp:8-31), and the framework functions
//   (framework.hpp:8-11).
// Same names, same argument meaning.  Every compute call runs on the GPU
// through the C ABI (include/parcd.h); a Graph is bound to one device
// context that keeps the graph resident.  Errors throw parcd::Error.
#ifndef PARCD_HPP
#define PARCD_HPP

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "parcd.h"
#include "rule_formulas.h"

namespace parcd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != PCD_OK) throw Error(rc, pcd_last_error());
}

// A device context (one GPU, one stream).
class Device {
  public:
    explicit Device(int device = 0) { check(pcd_ctx_create(device, &ctx_)); }
    ~Device() { pcd_ctx_destroy(ctx_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    pcd_ctx *get() const { return ctx_; }

  private:
    pcd_ctx *ctx_ = nullptr;
};

// graph.hpp:6-40
class Graph {
  public:
    int n = 0, m = 0;
    std::vector<int32_t> indices, adj, map;
    std::vector<int32_t> tri, tri_deg;
    double clustering_coef = 0.0;

    // graph.cpp:40-105; returns nullptr and prints to stderr on failure
    static Graph *read(const char *filename, Device &dev) {
        int n, m;
        int32_t *ind, *adj, *mp;
        int rc = pcd_read_pcd(filename, &n, &m, &ind, &adj, &mp);
        if (rc != PCD_OK) {
            fprintf(stderr, "%s\n", pcd_last_error());
            return nullptr;
        }
        Graph *g = new Graph();
        g->n = n;
        g->m = m;
        g->indices.assign(ind, ind + n + 1);
        g->adj.assign(adj, adj + 2 * (int64_t)m);
        g->map.assign(mp, mp + n);
        pcd_free(ind);
        pcd_free(adj);
        pcd_free(mp);
        g->bind(dev);
        return g;
    }

    void bind(Device &dev) {
        dev_ = &dev;
        check(pcd_graph_upload(dev.get(), n, m, indices.data(), adj.data()));
    }

    // graph.cpp:107-175 (on the device; num_threads is accepted and unused)
    void count_triangles(int num_threads = 1) {
        (void)num_threads;
        tri.resize(n);
        tri_deg.resize(n);
        check(pcd_count_triangles(ctx(), tri.data(), tri_deg.data(), &clustering_coef));
    }

    const int32_t *begin_neighbors(int i) const { return adj.data() + indices[i]; }
    const int32_t *end_neighbors(int i) const { return adj.data() + indices[i + 1]; }
    int degree(int i) const { return indices[i + 1] - indices[i]; }
    pcd_ctx *ctx() const {
        if (!dev_) throw Error(PCD_E_STATE, "graph not bound to a device");
        return dev_->get();
    }

  private:
    Device *dev_ = nullptr;
};

// common.hpp:26-37 (owning vectors instead of raw arrays)
class Partition {
  public:
    int num_labels = -1;
    std::vector<int32_t> labels, sizes, int_volume, ext_volume;
    Partition() = default;
    explicit Partition(std::vector<int32_t> l) : labels(std::move(l)) {}
};

// framework.hpp:8-11
inline void create_singleton_partition(const Graph &g, int32_t *labels) {
    // framework.cpp:14-24: identity shuffled with rand(): same labels as the
    // reference for the same glibc state
    for (int v = 0; v < g.n; v++) labels[v] = v;
    for (int v = 0; v < g.n; v++) {
        int j = rand() % g.n;
        int32_t t = labels[v];
        labels[v] = labels[j];
        labels[j] = t;
    }
}

// framework.cpp:27-59: vertices by (-cc, -deg, v), each unlabelled vertex
// opens a community with its unlabelled neighbours; computed on the device,
// triangles included (deg*(deg-1) is an int32 product there as well).
inline void create_clustering_partition(const Graph &g, int32_t *labels) {
    int nc = 0;
    check(pcd_clustering_labels(g.ctx(), &nc));
    check(pcd_labels_download(g.ctx(), labels));
}

inline int compress_labels(const Graph &g, int32_t *labels) {
    check(pcd_labels_upload(g.ctx(), labels));
    int k = 0;
    check(pcd_compress_labels(g.ctx(), &k));
    check(pcd_labels_download(g.ctx(), labels));
    return k;
}

inline void collect_statistics(const Graph &g, Partition &p, int num_threads = 1) {
    (void)num_threads;
    check(pcd_labels_upload(g.ctx(), p.labels.data()));
    int k = 0;
    check(pcd_collect_statistics(g.ctx(), &k));
    p.num_labels = k;
    p.sizes.resize(k);
    p.int_volume.resize(k);
    p.ext_volume.resize(k);
    check(pcd_stats_download(g.ctx(), p.sizes.data(), p.int_volume.data(), p.ext_volume.data()));
}

// rules.hpp:9-19.  The five methods keep the reference's signatures, so a
// rule written against rules.hpp compiles here unchanged.  The device
// evaluates a rule through its descriptor(): the four built-in rules below
// return their kind (and are final, so no override of their formulas can be
// silently bypassed by the device).  A rule of its own keeps the default
// descriptor, PCD_RULE_HOST: its methods work on the host, and every call
// that would run it on the device (refine, and calculate of the base class)
// throws parcd::Error with PCD_E_INVALID, because the device cannot run code
// it was not compiled with.
class Rule {
  public:
    virtual ~Rule() {}
    virtual double calculate(const Graph &g, const Partition &p) const {
        const pcd_rule d = device_rule();
        check(pcd_labels_upload(g.ctx(), p.labels.data()));
        int k = 0;
        check(pcd_collect_statistics(g.ctx(), &k));
        double s = 0;
        check(pcd_metric(g.ctx(), d, &s));
        return s;
    }
    virtual double score_merge(const Graph &g, const Partition &p, int c1, int c2, int num_links) const {
        return pcd::score_merge(params(g), stats(p, c1), stats(p, c2), num_links);
    }
    virtual double score_join(const Graph &g, const Partition &p, int v, int c, int num_neighbors) const {
        return pcd::score_join(params(g), g.degree(v), stats(p, c), num_neighbors);
    }
    virtual double score_leave(const Graph &g, const Partition &p, int v, int num_neighbors) const {
        return pcd::score_leave(params(g), g.degree(v), stats(p, p.labels[v]), num_neighbors);
    }
    virtual std::string get_name() const = 0;
    // the formula set the device evaluates for this rule
    virtual pcd_rule descriptor() const { return pcd_rule{PCD_RULE_HOST, 0.0}; }
    // descriptor(), or PCD_E_INVALID for a rule the device cannot run
    pcd_rule device_rule() const {
        const pcd_rule d = descriptor();
        if (d.kind < PCD_RULE_MODULARITY || d.kind > PCD_RULE_CONDUCTANCE)
            throw Error(PCD_E_INVALID, "rule '" + get_name() +
                                           "' has no device form: the device runs modularity, CPM, WCC and "
                                           "conductance");
        return d;
    }

  protected:
    pcd::RuleParams params(const Graph &g) const {
        pcd_rule d = device_rule();
        pcd::RuleParams r;
        r.kind = d.kind;
        r.lambda = d.lambda;
        pcd::set_edges(r, g.m);
        r.cc = g.clustering_coef;
        return r;
    }
    static pcd::CommStats stats(const Partition &p, int c) {
        pcd::CommStats s;
        s.size = p.sizes[c];
        s.int_vol = p.int_volume[c];
        s.ext_vol = p.ext_volume[c];
        s.pad = 0;
        return s;
    }
};

class ModularityRule final : public Rule {
  public:
    std::string get_name() const override { return "modularity"; }  // rules.cpp:72-74
    pcd_rule descriptor() const override { return pcd_rule{PCD_RULE_MODULARITY, 0.0}; }
};

class ConductanceRule final : public Rule {
  public:
    std::string get_name() const override { return "conductance"; }  // rules.cpp:116-118
    pcd_rule descriptor() const override { return pcd_rule{PCD_RULE_CONDUCTANCE, 0.0}; }
};

class CPMRule final : public Rule {
  public:
    explicit CPMRule(double l) : lambda(l) {}
    std::string get_name() const override {  // rules.cpp:146-148
        std::ostringstream s;
        s << lambda;
        return "CPM (lambda=" + s.str() + ")";
    }
    pcd_rule descriptor() const override { return pcd_rule{PCD_RULE_CPM, lambda}; }

  private:
    double lambda;
};

class WCCRule final : public Rule {
  public:
    std::string get_name() const override { return "WCC"; }  // rules.cpp:438-440
    pcd_rule descriptor() const override { return pcd_rule{PCD_RULE_WCC, 0.0}; }
};

// refinement.hpp:8-31.  refine() uploads p.labels, runs one device step and
// downloads new_labels.  The step index and seed drive the counter RNG.
class Refinement {
  public:
    virtual ~Refinement() {}
    virtual void refine(const Graph &g, const Partition &p, const Rule &r, int32_t *new_labels) = 0;
    uint64_t seed = 0;
    int step = 0;
    int tie = PCD_TIE_RANDOM;
};

class PropagationRefinement : public Refinement {
  public:
    explicit PropagationRefinement(double f, int n = 1) : frac(f), num_threads(n) {}
    void refine(const Graph &g, const Partition &p, const Rule &r, int32_t *new_labels) override {
        const pcd_rule d = r.device_rule();
        collect_statistics(g, const_cast<Partition &>(p));
        check(pcd_propagation_step(g.ctx(), d, frac, seed, step++, tie));
        check(pcd_labels_download(g.ctx(), new_labels));
    }

  private:
    double frac;
    int num_threads;
};

class MergingRefinement : public Refinement {
  public:
    explicit MergingRefinement(int r, int n = 1) : rounds(r), num_threads(n) {}
    void refine(const Graph &g, const Partition &p, const Rule &r, int32_t *new_labels) override {
        const pcd_rule d = r.device_rule();
        collect_statistics(g, const_cast<Partition &>(p));
        check(pcd_merging_step(g.ctx(), d, rounds, seed, step++, tie));
        check(pcd_labels_download(g.ctx(), new_labels));
    }

  private:
    int rounds;
    int num_threads;
};

}  // namespace parcd

#endif
