// pcd_main: drop-in for the reference `main` (main.cpp:55-287): same
// options, same stdout report, `orig_id label` lines on stderr, exit code 1.
// The refinement loop runs on the GPU with the partition resident in HBM;
// only one scalar (the metric) comes back per iteration.
//
// Extra options (no reference meaning changes):
//   -D            deterministic mode: lowest-label tie order
//   -S <seed>     seed of the counter-based vertex selection / tie keys
//   -g <device>   HIP device index
//   -T            print per-phase device times (HIP events)
#include <getopt.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "parcd.hpp"

using namespace std;

static double now() {
    return chrono::duration<double>(chrono::steady_clock::now().time_since_epoch()).count();
}

static void print_usage(const char *prog) {
    cerr << "usage: " << prog << " <graph> [options]                        \n"
         << "                                                               \n"
         << "options:                                                       \n"
         << "-m <metric>                                                    \n"
         << "  Set the metric to optimize for:                              \n"
         << "    mod: Modularity by Newman & Girvan                         \n"
         << "    wcc: Weighted Community Clustering by Prat et al.          \n"
         << "    cpm-<gamma>: Constant Potts Model by Traag et al.          \n"
         << "-p <#threads>                                                  \n"
         << "  Set the number of threads to use.                            \n"
         << "-f <fraction>                                                  \n"
         << "  Set the fraction of vertices to update for label propagation.\n"
         << "-r <#rounds>                                                   \n"
         << "  Set the number of rounds for parallel merging.               \n"
         << "-t <threshold>                                                 \n"
         << "  Set the iteration threshold.                                 \n"
         << "-n <#iterations>                                               \n"
         << "  Set the total number of iteration. This will disable the     \n"
         << "  iteration threshold set with -t.                             \n"
         << "-i 0/1                                                         \n"
         << "  Set the heuristic for creating the initial partitioning:     \n"
         << "    0: unique labels heuristic                                 \n"
         << "    1: triangle-based heuristic                                \n"
         << "-s <schedule>                                                  \n"
         << "  Set the schedule for the refinement phase:                   \n"
         << "    0: Semi-parallel label propagation                         \n"
         << "    1: Best-neighbor parallel merging                          \n"
         << "    2: Alternate the two strategies                            \n"
         << "-D  deterministic mode (lowest-label tie order)                \n"
         << "-S <seed>  seed of the counter-based random choices            \n"
         << "-g <device>  GPU index                                         \n"
         << "-T  print per-phase GPU times                                  \n"
         << "-h                                                             \n"
         << "  Print this help message.                                     \n"
         << endl;
}

int main(int argc, char *argv[]) {
    int c;
    int num_threads = 1;
    double frac = 0.5;
    int rounds = 1;
    int max_iter = -1;
    double threshold = 0;
    int initial = 0;
    int strategy = 0;
    int device = 0;
    uint64_t seed = 0;
    int tie = PCD_TIE_RANDOM;
    bool phase_times = false;
    parcd::Rule *r = new parcd::ModularityRule();
    (void)num_threads;

    while ((c = getopt(argc, argv, "?hr:f:m:p:t:s:i:m:n:DS:g:T")) != -1) {
        switch (c) {
        case 'p':
            num_threads = atoi(optarg);
            if (num_threads < 0) { cerr << "number of threads should be positive" << endl; exit(-1); }
            break;
        case 'r':
            rounds = atoi(optarg);
            if (rounds <= 0) { cerr << "number of rounds should be positive" << endl; exit(-1); }
            break;
        case 'n': max_iter = atoi(optarg); break;
        case 'i': initial = atoi(optarg); break;
        case 'f':
            frac = atof(optarg);
            if (frac <= 0 || frac > 1) { cerr << "fraction should be between 0.0 and 1.0" << endl; exit(-1); }
            break;
        case 't':
            threshold = atof(optarg);
            if (threshold < 0 || threshold > 1) { cerr << "threshold should be between 0.0 and 1.0" << endl; exit(-1); }
            break;
        case 's': strategy = atoi(optarg); break;
        case 'm':
            for (int i = 0; optarg[i]; i++) optarg[i] = tolower(optarg[i]);
            delete r;
            if (strcmp(optarg, "mod") == 0 || strcmp(optarg, "modularity") == 0) r = new parcd::ModularityRule();
            else if (strcmp(optarg, "wcc") == 0) r = new parcd::WCCRule();
            else if (strncmp(optarg, "cpm-", 4) == 0 && optarg[4] != 0) r = new parcd::CPMRule(atof(optarg + 4));
            else { cerr << "unknown metric: " << optarg << endl; exit(-1); }
            break;
        case 'D': tie = PCD_TIE_LOWEST; break;
        case 'S': seed = strtoull(optarg, nullptr, 10); break;
        case 'g': device = atoi(optarg); break;
        case 'T': phase_times = true; break;
        case 'h':
        case '?': print_usage(argv[0]); exit(-1);
        default:
            cerr << "unknown option" << endl << endl;
            print_usage(argv[0]);
            exit(-1);
        }
    }
    if (optind >= argc) { print_usage(argv[0]); exit(-1); }
    const char *graph_file = argv[optind];

    try {
        parcd::Device dev(device);
        pcd_ctx *ctx = dev.get();
        cout << "loading graph..." << endl;
        double t0 = now();
        parcd::Graph *g = parcd::Graph::read(graph_file, dev);
        double time_load = now() - t0;
        if (!g) { cerr << "error while loading graph" << endl; exit(-1); }
        cout << "-------------------------------------------------" << endl;
        cout << "loaded graph from " << graph_file << endl;
        cout << "num nodes: " << g->n << endl;
        cout << "num edges: " << g->m << endl;
        cout << "avg. degree: " << double(2 * g->m) / g->n << endl;
        cout << "max. degree: " << g->degree(0) << endl;
        cout << "-------------------------------------------------" << endl;
        const int n = g->n;
        vector<int32_t> labels(n);
        cout << "creating initial partitioning..." << endl;
        double t_init = now();
        const bool wcc = dynamic_cast<parcd::WCCRule *>(r) != nullptr;
        if (initial == 1 || wcc) {
            double tt = now();
            g->count_triangles();
            cout << "count triangles: " << now() - tt << endl;
        }
        switch (initial) {
        case 0: parcd::create_singleton_partition(*g, labels.data()); break;
        case 1: parcd::create_clustering_partition(*g, labels.data()); break;
        default: cerr << "unknown partitioning strategy" << endl; exit(-1);
        }
        double time_initial = now() - t_init;
        parcd::check(pcd_labels_upload(ctx, labels.data()));
        if (phase_times) parcd::check(pcd_enable_timing(ctx, 1));
        const pcd_rule rule = r->descriptor();
        double prev_score = 0.0;
        bool done = false;
        int it = 0, k = 0;
        cout << "refining partitioning..." << endl;
        double t_ref = now();
        while (!done) {
            double t_it = now();
            double tt = now();
            parcd::check(pcd_compress_labels(ctx, &k));
            parcd::check(pcd_collect_statistics(ctx, &k));
            parcd::check(pcd_synchronize(ctx));
            cout << " - collecting statistics: " << now() - tt << " sec" << endl;
            tt = now();
            double score = 0;
            parcd::check(pcd_metric(ctx, rule, &score));
            cout << " - calculate metric: " << now() - tt << " sec" << endl;
            cout << " - " << r->get_name() << ": " << score << endl;
            cout << " - improv. of prev. iteration: " << (score / prev_score - 1) * 100 << "%" << endl;
            done = max_iter != -1 ? it >= max_iter : (prev_score > 0 && score > 0 && score / prev_score - 1 <= threshold);
            if (!done) {
                tt = now();
                if (strategy == 0 || (strategy == 2 && it % 2 == 0)) {
                    parcd::check(pcd_propagation_step(ctx, rule, frac, seed, it, tie));
                    parcd::check(pcd_synchronize(ctx));
                    cout << " - label propagation: " << now() - tt << " sec" << endl;
                } else {
                    parcd::check(pcd_merging_step(ctx, rule, rounds, seed, it, tie));
                    parcd::check(pcd_synchronize(ctx));
                    cout << " - parallel merging: " << now() - tt << " sec" << endl;
                }
            }
            if (phase_times) {
                const char *names[PCD_PHASE_COUNT] = {"compress", "stats", "metric", "propagate", "merge", "triangles"};
                for (int ph = 0; ph < PCD_PHASE_COUNT; ph++) {
                    double ms = 0;
                    parcd::check(pcd_phase_time(ctx, ph, &ms));
                    if (ms > 0) cout << "   [gpu] " << names[ph] << ": " << ms << " ms" << endl;
                }
            }
            cout << "iteration " << (it + 1) << ": " << now() - t_it << " sec" << endl;
            it++;
            prev_score = score;
        }
        double time_refine = now() - t_ref;
        vector<int32_t> sizes(max(k, 1));
        parcd::check(pcd_stats_download(ctx, sizes.data(), nullptr, nullptr));
        parcd::check(pcd_labels_download(ctx, labels.data()));
        cout << "-------------------------------------------------" << endl;
        cout << "time read graph: " << time_load << " sec" << endl;
        cout << "time initial partitioning: " << time_initial << " sec" << endl;
        cout << "time refinement phase: " << time_refine << " sec" << endl;
        cout << "refinement iterations: " << it << endl;
        cout << "avg. time per iteration: " << time_refine / it << " sec" << endl;
        cout << r->get_name() << ": " << prev_score << endl;
        cout << "num. communities: " << k << endl;
        cout << "avg. community size: " << double(n) / k << endl;
        cout << "largest community size: " << *max_element(sizes.begin(), sizes.begin() + max(k, 1)) << endl;
        cout << "-------------------------------------------------" << endl;
        // main.cpp:282-284
        string out;
        out.reserve((size_t)n * 16);
        char buf[32];
        for (int i = 0; i < n; i++) {
            int len = snprintf(buf, sizeof(buf), "%d %d\n", g->map[i], labels[i]);
            out.append(buf, len);
        }
        fwrite(out.data(), 1, out.size(), stderr);
        delete g;
    } catch (const parcd::Error &e) {
        cerr << "error: " << e.what() << endl;
        exit(-1);
    }
    delete r;
    return 1;  // main.cpp:286
}
