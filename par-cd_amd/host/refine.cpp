// The refinement loop of main.cpp:226-266 as one library call, over the
// public entry points: compress, statistics, metric, stop rule, then one
// propagation (-s 0), merging (-s 1) or alternating (-s 2) step.  One read-back
// per phase as in pcd_main; no host allocation inside the loop.
#include "parcd.h"
#include "hostutil.hpp"

using namespace pcd;

extern "C" int pcd_refine(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, double *score_out,
                          int *iterations_out, int *k_out) {
    if (!ctx || !job) return fail(PCD_E_INVALID, "null argument");
    if (strategy < 0 || strategy > 2) return fail(PCD_E_INVALID, "strategy must be 0, 1 or 2");
    const int cap = job->iter_cap > 0 ? job->iter_cap : 1000;
    double prev = 0.0, score = 0.0;
    int it = 0, k = 0;
    for (;;) {
        int rc = pcd_compress_labels(ctx, &k);
        if (rc == PCD_OK) rc = pcd_collect_statistics(ctx, &k);
        if (rc == PCD_OK) rc = pcd_metric(ctx, job->rule, &score);
        if (rc != PCD_OK) return rc;
        bool done = job->max_iter != -1 ? it >= job->max_iter
                                        : (prev > 0 && score > 0 && score / prev - 1 <= job->threshold);
        done = done || it >= cap;
        if (!done) {
            if (strategy == 0 || (strategy == 2 && it % 2 == 0))
                rc = pcd_propagation_step(ctx, job->rule, job->frac, job->seed, it, job->tie);
            else
                rc = pcd_merging_step(ctx, job->rule, rounds, job->seed, it, job->tie);
            if (rc != PCD_OK) return rc;
        }
        it++;
        prev = score;
        if (done) break;
    }
    if (score_out) *score_out = score;
    if (iterations_out) *iterations_out = it;
    if (k_out) *k_out = k;
    return PCD_OK;
}
