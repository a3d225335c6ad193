// The refinement loop of main.cpp:226-266 as one library call: passes of
// compress, statistics, metric, stop rule, then one propagation (-s 0),
// merging (-s 1) or alternating (-s 2) step, through pcd_iterate_many
// (device-resident passes for modularity and CPM with -s 0).
#include "parcd.h"
#include "hostutil.hpp"

using namespace pcd;

extern "C" int pcd_refine(pcd_ctx *ctx, const pcd_job *job, int strategy, int rounds, double *score_out,
                          int *iterations_out, int *k_out) {
    if (!ctx || !job) return fail(PCD_E_INVALID, "null argument");
    if (strategy < 0 || strategy > 2) return fail(PCD_E_INVALID, "strategy must be 0, 1 or 2");
    // pcd_iterate_many applies -n and the iteration cap (done on the pass
    // that reaches them); it runs device-resident where it can
    double prev = 0.0, score = 0.0;
    int it = 0, k = 0;
    for (;;) {
        int done = 0, passes = 0;
        int rc = pcd_iterate_many(ctx, job, strategy, rounds, it, 1 << 20, prev, &passes, &score, &k, &done);
        if (rc != PCD_OK) return rc;
        it += passes;
        prev = score;
        if (done) break;
    }
    if (score_out) *score_out = score;
    if (iterations_out) *iterations_out = it;
    if (k_out) *k_out = k;
    return PCD_OK;
}
