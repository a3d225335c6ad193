"""The batched-form benchmark workload itself: 4,096 SBM graphs of
pcd_gen_sbm_batch, seed 4 (n log-uniform in [50, 20000], 2-64 blocks, p_in
in [0.05, 0.3], p_out in [0.001, 0.01], modularity or CPM with resolution
in [0.005, 0.05]), refined in one batch (one round, the bin-map autotuner,
rounds to the end) from the singleton partition, as the benchmark runs it.

* A stratified sample of graphs (every n decade, both rules) equals the
  oracle's serial replay (counter mode) label for label, with the same
  iterations, k and score (refinement.cpp:114-182 per step, main.cpp:226-266
  around it).
* Every graph equals the same graph run alone through a plain context."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_graphs import singleton_labels

pytestmark = pytest.mark.gpu
parcd = pytest.importorskip("parcd")

ITER_CAP = 100  # graphs whose score never turns positive stop here (the reference would loop)


@pytest.fixture(scope="module")
def workload():
    gs, prm = parcd.gen_sbm_batch(4, 4096)
    keep = [i for i, g in enumerate(gs) if g[1] > 0]  # a graph without edges has no score to refine
    gs = [gs[i] for i in keep]
    prm = prm[keep]
    labs = [singleton_labels(g[0]) for g in gs]
    jobs = [parcd.job(kind=int(p[4]), lam=float(p[5]) if int(p[4]) == parcd.RULE_CPM else 0.0, seed=i,
                      iter_cap=ITER_CAP) for i, p in zip(keep, prm)]
    arena = sum(parcd.batch_footprint(g[0], g[1]) for g in gs) + (256 << 20)
    print("batch workload: %d graphs, %d entries" % (len(gs), sum(2 * g[1] for g in gs)), flush=True)
    with parcd.Batch(0, max_graphs=len(gs), arena_bytes=arena) as b:
        slots = b.add_many([g[:4] for g in gs], labs, jobs)
        assert all(s is not None for s in slots)
        running = b.step(1)
        b.autotune(reps=1)
        while running:
            running = b.step(1)
        labels = [np.array(x) for x in b.labels_many(slots)]
        status = [b.status(s) for s in slots]
    return keep, gs, prm, labs, jobs, labels, status


def test_workload_sample_equals_replay(workload):
    keep, gs, prm, labs, jobs, labels, status = workload
    n_of = np.array([g[0] for g in gs])
    e_of = np.array([2 * g[1] for g in gs], np.int64)
    kind_of = np.array([int(p[4]) for p in prm])
    picked = []
    # per n decade and rule, the graphs with the fewest entries (oracle time)
    for lo, hi in ((50, 200), (200, 1000), (1000, 5000), (5000, 20001)):
        for kind in (parcd.RULE_MODULARITY, parcd.RULE_CPM):
            idx = np.nonzero((n_of >= lo) & (n_of < hi) & (kind_of == kind))[0]
            idx = idx[np.argsort(e_of[idx], kind="stable")]
            picked += list(idx[:5])
    assert len(picked) >= 32
    for i in picked:
        n, m, ind, adj = gs[i]
        st, iters, k, score = status[i]
        assert st == parcd.BATCH_DONE
        jb = jobs[i]
        ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=jb.rule.kind, lam=jb.rule.lam, frac=0.5,
                                    max_iter=ITER_CAP if iters >= ITER_CAP else -1, mode=1, seed=jb.seed,
                                    tie=O.TIE_RANDOM, init_labels=labs[i])
        assert np.array_equal(labels[i], ol), (i, n)
        assert (iters, k) == (oit, ok), i
        assert abs(score - osc) <= 1e-9 * max(1.0, abs(osc)), i


def test_workload_every_graph_equals_plain(workload):
    keep, gs, prm, labs, jobs, labels, status = workload
    same = 0
    with parcd.Context(0) as ctx:
        for i, (g, lab, jb) in enumerate(zip(gs, labs, jobs)):
            ctx.upload_graph(g[0], g[1], g[2], g[3])
            ctx.upload_labels(lab)
            ps, pit, pk = ctx.refine(jb)  # pcd_refine: the loop of main.cpp:226-266 in one call
            pl = ctx.labels()
            st, iters, k, score = status[i]
            same += int(np.array_equal(labels[i], pl) and (iters, k, score) == (pit, pk, ps))
    assert same == len(gs)
