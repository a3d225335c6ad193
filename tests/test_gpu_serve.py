"""Serving loop on the GPU: requests that arrive while others run are
admitted as room frees up, overlap on the device (batch rounds and worker
contexts), and each result equals the same request run alone."""
import time

import numpy as np
import pytest

import oracle_lib as O
from gpu_graphs import bin_zoo, rmat, sbm, singleton_labels

# A run that began with these tests ended with the machine's GPU faulted
# before any output, and reading the serving code found no cause: the tests
# stay skipped until a cause is shown.
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="serving path held off the GPU after an unexplained GPU fault")]
parcd = pytest.importorskip("parcd")


def alone(g, lab, jb, initial):
    n, m, ind, adj = g[:4]
    with parcd.Context(0) as ctx:
        ctx.upload_graph(n, m, ind, adj)
        if initial == 1 or jb.rule.kind == O.WCC:
            tri = ctx.triangles()[0]
            if initial == 1:
                lab = parcd.clustering_partition(n, ind, adj, tri)
        ctx.upload_labels(lab)
        score, iters, k = ctx.refine(jb)
        return ctx.labels(), score, iters, k


def requests():
    out = []
    for i in range(12):
        g = sbm(3 + i % 5, 40 + 23 * i, 0.15 + 0.01 * i, 0.004, seed=100 + i)
        kind = O.CPM if i % 3 == 1 else O.MOD
        out.append((g, dict(kind=kind, lam=0.01 if kind == O.CPM else 0.0, seed=i), 0))
    out.append((rmat(11, seed=7), dict(kind=O.MOD, seed=50), 0))
    out.append((bin_zoo(), dict(kind=O.MOD, seed=51), 0))                  # degrees above 4096: worker
    out.append((sbm(4, 60, 0.3, 0.01, seed=60), dict(kind=O.WCC, seed=52), 1))  # WCC -i 1: worker
    out.append((sbm(2, 3000, 0.3, 0.01, seed=61), dict(kind=O.MOD, seed=53), 0))  # larger than the arena: worker
    return out


def test_served_results_equal_alone():
    reqs = requests()
    labs = [singleton_labels(g[0]) for g, _, _ in reqs]
    jobs = [parcd.job(**jd) for _, jd, _ in reqs]
    expect = [alone(g, lab, jb, ini) for (g, _, ini), lab, jb in zip(reqs, labs, jobs)]
    with parcd.Server(0, batch_arena_bytes=8 << 20, batch_slots=4, plain_workers=2, plain_bytes=1 << 30) as srv:
        ids = {}
        for i, ((g, _, ini), lab, jb) in enumerate(zip(reqs, labs, jobs)):
            ids[srv.submit(g[0], g[1], g[2], g[3], None if ini else lab, jb, initial=ini)] = i
            if i % 4 == 3:
                time.sleep(0.002)  # later requests arrive while earlier ones run
        got = {}
        t0 = time.time()
        while len(got) < len(ids):
            for rid in srv.wait(200):
                got[ids[rid]] = srv.result(rid)
            assert time.time() - t0 < 60
        rounds = srv.info()["rounds"]
    routes = [got[i][5] for i in range(len(reqs))]
    assert routes[:13] == [0] * 13 and routes[13:] == [1, 1, 1]
    assert rounds > 0
    for i, (el, es, eit, ek) in enumerate(expect):
        gl, gs, git, gk, lat, route = got[i]
        assert np.array_equal(gl, el) and git == eit and gk == ek and gs == es, i
        assert lat > 0


def test_server_rejects_and_reports_errors():
    with parcd.Server(0, batch_arena_bytes=1 << 20, batch_slots=2, plain_workers=0, plain_bytes=0) as srv:
        g = bin_zoo()
        with pytest.raises(parcd.PcdError):   # needs a worker, none configured
            srv.submit(g[0], g[1], g[2], g[3], singleton_labels(g[0]), parcd.job())
        h = sbm(4, 50, 0.3, 0.01, seed=11)
        bad = singleton_labels(h[0])
        bad[3] = h[0] + 5                      # label out of range: the batch refuses it
        rid = srv.submit(h[0], h[1], h[2], h[3], bad, parcd.job())
        t0 = time.time()
        done = []
        while not done:
            done = srv.wait(200)
            assert time.time() - t0 < 30
        with pytest.raises(parcd.PcdError):
            srv.result(rid)
