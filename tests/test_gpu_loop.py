"""The device-resident loop (pcd_iterate_many): passes replayed from a
captured propagation pass and a captured merging pass in the order of the
schedule (-s 0, 1, 2), the stop rule of main.cpp:241-242 (and -n, the
iteration cap) on the device.  Its results must equal pcd_iterate called
pass by pass (graph_loop 0), and the serial replay of the oracle: labels,
scores, iteration counts, k and statistics."""
import numpy as np
import pytest

import oracle_lib as O
import gpu_graphs as GG

pytestmark = pytest.mark.gpu
parcd = pytest.importorskip("parcd")


def graphs():
    return {"sbm": GG.sbm(8, 120, 0.12, 0.004, seed=31), "bin_zoo": GG.bin_zoo(), "rmat12": GG.rmat(12),
            "unsorted": GG.unsorted_copy(GG.sbm(6, 90, 0.2, 0.01, seed=32))}


JOBS = [dict(kind=O.MOD, seed=3), dict(kind=O.CPM, lam=0.02, seed=4, tie=O.TIE_LOWEST),
        dict(kind=O.MOD, seed=5, max_iter=7), dict(kind=O.CPM, lam=0.3, seed=6, iter_cap=9)]


def per_pass(ctx, jb, it0, count, prev=0.0, strategy=0):
    """pcd_iterate pass by pass: (passes, last score, last k, done)."""
    out = (0, prev, 0, False)
    for i in range(count):
        s, k, d = ctx.iterate(jb, it0 + i, prev, strategy=strategy)
        prev = s
        out = (i + 1, s, k, d)
        if d:
            break
    return out


@pytest.mark.parametrize("name", ["sbm", "bin_zoo", "rmat12", "unsorted"])
@pytest.mark.parametrize("j", range(len(JOBS)))
def test_refine_equals_pass_by_pass_and_replay(name, j):
    g = graphs()[name]
    n, m, ind, adj, _ = g
    jd = JOBS[j]
    jb = parcd.job(**jd)
    lab = GG.singleton_labels(n)
    res = {}
    for loop in (1, 0):
        with parcd.Context(0) as ctx:
            ctx.set_tuning("graph_loop", loop)
            ctx.upload_graph(n, m, ind, adj)
            ctx.upload_labels(lab)
            score, iters, k = ctx.refine(jb)
            res[loop] = (ctx.labels(), score, iters, k, ctx.stats(k))
    a, b = res[1], res[0]
    assert np.array_equal(a[0], b[0]) and a[1:4] == b[1:4]
    for x, y in zip(a[4], b[4]):
        assert np.array_equal(x, y)
    # the oracle has no iteration cap: a capped run is replayed with -n cap
    capped = jd.get("iter_cap") and b[2] >= jd["iter_cap"]
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=jd["kind"], lam=jd.get("lam", 0.0),
                                max_iter=jd["iter_cap"] if capped else jd.get("max_iter", -1), mode=1,
                                seed=jd["seed"], tie=jd.get("tie", O.TIE_RANDOM), init_labels=lab)
    assert np.array_equal(a[0], ol) and (a[2], a[3]) == (oit, ok)
    assert abs(a[1] - osc) <= 1e-9 * max(1.0, abs(osc))


@pytest.mark.parametrize("name", ["bin_zoo", "rmat12"])
@pytest.mark.parametrize("j", [0, 1])
def test_bin_streams_equal_one_stream(name, j):
    """Propagation with the hub rows and degrees 1025-4096 on side streams
    (bin_streams 1): the captured loop, whose passes then hold parallel
    branches, and pass-by-pass calls both equal the one-stream run."""
    g = graphs()[name]
    n, m, ind, adj, _ = g
    jb = parcd.job(**JOBS[j])
    lab = GG.singleton_labels(n)
    res = []
    for st, loop in ((0, 1), (1, 1), (1, 0)):
        with parcd.Context(0) as ctx:
            ctx.set_tuning("bin_streams", st)
            ctx.set_tuning("bin_streams_min_edges", 0)  # the graphs are below the default threshold
            ctx.set_tuning("graph_loop", loop)
            ctx.upload_graph(n, m, ind, adj)
            ctx.upload_labels(lab)
            score, iters, k = ctx.refine(jb)
            res.append((ctx.labels(), score, iters, k))
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and r[1:] == res[0][1:]


@pytest.mark.parametrize("name", ["sbm", "bin_zoo"])
def test_ranges_of_passes_then_pass_by_pass(name):
    """pcd_iterate_many over a range that ends before the stop rule, then
    more passes: each range and the pass after it equal the per-pass loop;
    the partition left on the device is the compressed form of the same
    one."""
    g = graphs()[name]
    n, m, ind, adj, _ = g
    jb = parcd.job(kind=O.MOD, seed=9, max_iter=12)
    lab = GG.singleton_labels(n)
    with parcd.Context(0) as A, parcd.Context(0) as B:
        for c in (A, B):
            c.upload_graph(n, m, ind, adj)
            c.upload_labels(lab)
        it, prev = 0, 0.0
        for count in (1, 3, 2, 5):
            pa, sa, ka, da = A.iterate_many(jb, it, count, prev)
            pb, sb, kb, db = per_pass(B, jb, it, count, prev)
            assert (pa, sa, ka, da) == (pb, sb, kb, db), (it, count)
            # the same partition (A may hold it compressed)
            assert np.array_equal(O.compress_labels(A.labels())[0], O.compress_labels(B.labels())[0])
            it += pa
            prev = sa
            if da:
                break
        # one more pass each way
        if not da:
            s1, k1, d1 = A.iterate(jb, it, prev)
            s2, k2, d2 = B.iterate(jb, it, prev)
            assert (s1, k1, d1) == (s2, k2, d2)
            assert np.array_equal(A.labels(), B.labels())


def test_stop_rule_fires_inside_the_replays():
    """A run long enough that the host polls several batches of replays
    before the rule fires on the device."""
    g = GG.sbm(40, 50, 0.2, 0.002, seed=33)
    n, m, ind, adj, _ = g
    lab = GG.singleton_labels(n)
    jb = parcd.job(kind=O.MOD, seed=12, threshold=0.0)
    with parcd.Context(0) as ctx:
        ctx.upload_graph(n, m, ind, adj)
        ctx.upload_labels(lab)
        passes, score, k, done = ctx.iterate_many(jb, 0, 1000, 0.0)
        got = ctx.labels()
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=O.MOD, mode=1, seed=12, init_labels=lab)
    assert done and passes == oit and k == ok
    assert np.array_equal(got, ol)
    assert abs(score - osc) <= 1e-9 * max(1.0, abs(osc))


@pytest.mark.parametrize("name", ["sbm", "bin_zoo"])
def test_wcc_refine_equals_pass_by_pass_and_replay(name):
    """WCC from the triangle-based initial partition (-i 1): the captured
    pass runs the internal-triangle metric on the device too."""
    g = graphs()[name]
    n, m, ind, adj, _ = g
    jb = parcd.job(kind=O.WCC, seed=7, iter_cap=25)
    res = {}
    for loop in (1, 0):
        with parcd.Context(0) as ctx:
            ctx.set_tuning("graph_loop", loop)
            ctx.upload_graph(n, m, ind, adj)
            cc = ctx.triangles()[2]
            ctx.clustering()
            score, iters, k = ctx.refine(jb)
            res[loop] = (ctx.labels(), score, iters, k, cc)
    a, b = res[1], res[0]
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]
    # the device's clustering coefficient, as the gains read it
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=O.WCC, initial=1, mode=1, seed=7,
                                max_iter=25 if b[2] >= 25 else -1, cc_override=a[4])
    assert np.array_equal(a[0], ol) and (a[2], a[3]) == (oit, ok)
    assert abs(a[1] - osc) <= 1e-9 * max(1.0, abs(osc))


# merging can take many passes to stop on its own (each pass merges only
# mutual best pairs): every job carries an iteration cap, and the oracle is
# replayed with -n equal to it when the run reached it
MERGE_JOBS = [dict(kind=O.MOD, seed=21, iter_cap=40), dict(kind=O.CPM, lam=0.02, seed=22, tie=O.TIE_LOWEST, iter_cap=40),
              dict(kind=O.MOD, seed=23, max_iter=6, iter_cap=40), dict(kind=O.CPM, lam=0.3, seed=24, iter_cap=7)]


@pytest.mark.parametrize("strategy,rounds", [(1, 1), (1, 3), (2, 1), (2, 2)])
@pytest.mark.parametrize("name", ["sbm", "bin_zoo", "rmat12"])
@pytest.mark.parametrize("j", range(len(MERGE_JOBS)))
def test_merging_schedules_equal_pass_by_pass_and_replay(name, j, strategy, rounds):
    """-s 1 and -s 2 on the device: merging passes (community rows laid out
    on the device, hub rows of bin_zoo included) replayed alone or
    alternating with propagation passes."""
    g = graphs()[name]
    n, m, ind, adj, _ = g
    jd = MERGE_JOBS[j]
    jb = parcd.job(**jd)
    lab = GG.singleton_labels(n)
    res = {}
    for loop in (1, 0):
        with parcd.Context(0) as ctx:
            ctx.set_tuning("graph_loop", loop)
            ctx.upload_graph(n, m, ind, adj)
            ctx.upload_labels(lab)
            score, iters, k = ctx.refine(jb, strategy=strategy, rounds=rounds)
            res[loop] = (ctx.labels(), score, iters, k, ctx.stats(k))
    a, b = res[1], res[0]
    assert np.array_equal(a[0], b[0]) and a[1:4] == b[1:4]
    for x, y in zip(a[4], b[4]):
        assert np.array_equal(x, y)
    capped = jd.get("iter_cap") and b[2] >= jd["iter_cap"]
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=jd["kind"], lam=jd.get("lam", 0.0), strategy=strategy,
                                rounds=rounds, max_iter=jd["iter_cap"] if capped else jd.get("max_iter", -1), mode=1,
                                seed=jd["seed"], tie=jd.get("tie", O.TIE_RANDOM), init_labels=lab)
    assert np.array_equal(a[0], ol) and (a[2], a[3]) == (oit, ok)
    assert abs(a[1] - osc) <= 1e-9 * max(1.0, abs(osc))


@pytest.mark.parametrize("strategy", [1, 2])
def test_wcc_merging_schedules_on_device(strategy):
    """WCC from the triangle-based partition with -s 1 / -s 2 on the device."""
    g = graphs()["sbm"]
    n, m, ind, adj, _ = g
    jb = parcd.job(kind=O.WCC, seed=25, iter_cap=20)
    res = {}
    for loop in (1, 0):
        with parcd.Context(0) as ctx:
            ctx.set_tuning("graph_loop", loop)
            ctx.upload_graph(n, m, ind, adj)
            cc = ctx.triangles()[2]
            ctx.clustering()
            score, iters, k = ctx.refine(jb, strategy=strategy)
            res[loop] = (ctx.labels(), score, iters, k, cc)
    a, b = res[1], res[0]
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=O.WCC, initial=1, strategy=strategy, mode=1, seed=25,
                                max_iter=20 if b[2] >= 20 else -1, cc_override=a[4])
    assert np.array_equal(a[0], ol) and (a[2], a[3]) == (oit, ok)
    assert abs(a[1] - osc) <= 1e-9 * max(1.0, abs(osc))


def test_merging_pass_ranges_then_single_passes():
    """pcd_iterate_many over ranges of -s 2 passes (a range may end on either
    step kind), then single passes: each equals the per-pass loop."""
    g = graphs()["bin_zoo"]
    n, m, ind, adj, _ = g
    jb = parcd.job(kind=O.MOD, seed=26, max_iter=14)
    lab = GG.singleton_labels(n)
    with parcd.Context(0) as A, parcd.Context(0) as B:
        for c in (A, B):
            c.upload_graph(n, m, ind, adj)
            c.upload_labels(lab)
        it, prev, da = 0, 0.0, False
        for count in (1, 3, 2, 4, 3):
            pa, sa, ka, da = A.iterate_many(jb, it, count, prev, strategy=2)
            pb, sb, kb, db = per_pass(B, jb, it, count, prev, strategy=2)
            assert (pa, sa, ka, da) == (pb, sb, kb, db), (it, count)
            assert np.array_equal(O.compress_labels(A.labels())[0], O.compress_labels(B.labels())[0])
            it += pa
            prev = sa
            if da:
                break
