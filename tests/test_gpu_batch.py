"""Batched refinement on the GPU: every graph of a batch ends exactly where
the same graph run alone ends (plain context loop and the serial replay),
whatever the batch's composition, admission time or bin map."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_graphs import TINY, csr_direct, rmat, sbm, singleton_labels, unsorted_copy

pytestmark = pytest.mark.gpu
parcd = pytest.importorskip("parcd")


def graphs():
    """Mixed shapes: SBMs of several sizes, a small R-MAT (degrees up to
    a few hundred), a graph that is not degree-sorted."""
    out = [sbm(4, 50, 0.3, 0.01, seed=11), sbm(10, 100, 0.08, 0.002, seed=12), sbm(2, 400, 0.2, 0.01, seed=13),
           rmat(10, seed=5), unsorted_copy(sbm(6, 80, 0.2, 0.01, seed=14), seed=9), sbm(3, 20, 0.5, 0.05, seed=15)]
    return out


JOBS = [
    dict(kind=O.MOD, seed=1), dict(kind=O.CPM, lam=0.02, seed=2), dict(kind=O.MOD, seed=3, tie=O.TIE_LOWEST),
    dict(kind=O.CPM, lam=0.005, seed=4, max_iter=3), dict(kind=O.MOD, seed=5, frac=0.8),
    dict(kind=O.MOD, seed=6, threshold=0.01),
]


def plain(g, lab, jb):
    n, m, ind, adj, _ = g
    with parcd.Context(0) as ctx:
        return ctx.run(n, m, ind, adj, lab, jb)


def run_batch(items, map_=None, rounds=1):
    with parcd.Batch(0, max_graphs=16, arena_bytes=256 << 20) as b:
        if map_ is not None:
            b.set_map(map_)
        slots = [b.add(g[0], g[1], g[2], g[3], lab, jb) for g, lab, jb in items]
        assert all(s is not None for s in slots)
        while b.step(rounds):
            pass
        return [(b.labels(s),) + b.status(s)[1:] + (b.status(s)[0],) for s in slots]


def test_batch_equals_plain_and_replay():
    gs = graphs()
    items = [(g, singleton_labels(g[0]), parcd.job(**jb)) for g, jb in zip(gs, JOBS)]
    res = run_batch(items)
    for (g, lab, jb), (bl, iters, k, score, status), jd in zip(items, res, JOBS):
        assert status == parcd.BATCH_DONE
        pl, ps, pit, pk = plain(g, lab, jb)
        assert np.array_equal(bl, pl) and iters == pit and k == pk and score == ps
        n, m, ind, adj, _ = g
        ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=jd["kind"], lam=jd.get("lam", 0.0),
                                    frac=jd.get("frac", 0.5), max_iter=jd.get("max_iter", -1),
                                    threshold=jd.get("threshold", 0.0), mode=1, seed=jd["seed"],
                                    tie=jd.get("tie", O.TIE_RANDOM), init_labels=lab)
        assert np.array_equal(bl, ol) and iters == oit and k == ok
        assert abs(score - osc) <= 1e-9 * max(1.0, abs(osc))


def test_result_independent_of_the_mix():
    gs = graphs()
    lab = [singleton_labels(g[0]) for g in gs]
    jobs = [parcd.job(**jb) for jb in JOBS]
    alone = run_batch([(gs[1], lab[1], jobs[1])])[0]
    mixed = run_batch([(g, l, j) for g, l, j in zip(gs, lab, jobs)])[1]
    assert np.array_equal(alone[0], mixed[0]) and alone[1:] == mixed[1:]


def test_every_bin_map_gives_the_same_result():
    gs = graphs()
    items = [(g, singleton_labels(g[0]), parcd.job(**jb)) for g, jb in zip(gs, JOBS)]
    ref = run_batch(items, map_=0)
    for mp in range(1, len(parcd.batch_map_names())):
        got = run_batch(items, map_=mp, rounds=3)
        for a, b in zip(ref, got):
            assert np.array_equal(a[0], b[0]) and a[1:] == b[1:], parcd.batch_map_names()[mp]


def test_admission_between_rounds_and_slot_reuse():
    gs = graphs()
    lab = [singleton_labels(g[0]) for g in gs]
    jobs = [parcd.job(**jb) for jb in JOBS]
    expect = [plain(g, l, j) for g, l, j in zip(gs, lab, jobs)]
    with parcd.Batch(0, max_graphs=3, arena_bytes=64 << 20) as b:
        pending = list(range(len(gs)))
        live = {}
        done = {}
        rounds = 0
        while pending or live:
            while pending:
                i = pending[0]
                s = b.add(gs[i][0], gs[i][1], gs[i][2], gs[i][3], lab[i], jobs[i])
                if s is None:
                    break
                live[s] = i
                pending.pop(0)
            b.step(1)
            rounds += 1
            for s in b.finished():
                done[live.pop(s)] = (b.labels(s), b.status(s))
                b.release(s)
            assert rounds < 10000
    for i, (pl, ps, pit, pk) in enumerate(expect):
        bl, (st, it, k, sc) = done[i]
        assert np.array_equal(bl, pl) and it == pit and k == pk and sc == ps


def test_autotune_keeps_results():
    gs = graphs()
    items = [(g, singleton_labels(g[0]), parcd.job(**jb)) for g, jb in zip(gs, JOBS)]
    ref = run_batch(items)
    with parcd.Batch(0, max_graphs=16, arena_bytes=256 << 20) as b:
        slots = [b.add(g[0], g[1], g[2], g[3], lab, jb) for g, lab, jb in items]
        b.step(1)
        chosen, ms = b.autotune(reps=2)
        assert 0 <= chosen < len(ms) and all(t > 0 for t in ms)
        while b.step(2):
            pass
        for s, r in zip(slots, ref):
            assert np.array_equal(b.labels(s), r[0]) and b.status(s)[1:] == r[1:4]


@pytest.mark.parametrize("case", ["mod", "cpm_leave", "capped"])
def test_batch_statistics_match_plain(case):
    """Statistics at the end (updated from the moved vertices every round
    after the first) equal the oracle's full count; the CPM case starts from
    a coarse random partition at a resolution where vertices leave to
    NEW_LABEL singletons, the capped case stops after a few rounds while
    many vertices still move."""
    g = sbm(10, 100, 0.08, 0.002, seed=12)
    n = g[0]
    if case == "mod":
        lab, jb = singleton_labels(n), parcd.job(kind=O.MOD, seed=1)
    elif case == "cpm_leave":
        lab, jb = np.random.default_rng(4).integers(0, 4, n).astype(np.int32), parcd.job(kind=O.CPM, lam=0.3, seed=2)
    else:
        lab, jb = singleton_labels(n), parcd.job(kind=O.MOD, seed=3, max_iter=3)
    with parcd.Batch(0, max_graphs=2, arena_bytes=64 << 20) as b:
        s = b.add(g[0], g[1], g[2], g[3], lab, jb)
        while b.step(1):
            pass
        bs = b.stats(s)
        bl = b.labels(s)
        st = b.status(s)
    k, sz, iv, ev = O.collect_statistics(g[0], g[2], g[3], bl)
    assert k == len(bs[0]) == st[2]
    assert np.array_equal(bs[0], sz) and np.array_equal(bs[1], iv) and np.array_equal(bs[2], ev)
    pl, ps, pit, pk = plain(g, lab, jb)
    assert np.array_equal(bl, pl) and st[1:] == (pit, pk, ps)


def test_batch_rejections():
    with parcd.Batch(0, max_graphs=1, arena_bytes=1 << 20) as b:
        g = sbm(4, 50, 0.3, 0.01, seed=11)
        lab = singleton_labels(g[0])
        with pytest.raises(parcd.PcdError):
            b.add(g[0], g[1], g[2], g[3], lab, parcd.job(kind=O.WCC))
        bad = lab.copy()
        bad[0] = g[0]
        with pytest.raises(parcd.PcdError):
            b.add(g[0], g[1], g[2], g[3], bad, parcd.job())
        big = sbm(2, 2000, 0.3, 0.01, seed=3)  # ~64 MB footprint > 1 MB arena
        assert b.add(big[0], big[1], big[2], big[3], singleton_labels(big[0]), parcd.job()) is None
        s = b.add(g[0], g[1], g[2], g[3], lab, parcd.job())
        assert s == 0
        assert b.add(g[0], g[1], g[2], g[3], lab, parcd.job()) is None  # no free slot


def test_add_many_and_labels_many_equal_one_by_one():
    """pcd_batch_add_many admits what one pcd_batch_add per graph admits, in
    the same slots, with the same results; pcd_batch_labels_many returns
    every graph's labels in the order asked for."""
    gs = graphs() * 3  # 18 graphs: several runs of consecutive slots
    items = [(g, singleton_labels(g[0]), parcd.job(**JOBS[i % len(JOBS)])) for i, g in enumerate(gs)]
    ref = run_batch(items[: len(JOBS)])
    with parcd.Batch(0, max_graphs=32, arena_bytes=512 << 20) as b:
        # a hole at slot 1 makes the admitted slots two runs
        s0 = b.add(gs[0][0], gs[0][1], gs[0][2], gs[0][3], items[0][1], items[0][2])
        s1 = b.add(gs[1][0], gs[1][1], gs[1][2], gs[1][3], items[1][1], items[1][2])
        b.release(s1)
        slots = b.add_many([g[:4] for g, _, _ in items[1:]], [x[1] for x in items[1:]], [x[2] for x in items[1:]])
        assert slots[0] == s1 and all(s is not None for s in slots)
        slots = [s0] + slots
        assert len(set(slots)) == len(slots)
        while b.step(1):
            pass
        order = slots[::-1]
        many = b.labels_many(order)
        for s, lab in zip(order, many):
            assert np.array_equal(lab, b.labels(s))
        for i, s in enumerate(slots):
            r = ref[i % len(JOBS)]
            assert np.array_equal(b.labels(s), r[0]) and b.status(s)[1:] == r[1:4], i


def test_add_many_from_page_locked_memory():
    """Graphs in page-locked host memory are read by the device directly
    (pull pieces of 1 MiB); results equal the pageable-source admission."""
    gs = graphs() + [sbm(2, 3000, 0.1, 0.001, seed=31)]  # one adjacency above one piece
    jobs = [parcd.job(**JOBS[i % len(JOBS)]) for i in range(len(gs))]
    labs = [singleton_labels(g[0]) for g in gs]

    def pinned(a):
        out = parcd.pinned_empty(len(a), np.int32)
        out[:] = a
        return out

    res = []
    for use_pinned in (False, True):
        conv = pinned if use_pinned else (lambda a: np.ascontiguousarray(a, np.int32))
        with parcd.Batch(0, max_graphs=16, arena_bytes=512 << 20) as b:
            slots = b.add_many([(g[0], g[1], conv(g[2]), conv(g[3])) for g in gs], [conv(l) for l in labs], jobs)
            while b.step(1):
                pass
            res.append([(b.labels(s_), b.status(s_)) for s_ in slots])
    assert gs[-1][1] * 2 > (1 << 18)
    for (la, sa), (lb, sb) in zip(res[0], res[1]):
        assert np.array_equal(la, lb) and sa == sb


def test_add_many_rejections_leave_the_rest_admitted():
    g = sbm(4, 50, 0.3, 0.01, seed=11)
    lab = singleton_labels(g[0])
    bad = lab.copy()
    bad[0] = g[0]
    big = sbm(2, 2000, 0.3, 0.01, seed=3)  # footprint above the arena
    with parcd.Batch(0, max_graphs=3, arena_bytes=4 << 20) as b:
        gl = [g[:4], big[:4], g[:4], g[:4], g[:4]]
        ll = [lab, singleton_labels(big[0]), lab, lab, lab]
        slots = b.add_many(gl, ll, [parcd.job()] * 5)
        # the big graph finds no arena room; the fifth graph no slot
        assert slots[1] is None and slots[4] is None
        assert b.last_codes[1] == -3 and b.last_codes[4] == -3
        assert sorted(s for s in slots if s is not None) == [0, 1, 2]
        b.release(slots[2])
        with pytest.raises(parcd.PcdError):
            b.add_many([g[:4], g[:4]], [bad, lab], [parcd.job(), parcd.job()])
        assert b.last_codes[0] == -1  # PCD_E_INVALID
        assert b.last_codes[1] == 0
        with pytest.raises(parcd.PcdError):
            b.add_many([g[:4]], [lab], [parcd.job(kind=O.WCC)])
        with pytest.raises(parcd.PcdError):
            b.labels_many([slots[0]])  # still running
        while b.step(1):
            pass
        live = [s for s in b.finished()]
        assert len(live) == 3
        ref = run_batch([(g, lab, parcd.job())])[0]
        for s, l in zip(live, b.labels_many(live)):
            assert np.array_equal(l, ref[0])


def test_add_many_checks_rows_on_the_device():
    """pcd_batch_add_many checks the adjacency rows on the device after the
    upload: a descending row, a self loop and an id out of range are refused
    with PCD_E_FORMAT, and their slots and arena space come back."""
    g = sbm(4, 50, 0.3, 0.01, seed=11)
    lab = singleton_labels(g[0])
    r0, r1 = int(g[2][5]), int(g[2][6])
    assert r1 - r0 >= 2
    broken = []
    for kind_ in range(3):
        adj = g[3].copy()
        if kind_ == 0:
            adj[r0], adj[r0 + 1] = adj[r0 + 1], adj[r0]
        elif kind_ == 1:
            adj[r0] = 5
        else:
            adj[r1 - 1] = g[0]
        broken.append((g[0], g[1], g[2], adj))
    ref = run_batch([(g, lab, parcd.job())])[0]
    with parcd.Batch(0, max_graphs=4, arena_bytes=4 << 20) as b:
        s0 = b.add(g[0], g[1], g[2], g[3], lab, parcd.job())
        free0 = b.info()["free_bytes"]
        with pytest.raises(parcd.PcdError):
            b.add_many(broken, [lab] * 3, [parcd.job()] * 3)
        assert b.last_codes == [-5, -5, -5]  # PCD_E_FORMAT
        assert b.info()["free_bytes"] == free0 and b.info()["graphs"] == 1
        with pytest.raises(parcd.PcdError):
            b.add_many([g[:4], broken[1], g[:4]], [lab] * 3, [parcd.job()] * 3)
        ok = b.last_slots
        assert ok[1] is None and b.last_codes == [0, -5, 0]
        while b.step(1):
            pass
        for s_ in (s0, ok[0], ok[2]):
            assert np.array_equal(b.labels(s_), ref[0])
    # the single-graph call still checks on the host
    with parcd.Batch(0, max_graphs=2, arena_bytes=4 << 20) as b:
        with pytest.raises(parcd.PcdError):
            b.add(*broken[0], lab, parcd.job())


def _pinned(a):
    out = parcd.pinned_empty(len(a), np.int32)
    out[:] = a
    return out


def _run_to_end(b):
    while b.step(1):
        pass


def test_add_many_async_equals_add_many():
    """pcd_batch_add_many_async: graphs are pending until a step joins them,
    then end exactly where pcd_batch_add_many's do, over several chunks of
    256 graphs, from page-locked or pageable arrays (labels and degree
    orders staged), admitted before the first round or while others run."""
    base = graphs() + [sbm(2, 3000, 0.1, 0.001, seed=31)]
    gs = (base * 90)[:620]  # three chunks
    jobs = [parcd.job(**JOBS[i % len(JOBS)]) for i in range(len(gs))]
    labs = [singleton_labels(g[0]) for g in gs]
    arena = sum(parcd.batch_footprint(g[0], g[1]) for g in gs) + (64 << 20)
    with parcd.Batch(0, max_graphs=len(gs), arena_bytes=arena) as b:
        slots = b.add_many([g[:4] for g in gs], labs, jobs)
        _run_to_end(b)
        ref = [(b.labels(s_), b.status(s_)) for s_ in slots]
    pinned_csr = ([(g[0], g[1], _pinned(g[2]), _pinned(g[3])) for g in base] * 90)[:620]
    for case in ("pageable", "pinned", "pinned_all", "while_running"):
        print("async admission:", case, flush=True)
        with parcd.Batch(0, max_graphs=len(gs), arena_bytes=arena) as b:
            csr = [g[:4] for g in gs] if case == "pageable" else pinned_csr
            ll = [_pinned(l) for l in labs] if case == "pinned_all" else labs
            if case == "while_running":
                head = b.add_many(csr[:18], ll[:18], jobs[:18])
                b.step(1)
                b.step(1)
                tail = b.add_many(csr[18:], ll[18:], jobs[18:], wait=False)
                assert all(b.status(s_)[0] == parcd.BATCH_PENDING for s_ in tail)
                slots = head + tail
            else:
                slots = b.add_many(csr, ll, jobs, wait=False)
                assert all(s_ is not None for s_ in slots)
                assert all(b.status(s_)[0] == parcd.BATCH_PENDING for s_ in slots)
                assert b.info()["running"] == len(gs)
            _run_to_end(b)
            for i, s_ in enumerate(slots):
                la, sa = b.labels(s_), b.status(s_)
                assert np.array_equal(la, ref[i][0]) and sa == ref[i][1], (case, i)


def test_add_many_async_rejections_and_release():
    """Host refusals come back from the call; rows that fail the device
    check leave their graph PCD_BATCH_REJECTED at the join, without arena
    space; a pending slot can be released (its uploads are waited for); a
    second call joins the first's chunks."""
    g = sbm(4, 50, 0.3, 0.01, seed=11)
    lab = singleton_labels(g[0])
    r0 = int(g[2][5])
    adj = g[3].copy()
    adj[r0], adj[r0 + 1] = adj[r0 + 1], adj[r0]
    broken = (g[0], g[1], g[2], adj)
    ref = run_batch([(g, lab, parcd.job())])[0]
    with parcd.Batch(0, max_graphs=8, arena_bytes=16 << 20) as b:
        free0 = b.info()["free_bytes"]
        with pytest.raises(parcd.PcdError):
            b.add_many([g[:4], g[:4], broken, g[:4]], [lab] * 4,
                       [parcd.job(), parcd.job(kind=O.WCC), parcd.job(), parcd.job()], wait=False)
        sl = b.last_slots
        assert b.last_codes == [0, -1, 0, 0] and sl[1] is None
        assert [b.status(sl[i])[0] for i in (0, 2, 3)] == [parcd.BATCH_PENDING] * 3
        b.release(sl[3])  # pending: waited for, then freed
        assert b.info()["graphs"] == 2
        # a second call joins the first call's chunks before it queues
        more = b.add_many([g[:4]], [lab], [parcd.job()], wait=False)
        assert b.status(sl[0])[0] == parcd.BATCH_RUNNING
        assert b.status(sl[2])[0] == parcd.BATCH_REJECTED
        assert b.status(more[0])[0] == parcd.BATCH_PENDING
        _run_to_end(b)
        assert b.status(sl[2])[0] == parcd.BATCH_REJECTED
        for s_ in (sl[0], more[0]):
            assert b.status(s_)[0] == parcd.BATCH_DONE
            assert np.array_equal(b.labels(s_), ref[0])
        with pytest.raises(parcd.PcdError):
            b.labels(sl[2])
        for s_ in (sl[0], sl[2], more[0]):
            b.release(s_)
        assert b.info()["graphs"] == 0 and b.info()["free_bytes"] == free0


@pytest.mark.parametrize("strategy", [0, 1, 2])
def test_refine_call_equals_replay(strategy):
    g = sbm(6, 60, 0.25, 0.01, seed=21)
    n, m, ind, adj, _ = g
    lab = singleton_labels(n)
    jb = parcd.job(kind=O.MOD, seed=3, tie=O.TIE_LOWEST)
    with parcd.Context(0) as ctx:
        ctx.upload_graph(n, m, ind, adj)
        ctx.upload_labels(lab)
        score, iters, k = ctx.refine(jb, strategy=strategy)
        got = ctx.labels()
    ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=O.MOD, strategy=strategy, mode=1, seed=3, tie=O.TIE_LOWEST,
                                init_labels=lab)
    assert np.array_equal(got, ol) and iters == oit and k == ok
    assert abs(score - osc) <= 1e-9 * max(1.0, abs(osc))


def test_batch_tiny_graphs_with_isolated_vertices():
    """Graphs with isolated vertices, a star and a clique end where they end
    alone and in the replay; a graph without edges is refused (batches take
    m >= 1) and leaves the others running."""
    names = ["edge_and_isolated", "path_and_isolated", "star", "clique5"]
    gs = [csr_direct(*TINY[k]) for k in names]
    jobs = [dict(kind=O.MOD, seed=1, max_iter=4), dict(kind=O.CPM, lam=0.05, seed=2, max_iter=4),
            dict(kind=O.MOD, seed=3, tie=O.TIE_LOWEST, max_iter=4), dict(kind=O.CPM, lam=0.3, seed=4, max_iter=4)]
    items = [(g, singleton_labels(g[0]), parcd.job(**jb)) for g, jb in zip(gs, jobs)]
    with parcd.Batch(0, max_graphs=8, arena_bytes=64 << 20) as b:
        e = csr_direct(*TINY["no_edges"])
        with pytest.raises(parcd.PcdError):
            b.add(e[0], e[1], e[2], e[3], singleton_labels(e[0]), parcd.job(kind=O.MOD, max_iter=2))
        slots = [b.add(g[0], g[1], g[2], g[3], lab, jb) for g, lab, jb in items]
        while b.step(1):
            pass
        res = [(b.labels(s),) + b.status(s)[1:] + (b.status(s)[0],) for s in slots]
    for (g, lab, jb), (bl, iters, k, score, status), jd in zip(items, res, jobs):
        assert status == parcd.BATCH_DONE
        pl, ps, pit, pk = plain(g, lab, jb)
        assert np.array_equal(bl, pl) and iters == pit and k == pk and score == ps
        n, m, ind, adj, _ = g
        ol, osc, oit, ok, _ = O.run(n, m, ind, adj, kind=jd["kind"], lam=jd.get("lam", 0.0), max_iter=4, mode=1,
                                    seed=jd["seed"], tie=jd.get("tie", O.TIE_RANDOM), init_labels=lab)
        assert np.array_equal(bl, ol) and iters == oit and k == ok


def test_batch_admits_a_graph_set_in_place(tmp_path):
    """Graphs read from a graph set file (pointers into its read-only
    mapping, no host copies) are admitted with pcd_batch_add_many and end
    exactly where the same graphs admitted from their own arrays end."""
    gs = graphs()
    path = str(tmp_path / "batch.pcds")
    parcd.write_graph_set(path, [g[:4] for g in gs])
    items = [(g, singleton_labels(g[0]), parcd.job(**JOBS[i % len(JOBS)])) for i, g in enumerate(gs)]
    ref = run_batch(items)
    with parcd.GraphSet(path) as s, parcd.Batch(0, max_graphs=16, arena_bytes=256 << 20) as b:
        views = [s.graph(i)[:4] for i in range(len(s))]
        for v, g in zip(views, gs):
            assert v[:2] == g[:2] and np.array_equal(v[2], g[2]) and np.array_equal(v[3], g[3])
        slots = b.add_many(views, [x[1] for x in items], [x[2] for x in items])
        assert all(x is not None for x in slots)
        while b.step(1):
            pass
        for i, sl in enumerate(slots):
            assert np.array_equal(b.labels(sl), ref[i][0]) and b.status(sl)[1:] == ref[i][1:4], i
