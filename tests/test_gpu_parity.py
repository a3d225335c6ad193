"""GPU parity: the HIP path (through the C ABI) against the oracle.

Integer results (compressed labels, statistics, triangles, contraction
pairs, refinement decisions in counter mode) must be bit-exact; metrics
must agree to 1e-9 relative.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
import gpu_graphs as GG
from conftest import golden, load

pytestmark = pytest.mark.gpu

parcd = pytest.importorskip("parcd")


@pytest.fixture(scope="module")
def ctx():
    if parcd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    c = parcd.Context(0)
    yield c
    c.close()


def graphs():
    out = {}
    for p in golden("graph_*.npz"):
        d = load(p)
        out["golden_" + os.path.basename(p)[6:-4]] = O.parse_pcd(d["pcd"].tobytes())
    return out


GOLDEN = graphs()
RULES = [(O.MOD, 0.0), (O.CPM, 0.05), (O.CPM, 0.01)]
BIN_STREAMS_DEFAULT = 1  # the library's default of the bin_streams tuning key


def upload(ctx, g):
    n, m, ind, adj, mp = g
    ctx.upload_graph(n, m, ind, adj)
    return n, m, ind, adj


def oracle_stats(g, lab):
    n, m, ind, adj = g[:4]
    return O.collect_statistics(n, ind, adj, lab)


def check_partition(ctx, g, lab_in, lam=0.05):
    n, m, ind, adj = upload(ctx, g)
    ctx.upload_labels(lab_in)
    k = ctx.compress()
    lab_o, k_o = O.compress_labels(lab_in)
    assert k == k_o
    lab = ctx.labels()
    assert np.array_equal(lab, lab_o)
    k2 = ctx.statistics()
    assert k2 == k_o
    s, iv, ev = ctx.stats(k2)
    _, so, ivo, evo = oracle_stats(g, lab_o)
    assert np.array_equal(s, so) and np.array_equal(iv, ivo) and np.array_equal(ev, evo)
    for kind, lm in [(O.MOD, 0.0), (O.CPM, lam), (O.COND, 0.0)]:
        got = ctx.metric(kind, lm)
        want = O.calculate(kind, lm, n, m, ind, adj, lab_o, so, ivo, evo)
        if np.isnan(want):
            assert np.isnan(got)
        else:
            assert got == pytest.approx(want, rel=1e-9, abs=1e-12), kind
    return lab_o, so, ivo, evo


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_partition_golden_graphs(ctx, name):
    g = GOLDEN[name]
    n = g[0]
    rng = np.random.default_rng(1)
    for lab in [np.arange(n)[rng.permutation(n)], rng.integers(0, max(2, n // 7), n),
                np.where(rng.random(n) < 0.3, -1, rng.integers(0, n, n) * 2 + 5), np.zeros(n)]:
        check_partition(ctx, g, lab)


def test_partition_against_golden_fixtures(ctx):
    for path in golden("part_*.npz"):
        d = load(path)
        name = "golden_" + os.path.basename(path)[5:-4].rsplit("_", 1)[0]
        g = GOLDEN[name]
        upload(ctx, g)
        ctx.upload_labels(d["in_labels"])
        k = ctx.compress()
        assert k == int(d["k"][0])
        assert np.array_equal(ctx.labels(), d["labels"])
        ctx.statistics()
        s, iv, ev = ctx.stats(k)
        assert np.array_equal(s, d["sizes"]) and np.array_equal(iv, d["int_volume"])
        assert np.array_equal(ev, d["ext_volume"])
        lam = float(d["lam"][0])
        for kind, lm in [(0, 0.0), (1, lam), (3, 0.0)]:
            want = float(d["calculate"][kind])
            got = ctx.metric(kind, lm)
            assert (np.isnan(got) and np.isnan(want)) or got == pytest.approx(want, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("stats_cache", [0, 1])
def test_partition_bin_zoo_and_unsorted(ctx, stats_cache):
    """Statistics over long rows with and without the LDS label cache."""
    g = GG.bin_zoo()
    rng = np.random.default_rng(5)
    n = g[0]
    ctx.set_tuning("stats_cache", stats_cache)
    try:
        for gg in [g, GG.unsorted_copy(g), GG.dense_blocks()]:
            nn = gg[0]
            for lab in [GG.singleton_labels(nn), rng.integers(0, 300, nn), rng.integers(0, 5, nn)]:
                check_partition(ctx, gg, lab)
    finally:
        ctx.set_tuning("stats_cache", 1)


def _lp_compare(ctx, g, lab_in, kind, lam, frac, seed, step, tie, cc=None):
    n, m, ind, adj = g[:4]
    lab_o, so, ivo, evo = check_partition(ctx, g, lab_in)
    if cc is not None:
        ctx.set_cc(cc)
    ctx.propagate(kind, lam, frac=frac, seed=seed, step=step, tie=tie)
    got = ctx.labels()
    want = O.propagation_refine(kind, lam, n, m, ind, adj, lab_o, so, ivo, evo, cc or 0.0, frac, 1, seed, step, tie)
    assert np.array_equal(got, want), (np.nonzero(got != want)[0][:10], got[got != want][:10], want[got != want][:10])
    return got


@pytest.mark.parametrize("tie", [O.TIE_RANDOM, O.TIE_LOWEST])
@pytest.mark.parametrize("kind,lam", RULES)
def test_propagation_golden(ctx, kind, lam, tie):
    rng = np.random.default_rng(2)
    for name, g in sorted(GOLDEN.items()):
        n = g[0]
        for lab in [GG.singleton_labels(n), rng.integers(0, max(2, n // 9), n)]:
            for frac in (0.5, 1.0):
                _lp_compare(ctx, g, lab, kind, lam, frac, seed=1234, step=3, tie=tie)


@pytest.mark.parametrize("aux", [0, 1])
@pytest.mark.parametrize("tie", [O.TIE_RANDOM, O.TIE_LOWEST])
@pytest.mark.parametrize("kind,lam", RULES)
def test_propagation_bin_zoo(ctx, kind, lam, tie, aux):
    """Every degree bin; aux=1 switches the small, wave and block bins to
    {label, aux} record gathers (modularity and CPM)."""
    g = GG.bin_zoo()
    n = g[0]
    rng = np.random.default_rng(3)
    ctx.set_tuning("small_aux", aux)
    ctx.set_tuning("block_aux", aux)
    try:
        for lab in [GG.singleton_labels(n), rng.integers(0, 200, n), rng.integers(0, 20000, n)]:
            _lp_compare(ctx, g, lab, kind, lam, 0.5, seed=99, step=1, tie=tie)
            _lp_compare(ctx, g, lab, kind, lam, 1.0, seed=7, step=0, tie=tie)
        gu = GG.unsorted_copy(g)
        _lp_compare(ctx, gu, rng.integers(0, 500, n), kind, lam, 0.7, seed=5, step=2, tie=tie)
    finally:
        ctx.set_tuning("small_aux", 0)
        ctx.set_tuning("block_aux", 0)


@pytest.mark.parametrize("kind,lam", RULES)
@pytest.mark.parametrize("knob", ["small_fused", "small_occ8", "b1024_u16", "b4096_u16"])
def test_propagation_other_launch_forms(ctx, kind, lam, knob):
    """The small-degree bins launched one kernel each (small_fused 0), the
    fused launch built for the compiler's own occupancy (small_occ8 0), and
    the degree 257-512 (b1024_u16 0) and 1025-4096 (b4096_u16 0) tables with
    32-bit counts agree with the oracle as the default launches do."""
    g = GG.bin_zoo()
    n = g[0]
    rng = np.random.default_rng(9)
    # every degree 257-512 (SBM blocks of 600 at p_in 0.6): block labels,
    # one label and two alternating ones put counts of up to the degree into
    # both halves of the 16-bit count words
    big = GG.sbm(4, 600, 0.6, 0.01, seed=21)
    nb = big[0]
    assert np.all((np.diff(big[2]) > 256) & (np.diff(big[2]) <= 512))
    cases = [(g, [GG.singleton_labels(n), rng.integers(0, 200, n)]),
             (big, [big[4] // 600, np.zeros(nb, np.int64), np.arange(nb) % 2])]
    ctx.set_tuning(knob, 0)
    try:
        for gr, labs in cases:
            for lab in labs:
                _lp_compare(ctx, gr, lab, kind, lam, 0.5, seed=99, step=1, tie=O.TIE_RANDOM)
                _lp_compare(ctx, gr, lab, kind, lam, 1.0, seed=7, step=0, tie=O.TIE_LOWEST)
    finally:
        ctx.set_tuning(knob, 1)


@pytest.mark.parametrize("kind,lam", RULES)
def test_propagation_degree_257_512_rows(ctx, kind, lam):
    """The same rows under the default launches (16-bit counts)."""
    big = GG.sbm(4, 600, 0.6, 0.01, seed=21)
    nb = big[0]
    for lab in [big[4] // 600, np.zeros(nb, np.int64), np.arange(nb) % 2, GG.singleton_labels(nb)]:
        _lp_compare(ctx, big, lab, kind, lam, 0.5, seed=99, step=1, tie=O.TIE_RANDOM)
        _lp_compare(ctx, big, lab, kind, lam, 1.0, seed=7, step=0, tie=O.TIE_LOWEST)


@pytest.mark.parametrize("inc", [1, 0])
def test_statistics_follow_propagation(ctx, inc):
    """After each propagation step the statistics are updated from the moved
    vertices' rows (inc=1) or recounted (inc=0); after the next compression
    they must equal the oracle's full count of the new labels, including
    vertices that left to NEW_LABEL singletons (CPM at a high resolution)."""
    ctx.set_tuning("inc_stats", inc)
    left = 0
    try:
        for g in [GOLDEN["golden_rmat9"], GG.bin_zoo(), GG.unsorted_copy(GG.bin_zoo()), GG.dense_blocks(),
                  GG.rmat(12)]:
            n, m, ind, adj = upload(ctx, g)
            rng = np.random.default_rng(n)
            for kind, lam, init in [(O.MOD, 0.0, GG.singleton_labels(n)), (O.CPM, 0.3, rng.integers(0, 40, n)),
                                    (O.CPM, 0.01, GG.singleton_labels(n))]:
                ctx.upload_labels(init)
                for it in range(6):
                    k = ctx.compress()
                    lab = ctx.labels()
                    assert ctx.statistics() == k
                    s, iv, ev = ctx.stats(k)
                    ko, so, ivo, evo = O.collect_statistics(n, ind, adj, lab)
                    assert ko == k
                    assert np.array_equal(s, so) and np.array_equal(iv, ivo) and np.array_equal(ev, evo), it
                    assert ctx.metric(kind, lam) == pytest.approx(
                        O.calculate(kind, lam, n, m, ind, adj, lab, so, ivo, evo), rel=1e-9, abs=1e-12)
                    ctx.propagate(kind, lam, frac=0.6, seed=3, step=it)
                    left += int((ctx.labels() < 0).sum())
    finally:
        ctx.set_tuning("inc_stats", 1)
    assert left > 0  # NEW_LABEL moves were exercised


HUB_CASES = [(4096, 2048, 512, 0), (64, 2048, 512, 0), (1024, 512, 512, 0), (64, 100000, 512, 0),
             (4096, 100000, 512, 0), (2048, 1024, 256, 0), (64, 1024, 256, 0), (2048, 100000, 256, 0),
             (4096, 2048, 512, 1), (64, 2048, 512, 1), (1024, 512, 512, 1), (64, 100000, 512, 1),
             (2048, 1024, 256, 1), (64, 1024, 256, 1)]


@pytest.mark.parametrize("slots,part_entries,hub_nt,hub_aux", HUB_CASES)
def test_propagation_hub_paths(ctx, slots, part_entries, hub_nt, hub_aux):
    """Hub rows: partitioned LDS counting of labels (hub_aux 0; workgroups
    of 512 threads with 4096-slot tables or of 256 with 2048) or of
    {label, aux} records (hub_aux 1, modularity and CPM), and the
    global-table fallback when a partition overflows its table (forced by
    tiny tables or huge partitions).  WCC and conductance gains read the
    whole community record and always count labels."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    cc = O.count_triangles(n, ind, adj)[3]
    rng = np.random.default_rng(6)
    rules = [(O.MOD, 0.0, None), (O.CPM, 0.02, None)]
    if hub_aux == 0:
        rules += [(O.WCC, 0.0, cc), (O.COND, 0.0, None)]
    try:
        ctx.set_tuning("hub_part_entries", part_entries)
        ctx.set_tuning("hub_table_slots", slots)
        ctx.set_tuning("hub_aux", hub_aux)
        ctx.set_tuning("hub_nt", hub_nt)
        for lab in [GG.singleton_labels(n), rng.integers(0, 40, n), rng.integers(0, 3000, n)]:
            for kind, lam, c in rules:
                _lp_compare(ctx, g, lab, kind, lam, 1.0, seed=4, step=3, tie=O.TIE_RANDOM, cc=c)
                _lp_compare(ctx, g, lab, kind, lam, 0.5, seed=4, step=4, tie=O.TIE_LOWEST, cc=c)
    finally:
        ctx.set_tuning("hub_part_entries", 2048)
        ctx.set_tuning("hub_table_slots", 4096)
        ctx.set_tuning("hub_aux", 0)
        ctx.set_tuning("hub_nt", 512)


@pytest.mark.parametrize("streams", [0, 1])
@pytest.mark.parametrize("kind,lam", [(O.MOD, 0.0), (O.CPM, 0.02), (O.WCC, 0.0)])
def test_propagation_bin_streams(ctx, kind, lam, streams):
    """Hub rows and degrees 1025-4096 forked onto side streams while the other
    bins run on the context's stream (bin_streams 1, the default), and every
    bin on one stream (0): the step equals the oracle's either way."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    cc = O.count_triangles(n, ind, adj)[3] if kind == O.WCC else None
    rng = np.random.default_rng(8)
    ctx.set_tuning("bin_streams", streams)
    ctx.set_tuning("bin_streams_min_edges", 0)  # the graph is below the default threshold
    try:
        for lab in [GG.singleton_labels(n), rng.integers(0, 40, n), rng.integers(0, 3000, n)]:
            _lp_compare(ctx, g, lab, kind, lam, 1.0, seed=4, step=3, tie=O.TIE_RANDOM, cc=cc)
            _lp_compare(ctx, g, lab, kind, lam, 0.5, seed=4, step=4, tie=O.TIE_LOWEST, cc=cc)
    finally:
        ctx.set_tuning("bin_streams", BIN_STREAMS_DEFAULT)
        ctx.set_tuning("bin_streams_min_edges", 1 << 19)


def test_gather_probe(ctx):
    """pcd_gather_probe runs over the uploaded graph at one and two levels
    (labels outside [0, n), NEW_LABEL here, are not followed) and refuses
    other levels."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    ctx.upload_graph(n, m, ind, adj)
    lab = GG.singleton_labels(n)
    lab[::7] = -1
    ctx.upload_labels(lab)
    assert ctx.gather_probe(1, 2) > 0 and ctx.gather_probe(2, 2) > 0
    with pytest.raises(Exception):
        ctx.gather_probe(3, 2)


def test_hub_partitions_fit_with_default_tuning():
    """A fresh context's default tuning counts hub partitions in LDS (no
    overflow into the global-table fallback) on hubs of 5k-14k entries."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    with parcd.Context(0) as c:
        c.upload_graph(n, m, ind, adj)
        assert c.bin_sizes()["hub"][0] == 3
        lab_o, so, ivo, evo = check_partition(c, g, GG.singleton_labels(n))
        c.hub_overflows()
        c.propagate(O.MOD, 0.0, frac=1.0, seed=1, step=0, tie=O.TIE_RANDOM)
        assert c.hub_overflows() == 0
        want = O.propagation_refine(O.MOD, 0.0, n, m, ind, adj, lab_o, so, ivo, evo, 0.0, 1.0, 1, 1, 0, O.TIE_RANDOM)
        assert np.array_equal(c.labels(), want)


def test_propagation_wcc(ctx):
    for name in ["golden_sbm200", "golden_sbm210", "golden_rmat9"]:
        g = GOLDEN[name]
        n, m, ind, adj = g[:4]
        upload(ctx, g)
        tri, td, cc = ctx.triangles()
        et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
        assert cc == pytest.approx(cc_o, rel=1e-12)
        rng = np.random.default_rng(4)
        lab = rng.integers(0, max(2, n // 10), n)
        for tie in (O.TIE_RANDOM, O.TIE_LOWEST):
            _lp_compare(ctx, g, lab, O.WCC, 0.0, 1.0, seed=3, step=1, tie=tie, cc=cc)


def _path(n):
    a = np.arange(n - 1)
    return GG.csr_from_edges(a, a + 1)


def _star_with_triangles(leaves):
    """A hub above 46,341 neighbours (the reference's int32 deg * (deg - 1)
    wraps) whose leaves pair up into triangles, plus a clique."""
    hub = np.zeros(leaves, np.int64)
    lv = np.arange(1, leaves + 1)
    a = [hub, lv[0::2][: leaves // 2]]
    b = [lv, lv[1::2][: leaves // 2]]
    i, j = np.triu_indices(12, 1)
    a.append(i + leaves + 1)
    b.append(j + leaves + 1)
    return GG.csr_from_edges(np.concatenate(a), np.concatenate(b))


@pytest.mark.parametrize("rounds", [128, 1, 0])
def test_clustering_partition_on_device(ctx, rounds):
    """pcd_clustering_labels equals the reference's serial greedy
    (framework.cpp:27-59, as restated on the host and pinned by the -i 1 run
    fixtures), also when the rounds stop early and the host finishes the
    chains (rounds 1 and 0), on a path numbered in order and on a hub whose
    cc divides by a wrapped int32."""
    ctx.set_tuning("cluster_rounds", rounds)
    try:
        graphs = [GOLDEN[k] for k in sorted(GOLDEN)] + [GG.bin_zoo(), GG.unsorted_copy(GG.bin_zoo()),
                                                       GG.dense_blocks(), GG.rmat(12), _path(3000),
                                                       _star_with_triangles(50000)]
        for g in graphs:
            n, m, ind, adj = upload(ctx, g)
            nc = ctx.clustering()
            got = ctx.labels()
            tri = ctx.triangles()[0]
            want = parcd.clustering_partition(n, ind, adj, tri)
            assert nc == int(want.max()) + 1
            assert np.array_equal(got, want), (n, np.nonzero(got != want)[0][:10])
        # the oracle's restatement on one graph as a second check
        g = GOLDEN["golden_sbm200"]
        n, m, ind, adj = upload(ctx, g)
        ctx.clustering()
        _, tri_o, _, _ = O.count_triangles(n, ind, adj)
        assert np.array_equal(ctx.labels(), O.clustering_partition(n, ind, adj, tri_o))
    finally:
        ctx.set_tuning("cluster_rounds", 128)


def test_triangles(ctx):
    cases = list(GOLDEN.items()) + [("zoo", GG.bin_zoo()), ("rmat14", GG.rmat(14)), ("dense", GG.dense_blocks())]
    for name, g in cases:
        n, m, ind, adj = upload(ctx, g)
        tri, td, cc = ctx.triangles()
        et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
        assert np.array_equal(tri, tri_o), name
        assert np.array_equal(td, td_o), name
        assert cc == pytest.approx(cc_o, rel=1e-12, abs=1e-15), name


def test_triangles_golden_fixture(ctx):
    for p in golden("graph_*.npz"):
        d = load(p)
        g = O.parse_pcd(d["pcd"].tobytes())
        upload(ctx, g)
        tri, td, cc = ctx.triangles()
        assert np.array_equal(tri, d["tri"]) and np.array_equal(td, d["tri_deg"])
        assert cc == pytest.approx(float(d["cc"][0]), rel=1e-12, abs=1e-15)


def test_wcc_metric(ctx):
    cases = [GOLDEN["golden_sbm200"], GOLDEN["golden_rmat9"], GG.bin_zoo(), GG.dense_blocks()]
    rng = np.random.default_rng(8)
    for g in cases:
        n, m, ind, adj = g[:4]
        et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
        for lab in [rng.integers(0, max(2, n // 20), n), np.zeros(n, np.int32)]:
            lab_o, so, ivo, evo = check_partition(ctx, g, lab)
            got = ctx.metric(O.WCC)
            want = O.calculate(O.WCC, 0, n, m, ind, adj, lab_o, so, ivo, evo, tri_o, td_o)
            assert got == pytest.approx(want, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("onesweep", [0, 1])
def test_contraction_pairs(ctx, onesweep):
    """Both radix-sort forms: three launches per pass, and one launch per
    pass with decoupled look-back (R-MAT-16: about 500 tiles per pass)."""
    cases = [GOLDEN["golden_sbm200"], GOLDEN["golden_rmat9"], GG.bin_zoo(), GG.rmat(13), GG.rmat(16)]
    rng = np.random.default_rng(9)
    ctx.set_tuning("onesweep", onesweep)
    try:
        for g in cases:
            n, m, ind, adj = g[:4]
            for lab in [GG.singleton_labels(n), rng.integers(0, max(2, n // 30), n)]:
                lab_o, so, ivo, evo = check_partition(ctx, g, lab)
                l, k, w = ctx.contract()
                lo, ko, wo = O.community_pairs(n, ind, adj, lab_o, len(so))
                assert np.array_equal(l, lo) and np.array_equal(k, ko) and np.array_equal(w, wo)
    finally:
        ctx.set_tuning("onesweep", 0)


def test_contraction_pairs_one_sweep_fallback(ctx):
    """One-sweep sort whose second pass gives up (spin limit 0, set from the
    host: every wait still ends): the pass records itself, and three-launch
    passes redo the sort from that pass's input (the second buffer), so the
    pairs stay exact."""
    cases = [GG.rmat(13), GG.rmat(16)]
    rng = np.random.default_rng(19)
    ctx.set_tuning("onesweep", 1)
    ctx.set_tuning("os_spin_limit", 0)
    try:
        ctx.sort_fallbacks()
        for g in cases:
            n, m, ind, adj = g[:4]
            for lab in [GG.singleton_labels(n), rng.integers(0, max(2, n // 30), n)]:
                lab_o, so, ivo, evo = check_partition(ctx, g, lab)
                l, k, w = ctx.contract()
                lo, ko, wo = O.community_pairs(n, ind, adj, lab_o, len(so))
                assert np.array_equal(l, lo) and np.array_equal(k, ko) and np.array_equal(w, wo)
        assert ctx.sort_fallbacks() > 0
    finally:
        ctx.set_tuning("os_spin_limit", 1 << 20)
        ctx.set_tuning("onesweep", 0)


@pytest.mark.parametrize("merge_sort", [0, 1])
@pytest.mark.parametrize("tie", [O.TIE_RANDOM, O.TIE_LOWEST])
@pytest.mark.parametrize("kind,lam", RULES + [(O.WCC, 0.0)])
def test_merging(ctx, kind, lam, tie, merge_sort):
    """Both partner choices: counting over community rows (merge_sort 0)
    and the sorted unique community pairs (1)."""
    cases = [GOLDEN["golden_sbm200"], GOLDEN["golden_rmat9"], GOLDEN["golden_starclique"], GG.bin_zoo()]
    rng = np.random.default_rng(10)
    ctx.set_tuning("merge_sort", merge_sort)
    try:
        for g in cases:
            n, m, ind, adj = g[:4]
            cc = 0.0
            if kind == O.WCC:
                upload(ctx, g)
                cc = ctx.triangles()[2]
            for lab in [GG.singleton_labels(n), rng.integers(0, max(2, n // 5), n)]:
                for rounds in (1, 3):
                    lab_o, so, ivo, evo = check_partition(ctx, g, lab)
                    if kind == O.WCC:
                        ctx.set_cc(cc)
                    ctx.merge(kind, lam, rounds=rounds, seed=77, step=2, tie=tie)
                    got = ctx.labels()
                    want, bp, fp = O.merging_refine(kind, lam, n, m, ind, adj, lab_o, so, ivo, evo, cc, rounds, 1, 77,
                                                    2, tie)
                    assert np.array_equal(got, want)
    finally:
        ctx.set_tuning("merge_sort", 0)


@pytest.mark.parametrize("inc", [1, 0])
def test_statistics_follow_merging(ctx, inc):
    """After a merging step, the statistics of the merged labels come from
    the pairs' records and links (inc_stats 1; the compression renumbers
    them) or from a full recount (0).  Either way, compress + statistics
    equal the oracle's count of the new labels, over several merging steps
    (hub rows included)."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    rng = np.random.default_rng(15)
    ctx.set_tuning("inc_stats", inc)
    try:
        for lab in [GG.singleton_labels(n), rng.integers(0, 40, n), rng.integers(0, 3000, n)]:
            lab_o, so, ivo, evo = check_partition(ctx, g, lab)
            for it in range(3):
                ctx.merge(O.MOD, 0.0, rounds=1 + it % 2, seed=16, step=it, tie=O.TIE_RANDOM)
                want = O.merging_refine(O.MOD, 0.0, n, m, ind, adj, lab_o, so, ivo, evo, 0.0, 1 + it % 2, 1, 16, it,
                                        O.TIE_RANDOM)[0]
                k = ctx.compress()
                lab_o, k_o = O.compress_labels(want)
                assert k == k_o and np.array_equal(ctx.labels(), lab_o)
                assert ctx.statistics() == k_o
                s_, iv, ev = ctx.stats(k)
                _, so, ivo, evo = O.collect_statistics(n, ind, adj, lab_o)
                assert np.array_equal(s_, so) and np.array_equal(iv, ivo) and np.array_equal(ev, evo), it
                assert ctx.metric(O.MOD) == pytest.approx(O.calculate(O.MOD, 0.0, n, m, ind, adj, lab_o, so, ivo, evo),
                                                          rel=1e-9, abs=1e-12)
    finally:
        ctx.set_tuning("inc_stats", 1)


@pytest.mark.parametrize("slots,part_entries", [(4096, 2048), (64, 100000), (1024, 512)])
def test_merging_community_rows_hub_paths(ctx, slots, part_entries):
    """Community rows above 4096 entries go through the hub path (chunks,
    label partitions, LDS counting) and, when a partition overflows its
    table (forced by tiny tables or huge partitions), the global-table
    fallback.  Partitions of 40 communities make every row a hub row; of
    3,000 communities, rows in every bin; uncompressed labels leave
    communities empty."""
    g = GG.bin_zoo()
    n, m, ind, adj = g[:4]
    rng = np.random.default_rng(12)
    try:
        ctx.set_tuning("hub_part_entries", part_entries)
        ctx.set_tuning("hub_table_slots", slots)
        for lab in [GG.singleton_labels(n), rng.integers(0, 40, n), rng.integers(0, 3000, n)]:
            for kind, lam in [(O.MOD, 0.0), (O.CPM, 0.02)]:
                for rounds in (1, 2):
                    lab_o, so, ivo, evo = check_partition(ctx, g, lab)
                    ctx.merge(kind, lam, rounds=rounds, seed=13, step=1, tie=O.TIE_RANDOM)
                    want = O.merging_refine(kind, lam, n, m, ind, adj, lab_o, so, ivo, evo, 0.0, rounds, 1, 13, 1,
                                            O.TIE_RANDOM)[0]
                    assert np.array_equal(ctx.labels(), want)
        # statistics of labels that skip ids (k = max label + 1, empty communities)
        upload(ctx, g)
        lab = 2 * rng.integers(0, 500, n)
        ctx.upload_labels(lab)
        k = ctx.statistics()
        assert k == int(lab.max()) + 1
        _, so, ivo, evo = O.collect_statistics(n, ind, adj, lab)
        ctx.merge(O.MOD, 0.0, rounds=1, seed=14, step=0, tie=O.TIE_LOWEST)
        want = O.merging_refine(O.MOD, 0.0, n, m, ind, adj, lab, so, ivo, evo, 0.0, 1, 1, 14, 0, O.TIE_LOWEST)[0]
        assert np.array_equal(ctx.labels(), want)
    finally:
        ctx.set_tuning("hub_part_entries", 2048)
        ctx.set_tuning("hub_table_slots", 4096)


def gpu_run(ctx, g, kind, lam, strategy, max_iter, seed, tie, frac=0.5, rounds=1, init=None):
    """Outer loop of main.cpp:226-266 driven through the C ABI."""
    n, m, ind, adj = upload(ctx, g)
    ctx.upload_labels(GG.singleton_labels(n) if init is None else init)
    prev, it, trace = 0.0, 0, []
    while True:
        ctx.compress()
        k = ctx.statistics()
        score = ctx.metric(kind, lam)
        trace.append(score)
        done = it >= max_iter if max_iter != -1 else (prev > 0 and score > 0 and score / prev - 1 <= 0)
        if done:
            break
        if strategy == 0 or (strategy == 2 and it % 2 == 0):
            ctx.propagate(kind, lam, frac=frac, seed=seed, step=it, tie=tie)
        else:
            ctx.merge(kind, lam, rounds=rounds, seed=seed, step=it, tie=tie)
        it += 1
        prev = score
    return ctx.labels(), trace[-1], it + 1, k, trace


@pytest.mark.parametrize("strategy", [0, 1, 2])
@pytest.mark.parametrize("kind,lam", [(O.MOD, 0.0), (O.CPM, 0.01)])
def test_full_run_replay(ctx, strategy, kind, lam):
    """Deterministic mode: the GPU loop equals the serial replay label for label."""
    for g in [GOLDEN["golden_sbm200"], GOLDEN["golden_rmat9"], GG.sbm(20, 100, 0.1, 0.002)]:
        n, m, ind, adj = g[:4]
        init = GG.singleton_labels(n)
        for tie in (O.TIE_LOWEST, O.TIE_RANDOM):
            lab, score, iters, k, trace = gpu_run(ctx, g, kind, lam, strategy, -1, 11, tie, init=init)
            lab_o, score_o, iters_o, k_o, trace_o = O.run(n, m, ind, adj, kind=kind, lam=lam, strategy=strategy,
                                                          mode=1, seed=11, tie=tie, init_labels=init)
            assert iters == iters_o and k == k_o
            assert np.array_equal(lab, lab_o)
            assert np.allclose(trace, trace_o, rtol=1e-9, atol=1e-12)


def test_rmat_large_properties(ctx):
    """R-MAT scale 18: statistics exact vs oracle; LP step exact vs replay."""
    g = GG.rmat(18)
    n, m, ind, adj = g[:4]
    init = GG.singleton_labels(n)
    lab = _lp_compare(ctx, g, init, O.MOD, 0.0, 0.5, seed=2, step=0, tie=O.TIE_RANDOM)
    assert lab.min() >= -1
    rng = np.random.default_rng(1)
    _lp_compare(ctx, g, rng.integers(0, 5000, n), O.CPM, 0.01, 0.5, seed=2, step=5, tie=O.TIE_LOWEST)


@pytest.mark.parametrize("kind,lam", [(O.MOD, 0.0), (O.CPM, 0.01)])
def test_rmat22_benchmark_iterations(ctx, kind, lam):
    """The benchmark's own workloads at full size (R-MAT scale 22 seed 2,
    2.4 M vertices, 128 M entries; modularity and CPM 0.01): their first three
    refinement iterations from the singleton partition (frac 0.5, seed 1,
    random ties).  Each compression, each statistics table (recounted in the
    first iteration, then updated from the moved rows and renumbered by the
    compression), each metric value and each propagation step equals the
    oracle's on the same inputs."""
    g = GG.rmat(22)
    n, m, ind, adj = g[:4]
    ctx.upload_graph(n, m, ind, adj)
    prev = GG.singleton_labels(n)
    ctx.upload_labels(prev)
    for it in range(3):
        k = ctx.compress()
        lab_o, k_o = O.compress_labels(prev)
        assert k == k_o
        assert np.array_equal(ctx.labels(), lab_o)
        assert ctx.statistics() == k_o
        s, iv, ev = ctx.stats(k)
        _, so, ivo, evo = O.collect_statistics(n, ind, adj, lab_o)
        assert np.array_equal(s, so) and np.array_equal(iv, ivo) and np.array_equal(ev, evo), it
        assert int(so.sum()) == n and int(ivo.sum() + evo.sum()) == 2 * m
        q = ctx.metric(kind, lam)
        assert q == pytest.approx(O.calculate(kind, lam, n, m, ind, adj, lab_o, so, ivo, evo), rel=1e-9, abs=1e-12)
        ctx.propagate(kind, lam, frac=0.5, seed=1, step=it, tie=O.TIE_RANDOM)
        got = ctx.labels()
        want = O.propagation_refine(kind, lam, n, m, ind, adj, lab_o, so, ivo, evo, 0.0, 0.5, 1, 1, it, O.TIE_RANDOM)
        assert np.array_equal(got, want), (it, int(np.count_nonzero(got != want)))
        prev = got


def test_rmat22_merging_first_iterations(ctx):
    """The merging leg of the benchmark at full size (R-MAT scale 22 seed 2,
    -s 1): in each of the first two iterations from the singleton partition,
    the contraction's community pairs (l, k, links) -- 128 M of them at first,
    sorted by the radix sort -- equal the oracle's table
    (refinement.cpp:228-238), and the merging step's labels equal the
    oracle's replay (refinement.cpp:185-288)."""
    g = GG.rmat(22)
    n, m, ind, adj = g[:4]
    ctx.upload_graph(n, m, ind, adj)
    prev = GG.singleton_labels(n)
    for it in range(2):
        lab_o, so, ivo, evo = check_partition(ctx, g, prev)
        l, k, w = ctx.contract()
        lo, ko, wo = O.community_pairs(n, ind, adj, lab_o, len(so))
        assert len(l) == len(lo), it
        assert np.array_equal(l, lo) and np.array_equal(k, ko) and np.array_equal(w, wo), it
        del l, k, w, lo, ko, wo
        ctx.merge(O.MOD, 0.0, rounds=1, seed=5, step=it, tie=O.TIE_RANDOM)
        got = ctx.labels()
        want, _, _ = O.merging_refine(O.MOD, 0.0, n, m, ind, adj, lab_o, so, ivo, evo, 0.0, 1, 1, 5, it,
                                      O.TIE_RANDOM)
        assert np.array_equal(got, want), (it, int(np.count_nonzero(got != want)))
        prev = got


def test_chung_lu_wcc_benchmark_iterations(ctx):
    """The benchmark's WCC workload at full size (Chung-Lu seed 3: 10^6
    vertices, degrees up to 8,415, so rows in the hub path), run as
    `-m wcc -i 1`: triangles, the triangle-based initial partition on the
    device, then two refinement iterations whose compression, statistics,
    WCC metric and propagation step each equal the oracle's."""
    import parcd as P
    g = P.make_graph("chung_lu", 3, n=1000000, exponent=2.3, dmin=4, dmax=5000, num_comms=32, cmin=100,
                     cmax=50000, boost=20.0)
    n, m, ind, adj = g[:4]
    assert int(np.diff(ind).max()) > 4096  # hub rows present
    ctx.upload_graph(n, m, ind, adj)
    tri, td, cc = ctx.triangles()
    et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
    assert np.array_equal(tri, tri_o) and np.array_equal(td, td_o)
    assert cc == pytest.approx(cc_o, rel=1e-12)
    ctx.clustering()
    prev = ctx.labels()
    assert np.array_equal(prev, O.clustering_partition(n, ind, adj, tri_o))
    for it in range(2):
        k = ctx.compress()
        lab_o, k_o = O.compress_labels(prev)
        assert k == k_o and np.array_equal(ctx.labels(), lab_o)
        ctx.statistics()
        s, iv, ev = ctx.stats(k)
        _, so, ivo, evo = O.collect_statistics(n, ind, adj, lab_o)
        assert np.array_equal(s, so) and np.array_equal(iv, ivo) and np.array_equal(ev, evo), it
        q = ctx.metric(O.WCC)
        q_o = O.calculate(O.WCC, 0, n, m, ind, adj, lab_o, so, ivo, evo, tri_o, td_o)
        assert q == pytest.approx(q_o, rel=1e-9, abs=1e-12)
        ctx.propagate(O.WCC, 0.0, frac=0.5, seed=1, step=it, tie=O.TIE_RANDOM)
        got = ctx.labels()
        want = O.propagation_refine(O.WCC, 0.0, n, m, ind, adj, lab_o, so, ivo, evo, cc, 0.5, 1, 1, it, O.TIE_RANDOM)
        assert np.array_equal(got, want), (it, int(np.count_nonzero(got != want)))
        prev = got


@pytest.mark.parametrize("kind,lam,max_iter", [(O.MOD, 0.0, -1), (O.CPM, 0.02, -1), (O.WCC, 0.0, -1),
                                               (O.MOD, 0.0, 3)])
def test_iterate_equals_stepwise_calls(kind, lam, max_iter):
    """pcd_iterate (the loop body with the propagation queued before the
    score is read back and dropped when the stop rule ends the loop) gives
    the scores, k and stop decision of the separate compress / statistics /
    metric / propagation calls, pass by pass; after the loop the labels and
    statistics are those of the stepwise loop, and one more step from there
    agrees too."""
    for g in [GG.bin_zoo(), GOLDEN["golden_sbm200"]]:
        n, m, ind, adj = g[:4]
        cap = 25  # the iteration cap ends runs that keep improving (WCC here)
        jb = parcd.job(kind=kind, lam=lam, frac=0.5, seed=5, tie=O.TIE_RANDOM, max_iter=max_iter, iter_cap=cap)
        with parcd.Context(0) as a, parcd.Context(0) as b:
            for c in (a, b):
                c.upload_graph(n, m, ind, adj)
                c.upload_labels(GG.singleton_labels(n))
            if kind == O.WCC:
                b.triangles()
            prev_a = prev_b = 0.0
            it = 0
            while True:
                sa, ka, da = a.iterate(jb, it, prev_a)
                kb = b.compress()
                b.statistics()
                sb = b.metric(kind, lam)
                db = (max_iter != -1 and it >= max_iter) or it >= cap or (
                    max_iter == -1 and prev_b > 0 and sb > 0 and sb / prev_b - 1 <= 0.0)
                assert (sa, ka, da) == (sb, kb, db), it
                if not db:
                    b.propagate(kind, lam, frac=0.5, seed=5, step=it, tie=O.TIE_RANDOM)
                it += 1
                prev_a, prev_b = sa, sb
                if da:
                    break
            assert np.array_equal(a.labels(), b.labels())
            k = a.statistics()
            assert k == b.statistics()
            for x, y in zip(a.stats(k), b.stats(k)):
                assert np.array_equal(x, y)
            for c in (a, b):
                c.propagate(kind, lam, frac=0.5, seed=5, step=it, tie=O.TIE_RANDOM)
            assert np.array_equal(a.labels(), b.labels())


@pytest.mark.parametrize("name", sorted(GG.TINY))
def test_degenerate_graphs(ctx, name):
    """Graphs without edges, a single vertex, isolated vertices, a star and
    a clique: compression, statistics, every metric (NaN and -inf where m or
    a community volume is 0, as the reference divides), one propagation step
    of every rule, the bounded -s 0 / -s 1 loops (stepwise and pcd_refine),
    triangles and the -i 1 partition equal the oracle."""
    g = GG.csr_direct(*GG.TINY[name])
    n, m, ind, adj = g[:4]
    et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
    rng = np.random.default_rng(1)
    for lab in [GG.singleton_labels(n), np.zeros(n, np.int32), rng.integers(0, 3, n).astype(np.int32)]:
        lab_o, so, ivo, evo = check_partition(ctx, g, lab)
        got = ctx.metric(O.WCC)
        want = O.calculate(O.WCC, 0, n, m, ind, adj, lab_o, so, ivo, evo, tri_o, td_o)
        assert (np.isnan(got) and np.isnan(want)) or got == pytest.approx(want, rel=1e-9, abs=1e-12)
        for kind, lam, c in [(O.MOD, 0.0, None), (O.CPM, 0.05, None), (O.WCC, 0.0, cc_o), (O.COND, 0.0, None)]:
            _lp_compare(ctx, g, lab, kind, lam, 1.0, seed=1, step=0, tie=O.TIE_RANDOM, cc=c)
            _lp_compare(ctx, g, lab, kind, lam, 0.5, seed=2, step=1, tie=O.TIE_LOWEST, cc=c)
    upload(ctx, g)
    tri, td, cc = ctx.triangles()
    assert np.array_equal(tri, tri_o[:n]) and np.array_equal(td, td_o[:n]) and cc == pytest.approx(cc_o, abs=1e-12)
    ctx.clustering()
    assert np.array_equal(ctx.labels(), O.clustering_partition(n, ind, adj, tri_o))
    for strategy in (0, 1):
        init = GG.singleton_labels(n)
        lab, score, iters, k, trace = gpu_run(ctx, g, O.MOD, 0.0, strategy, 4, 3, O.TIE_LOWEST, init=init)
        lab_o, score_o, iters_o, k_o, trace_o = O.run(n, m, ind, adj, kind=O.MOD, strategy=strategy, mode=1, seed=3,
                                                      tie=O.TIE_LOWEST, init_labels=init, max_iter=4)
        assert np.array_equal(lab, lab_o) and iters == iters_o and k == k_o
        upload(ctx, g)
        ctx.upload_labels(init)
        sc, it, kk = ctx.refine(parcd.job(kind=O.MOD, seed=3, tie=O.TIE_LOWEST, max_iter=4), strategy=strategy)
        assert np.array_equal(ctx.labels(), lab_o) and it == iters_o and kk == k_o


def test_wcc_int32_wrap_on_device(ctx):
    """WCC gains of a community above 46,341 vertices: the reference's int32
    products (n_x - 1) * (n_x - 2) / 2, (c_x - 1) * c_y, c_y * (c_y - 1) / 2
    and c_x * c_y wrap there (rules.cpp:243, 265, 267-268, 297, 326, 328-329,
    358).  One propagation step and one merging step from the planted
    blocks (a 50,000-vertex community) equal the oracle, whose wrapping is
    pinned against the reference's own gains."""
    g = parcd.make_graph("sbm", 11, block_sizes=[50000, 8000], p_in=0.0005, p_out=0.00005)
    n, m, ind, adj, mp = g
    upload(ctx, g)
    tri, td, cc = ctx.triangles()
    et, tri_o, td_o, cc_o = O.count_triangles(n, ind, adj)
    assert np.array_equal(tri, tri_o) and np.array_equal(td, td_o)
    assert cc == pytest.approx(cc_o, rel=1e-12)
    rng = np.random.default_rng(12)
    lab = (mp >= 50000).astype(np.int32)
    moved = rng.random(n) < 0.05        # a few vertices elsewhere, so moves go both ways
    lab[moved] = rng.integers(2, 400, int(moved.sum()))
    lab_o, so, ivo, evo = check_partition(ctx, g, lab)
    assert so.max() >= 46342            # the wrapping sizes are reached
    got = ctx.metric(O.WCC)
    want = O.calculate(O.WCC, 0, n, m, ind, adj, lab_o, so, ivo, evo, tri_o, td_o)
    assert got == pytest.approx(want, rel=1e-9, abs=1e-12)
    for tie in (O.TIE_RANDOM, O.TIE_LOWEST):
        _lp_compare(ctx, g, lab, O.WCC, 0.0, 1.0, seed=5, step=3, tie=tie, cc=cc)
        lab_o, so, ivo, evo = check_partition(ctx, g, lab)
        ctx.set_cc(cc)
        ctx.merge(O.WCC, 0.0, rounds=2, seed=6, step=4, tie=tie)
        got = ctx.labels()
        want, bp, fp = O.merging_refine(O.WCC, 0.0, n, m, ind, adj, lab_o, so, ivo, evo, cc, 2, 1, 6, 4, tie)
        assert np.array_equal(got, want)
