"""Pins the oracle (oracle/oracle.c) against the reference's golden fixtures,
brute force and closed forms.  CPU only."""
import itertools
import os
import tempfile

import numpy as np
import pytest

import oracle_lib as O
from conftest import golden, load


# ------------------------------------------------------------------ RNG spec
def test_mix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014): first outputs from state 0 are
    # mix64(k * 0x9E3779B97F4A7C15) for k = 1, 2, 3.
    L = O.lib()
    golden_ = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for k, want in enumerate(golden_, 1):
        assert L.orc_mix64((k * 0x9E3779B97F4A7C15) & (2**64 - 1)) == want


def test_hash3_and_selection_spec():
    L = O.lib()
    # hash3(seed, a, b) = mix64(seed + golden * ((a << 32 | b) + 1))
    for seed, a, b in [(0, 0, 0), (7, 3, 11), (2**63 + 5, 2**32 - 1, 2**31)]:
        x = ((a << 32) | b) + 1
        want = L.orc_mix64((seed + 0x9E3779B97F4A7C15 * x) & (2**64 - 1))
        assert L.orc_hash3(seed, a, b) == want
    # selection rate tracks frac
    f = L.orc_frac_threshold(0.5)
    assert f == int(0.5 * 2147483647)
    sel = np.array([L.orc_select(42, 3, v, f) for v in range(20000)])
    assert abs(sel.mean() - 0.5) < 0.02
    assert L.orc_frac_threshold(1.0) == 2147483647
    assert all(L.orc_select(1, 1, v, 2147483647) for v in range(100))
    # lowest tie order is the label itself
    assert L.orc_tie_key(O.TIE_LOWEST, 9, 1, 2, 12345) == 12345


# ------------------------------------------------------------------ convert
@pytest.mark.parametrize("path", golden("convert_*.npz"), ids=os.path.basename)
def test_convert_matches_reference_bytes(path):
    d = load(path)
    with tempfile.TemporaryDirectory() as tmp:
        tin = os.path.join(tmp, "in.txt")
        open(tin, "wb").write(d["text"].tobytes())
        rc, a, b, nn, _ = O.read_edges_text(tin)
        assert rc == 0
        n, m, ind, adj, mp = O.convert_edges(nn, a, b)
        assert O.pcd_bytes(n, m, ind, adj, mp) == d["pcd"].tobytes()
        out = os.path.join(tmp, "o.pcd")
        assert O.lib().orc_write_pcd(out.encode(), n, m, ind, adj, mp) == 0
        assert open(out, "rb").read() == d["pcd"].tobytes()


def _brute_convert(a, b):
    """numpy statement of convert.cpp:76-196 for loop-free input."""
    nn = int(max(a.max(), b.max())) + 1
    und = {(min(x, y), max(x, y)) for x, y in zip(a.tolist(), b.tolist()) if x != y}
    nbr = [set() for _ in range(nn)]
    for x, y in und:
        nbr[x].add(y)
        nbr[y].add(x)
    order = sorted(range(nn), key=lambda v: (len(nbr[v]), v), reverse=True)
    new = {v: i for i, v in enumerate(order)}
    n = sum(1 for v in order if nbr[v])
    ind = [0]
    adj = []
    for v in order[:n]:
        adj += sorted(new[u] for u in nbr[v])
        ind.append(len(adj))
    return n, len(adj) // 2, np.array(ind), np.array(adj), np.array(order[:n])


def test_convert_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(20):
        ne = int(rng.integers(1, 300))
        nn = int(rng.integers(2, 60))
        a = rng.integers(0, nn, ne)
        b = rng.integers(0, nn, ne)
        keep = a != b
        a, b = a[keep], b[keep]
        if a.size == 0:
            continue
        got = O.convert_edges(int(max(a.max(), b.max())) + 1, a, b)
        want = _brute_convert(a, b)
        for g, w in zip(got, want):
            assert np.array_equal(np.asarray(g), np.asarray(w))


def test_read_edges_rejects_corrupt(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1 2\n3 x\n")
    rc, *_ , bad = O.read_edges_text(str(p))
    assert rc == -1 and bad == 2
    p.write_text("1 2\n" + "9" * 2000 + " 3\n")
    rc, *_ = O.read_edges_text(str(p))
    assert rc == -2


# ------------------------------------------------------------------ triangles
def test_triangles_match_reference(golden_graphs):
    for name, (n, m, ind, adj, mp, d) in golden_graphs.items():
        et, tri, td, cc = O.count_triangles(n, ind, adj)
        assert np.array_equal(tri, d["tri"]), name
        assert np.array_equal(td, d["tri_deg"]), name
        assert cc == pytest.approx(float(d["cc"][0]), rel=1e-12, abs=0), name


def test_triangles_brute_force_and_networkx(golden_graphs):
    nx = pytest.importorskip("networkx")
    for name, (n, m, ind, adj, mp, d) in golden_graphs.items():
        A = np.zeros((n, n), bool)
        for v in range(n):
            A[v, adj[ind[v]:ind[v + 1]]] = True
        et, tri, td, cc = O.count_triangles(n, ind, adj)
        # per-edge triangle count = common neighbours (stored on the higher endpoint)
        Ai = A.astype(np.int64)
        common = Ai @ Ai
        for v in range(n):
            for j in range(ind[v], ind[v + 1]):
                u = adj[j]
                assert et[j] == (common[v, u] if u < v else 0)
        assert np.array_equal(tri, 2 * (np.diag(Ai @ Ai @ Ai) // 2))
        G = nx.Graph()
        G.add_nodes_from(range(n))
        G.add_edges_from((v, u) for v in range(n) for u in adj[ind[v]:ind[v + 1]])
        t = nx.triangles(G)
        assert all(tri[v] == 2 * t[v] for v in range(n))
        assert cc == pytest.approx(nx.average_clustering(G), rel=1e-12, abs=1e-15)


# ------------------------------------------------------------------ partitions
PARTS = golden("part_*.npz")


@pytest.mark.parametrize("path", PARTS, ids=os.path.basename)
def test_part_matches_reference(path, golden_graphs):
    d = load(path)
    name = os.path.basename(path)[5:-4].rsplit("_", 1)[0]
    n, m, ind, adj, mp, g = golden_graphs[name]
    lab, k = O.compress_labels(d["in_labels"])
    assert k == int(d["k"][0])
    assert np.array_equal(lab, d["labels"])
    k2, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
    assert k2 == int(d["num_labels"][0])
    assert np.array_equal(s, d["sizes"]) and np.array_equal(iv, d["int_volume"]) and np.array_equal(ev, d["ext_volume"])
    lam = float(d["lam"][0])
    tri, td, cc = g["tri"], g["tri_deg"], float(g["cc"][0])
    for kind in range(4):
        got = O.calculate(kind, lam, n, m, ind, adj, lab, s, iv, ev, tri, td)
        want = float(d["calculate"][kind])
        if np.isnan(want):
            assert np.isnan(got)
        else:
            assert got == pytest.approx(want, rel=1e-12, abs=1e-15), kind
    L = O.lib()
    deg = np.diff(ind)
    for kind in range(4):
        js = d["join_s%d" % kind]
        for i, (v, c, num) in enumerate(zip(d["join_v"], d["join_c"], d["join_n"])):
            got = L.orc_score_join(kind, lam, m, cc, int(deg[v]), int(s[c]), int(iv[c]), int(ev[c]), int(num))
            assert got == js[i] or (np.isnan(got) and np.isnan(js[i])), (kind, v, c)
        ls = d["leave_s%d" % kind]
        for i, (v, num) in enumerate(zip(d["leave_v"], d["leave_n"])):
            c = lab[v]
            got = L.orc_score_leave(kind, lam, m, cc, int(deg[v]), int(s[c]), int(iv[c]), int(ev[c]), int(num))
            assert got == ls[i] or (np.isnan(got) and np.isnan(ls[i])), (kind, v)
        ms = d["merge_s%d" % kind]
        for i, (a, b, links) in enumerate(zip(d["merge_i"], d["merge_j"], d["merge_links"])):
            got = L.orc_score_merge(kind, lam, m, cc, int(s[a]), int(iv[a]), int(ev[a]), int(s[b]), int(iv[b]),
                                    int(ev[b]), int(links))
            assert got == ms[i] or (np.isnan(got) and np.isnan(ms[i])), (kind, a, b)
    # community-pair table = contraction content
    pl, pk, pc = O.community_pairs(n, ind, adj, lab, k)
    assert np.array_equal(pl, d["merge_i"]) and np.array_equal(pk, d["merge_j"]) and np.array_equal(pc, d["merge_links"])


def test_stats_and_metrics_closed_forms(golden_graphs):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(3)
    for name, (n, m, ind, adj, mp, g) in golden_graphs.items():
        lab = rng.integers(0, max(2, n // 6), n)
        lab, k = O.compress_labels(lab)
        _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
        src = np.repeat(np.arange(n), np.diff(ind))
        same = lab[src] == lab[adj]
        assert np.array_equal(iv, np.bincount(lab[src][same], minlength=k))
        assert np.array_equal(ev, np.bincount(lab[src][~same], minlength=k))
        assert np.array_equal(s, np.bincount(lab, minlength=k))
        q = O.calculate(O.MOD, 0, n, m, ind, adj, lab, s, iv, ev)
        G = nx.Graph()
        G.add_nodes_from(range(n))
        G.add_edges_from(zip(src.tolist(), adj.tolist()))
        comms = [set(np.nonzero(lab == c)[0].tolist()) for c in range(k)]
        assert q == pytest.approx(nx.community.modularity(G, comms), rel=1e-10, abs=1e-12)
        vol = (iv + ev).astype(np.int64)
        exact = iv.sum() / (2.0 * m) - (vol * vol).sum() / (4.0 * m * m)
        assert q == pytest.approx(exact, rel=1e-12, abs=1e-14)
        lam = 0.037
        cpm = O.calculate(O.CPM, lam, n, m, ind, adj, lab, s, iv, ev)
        s64 = s.astype(np.int64)
        assert cpm == pytest.approx((iv.sum() - lam * (s64 * (s64 - 1)).sum()) / (2 * m), rel=1e-12, abs=1e-14)


def test_join_plus_leave_is_delta_modularity(golden_graphs):
    """score_join(v,b,k_b) + score_leave(v,k_a) == Q(after move) - Q(before)."""
    rng = np.random.default_rng(11)
    L = O.lib()
    for name, (n, m, ind, adj, mp, g) in golden_graphs.items():
        if n < 10:
            continue
        lab = rng.integers(0, max(2, n // 10), n)
        lab, k = O.compress_labels(lab)
        _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
        q0 = O.calculate(O.MOD, 0, n, m, ind, adj, lab, s, iv, ev)
        c0 = O.calculate(O.CPM, 0.05, n, m, ind, adj, lab, s, iv, ev)
        for v in rng.integers(0, n, 12):
            nb = adj[ind[v]:ind[v + 1]]
            counts = {}
            for u in nb:
                counts[lab[u]] = counts.get(lab[u], 0) + 1
            own = lab[v]
            cands = [c for c in counts if c != own]
            if not cands:
                continue
            b = cands[0]
            lab2 = lab.copy()
            lab2[v] = b
            _, s2, iv2, ev2 = O.collect_statistics(n, ind, adj, lab2)
            if len(s2) < k:   # the move emptied the highest label
                s2, iv2, ev2 = (np.pad(x, (0, k - len(x))) for x in (s2, iv2, ev2))
            q1 = O.calculate(O.MOD, 0, n, m, ind, adj, lab2, s2, iv2, ev2)
            c1 = O.calculate(O.CPM, 0.05, n, m, ind, adj, lab2, s2, iv2, ev2)
            d = int(ind[v + 1] - ind[v])
            j = L.orc_score_join(O.MOD, 0, m, 0, d, int(s[b]), int(iv[b]), int(ev[b]), counts[b])
            l_ = L.orc_score_leave(O.MOD, 0, m, 0, d, int(s[own]), int(iv[own]), int(ev[own]), counts.get(own, 0))
            assert j + l_ == pytest.approx(q1 - q0, abs=1e-12)
            jc = L.orc_score_join(O.CPM, 0.05, m, 0, d, int(s[b]), int(iv[b]), int(ev[b]), counts[b])
            lc = L.orc_score_leave(O.CPM, 0.05, m, 0, d, int(s[own]), int(iv[own]), int(ev[own]), counts.get(own, 0))
            assert (jc + lc) / m == pytest.approx(c1 - c0, abs=1e-12)


def test_merge_score_is_delta(golden_graphs):
    rng = np.random.default_rng(12)
    L = O.lib()
    for name, (n, m, ind, adj, mp, g) in golden_graphs.items():
        if n < 10:
            continue
        lab = rng.integers(0, max(3, n // 12), n)
        lab, k = O.compress_labels(lab)
        _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
        q0 = O.calculate(O.MOD, 0, n, m, ind, adj, lab, s, iv, ev)
        c0 = O.calculate(O.CPM, 0.02, n, m, ind, adj, lab, s, iv, ev)
        pl, pk, pc = O.community_pairs(n, ind, adj, lab, k)
        for t in rng.integers(0, len(pl), 6):
            a, b, links = int(pl[t]), int(pk[t]), int(pc[t])
            lab2 = np.where(lab == b, a, lab)
            lab2c, _ = O.compress_labels(lab2)
            _, s2, iv2, ev2 = O.collect_statistics(n, ind, adj, lab2c)
            q1 = O.calculate(O.MOD, 0, n, m, ind, adj, lab2c, s2, iv2, ev2)
            c1 = O.calculate(O.CPM, 0.02, n, m, ind, adj, lab2c, s2, iv2, ev2)
            sm = L.orc_score_merge(O.MOD, 0, m, 0, int(s[a]), int(iv[a]), int(ev[a]), int(s[b]), int(iv[b]),
                                   int(ev[b]), links)
            assert sm == pytest.approx(q1 - q0, abs=1e-12)
            sc = L.orc_score_merge(O.CPM, 0.02, m, 0, int(s[a]), int(iv[a]), int(ev[a]), int(s[b]), int(iv[b]),
                                   int(ev[b]), links)
            assert sc / m == pytest.approx(c1 - c0, abs=1e-12)


def test_wcc_calculate_brute_force(golden_graphs):
    """WCC (Prat et al.) by its definition: for v with t(v) > 0,
    WCC(v) = t(v,C)/t(v) * vt(v)/(|C|-1 + vt(v) - vt(v,C)), averaged over all n,
    where t counts triangles and vt counts vertices sharing a triangle with v."""
    rng = np.random.default_rng(13)
    for name, (n, m, ind, adj, mp, g) in golden_graphs.items():
        if n > 250:
            continue
        A = np.zeros((n, n), bool)
        for v in range(n):
            A[v, adj[ind[v]:ind[v + 1]]] = True
        et, tri, td, cc = O.count_triangles(n, ind, adj)
        for trial in range(3):
            lab = rng.integers(0, max(2, n // 15), n)
            lab, k = O.compress_labels(lab)
            _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
            score, it, itd = O.wcc_calculate(n, ind, adj, lab, s, tri, td)
            want = 0.0
            for v in range(n):
                nb = np.nonzero(A[v])[0]
                t_all = t_in = 0
                vt_all, vt_in = set(), set()
                for a_, b_ in itertools.combinations(nb, 2):
                    if A[a_, b_]:
                        t_all += 1
                        vt_all |= {a_, b_}
                        if lab[a_] == lab[v] and lab[b_] == lab[v]:
                            t_in += 1
                            vt_in |= {a_, b_}
                assert 2 * t_all == tri[v] and 2 * t_in == it[v]
                # tri_deg counts incident edges on >=1 triangle (= |vt_all|)
                assert td[v] == len(vt_all) and itd[v] == len(vt_in)
                if t_all > 0:
                    c = s[lab[v]]
                    want += (t_in / t_all) * (len(vt_all) / (c - 1 + len(vt_all) - len(vt_in))) / n
            assert score == pytest.approx(want, rel=1e-12, abs=1e-15)


def test_synthetic_scores_match_reference():
    d = load(golden("synth_scores.npz")[0])
    L = O.lib()
    m = int(d["m"][0]); cc = float(d["cc"][0]); lam = float(d["lambda"][0])
    s, iv, ev, deg, own = d["sizes"], d["int_volume"], d["ext_volume"], d["deg"], d["own"]
    for kind in range(4):
        for i, (v, c, num) in enumerate(zip(d["q_v"], d["q_c"], d["q_n"])):
            j = L.orc_score_join(kind, lam, m, cc, int(deg[v]), int(s[c]), int(iv[c]), int(ev[c]), int(num))
            o = own[v]
            l_ = L.orc_score_leave(kind, lam, m, cc, int(deg[v]), int(s[o]), int(iv[o]), int(ev[o]), int(num))
            wj, wl = d["join_s%d" % kind][i], d["leave_s%d" % kind][i]
            assert j == wj or (np.isnan(j) and np.isnan(wj)), (kind, i)
            assert l_ == wl or (np.isnan(l_) and np.isnan(wl)), (kind, i)
        for i, (a, b, links) in enumerate(zip(d["m_i"], d["m_j"], d["m_l"])):
            got = L.orc_score_merge(kind, lam, m, cc, int(s[a]), int(iv[a]), int(ev[a]), int(s[b]), int(iv[b]),
                                    int(ev[b]), int(links))
            want = d["merge_s%d" % kind][i]
            assert got == want or (np.isnan(got) and np.isnan(want)), (kind, i)


# ------------------------------------------------------------------ refinement
@pytest.mark.parametrize("path", golden("refine_*.npz"), ids=os.path.basename)
def test_refine_step_matches_reference(path, golden_graphs):
    d = load(path)
    name = os.path.basename(path)[7:].split("_k")[0]
    n, m, ind, adj, mp, g = golden_graphs[name]
    kind, lam = int(d["kind"]), float(d["lam"])
    lab, k = O.compress_labels(d["in_labels"])
    assert np.array_equal(lab, d["labels"])
    _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
    cc = float(g["cc"][0])
    O.lib().orc_srand(int(d["srand"]))
    if int(d["strategy"]) == 0:
        out = O.propagation_refine(kind, lam, n, m, ind, adj, lab, s, iv, ev, cc, float(d["frac"]), 0, 0, 0, 0)
    else:
        out, _, _ = O.merging_refine(kind, lam, n, m, ind, adj, lab, s, iv, ev, cc, int(d["rounds"]), 0, 0, 0, 0)
    assert np.array_equal(out, d["new_labels"])


def _parse_args(s):
    a = s.split()
    opts = dict(kind=O.MOD, lam=0.0, frac=0.5, rounds=1, max_iter=-1, initial=0, strategy=0)
    i = 0
    while i < len(a):
        f, v = a[i], a[i + 1]
        if f == "-m":
            if v in ("mod", "modularity"):
                opts["kind"] = O.MOD
            elif v == "wcc":
                opts["kind"] = O.WCC
            else:
                opts["kind"], opts["lam"] = O.CPM, float(v[4:])
        elif f == "-s":
            opts["strategy"] = int(v)
        elif f == "-n":
            opts["max_iter"] = int(v)
        elif f == "-i":
            opts["initial"] = int(v)
        elif f == "-r":
            opts["rounds"] = int(v)
        elif f == "-f":
            opts["frac"] = float(v)
        i += 2
    return opts


@pytest.mark.parametrize("path", golden("run_*.npz"), ids=os.path.basename)
def test_full_run_matches_reference_p1(path, golden_graphs):
    d = load(path)
    name = os.path.basename(path)[7:-4]
    n, m, ind, adj, mp, g = golden_graphs[name]
    opts = _parse_args(d["args"].tobytes().decode())
    labels, score, iters, k, trace = O.run(n, m, ind, adj, mode=0, **opts)
    assert iters == int(d["iters"]) and k == int(d["k"])
    assert np.array_equal(labels, d["labels"])
    assert score == pytest.approx(float(d["score"]), rel=1e-5)
    assert np.allclose(trace, d["trace"], rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------------ NMI
def _same_float(x, y):
    return (np.isnan(x) and np.isnan(y)) or x == pytest.approx(y, rel=1e-13, abs=1e-15)


@pytest.mark.parametrize("path", golden("nmi_*.npz"))
def test_nmi_matches_reference(path, tmp_path):
    """compare_communities of the compiled reference (common.cpp:75-173),
    including the int32 wrap of ni * nj (values above 1, and NaN)."""
    d = load(path)
    want = float(d["nmi"][0])
    if "truth" in d:
        t = tmp_path / "truth.txt"
        t.write_bytes(d["truth"].tobytes())
        got = O.nmi_file(int(d["n"]), d["labels"], str(t))
    else:
        got = O.nmi(d["a"], d["b"])
    assert got is not None and _same_float(got, want), (got, want)


def test_nmi_against_sklearn_and_invariants():
    sk = pytest.importorskip("sklearn.metrics")
    rng = np.random.default_rng(11)
    for n, ka, kb in [(50, 3, 4), (1000, 10, 25), (3000, 200, 7), (2000, 2000, 40)]:
        a, b = rng.integers(0, ka, n), rng.integers(0, kb, n)
        got = O.nmi(a, b)
        assert got == pytest.approx(sk.normalized_mutual_info_score(a, b), rel=1e-12, abs=1e-14)
        # symmetric, invariant under renaming communities, 1 against itself
        assert got == pytest.approx(O.nmi(b, a), rel=1e-12)
        perm = rng.permutation(ka) * 3 + 1
        assert got == pytest.approx(O.nmi(perm[a], b), rel=1e-12)
        assert O.nmi(a, a) == pytest.approx(1.0, rel=1e-12)
    # a planted partition recovered up to names: 1; one community on either side: 0
    a = np.repeat(np.arange(10), 7)
    assert O.nmi(a, (a * 13 + 5) % 97) == pytest.approx(1.0, rel=1e-12)
    assert O.nmi(np.zeros(70, np.int32), a) == 0.0
    assert O.nmi(np.full(5, -1), np.zeros(5)) is None


def test_nmi_file_matches_pairs(tmp_path):
    """A ground-truth file covering every vertex once equals the two-labeling
    form with the file's line numbers (from 1) as labels."""
    rng = np.random.default_rng(12)
    n = 300
    truth = rng.integers(0, 9, n)
    lab = rng.integers(0, 14, n)
    lines = ["# generated", ""] + [" ".join(str(v) for v in np.nonzero(truth == c)[0]) for c in range(9)]
    p = tmp_path / "t.txt"
    p.write_text("\n".join(lines) + "\n")
    assert O.nmi_file(n, lab, str(p)) == pytest.approx(O.nmi(truth + 1, lab), rel=1e-13)
    p.write_text("1 2 %d\n" % n)
    assert O.nmi_file(n, lab, str(p)) is None


def _distinct_size_partition(n, rng):
    """Communities of pairwise different sizes 1, 2, 3, ... over a random
    vertex order (the remainder joins the largest)."""
    sizes, tot = [], 0
    while tot + len(sizes) + 1 <= n:
        sizes.append(len(sizes) + 1)
        tot += sizes[-1]
    sizes[-1] += n - tot
    lab = np.repeat(np.arange(len(sizes)), sizes)[np.argsort(rng.permutation(n))]
    return lab.astype(np.int32)


@pytest.mark.parametrize("name", ["golden_gnp150", "golden_sbm200", "golden_rmat9", "golden_starclique"])
def test_counter_mode_equals_libc_mode_without_ties(golden_graphs, name):
    """At frac = 1 every vertex is selected in both modes, so the two
    replays differ only in how they break equal gains.  With CPM at
    resolution 0.0137 = 137/10^4 over communities of pairwise different
    sizes, two candidates' gains k - lambda s (joins) or links - lambda s_i s_j
    (merges) can only be equal for equal sizes, and none is 0: the inputs
    have no ties, and counter mode (hash selection, tie keys, its own
    comparison branches) must give the labels of libc mode, which the
    reference's refine fixtures pin (refinement.cpp:114-182, 185-288)."""
    n, m, ind, adj, _, _ = golden_graphs[name[7:]]
    rng = np.random.default_rng(21)
    lam = 0.0137
    for trial in range(3):
        lab, k = O.compress_labels(_distinct_size_partition(n, rng))
        _, s, iv, ev = O.collect_statistics(n, ind, adj, lab)
        assert len(set(s.tolist())) == len(s)
        for tie in (O.TIE_RANDOM, O.TIE_LOWEST):
            libc = O.propagation_refine(O.CPM, lam, n, m, ind, adj, lab, s, iv, ev, 0.0, 1.0, 0, 5 + trial, 0, tie)
            ctr = O.propagation_refine(O.CPM, lam, n, m, ind, adj, lab, s, iv, ev, 0.0, 1.0, 1, 5 + trial, 0, tie)
            assert np.array_equal(libc, ctr)
            assert (libc != lab).any()
            for rounds in (1, 3):
                ml, _, _ = O.merging_refine(O.CPM, lam, n, m, ind, adj, lab, s, iv, ev, 0.0, rounds, 0, 7, 1, tie)
                mc, _, _ = O.merging_refine(O.CPM, lam, n, m, ind, adj, lab, s, iv, ev, 0.0, rounds, 1, 7, 1, tie)
                assert np.array_equal(ml, mc)


def test_quality_fixture_is_the_oracle():
    """tests/golden/quality_sbm1.npz holds the oracle's glibc-mode runs on
    SBM-1; re-running a few seeds of every case reproduces it exactly."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_quality as MQ
    d = load(golden("quality_sbm1.npz")[0])
    g = MQ.sbm1()
    n, m, ind, adj, _ = g
    assert tuple(d["graph_n_m"]) == (n, m)
    for c, (_, kind, lam, ini) in enumerate(MQ.CASES):
        for side, st in enumerate((0, 2)):
            for j in (0, 1, 255):
                sd = int(d["seeds"][j])
                lab, score, iters = MQ.reference_run(n, m, ind, adj, kind, lam, ini, st, sd)
                assert score == d["scores"][c, side, j] and iters == d["iterations"][c, side, j]
