"""CPU tests of the C-ABI library: exports, PCD I/O, converter bytes against
the reference's golden outputs, generators.  No GPU calls."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT, golden, load

parcd = pytest.importorskip("parcd")
HEADER = os.path.join(ROOT, "include", "parcd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = parcd.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(parcd.EXPORTS) == syms


def test_no_device_is_an_error_not_a_fallback():
    if parcd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(parcd.PcdError):
        parcd.Context(0)


@pytest.mark.parametrize("path", golden("convert_*.npz"), ids=os.path.basename)
def test_convert_bytes_match_reference(path, tmp_path):
    d = load(path)
    tin = tmp_path / "in.txt"
    tin.write_bytes(d["text"].tobytes())
    a, b, nn = parcd.read_edge_text(str(tin))
    n, m, ind, adj, mp = parcd.convert_edges(nn, a, b)
    assert O.pcd_bytes(n, m, ind, adj, mp) == d["pcd"].tobytes()
    out = tmp_path / "o.pcd"
    parcd.write_pcd(str(out), n, m, ind, adj, mp)
    assert out.read_bytes() == d["pcd"].tobytes()
    n2, m2, ind2, adj2, mp2 = parcd.read_pcd(str(out))
    assert (n2, m2) == (n, m) and np.array_equal(ind2, ind) and np.array_equal(adj2, adj)
    # the pcd_convert tool writes the same bytes
    tool = os.path.join(ROOT, "par-cd_amd", "bin", "pcd_convert")
    out2 = tmp_path / "o2.pcd"
    r = subprocess.run([tool, str(tin), str(out2)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out2.read_bytes() == d["pcd"].tobytes()


def test_convert_matches_oracle_on_random_multigraphs():
    rng = np.random.default_rng(17)
    for trial in range(30):
        nn = int(rng.integers(2, 400))
        ne = int(rng.integers(1, 3000))
        a = rng.integers(0, nn, ne)
        b = rng.integers(0, nn, ne)
        keep = a != b
        a, b = a[keep], b[keep]
        if a.size == 0:
            continue
        nn = int(max(a.max(), b.max())) + 1
        got = parcd.convert_edges(nn, a, b)
        want = O.convert_edges(nn, a, b)
        for x, y in zip(got, want):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_read_pcd_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.pcd"
    p.write_bytes(b"XYZ" + b"\0" * 20)
    with pytest.raises(parcd.PcdError):
        parcd.read_pcd(str(p))
    d = load(golden("convert_sbm200.npz")[0])
    p.write_bytes(d["pcd"].tobytes() + b"\0")   # trailing byte (graph.cpp:90-93)
    with pytest.raises(parcd.PcdError):
        parcd.read_pcd(str(p))
    p.write_bytes(d["pcd"].tobytes()[:-3])      # truncated
    with pytest.raises(parcd.PcdError):
        parcd.read_pcd(str(p))
    with pytest.raises(parcd.PcdError):
        parcd.read_pcd(str(tmp_path / "missing.pcd"))


def test_read_edge_text_rejects_corrupt(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text("# c\n\n1 2\n3 4 \n")
    a, b, nn = parcd.read_edge_text(str(p))
    assert list(a) == [1, 3] and list(b) == [2, 4] and nn == 5
    p.write_text("1 2\nx 3\n")
    with pytest.raises(parcd.PcdError):
        parcd.read_edge_text(str(p))


def test_generators_are_deterministic_and_shaped():
    a1, b1, n1 = parcd.gen_sbm(1, [100] * 10, 0.08, 0.0005)
    a2, b2, n2 = parcd.gen_sbm(1, [100] * 10, 0.08, 0.0005)
    assert n1 == 1000 and np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert np.all(a1 != b1)
    same = (a1 // 100) == (b1 // 100)
    # expected intra edges 10 * C(100,2) * 0.08 = 3960, inter 45*100*100*0.0005 = 225
    assert abs(same.sum() - 3960) < 4 * np.sqrt(3960)
    assert abs((~same).sum() - 225) < 5 * np.sqrt(225)
    a3, b3, n3 = parcd.gen_sbm(2, [100] * 10, 0.08, 0.0005)
    assert not (a3.size == a1.size and np.array_equal(a3, a1))
    ra, rb, rn = parcd.gen_rmat(2, 10, 16 << 10)
    ra2, rb2, _ = parcd.gen_rmat(2, 10, 16 << 10)
    assert np.array_equal(ra, ra2) and rn <= 1024 and np.all(ra != rb)
    # R-MAT skew: low ids (quadrant a) dominate
    assert (ra < 512).mean() > 0.7
    ca, cb, cn = parcd.gen_chung_lu(3, 5000, 2.3, 4, 500, 8, 20, 400, 20.0)
    assert cn == 5000 and np.all(ca != cb)
    deg = np.bincount(np.concatenate([ca, cb]), minlength=cn)
    assert 4 <= deg.mean() < 40


def test_graft_entry_imports():
    import importlib.util
    spec = importlib.util.spec_from_file_location("ge", os.path.join(ROOT, "__graft_entry__.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    assert callable(m.build) and callable(m.smoke)


@pytest.mark.parametrize("path", golden("nmi_*.npz"))
def test_nmi_matches_reference(path, tmp_path):
    """pcd_nmi / pcd_nmi_file (host code of the library) against the compiled
    reference's compare_communities, NaN and int32-wrap cases included."""
    d = load(path)
    want = float(d["nmi"][0])
    if "truth" in d:
        t = tmp_path / "truth.txt"
        t.write_bytes(d["truth"].tobytes())
        got = parcd.nmi_file(d["labels"], str(t))
    else:
        got = parcd.nmi(d["a"], d["b"])
    assert (np.isnan(got) and np.isnan(want)) or got == pytest.approx(want, rel=1e-13, abs=1e-15), (got, want)


def test_nmi_random_and_errors(tmp_path):
    rng = np.random.default_rng(21)
    for n, ka, kb in [(10, 2, 3), (5000, 50, 8), (20000, 3000, 3000)]:
        a, b = rng.integers(0, ka, n), rng.integers(0, kb, n)
        assert parcd.nmi(a, b) == pytest.approx(O.nmi(a, b), rel=1e-13)
    with pytest.raises(parcd.PcdError):
        parcd.nmi(np.array([0, -1]), np.array([0, 0]))
    p = tmp_path / "t.txt"
    p.write_text("0 1\n2 9\n")
    with pytest.raises(parcd.PcdError):  # vertex outside [0, n)
        parcd.nmi_file(np.zeros(5, np.int32), str(p))
    with pytest.raises(parcd.PcdError):
        parcd.nmi_file(np.zeros(5, np.int32), str(tmp_path / "missing.txt"))
