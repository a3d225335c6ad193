"""CPU tests of graph sets (many PCD graphs in one file): layout, canonical
bytes, member bytes equal to the reference `convert` output, zero-copy
views, rejections and the pcd_graphset tool.  No GPU calls."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT, golden, load

parcd = pytest.importorskip("parcd")
TOOL = os.path.join(ROOT, "par-cd_amd", "bin", "pcd_graphset")


def golden_pcds():
    """The reference `convert` outputs held as golden fixtures."""
    return [load(p)["pcd"].tobytes() for p in golden("convert_*.npz")]


def expected_set_bytes(members):
    """The set layout written out by hand: magic, count, {offset, bytes}
    table, each member 11 bytes before a 64-byte boundary, zero gaps."""
    at = 16 + 16 * len(members)
    table, body = [], b""
    for mb in members:
        off = at
        while (off + 11) % 64:
            off += 1
        body += b"\0" * (off - at) + mb
        table.append((off, len(mb)))
        at = off + len(mb)
    head = b"PCDSET1\0" + struct.pack("<q", len(members))
    head += b"".join(struct.pack("<qq", o, n) for o, n in table)
    return head + body


def small_graphs():
    out = []
    for seed in range(4):
        n, m, ind, adj, mp = parcd.make_graph("sbm", seed, block_sizes=[20 + 7 * seed] * 3, p_in=0.3, p_out=0.02)
        out.append((n, m, ind, adj, mp))
    # an isolated vertex and an empty graph of one vertex
    out.append((3, 1, np.array([0, 1, 2, 2]), np.array([1, 0]), np.array([5, 6, 7])))
    out.append((1, 0, np.array([0, 0]), np.array([], np.int32), np.array([0])))
    return out


def test_members_are_the_reference_convert_bytes(tmp_path):
    pcds = golden_pcds()
    assert len(pcds) >= 5
    paths = []
    for i, b in enumerate(pcds):
        p = tmp_path / ("g%d.pcd" % i)
        p.write_bytes(b)
        paths.append(str(p))
    out = tmp_path / "set.pcds"
    parcd.pack_graph_set(str(out), paths)
    assert out.read_bytes() == expected_set_bytes(pcds)
    with parcd.GraphSet(str(out)) as s:
        assert len(s) == len(pcds)
        for i, b in enumerate(pcds):
            n, m, ind, adj, mp = O.parse_pcd(b)
            g = s.graph(i)
            assert g[:2] == (n, m)
            assert np.array_equal(g[2], ind) and np.array_equal(g[3], adj) and np.array_equal(g[4], mp)
            e = tmp_path / ("x%d.pcd" % i)
            s.extract(i, str(e))
            assert e.read_bytes() == b
            assert parcd.read_pcd(str(e))[0] == n


def test_writer_is_canonical_and_views_are_aligned(tmp_path):
    graphs = small_graphs()
    a = tmp_path / "a.pcds"
    parcd.write_graph_set(str(a), graphs)
    members = [O.pcd_bytes(n, m, ind, adj, mp) for n, m, ind, adj, mp in graphs]
    assert a.read_bytes() == expected_set_bytes(members)
    with parcd.GraphSet(str(a)) as s:
        views = [s.graph(i) for i in range(len(s))]
        for g, v in zip(graphs, views):
            assert v[:2] == g[:2]
            for x, y in zip(v[2:], g[2:]):
                assert np.array_equal(x, y)
            assert v[2].ctypes.data % 64 == 0 and v[3].ctypes.data % 4 == 0
            assert not v[2].flags.writeable
        # writing what the set holds gives the same file
        b = tmp_path / "b.pcds"
        parcd.write_graph_set(str(b), views)
    assert b.read_bytes() == a.read_bytes()
    # packing the members' PCD files gives the same file too
    paths = []
    for i, g in enumerate(graphs):
        p = tmp_path / ("m%d.pcd" % i)
        parcd.write_pcd(str(p), *g)
        paths.append(str(p))
    c = tmp_path / "c.pcds"
    parcd.pack_graph_set(str(c), paths)
    assert c.read_bytes() == a.read_bytes()


def test_map_defaults_to_identity_and_empty_set(tmp_path):
    n, m, ind, adj, _ = small_graphs()[0]
    p = tmp_path / "s.pcds"
    parcd.write_graph_set(str(p), [(n, m, ind, adj)])
    with parcd.GraphSet(str(p)) as s:
        assert np.array_equal(s.graph(0)[4], np.arange(n))
    e = tmp_path / "e.pcds"
    parcd.write_graph_set(str(e), [])
    assert e.read_bytes() == b"PCDSET1\0" + struct.pack("<q", 0)
    with parcd.GraphSet(str(e)) as s:
        assert len(s) == 0
        with pytest.raises(parcd.PcdError):
            s.graph(0)


def test_views_keep_the_set_open(tmp_path):
    p = tmp_path / "s.pcds"
    parcd.write_graph_set(str(p), small_graphs()[:2])
    s = parcd.GraphSet(str(p))
    ind = s.graph(1)[2]
    want = ind.copy()
    del s
    assert np.array_equal(ind, want)


def corrupt(data, off, new):
    b = bytearray(data)
    b[off:off + len(new)] = new
    return bytes(b)


def test_rejections(tmp_path):
    graphs = small_graphs()[:3]
    good = tmp_path / "g.pcds"
    parcd.write_graph_set(str(good), graphs)
    data = good.read_bytes()
    off1, len1 = struct.unpack_from("<qq", data, 16 + 16)
    cases = {
        "magic": corrupt(data, 0, b"PCDSET2"),
        "short": data[:12],
        "count": corrupt(data, 8, struct.pack("<q", 1 << 40)),
        "negative count": corrupt(data, 8, struct.pack("<q", -1)),
        "misplaced member": corrupt(data, 32, struct.pack("<q", off1 + 64)),
        "length": corrupt(data, 40, struct.pack("<q", len1 + 4)),
        "member magic": corrupt(data, off1, b"XCD"),
        "member size": corrupt(data, off1 + 3, struct.pack("<i", 1)),
        "gap byte": corrupt(data, off1 - 1, b"\1"),
        "trailing": data + b"\0",
        "truncated": data[:-4],
    }
    for name, b in cases.items():
        p = tmp_path / ("bad_%s.pcds" % name.replace(" ", "_"))
        p.write_bytes(b)
        with pytest.raises(parcd.PcdError, match="graph set"):
            parcd.GraphSet(str(p))
    with pytest.raises(parcd.PcdError, match="failed to open"):
        parcd.GraphSet(str(tmp_path / "missing.pcds"))
    # a member whose arrays break the CSR rules opens, and reports its error
    n = graphs[1][0]
    adj0 = off1 + 11 + 4 * (n + 1)
    p = tmp_path / "bad_adj.pcds"
    p.write_bytes(corrupt(data, adj0, struct.pack("<i", n)))
    with parcd.GraphSet(str(p)) as s:
        s.graph(0)
        s.graph(2)
        for _ in range(2):
            with pytest.raises(parcd.PcdError, match="member 1"):
                s.graph(1)
        with pytest.raises(parcd.PcdError):
            s.extract(1, str(tmp_path / "x.pcd"))


def test_writer_rejections(tmp_path):
    L = parcd.lib()
    import ctypes as C
    n, m, ind, adj, mp = small_graphs()[0]
    # fewer graphs than announced: no file is left
    p = tmp_path / "few.pcds"
    w = C.c_void_p()
    assert L.pcd_graphset_writer_open(str(p).encode(), 2, C.byref(w)) == 0
    ind, adj = parcd._i32(ind), parcd._i32(adj)
    assert L.pcd_graphset_writer_add(w, n, m, parcd._p(ind), parcd._p(adj), None) == 0
    assert L.pcd_graphset_writer_close(w) == -6
    assert not p.exists()
    # more graphs than announced, and a graph that breaks the CSR rules
    q = tmp_path / "one.pcds"
    assert L.pcd_graphset_writer_open(str(q).encode(), 1, C.byref(w)) == 0
    bad = adj.copy()
    bad[0] = n
    assert L.pcd_graphset_writer_add(w, n, m, parcd._p(ind), parcd._p(bad), None) == -1
    assert L.pcd_graphset_writer_add(w, n, m, parcd._p(ind), parcd._p(adj), None) == 0
    assert L.pcd_graphset_writer_add(w, n, m, parcd._p(ind), parcd._p(adj), None) == -6
    assert L.pcd_graphset_writer_close(w) == 0
    with parcd.GraphSet(str(q)) as s:
        assert len(s) == 1
    # packing stops at an unreadable file and leaves nothing behind
    r = tmp_path / "pack.pcds"
    good = tmp_path / "ok.pcd"
    parcd.write_pcd(str(good), n, m, ind, adj, mp)
    junk = tmp_path / "junk.pcd"
    junk.write_bytes(b"PCD" + b"\0" * 3)
    with pytest.raises(parcd.PcdError):
        parcd.pack_graph_set(str(r), [str(good), str(junk)])
    assert not r.exists()


def test_tool_pack_list_extract(tmp_path):
    pcds = golden_pcds()[:3]
    paths = []
    for i, b in enumerate(pcds):
        p = tmp_path / ("g%d.pcd" % i)
        p.write_bytes(b)
        paths.append(str(p))
    out = tmp_path / "t.pcds"
    r = subprocess.run([TOOL, "pack", str(out)] + paths, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == expected_set_bytes(pcds)
    r = subprocess.run([TOOL, "list", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    want = ["%d %d %d" % ((i,) + O.parse_pcd(b)[:2]) for i, b in enumerate(pcds)]
    assert r.stdout.split("\n")[:-1] == want
    x = tmp_path / "x.pcd"
    r = subprocess.run([TOOL, "extract", str(out), "2", str(x)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert x.read_bytes() == pcds[2]
    for bad in (["extract", str(out), "3", str(x)], ["extract", str(out), "-1", str(x)],
                ["list", str(tmp_path / "missing")], ["bogus", "a"]):
        r = subprocess.run([TOOL] + bad, capture_output=True, text=True)
        assert r.returncode == 1
