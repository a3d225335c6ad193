"""The drop-in tools end to end on the GPU: pcd_main against the serial
replay (labels) and against the reference binary (report layout)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT, golden, load

pytestmark = pytest.mark.gpu
parcd = pytest.importorskip("parcd")

MAIN = os.path.join(ROOT, "par-cd_amd", "bin", "pcd_main")
REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "main_pin")


def write_graph(tmp_path, name):
    d = load(golden("graph_%s.npz" % name)[0])
    p = tmp_path / (name + ".pcd")
    p.write_bytes(d["pcd"].tobytes())
    return str(p), O.parse_pcd(d["pcd"].tobytes())


def run_main(exe, path, args):
    r = subprocess.run([exe, path] + args, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def labels_from_stderr(err, mp):
    lab = {}
    for line in err.strip().splitlines():
        a, b = line.split()
        lab[int(a)] = int(b)
    return np.array([lab[int(o)] for o in mp], np.int32)


@pytest.mark.parametrize("name,args,kind,lam,strategy", [
    ("sbm200", ["-m", "mod"], O.MOD, 0.0, 0),
    ("sbm200", ["-m", "cpm-0.05", "-s", "2"], O.CPM, 0.05, 2),
    ("rmat9", ["-m", "mod", "-s", "1", "-n", "6"], O.MOD, 0.0, 1),
    ("sbm210", ["-m", "modularity", "-s", "2", "-r", "3"], O.MOD, 0.0, 2),
])
def test_pcd_main_equals_replay(tmp_path, name, args, kind, lam, strategy):
    path, (n, m, ind, adj, mp) = write_graph(tmp_path, name)
    rc, out, err = run_main(MAIN, path, args + ["-D"])
    assert rc == 1, err[-500:]   # main.cpp:286 returns 1
    got = labels_from_stderr(err, mp)
    max_iter = int(args[args.index("-n") + 1]) if "-n" in args else -1
    rounds = int(args[args.index("-r") + 1]) if "-r" in args else 1
    lab, score, iters, k, trace = O.run(n, m, ind, adj, kind=kind, lam=lam, strategy=strategy, rounds=rounds,
                                        max_iter=max_iter, mode=1, seed=0, tie=O.TIE_LOWEST)
    assert np.array_equal(got, lab)
    assert int(re.search(r"refinement iterations: (\d+)", out).group(1)) == iters
    assert int(re.search(r"num\. communities: (\d+)", out).group(1)) == k


def test_pcd_main_wcc_with_triangle_partition(tmp_path):
    path, (n, m, ind, adj, mp) = write_graph(tmp_path, "sbm200")
    rc, out, err = run_main(MAIN, path, ["-m", "wcc", "-i", "1", "-n", "5", "-D"])
    assert rc == 1, err[-500:]
    got = labels_from_stderr(err, mp)
    with parcd.Context(0) as ctx:
        ctx.upload_graph(n, m, ind, adj)
        cc = ctx.triangles()[2]
    lab, score, iters, k, trace = O.run(n, m, ind, adj, kind=O.WCC, initial=1, max_iter=5, mode=1, seed=0,
                                        tie=O.TIE_LOWEST, cc_override=cc)
    assert np.array_equal(got, lab)
    assert "count triangles:" in out


def _layout(text):
    """Report lines with every number replaced, in order."""
    return [re.sub(r"-?[0-9][0-9.e+-]*|nan|inf", "#", l) for l in text.strip().splitlines()]


def test_report_layout_matches_reference(tmp_path):
    if not os.path.exists(REF_MAIN):
        pytest.skip("reference binary not built")
    path, _ = write_graph(tmp_path, "sbm200")
    for args in (["-m", "mod", "-n", "3"], ["-m", "cpm-0.05", "-s", "1", "-n", "2"]):
        rc1, out1, _ = run_main(MAIN, path, args)
        rc2, out2, _ = run_main(REF_MAIN, path, args + ["-p", "1"])
        assert rc1 == rc2 == 1
        assert _layout(out1) == _layout(out2)


def test_pcd_main_errors(tmp_path):
    rc, out, err = run_main(MAIN, str(tmp_path / "missing.pcd"), [])
    assert rc != 0 and rc != 1 and "error while loading graph" in err
    path, _ = write_graph(tmp_path, "sbm200")
    rc, out, err = run_main(MAIN, path, ["-m", "bogus"])
    assert rc != 0 and "unknown metric" in err
