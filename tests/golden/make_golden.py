#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Runs in the development container only (it needs oracle/_ref/, built from the
reference's own CPU sources by oracle/Makefile).  Each fixture is an .npz of
inputs and the reference's outputs:

  convert_<name>.npz  text edge list -> PCD bytes (reference `convert`)
  graph_<name>.npz    PCD + Graph::count_triangles (tri, tri_deg, cc)
  part_<name>_<p>.npz labels -> compress_labels, collect_statistics, every
                      rule's calculate and score_join/leave/merge values
  refine_<...>.npz    one refinement step at one thread after srand(seed)
  run_<...>.npz       full `main -p 1` runs: labels, score, iterations
  synth_scores.npz    score functions on fabricated statistics, including
                      the int32-overflow region of the WCC gains
  nmi_<case>.npz      compare_communities on two labelings or on a labeling
                      and a ground-truth community file (written by
                      `make_golden.py nmi`, which leaves the other fixtures
                      alone)

Input graphs are seeded numpy draws (small SBM, G(n,p), R-MAT, star+clique
and hand-written edge lists).
"""
import os
import re
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")


def write_recs(path, recs):
    with open(path, "wb") as f:
        for tag, arr in recs.items():
            arr = np.asarray(arr)
            kind = 1 if arr.dtype.kind == "f" else 0
            arr = arr.astype(np.float64 if kind else np.int32).ravel()
            t = tag.encode()
            f.write(struct.pack("<i", len(t)) + t + struct.pack("<iq", kind, arr.size))
            f.write(arr.tobytes())


def read_recs(path):
    out = {}
    data = open(path, "rb").read()
    off = 0
    while off < len(data):
        (tl,) = struct.unpack_from("<i", data, off); off += 4
        tag = data[off:off + tl].decode(); off += tl
        kind, cnt = struct.unpack_from("<iq", data, off); off += 12
        dt = np.float64 if kind else np.int32
        nb = cnt * (8 if kind else 4)
        out[tag] = np.frombuffer(data[off:off + nb], dtype=dt).copy(); off += nb
    return out


def read_pcd(path):
    b = open(path, "rb").read()
    assert b[:3] == b"PCD"
    n, m = struct.unpack_from("<ii", b, 3)
    a = np.frombuffer(b[11:], dtype=np.int32)
    indices = a[: n + 1].copy()
    adj = a[n + 1: n + 1 + 2 * m].copy()
    mp = a[n + 1 + 2 * m: n + 1 + 2 * m + n].copy()
    return n, m, indices, adj, mp, b


# ---------------------------------------------------------------- generators
def gen_sbm(rng, sizes, p_in, p_out):
    n = int(sum(sizes))
    block = np.repeat(np.arange(len(sizes)), sizes)
    iu, ju = np.triu_indices(n, 1)
    p = np.where(block[iu] == block[ju], p_in, p_out)
    keep = rng.random(iu.size) < p
    return np.stack([iu[keep], ju[keep]], 1), block


def gen_gnp(rng, n, p):
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.size) < p
    return np.stack([iu[keep], ju[keep]], 1)


def gen_rmat(rng, scale, ne, a=0.57, b=0.19, c=0.19):
    src = np.zeros(ne, np.int64)
    dst = np.zeros(ne, np.int64)
    for _ in range(scale):
        r = rng.random(ne)
        bi = (r >= a) & (r < a + b)
        ci = (r >= a + b) & (r < a + b + c)
        di = r >= a + b + c
        src = 2 * src + (ci | di)
        dst = 2 * dst + (bi | di)
    e = np.stack([src, dst], 1)
    return e[e[:, 0] != e[:, 1]]


def star_clique():
    e = [(0, i) for i in range(1, 40)]
    e += [(40 + i, 40 + j) for i in range(12) for j in range(i + 1, 12)]
    e += [(5, 41), (7, 45), (52, 53), (53, 54), (54, 52), (54, 55)]
    return np.array(e)


def to_text(edges, rng=None, noise=False):
    lines = []
    if noise:
        lines.append("# comment line")
        lines.append("")
    for (x, y) in edges:
        if noise and rng.random() < 0.3:
            x, y = y, x
        if noise and rng.random() < 0.1:
            lines.append("  %d\t%d  " % (x, y))
        else:
            lines.append("%d %d" % (x, y))
        if noise and rng.random() < 0.1:
            lines.append("%d %d" % (y, x))
    return "\n".join(lines) + "\n"


def run(cmd, **kw):
    return subprocess.run(cmd, check=False, capture_output=True, text=True, **kw)


def convert(text, tmp):
    tin = os.path.join(tmp, "in.txt")
    tout = os.path.join(tmp, "out.pcd")
    open(tin, "w").write(text)
    r = run([os.path.join(REF, "convert"), tin, tout])
    assert r.returncode == 0, r.stderr
    return open(tout, "rb").read(), tout


def main_run(pcd, args):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = run([os.path.join(REF, "main_pin"), pcd, "-p", "1"] + args, env=env)
    assert r.returncode == 1, (r.returncode, r.stderr[-500:])
    lab = {}
    for line in r.stderr.splitlines():
        a, b = line.split()
        lab[int(a)] = int(b)
    it = int(re.search(r"refinement iterations: (\d+)", r.stdout).group(1))
    k = int(re.search(r"num\. communities: (\d+)", r.stdout).group(1))
    score_line = r.stdout.strip().splitlines()[-5]
    score = float(score_line.rsplit(":", 1)[1])
    trace = [float(x) for x in re.findall(r" - [^\n]*?: ([-0-9.e+naif]+)\n - improv", r.stdout)]
    return lab, it, k, score, trace, r.stdout


def main():
    if not os.path.exists(os.path.join(REF, "ref_harness")):
        sys.exit("build oracle/_ref first (make -C oracle ref)")
    rng = np.random.default_rng(20240601)
    graphs = {}
    graphs["readme"] = (np.array([[1, 4], [4, 3], [3, 5], [1, 5]]), False)
    e, _ = gen_sbm(rng, [25] * 8, 0.3, 0.01)
    graphs["sbm200"] = (e, True)
    graphs["gnp150"] = (gen_gnp(rng, 150, 0.06), True)
    graphs["rmat9"] = (gen_rmat(rng, 9, 4096), True)
    graphs["starclique"] = (star_clique(), True)
    e, _ = gen_sbm(rng, [60, 50, 40, 30, 20, 10], 0.25, 0.02)
    graphs["sbm210"] = (e, False)

    with tempfile.TemporaryDirectory() as tmp:
        pcds = {}
        for name, (edges, noise) in graphs.items():
            text = to_text(edges, rng, noise)
            pcd_bytes, path = convert(text, tmp)
            keep = os.path.join(tmp, name + ".pcd")
            open(keep, "wb").write(pcd_bytes)
            pcds[name] = keep
            np.savez_compressed(os.path.join(HERE, "convert_%s.npz" % name),
                                text=np.frombuffer(text.encode(), np.uint8),
                                pcd=np.frombuffer(pcd_bytes, np.uint8))
            out = os.path.join(tmp, "g.out")
            r = run([os.path.join(REF, "ref_harness"), "graph", keep, out])
            assert r.returncode == 0, r.stderr
            g = read_recs(out)
            np.savez_compressed(os.path.join(HERE, "graph_%s.npz" % name),
                                pcd=np.frombuffer(pcd_bytes, np.uint8), **g)
            print("graph", name, read_pcd(keep)[:2])

        # partitions
        for name in ["readme", "sbm200", "gnp150", "rmat9", "starclique", "sbm210"]:
            n, m, indices, adj, mp, _ = read_pcd(pcds[name])
            parts = {
                "single": np.arange(n)[rng.permutation(n)],
                "coarse": rng.integers(0, max(2, n // 20), n),
                "holes": np.where(rng.random(n) < 0.2, -1, rng.integers(0, max(2, n // 5), n) * 3 + 7),
                "one": np.zeros(n, np.int64),
            }
            for pname, lab in parts.items():
                inp = os.path.join(tmp, "p.in")
                out = os.path.join(tmp, "p.out")
                write_recs(inp, {"labels": lab, "lambda": np.array([0.05])})
                r = run([os.path.join(REF, "ref_harness"), "part", pcds[name], inp, out])
                assert r.returncode == 0, r.stderr
                res = read_recs(out)
                np.savez_compressed(os.path.join(HERE, "part_%s_%s.npz" % (name, pname)),
                                    in_labels=lab.astype(np.int32), lam=np.array([0.05]), **res)

        # single refinement steps (reference order of rand()/rand_r())
        for name in ["sbm200", "rmat9", "starclique", "gnp150"]:
            n = read_pcd(pcds[name])[0]
            for kind, lam in [(0, 0.0), (1, 0.05), (2, 0.0)]:
                for strategy in (0, 1):
                    for pname, lab in [("single", np.arange(n)[rng.permutation(n)]),
                                       ("coarse", rng.integers(0, max(2, n // 8), n))]:
                        inp = os.path.join(tmp, "r.in")
                        out = os.path.join(tmp, "r.out")
                        seed = int(rng.integers(1, 1 << 30))
                        write_recs(inp, {"labels": lab, "kind": np.array([kind]),
                                         "strategy": np.array([strategy]), "rounds": np.array([2]),
                                         "srand": np.array([seed]), "lambda": np.array([lam]),
                                         "frac": np.array([0.5 if strategy == 0 else 1.0])})
                        r = run([os.path.join(REF, "ref_harness"), "refine", pcds[name], inp, out])
                        assert r.returncode == 0, r.stderr
                        res = read_recs(out)
                        np.savez_compressed(
                            os.path.join(HERE, "refine_%s_k%d_s%d_%s.npz" % (name, kind, strategy, pname)),
                            in_labels=lab.astype(np.int32), kind=kind, lam=lam, strategy=strategy,
                            rounds=2, srand=seed, frac=0.5 if strategy == 0 else 1.0, **res)

        # full runs of the reference driver at one thread
        runs = [
            ("sbm200", ["-m", "mod"]),
            ("sbm200", ["-m", "mod", "-s", "1"]),
            ("sbm200", ["-m", "mod", "-s", "2"]),
            ("sbm200", ["-m", "cpm-0.05"]),
            ("sbm200", ["-m", "cpm-0.05", "-s", "2", "-r", "3"]),
            ("sbm200", ["-m", "wcc", "-i", "1", "-n", "6"]),
            ("sbm210", ["-m", "mod", "-f", "0.7"]),
            ("sbm210", ["-m", "wcc", "-i", "1", "-s", "2", "-n", "8"]),
            ("rmat9", ["-m", "mod", "-n", "8"]),
            ("rmat9", ["-m", "cpm-0.01", "-s", "1", "-n", "5"]),
            ("gnp150", ["-m", "mod", "-i", "1", "-s", "2"]),
            ("starclique", ["-m", "mod", "-s", "2"]),
        ]
        for i, (name, args) in enumerate(runs):
            lab, it, k, score, trace, stdout = main_run(pcds[name], args)
            mp = read_pcd(pcds[name])[4]
            labels = np.array([lab[int(o)] for o in mp], np.int32)
            np.savez_compressed(os.path.join(HERE, "run_%02d_%s.npz" % (i, name)),
                                args=np.frombuffer(" ".join(args).encode(), np.uint8),
                                labels=labels, iters=it, k=k, score=score,
                                trace=np.array(trace))
            print("run", name, args, it, k, score)

        # fabricated statistics (int32 overflow region of rules.cpp:265-358)
        K = 64
        sizes = np.concatenate([rng.integers(1, 60, K // 2), rng.integers(30000, 120000, K // 2)])
        intv = (sizes * rng.integers(1, 40, K)).astype(np.int64)
        intv = np.minimum(intv, 2**30) // 2 * 2
        extv = rng.integers(0, 2**28, K)
        nv = 48
        deg = rng.integers(1, 200000, nv)
        own = rng.integers(0, K, nv)
        qv = rng.integers(0, nv, 400)
        qc = rng.integers(0, K, 400)
        qn = np.minimum(rng.integers(0, 150000, 400), deg[qv])
        mi = rng.integers(0, K, 400)
        mj = rng.integers(0, K, 400)
        ml = rng.integers(0, 200000, 400)
        inp = os.path.join(tmp, "s.in")
        out = os.path.join(tmp, "s.out")
        rec = {"deg": deg, "m": np.array([2**29]), "cc": np.array([0.137]), "sizes": sizes,
               "int_volume": intv, "ext_volume": extv, "own": own, "lambda": np.array([0.013]),
               "q_v": qv, "q_c": qc, "q_n": qn, "m_i": mi, "m_j": mj, "m_l": ml}
        write_recs(inp, rec)
        r = run([os.path.join(REF, "ref_harness"), "synth", inp, out])
        assert r.returncode == 0, r.stderr
        res = read_recs(out)
        np.savez_compressed(os.path.join(HERE, "synth_scores.npz"),
                            **{k: np.asarray(v) for k, v in rec.items()}, **res)
    print("done")


def make_nmi():
    """compare_communities (common.cpp:75-173) on seeded labelings of the
    golden graphs and of a 150,000-vertex path, where the int32 product
    ni * nj of norm_mutual_info wraps (finite but shifted values, and NaN
    where the wrapped product is negative)."""
    rng = np.random.default_rng(20240917)
    with tempfile.TemporaryDirectory() as tmp:
        pcds = {}
        for name in ["sbm200", "rmat9", "gnp150"]:
            d = np.load(os.path.join(HERE, "graph_%s.npz" % name))
            path = os.path.join(tmp, name + ".pcd")
            open(path, "wb").write(d["pcd"].tobytes())
            pcds[name] = path
        npath = 150000
        text = "".join("%d %d\n" % (i, i + 1) for i in range(npath - 1))
        _, out = convert(text, tmp)
        pcds["path150k"] = os.path.join(tmp, "path.pcd")
        os.rename(out, pcds["path150k"])

        def pair(case, name, a, b):
            n = read_pcd(pcds[name])[0]
            assert len(a) == n and len(b) == n
            inp, out = os.path.join(tmp, "n.in"), os.path.join(tmp, "n.out")
            write_recs(inp, {"a": np.asarray(a), "b": np.asarray(b)})
            r = run([os.path.join(REF, "ref_harness"), "nmi", pcds[name], inp, out])
            assert r.returncode == 0, r.stderr
            res = read_recs(out)
            np.savez_compressed(os.path.join(HERE, "nmi_%s.npz" % case), n=n, a=np.asarray(a, np.int32),
                                b=np.asarray(b, np.int32), nmi=res["nmi"])
            print("nmi", case, res["nmi"])

        def truth(case, name, labels, text):
            n = read_pcd(pcds[name])[0]
            inp, out = os.path.join(tmp, "n.in"), os.path.join(tmp, "n.out")
            tpath = os.path.join(tmp, "truth.txt")
            open(tpath, "w").write(text)
            write_recs(inp, {"labels": np.asarray(labels)})
            r = run([os.path.join(REF, "ref_harness"), "nmifile", pcds[name], inp, tpath, out])
            assert r.returncode == 0, r.stderr
            res = read_recs(out)
            np.savez_compressed(os.path.join(HERE, "nmi_%s.npz" % case), n=n, labels=np.asarray(labels, np.int32),
                                truth=np.frombuffer(text.encode(), np.uint8), nmi=res["nmi"])
            print("nmi", case, res["nmi"])

        for name in ["sbm200", "rmat9", "gnp150"]:
            n = read_pcd(pcds[name])[0]
            c1, c2 = rng.integers(0, 8, n), rng.integers(0, 12, n)
            pair("%s_random" % name, name, c1, c2)
            pair("%s_same" % name, name, c1, c1)
            pair("%s_singletons" % name, name, rng.permutation(n), c2)
            pair("%s_one" % name, name, np.zeros(n, np.int64), c1)
            pair("%s_sparse_ids" % name, name, c1 * 1000 + 999983, c2 * 7 + 5)
        n = read_pcd(pcds["sbm200"])[0]
        pair("sbm200_both_one", "sbm200", np.zeros(n, np.int64), np.zeros(n, np.int64))
        ids = np.arange(npath)
        pair("path150k_wrap", "path150k", (ids >= 80000).astype(np.int64), (ids >= 70000).astype(np.int64))
        pair("path150k_nan", "path150k", ids // 50000, ids // 50000)
        pair("path150k_random", "path150k", rng.integers(0, 40, npath), rng.integers(0, 60, npath))
        # ground-truth files: comments, blank lines, tabs, a line cut by a
        # non-number, vertices listed twice or never
        n = read_pcd(pcds["sbm200"])[0]
        lab = rng.integers(0, 10, n)
        groups = [[v for v in range(n) if v % 8 == g] for g in range(8)]
        lines = ["# planted blocks", ""]
        for g, vs in enumerate(groups):
            sep = "\t" if g % 3 == 0 else " "
            lines.append(sep.join(str(v) for v in vs))
        lines.append("3 4 x 5")
        lines.append("")
        lines.append("7 7 9")
        truth("sbm200_file", "sbm200", lab, "\n".join(lines) + "\n")
        truth("sbm200_file_partial", "sbm200", lab, "\n".join(" ".join(str(v) for v in vs[: len(vs) // 2])
                                                                for vs in groups[:5]) + "\n")
    print("done")


if __name__ == "__main__":
    if sys.argv[1:] == ["nmi"]:
        make_nmi()
    else:
        main()
