"""Writes tests/golden/quality_sbm1.npz: the reference side of the quality
comparison on SBM-1 (100 blocks of 100, p_in 0.08, p_out 0.0005, generator
seed 1).  For each case (modularity, CPM 0.01, WCC with the triangle-based
initial partition) and schedule (-s 0, -s 2), the oracle in glibc mode --
label for label the reference binary at -p 1 (pinned by the run fixtures) --
is reset with srand(s) for s = 1..SEEDS and run to convergence; the final
score, iterations and the NMI against the planted blocks are stored.

Run from the repository root after the build:
    python tests/golden/make_quality.py
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "par-cd_amd"))

import oracle_lib as O  # noqa: E402
import parcd  # noqa: E402

SEEDS = 256
CASES = [("modularity", O.MOD, 0.0, 0), ("cpm-0.01", O.CPM, 0.01, 0), ("wcc -i 1", O.WCC, 0.0, 1)]


def sbm1():
    return parcd.make_graph("sbm", 1, block_sizes=[100] * 100, p_in=0.08, p_out=0.0005)


def planted(g):
    """SBM-1's generator numbers block b's vertices 100 b .. 100 b + 99; the
    PCD map holds each vertex's generator id."""
    return (np.asarray(g[4]) // 100).astype(np.int32)


def reference_run(n, m, ind, adj, kind, lam, initial, strategy, seed):
    lab, score, iters, k, _ = O.run(n, m, ind, adj, kind=kind, lam=lam, initial=initial, strategy=strategy, mode=0,
                                    seed=seed)
    return lab, score, iters


def _work(a):
    g, kind, lam, ini, st, sd = a
    n, m, ind, adj, _ = g
    lab, score, iters = reference_run(n, m, ind, adj, kind, lam, ini, st, sd)
    return score, iters, parcd.nmi(lab, planted(g))


def main():
    g = sbm1()
    work = [(g, kind, lam, ini, st, sd) for _, kind, lam, ini in CASES for st in (0, 2) for sd in range(1, SEEDS + 1)]
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(_work, work, chunksize=8)
    shape = (len(CASES), 2, SEEDS)
    out = {
        "scores": np.array([r[0] for r in res]).reshape(shape),
        "iterations": np.array([r[1] for r in res], np.int32).reshape(shape),
        "nmi": np.array([r[2] for r in res]).reshape(shape),
        "seeds": np.arange(1, SEEDS + 1, dtype=np.int32),
        "graph_n_m": np.array([g[0], g[1]], np.int64),
    }
    np.savez_compressed(os.path.join(HERE, "quality_sbm1.npz"), **out)
    for c, (name, *_rest) in enumerate(CASES):
        for s in (0, 1):
            print("%s -s %d: score %.6f +- %.6f, nmi %.5f" % (name, 2 * s, out["scores"][c, s].mean(),
                                                             out["scores"][c, s].std(), out["nmi"][c, s].mean()))


if __name__ == "__main__":
    main()
