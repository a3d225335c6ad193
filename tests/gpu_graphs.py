"""Seeded test graphs for the GPU parity tests (CPU-side construction)."""
import functools

import numpy as np

import oracle_lib as O


def csr_from_edges(a, b):
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    keep = a != b
    a, b = a[keep], b[keep]
    nn = int(max(a.max(), b.max())) + 1
    return O.convert_edges(nn, a, b)


def csr_direct(n, edges):
    """CSR with the given vertex count: isolated vertices stay (what a PCD
    file may hold; the edge-list converter drops them)."""
    nb = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    ind = np.zeros(n + 1, np.int32)
    for v in range(n):
        ind[v + 1] = ind[v] + len(nb[v])
    adj = np.array([u for v in range(n) for u in sorted(nb[v])], np.int32)
    return n, int(ind[-1]) // 2, ind, adj, np.arange(n, dtype=np.int32)


# small graphs at the edges of the input domain: (n, edge list)
TINY = {
    "no_edges": (6, []),
    "single_vertex": (1, []),
    "edge_and_isolated": (4, [(0, 1)]),
    "path_and_isolated": (7, [(0, 1), (1, 2), (2, 3)]),
    "star": (9, [(0, i) for i in range(1, 9)]),
    "clique5": (5, [(i, j) for i in range(5) for j in range(i + 1, 5)]),
}


@functools.lru_cache(maxsize=None)
def bin_zoo(seed=7):
    """Degrees spanning every bin: hubs above 4096, a 1024-4096 block range,
    wave ranges, and the small group bins; sparse random background."""
    rng = np.random.default_rng(seed)
    n = 30000
    a = [rng.integers(0, n, 120000)]
    b = [rng.integers(0, n, 120000)]
    # planted communities so labels repeat inside rows
    for c in range(40):
        mem = rng.choice(n, 60, replace=False)
        i, j = np.triu_indices(60, 1)
        keep = rng.random(i.size) < 0.5
        a.append(mem[i[keep]]); b.append(mem[j[keep]])
    for hub, deg in [(0, 14000), (1, 9000), (2, 5000), (3, 3000), (4, 1500), (5, 700), (6, 300), (7, 100)]:
        nb = rng.choice(np.arange(8, n), deg, replace=False)
        a.append(np.full(deg, hub)); b.append(nb)
    return csr_from_edges(np.concatenate(a), np.concatenate(b))


@functools.lru_cache(maxsize=None)
def rmat(scale, seed=2, factor=16):
    import parcd
    return parcd.make_graph("rmat", seed, scale=scale, draws=factor << scale)


@functools.lru_cache(maxsize=None)
def sbm(blocks, size, p_in, p_out, seed=1):
    import parcd
    return parcd.make_graph("sbm", seed, block_sizes=[size] * blocks, p_in=p_in, p_out=p_out)


def unsorted_copy(g, seed=3):
    """Same graph with vertex ids shuffled (not degree-sorted)."""
    n, m, ind, adj, mp = g
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)          # old -> new
    deg = np.diff(ind)
    inv = np.argsort(perm)             # new -> old
    new_deg = deg[inv]
    new_ind = np.concatenate([[0], np.cumsum(new_deg)]).astype(np.int32)
    new_adj = np.empty_like(adj)
    for v_new in range(n):
        v_old = inv[v_new]
        row = np.sort(perm[adj[ind[v_old]:ind[v_old + 1]]])
        new_adj[new_ind[v_new]:new_ind[v_new + 1]] = row
    return n, m, new_ind, new_adj, mp[inv]


def singleton_labels(n):
    L = O.lib()
    L.orc_srand(1)
    lab = np.zeros(n, np.int32)
    L.orc_singleton_partition(n, lab)
    return lab


@functools.lru_cache(maxsize=None)
def dense_blocks(seed=5):
    """A near-clique of 2200 vertices beside 30 dense blocks of 100: vertex
    prefixes N(v) ∩ [0, v) from a few up to above 2048 entries."""
    import parcd
    return parcd.make_graph("sbm", seed, block_sizes=[2200] + [100] * 30, p_in=0.97, p_out=0.001)
