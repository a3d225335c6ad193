"""ctypes access to oracle/liboracle.so (the CPU restatement used as checker).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

MOD, CPM, WCC, COND = 0, 1, 2, 3
TIE_RANDOM, TIE_LOWEST = 0, 1

_lib = None
IP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       capture_output=True)
    L = C.CDLL(ORACLE_SO)
    L.orc_mix64.restype = C.c_uint64
    L.orc_mix64.argtypes = [C.c_uint64]
    L.orc_hash3.restype = C.c_uint64
    L.orc_hash3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
    L.orc_frac_threshold.restype = C.c_int
    L.orc_frac_threshold.argtypes = [C.c_double]
    L.orc_select.restype = C.c_int
    L.orc_select.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int]
    L.orc_tie_key.restype = C.c_uint32
    L.orc_tie_key.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int]
    L.orc_merge_seed.restype = C.c_uint64
    L.orc_merge_seed.argtypes = [C.c_uint64, C.c_int]
    L.orc_read_edges_text.restype = C.c_int
    L.orc_read_edges_text.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                      C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_long)]
    L.orc_free.argtypes = [C.c_void_p]
    L.orc_convert_edges.restype = C.c_int
    L.orc_convert_edges.argtypes = [C.c_int, C.c_long, IP, IP, IP, IP, IP, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_write_pcd.restype = C.c_int
    L.orc_write_pcd.argtypes = [C.c_char_p, C.c_int, C.c_int, IP, IP, IP]
    L.orc_count_triangles.argtypes = [C.c_int, IP, IP, IP, IP, IP, C.POINTER(C.c_double)]
    L.orc_srand.argtypes = [C.c_uint]
    L.orc_rand.restype = C.c_int
    L.orc_singleton_partition.argtypes = [C.c_int, IP]
    L.orc_clustering_partition.argtypes = [C.c_int, IP, IP, IP, IP]
    L.orc_compress_labels.restype = C.c_int
    L.orc_compress_labels.argtypes = [C.c_int, IP]
    L.orc_collect_statistics.restype = C.c_int
    L.orc_collect_statistics.argtypes = [C.c_int, IP, IP, IP, IP, IP, IP]
    L.orc_wcc_improvement.restype = C.c_double
    L.orc_wcc_improvement.argtypes = [C.c_int] * 9 + [C.c_double]
    L.orc_score_join.restype = C.c_double
    L.orc_score_join.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_score_leave.restype = C.c_double
    L.orc_score_leave.argtypes = L.orc_score_join.argtypes
    L.orc_score_merge.restype = C.c_double
    L.orc_score_merge.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double] + [C.c_int] * 7
    L.orc_wcc_calculate.restype = C.c_double
    L.orc_wcc_calculate.argtypes = [C.c_int, IP, IP, IP, IP, IP, IP, C.c_void_p, C.c_void_p]
    L.orc_calculate.restype = C.c_double
    L.orc_calculate.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, IP, IP, IP, C.c_int, IP, IP, IP,
                                C.c_void_p, C.c_void_p]
    L.orc_propagation_refine.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, IP, IP, IP, IP, IP, IP,
                                         C.c_double, C.c_double, C.c_int, C.c_uint64, C.c_int, C.c_int, IP]
    L.orc_merging_refine.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, IP, IP, IP, IP, IP, IP,
                                     C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, IP,
                                     C.c_void_p, C.c_void_p]
    L.orc_community_pairs.restype = C.c_long
    L.orc_community_pairs.argtypes = [C.c_int, IP, IP, IP, C.c_int, IP, IP, IP]
    L.orc_run.restype = C.c_int
    L.orc_run.argtypes = [C.c_int, C.c_int, IP, IP, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                          C.c_double, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_void_p, C.c_double,
                          IP, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_int]
    L.orc_nmi.restype = C.c_int
    L.orc_nmi.argtypes = [C.c_int, IP, IP, C.POINTER(C.c_double)]
    L.orc_nmi_file.restype = C.c_int
    L.orc_nmi_file.argtypes = [C.c_int, IP, C.c_char_p, C.POINTER(C.c_double)]
    _lib = L
    return L


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def read_edges_text(path):
    L = lib()
    pa, pb = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    ne, nn, bad = C.c_long(), C.c_int(), C.c_long(0)
    rc = L.orc_read_edges_text(path.encode(), C.byref(pa), C.byref(pb), C.byref(ne), C.byref(nn), C.byref(bad))
    a = np.ctypeslib.as_array(pa, (max(ne.value, 1),))[: ne.value].copy()
    b = np.ctypeslib.as_array(pb, (max(ne.value, 1),))[: ne.value].copy()
    L.orc_free(C.cast(pa, C.c_void_p)); L.orc_free(C.cast(pb, C.c_void_p))
    return rc, a, b, nn.value, bad.value


def convert_edges(num_nodes, a, b):
    L = lib()
    a, b = i32(a), i32(b)
    ind = np.zeros(num_nodes + 1, np.int32)
    adj = np.zeros(max(1, 2 * len(a)), np.int32)
    mp = np.zeros(max(1, num_nodes), np.int32)
    n, m = C.c_int(), C.c_int()
    rc = L.orc_convert_edges(num_nodes, len(a), a, b, ind, adj, mp, C.byref(n), C.byref(m))
    if rc != 0:
        raise ValueError("convert rc=%d" % rc)
    n, m = n.value, m.value
    return n, m, ind[: n + 1].copy(), adj[: 2 * m].copy(), mp[:n].copy()


def pcd_bytes(n, m, ind, adj, mp):
    import struct
    return b"PCD" + struct.pack("<ii", n, m) + i32(ind).tobytes() + i32(adj).tobytes() + i32(mp).tobytes()


def parse_pcd(b):
    import struct
    b = bytes(b)
    assert b[:3] == b"PCD"
    n, m = struct.unpack_from("<ii", b, 3)
    a = np.frombuffer(b[11:], dtype=np.int32)
    assert a.size == (n + 1) + 2 * m + n
    return n, m, a[: n + 1].copy(), a[n + 1: n + 1 + 2 * m].copy(), a[n + 1 + 2 * m:].copy()


def count_triangles(n, ind, adj):
    L = lib()
    ind, adj = i32(ind), i32(adj)
    et = np.zeros(max(1, adj.size), np.int32)
    tri = np.zeros(max(1, n), np.int32)
    td = np.zeros(max(1, n), np.int32)
    cc = C.c_double()
    L.orc_count_triangles(n, ind, adj, et, tri, td, C.byref(cc))
    return et[: adj.size], tri[:n], td[:n], cc.value


def compress_labels(labels):
    L = lib()
    lab = i32(labels).copy()
    k = L.orc_compress_labels(lab.size, lab)
    return lab, k


def collect_statistics(n, ind, adj, labels):
    L = lib()
    lab = i32(labels)
    cap = int(lab.max()) + 1 if lab.size else 1
    s, iv, ev = (np.zeros(max(1, cap), np.int32) for _ in range(3))
    k = L.orc_collect_statistics(n, i32(ind), i32(adj), lab, s, iv, ev)
    return k, s[:k], iv[:k], ev[:k]


def calculate(kind, lam, n, m, ind, adj, labels, sizes, intv, extv, tri=None, tri_deg=None):
    L = lib()
    tri = None if tri is None else i32(tri)
    tri_deg = None if tri_deg is None else i32(tri_deg)
    return L.orc_calculate(kind, lam, n, m, i32(ind), i32(adj), i32(labels), len(sizes), i32(sizes),
                           i32(intv), i32(extv), _ptr(tri), _ptr(tri_deg))


def wcc_calculate(n, ind, adj, labels, sizes, tri, tri_deg):
    L = lib()
    it = np.zeros(max(1, n), np.int32)
    itd = np.zeros(max(1, n), np.int32)
    s = L.orc_wcc_calculate(n, i32(ind), i32(adj), i32(labels), i32(sizes), i32(tri), i32(tri_deg),
                            _ptr(it), _ptr(itd))
    return s, it[:n], itd[:n]


def propagation_refine(kind, lam, n, m, ind, adj, labels, sizes, intv, extv, cc, frac, mode, seed, step, tie):
    L = lib()
    out = np.zeros(max(1, n), np.int32)
    L.orc_propagation_refine(kind, lam, n, m, i32(ind), i32(adj), i32(labels), i32(sizes), i32(intv), i32(extv),
                             cc, frac, mode, seed, step, tie, out)
    return out[:n]


def merging_refine(kind, lam, n, m, ind, adj, labels, sizes, intv, extv, cc, rounds, mode, seed, step, tie):
    L = lib()
    k = len(sizes)
    out = np.zeros(max(1, n), np.int32)
    bp = np.zeros(max(1, k), np.int32)
    fp = np.zeros(max(1, k), np.int32)
    L.orc_merging_refine(kind, lam, n, m, k, i32(ind), i32(adj), i32(labels), i32(sizes), i32(intv), i32(extv),
                         cc, rounds, mode, seed, step, tie, out, _ptr(bp), _ptr(fp))
    return out[:n], bp[:k], fp[:k]


def community_pairs(n, ind, adj, labels, k):
    L = lib()
    E = int(ind[n])
    a, b, c = (np.zeros(max(1, E), np.int32) for _ in range(3))
    U = L.orc_community_pairs(n, i32(ind), i32(adj), i32(labels), k, a, b, c)
    return a[:U], b[:U], c[:U]


def run(n, m, ind, adj, kind=MOD, lam=0.0, frac=0.5, rounds=1, max_iter=-1, threshold=0.0, initial=0,
        strategy=0, mode=0, seed=0, tie=TIE_RANDOM, init_labels=None, cc_override=-1.0, trace_cap=4096):
    L = lib()
    labels = np.zeros(max(1, n), np.int32)
    score, iters, k = C.c_double(), C.c_int(), C.c_int()
    trace = np.zeros(trace_cap, np.float64)
    il = None if init_labels is None else i32(init_labels)
    L.orc_run(n, m, i32(ind), i32(adj), kind, lam, frac, rounds, max_iter, threshold, initial, strategy, mode,
              seed, tie, _ptr(il), cc_override, labels, C.byref(score), C.byref(iters), C.byref(k),
              trace.ctypes.data_as(C.c_void_p), trace_cap)
    return labels[:n], score.value, iters.value, k.value, trace[: min(iters.value, trace_cap)]


def nmi(a, b):
    """compare_communities(g, a, b) with g.n = len(a); None on negative labels."""
    a, b = i32(a), i32(b)
    out = C.c_double()
    if lib().orc_nmi(len(a), a, b, C.byref(out)) != 0:
        return None
    return out.value


def nmi_file(n, labels, path):
    """compare_communities(g, labels, path); None on a bad file."""
    lab = i32(labels)
    out = C.c_double()
    if lib().orc_nmi_file(int(n), lab, path.encode(), C.byref(out)) != 0:
        return None
    return out.value


def clustering_partition(n, ind, adj, tri):
    """framework.cpp:27-59 create_clustering_partition."""
    out = np.zeros(n, np.int32)
    lib().orc_clustering_partition(int(n), i32(ind), i32(adj), i32(tri), out)
    return out
