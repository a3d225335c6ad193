"""ctypes binding of libparcd.so (the C ABI in include/parcd.h).

Used by the tests and the benchmark to drive the HIP path.  There is no
fallback: if the library is missing or cannot run, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libparcd.so")

RULE_MODULARITY, RULE_CPM, RULE_WCC, RULE_CONDUCTANCE = 0, 1, 2, 3
TIE_RANDOM, TIE_LOWEST = 0, 1
NEW_LABEL = -1
PHASES = ["compress", "stats", "metric", "propagate", "merge", "triangles"]

EXPORTS = [
    "pcd_ctx_create", "pcd_ctx_destroy", "pcd_last_error", "pcd_device_count", "pcd_synchronize",
    "pcd_graph_upload", "pcd_graph_info", "pcd_labels_upload", "pcd_labels_download", "pcd_compress_labels",
    "pcd_collect_statistics", "pcd_stats_download", "pcd_metric", "pcd_propagation_step", "pcd_merging_step",
    "pcd_contract", "pcd_count_triangles", "pcd_set_clustering_coef", "pcd_enable_timing", "pcd_phase_time",
    "pcd_enable_counters", "pcd_propagation_counters", "pcd_read_pcd", "pcd_write_pcd", "pcd_read_edge_text",
    "pcd_convert_edges", "pcd_free", "pcd_gen_sbm", "pcd_gen_rmat", "pcd_gen_chung_lu",
    "pcd_kernel_slot_count", "pcd_kernel_slot_name", "pcd_kernel_timer_reset", "pcd_kernel_timer",
    "pcd_bin_count", "pcd_bin_name", "pcd_bin_counters_reset", "pcd_bin_counters", "pcd_bin_size",
    "pcd_singleton_partition", "pcd_set_tuning", "pcd_hub_overflows", "pcd_copy_probe", "pcd_gather_probe",
    "pcd_batch_create", "pcd_batch_destroy", "pcd_batch_footprint", "pcd_batch_add", "pcd_batch_add_many",
    "pcd_batch_labels_many", "pcd_batch_step", "pcd_nmi", "pcd_nmi_file", "pcd_clustering_labels",
    "pcd_kernel_timer_select",
    "pcd_batch_status", "pcd_batch_finished", "pcd_batch_labels", "pcd_batch_stats", "pcd_batch_release",
    "pcd_batch_info", "pcd_batch_map_count", "pcd_batch_map_name", "pcd_batch_set_map", "pcd_batch_autotune",
    "pcd_gen_sbm_batch", "pcd_refine", "pcd_iterate", "pcd_host_alloc", "pcd_host_free", "pcd_clustering_partition",
    "pcd_server_create", "pcd_server_destroy", "pcd_server_submit", "pcd_server_wait", "pcd_server_result",
    "pcd_server_info", "pcd_gen_requests",
    "pcd_graphset_open", "pcd_graphset_close", "pcd_graphset_count", "pcd_graphset_graph", "pcd_graphset_extract",
    "pcd_graphset_writer_open", "pcd_graphset_writer_add", "pcd_graphset_writer_close", "pcd_graphset_pack",
    "pcd_compress_labels_noalloc", "pcd_iterate_many", "pcd_batch_add_many_async", "pcd_sort_fallbacks",
]
BATCH_RUNNING, BATCH_DONE, BATCH_FREE, BATCH_PENDING, BATCH_REJECTED = 0, 1, 2, 3, 4


class PcdError(RuntimeError):
    pass


class Rule(C.Structure):
    _fields_ = [("kind", C.c_int), ("lam", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("batch_arena_bytes", C.c_int64), ("batch_slots", C.c_int),
                ("plain_workers", C.c_int), ("plain_bytes", C.c_int64), ("batch_max_edges", C.c_int64)]


class Job(C.Structure):
    """pcd_job: the ./main options of one graph in a batch."""
    _fields_ = [("rule", Rule), ("frac", C.c_double), ("threshold", C.c_double), ("max_iter", C.c_int),
                ("tie", C.c_int), ("seed", C.c_uint64), ("iter_cap", C.c_int)]


class GraphRef(C.Structure):
    """pcd_graph_ref: one graph of pcd_batch_add_many."""
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("indices", C.c_void_p), ("adj", C.c_void_p),
                ("labels", C.c_void_p)]


def job(kind=0, lam=0.0, frac=0.5, threshold=0.0, max_iter=-1, tie=0, seed=0, iter_cap=0):
    return Job(Rule(int(kind), float(lam)), float(frac), float(threshold), int(max_iter), int(tie), int(seed),
               int(iter_cap))


_lib = None
I32P = C.POINTER(C.c_int32)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PcdError("libparcd.so not built (run __graft_entry__.build() or make -C par-cd_amd)")
    L = C.CDLL(LIB_PATH)
    L.pcd_last_error.restype = C.c_char_p
    vp = C.c_void_p
    L.pcd_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pcd_ctx_destroy.argtypes = [vp]
    L.pcd_device_count.argtypes = [C.POINTER(C.c_int)]
    L.pcd_synchronize.argtypes = [vp]
    L.pcd_graph_upload.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.pcd_graph_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pcd_labels_upload.argtypes = [vp, vp]
    L.pcd_labels_download.argtypes = [vp, vp]
    L.pcd_compress_labels.argtypes = [vp, C.POINTER(C.c_int)]
    L.pcd_compress_labels_noalloc.argtypes = [C.c_int, vp, vp, vp, vp]
    L.pcd_iterate_many.argtypes = [vp, C.POINTER(Job), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    L.pcd_collect_statistics.argtypes = [vp, C.POINTER(C.c_int)]
    L.pcd_stats_download.argtypes = [vp, vp, vp, vp]
    L.pcd_metric.argtypes = [vp, Rule, C.POINTER(C.c_double)]
    L.pcd_propagation_step.argtypes = [vp, Rule, C.c_double, C.c_uint64, C.c_int, C.c_int]
    L.pcd_merging_step.argtypes = [vp, Rule, C.c_int, C.c_uint64, C.c_int, C.c_int]
    L.pcd_contract.argtypes = [vp, C.POINTER(C.c_int64), vp, vp, vp]
    L.pcd_count_triangles.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
    L.pcd_set_clustering_coef.argtypes = [vp, C.c_double]
    L.pcd_enable_timing.argtypes = [vp, C.c_int]
    L.pcd_phase_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.pcd_enable_counters.argtypes = [vp, C.c_int]
    L.pcd_propagation_counters.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pcd_read_pcd.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(I32P),
                               C.POINTER(I32P), C.POINTER(I32P)]
    L.pcd_write_pcd.argtypes = [C.c_char_p, C.c_int, C.c_int, vp, vp, vp]
    L.pcd_graphset_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pcd_graphset_close.argtypes = [vp]
    L.pcd_graphset_count.argtypes = [vp, C.POINTER(C.c_int)]
    L.pcd_graphset_graph.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(I32P),
                                     C.POINTER(I32P), C.POINTER(I32P)]
    L.pcd_graphset_extract.argtypes = [vp, C.c_int, C.c_char_p]
    L.pcd_graphset_writer_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
    L.pcd_graphset_writer_add.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    L.pcd_graphset_writer_close.argtypes = [vp]
    L.pcd_graphset_pack.argtypes = [C.c_char_p, C.c_int, vp]
    L.pcd_read_edge_text.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(I32P), C.POINTER(I32P),
                                     C.POINTER(C.c_int)]
    L.pcd_convert_edges.argtypes = [C.c_int, C.c_int64, vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(I32P), C.POINTER(I32P), C.POINTER(I32P)]
    L.pcd_free.argtypes = [vp]
    L.pcd_free.restype = None
    L.pcd_gen_sbm.argtypes = [C.c_uint64, C.c_int, vp, C.c_double, C.c_double, C.POINTER(C.c_int64),
                              C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.c_int)]
    L.pcd_gen_rmat.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                               C.POINTER(C.c_int64), C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.c_int)]
    L.pcd_gen_chung_lu.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.POINTER(C.c_int64), C.POINTER(I32P), C.POINTER(I32P),
                                   C.POINTER(C.c_int)]
    L.pcd_kernel_slot_name.restype = C.c_char_p
    L.pcd_kernel_slot_name.argtypes = [C.c_int]
    L.pcd_kernel_timer_reset.argtypes = [vp]
    L.pcd_kernel_timer.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.pcd_bin_name.restype = C.c_char_p
    L.pcd_bin_name.argtypes = [C.c_int]
    L.pcd_bin_counters_reset.argtypes = [vp]
    L.pcd_bin_counters.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pcd_bin_size.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pcd_singleton_partition.argtypes = [C.c_int, vp]
    L.pcd_set_tuning.argtypes = [vp, C.c_char_p, C.c_int]
    L.pcd_hub_overflows.argtypes = [vp, C.POINTER(C.c_int64)]
    L.pcd_sort_fallbacks.argtypes = [vp, C.POINTER(C.c_int64)]
    L.pcd_copy_probe.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.pcd_gather_probe.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    ip, dp, lp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    L.pcd_batch_create.argtypes = [C.c_int, C.c_int, C.c_int64, C.POINTER(vp)]
    L.pcd_batch_destroy.argtypes = [vp]
    L.pcd_batch_footprint.argtypes = [C.c_int, C.c_int, lp]
    L.pcd_batch_add.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, C.POINTER(Job), ip]
    L.pcd_batch_add_many.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.pcd_batch_add_many_async.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.pcd_batch_labels_many.argtypes = [vp, C.c_int, vp, vp]
    L.pcd_nmi.argtypes = [C.c_int, vp, vp, dp]
    L.pcd_clustering_labels.argtypes = [vp, ip]
    L.pcd_kernel_timer_select.argtypes = [vp, C.c_uint32]
    L.pcd_nmi_file.argtypes = [C.c_int, vp, C.c_char_p, dp]
    L.pcd_batch_step.argtypes = [vp, C.c_int, ip]
    L.pcd_batch_status.argtypes = [vp, C.c_int, ip, ip, ip, dp]
    L.pcd_batch_finished.argtypes = [vp, vp, C.c_int, ip]
    L.pcd_batch_labels.argtypes = [vp, C.c_int, vp]
    L.pcd_batch_stats.argtypes = [vp, C.c_int, vp, vp, vp]
    L.pcd_batch_release.argtypes = [vp, C.c_int]
    L.pcd_batch_info.argtypes = [vp, lp, ip, ip, lp]
    L.pcd_batch_map_name.argtypes = [C.c_int]
    L.pcd_batch_map_name.restype = C.c_char_p
    L.pcd_batch_set_map.argtypes = [vp, C.c_int]
    L.pcd_batch_autotune.argtypes = [vp, C.c_int, ip, dp]
    L.pcd_host_alloc.argtypes = [C.c_int64, C.POINTER(vp)]
    L.pcd_host_free.argtypes = [vp]
    L.pcd_clustering_partition.argtypes = [C.c_int, vp, vp, vp, vp]
    L.pcd_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
    L.pcd_server_destroy.argtypes = [vp]
    L.pcd_server_submit.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, C.POINTER(Job), C.c_int,
                                    C.POINTER(C.c_int64)]
    L.pcd_server_wait.argtypes = [vp, C.c_int, vp, C.c_int, ip]
    L.pcd_server_result.argtypes = [vp, C.c_int64, vp, dp, ip, ip, dp, ip]
    L.pcd_server_info.argtypes = [vp, ip, ip, ip, lp]
    L.pcd_gen_requests.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_int64] + \
        [C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.POINTER(C.c_int64)),
         C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.POINTER(C.c_double))]
    L.pcd_refine.argtypes = [vp, C.POINTER(Job), C.c_int, C.c_int, dp, ip, ip]
    L.pcd_iterate.argtypes = [vp, C.POINTER(Job), C.c_int, C.c_int, C.c_int, C.c_double, dp, ip, ip]
    L.pcd_gen_sbm_batch.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + \
        [C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.POINTER(C.c_int64)),
         C.POINTER(I32P), C.POINTER(I32P), C.POINTER(C.POINTER(C.c_double))]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise PcdError("pcd error %d: %s" % (rc, lib().pcd_last_error().decode(errors="replace")))


NOALLOC_MAX_VERTICES = 4194304


def compress_labels_noalloc(n, labels_dev, out_dev, num_labels_dev, stream=None):
    """pcd_compress_labels_noalloc on device addresses (ints) allocated with
    the HIP runtime libparcd.so uses: enqueues the compression on `stream` (a
    hipStream_t as an int, None for the null stream) and returns without
    waiting."""
    check(lib().pcd_compress_labels_noalloc(int(n), C.c_void_p(labels_dev), C.c_void_p(out_dev),
                                            C.c_void_p(num_labels_dev), C.c_void_p(stream)))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _take(ptr, count):
    out = np.ctypeslib.as_array(ptr, (max(int(count), 1),))[: int(count)].copy()
    lib().pcd_free(C.cast(ptr, C.c_void_p))
    return out


def device_count():
    c = C.c_int()
    lib().pcd_device_count(C.byref(c))
    return c.value


def rule(kind, lam=0.0):
    return Rule(int(kind), float(lam))


# ------------------------------------------------------------ host helpers
def read_pcd(path):
    n, m = C.c_int(), C.c_int()
    pi, pa, pm = I32P(), I32P(), I32P()
    check(lib().pcd_read_pcd(path.encode(), C.byref(n), C.byref(m), C.byref(pi), C.byref(pa), C.byref(pm)))
    return n.value, m.value, _take(pi, n.value + 1), _take(pa, 2 * m.value), _take(pm, n.value)


def write_pcd(path, n, m, indices, adj, mp):
    check(lib().pcd_write_pcd(path.encode(), n, m, _p(_i32(indices)), _p(_i32(adj)), _p(_i32(mp))))


class GraphSet:
    """A graph set file opened read-only (pcd_graphset_open): many PCD graphs
    in one mapping.  graph(i) returns (n, m, indices, adj, map) as read-only
    numpy views of the mapping; they keep the set open while they live."""

    def __init__(self, path):
        h = C.c_void_p()
        check(lib().pcd_graphset_open(path.encode(), C.byref(h)))
        self._h = h
        c = C.c_int()
        check(lib().pcd_graphset_count(h, C.byref(c)))
        self.count = c.value

    def __len__(self):
        return self.count

    def _view(self, ptr, count):
        arr = (C.c_int32 * max(count, 1)).from_address(C.cast(ptr, C.c_void_p).value)
        arr._owner = self  # the mapping stays while a view refers to it
        out = np.frombuffer(arr, dtype=np.int32, count=count)
        out.flags.writeable = False
        return out

    def graph(self, i):
        if self._h is None:
            raise PcdError("graph set is closed")
        n, m = C.c_int(), C.c_int()
        pi, pa, pm = I32P(), I32P(), I32P()
        check(lib().pcd_graphset_graph(self._h, int(i), C.byref(n), C.byref(m), C.byref(pi), C.byref(pa),
                                       C.byref(pm)))
        n, m = n.value, m.value
        return n, m, self._view(pi, n + 1), self._view(pa, 2 * m), self._view(pm, n)

    def extract(self, i, path):
        check(lib().pcd_graphset_extract(self._h, int(i), path.encode()))

    def close(self):
        """Unmaps the file; only call it when no view of it is in use."""
        if self._h is not None:
            lib().pcd_graphset_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_graph_set(path, graphs):
    """Writes graphs, each (n, m, indices, adj) or (n, m, indices, adj, map),
    as one graph set (pcd_graphset_writer_*)."""
    graphs = list(graphs)
    w = C.c_void_p()
    check(lib().pcd_graphset_writer_open(path.encode(), len(graphs), C.byref(w)))
    try:
        for g in graphs:
            mp = _i32(g[4]) if len(g) > 4 and g[4] is not None else None
            rc = lib().pcd_graphset_writer_add(w, int(g[0]), int(g[1]), _p(_i32(g[2])), _p(_i32(g[3])),
                                              _p(mp) if mp is not None else None)
            check(rc)
    finally:
        rc = lib().pcd_graphset_writer_close(w)
    check(rc)


def pack_graph_set(path, pcd_paths):
    """Packs PCD files into one graph set (pcd_graphset_pack)."""
    arr = (C.c_char_p * max(len(pcd_paths), 1))(*[p.encode() for p in pcd_paths])
    check(lib().pcd_graphset_pack(path.encode(), len(pcd_paths), C.cast(arr, C.c_void_p)))


def read_edge_text(path):
    ne, nn = C.c_int64(), C.c_int()
    pa, pb = I32P(), I32P()
    check(lib().pcd_read_edge_text(path.encode(), C.byref(ne), C.byref(pa), C.byref(pb), C.byref(nn)))
    return _take(pa, ne.value), _take(pb, ne.value), nn.value


def convert_edges(num_nodes, a, b):
    a, b = _i32(a), _i32(b)
    n, m = C.c_int(), C.c_int()
    pi, pa, pm = I32P(), I32P(), I32P()
    check(lib().pcd_convert_edges(num_nodes, a.size, _p(a), _p(b), C.byref(n), C.byref(m), C.byref(pi),
                                  C.byref(pa), C.byref(pm)))
    return n.value, m.value, _take(pi, n.value + 1), _take(pa, 2 * m.value), _take(pm, n.value)


def gen_sbm(seed, block_sizes, p_in, p_out):
    bs = _i32(block_sizes)
    ne, nn = C.c_int64(), C.c_int()
    pa, pb = I32P(), I32P()
    check(lib().pcd_gen_sbm(seed, bs.size, _p(bs), p_in, p_out, C.byref(ne), C.byref(pa), C.byref(pb), C.byref(nn)))
    return _take(pa, ne.value), _take(pb, ne.value), nn.value


def gen_rmat(seed, scale, draws, a=0.57, b=0.19, c=0.19):
    ne, nn = C.c_int64(), C.c_int()
    pa, pb = I32P(), I32P()
    check(lib().pcd_gen_rmat(seed, scale, draws, a, b, c, C.byref(ne), C.byref(pa), C.byref(pb), C.byref(nn)))
    return _take(pa, ne.value), _take(pb, ne.value), nn.value


def gen_chung_lu(seed, n, exponent, dmin, dmax, num_comms, cmin, cmax, boost):
    ne, nn = C.c_int64(), C.c_int()
    pa, pb = I32P(), I32P()
    check(lib().pcd_gen_chung_lu(seed, n, exponent, dmin, dmax, num_comms, cmin, cmax, boost, C.byref(ne),
                                 C.byref(pa), C.byref(pb), C.byref(nn)))
    return _take(pa, ne.value), _take(pb, ne.value), nn.value


def make_graph(kind, seed, **kw):
    """Generator + converter: returns (n, m, indices, adj, map)."""
    if kind == "sbm":
        a, b, nn = gen_sbm(seed, kw["block_sizes"], kw["p_in"], kw["p_out"])
    elif kind == "rmat":
        a, b, nn = gen_rmat(seed, kw["scale"], kw.get("draws", 16 << kw["scale"]), kw.get("a", 0.57),
                            kw.get("b", 0.19), kw.get("c", 0.19))
    elif kind == "chung_lu":
        a, b, nn = gen_chung_lu(seed, kw["n"], kw.get("exponent", 2.3), kw.get("dmin", 4), kw.get("dmax", 5000),
                                kw.get("num_comms", 32), kw.get("cmin", 100), kw.get("cmax", 50000),
                                kw.get("boost", 20.0))
    else:
        raise ValueError(kind)
    return convert_edges(nn, a, b)


def _unpack_many(count, pn, pm, pio, pao, pi, pa, pp, pinned):
    n = _take(pn, count)
    m = _take(pm, count)
    io = _take(pio, count)
    ao = _take(pao, count)
    params = _take(pp, 6 * count).reshape(count, 6)
    tn = int(io[-1]) + int(n[-1]) + 1
    te = int(ao[-1]) + 2 * int(m[-1])
    if pinned:
        ind, adj = pinned_empty(tn), pinned_empty(te)
        ind[:] = np.ctypeslib.as_array(pi, (tn,))
        adj[:] = np.ctypeslib.as_array(pa, (max(te, 1),))[:te]
        lib().pcd_free(C.cast(pi, C.c_void_p))
        lib().pcd_free(C.cast(pa, C.c_void_p))
    else:
        ind = _take(pi, tn)
        adj = _take(pa, te)
    graphs = [(int(n[i]), int(m[i]), ind[io[i]: io[i] + n[i] + 1], adj[ao[i]: ao[i] + 2 * m[i]]) for i in range(count)]
    return graphs, params


def _many_ptrs():
    return (I32P(), I32P(), C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)(), I32P(), I32P(),
            C.POINTER(C.c_double)())


def gen_sbm_batch(seed, count, nmin=50, nmax=20000, bmin=2, bmax=64, pin=(0.05, 0.3), pout=(0.001, 0.01),
                  lam=(0.005, 0.05), pinned=False):
    """The batched-form workload (pcd_gen_sbm_batch): a list of
    (n, m, indices, adj) views into two concatenated arrays (page-locked
    with pinned=True), and the per-graph parameter rows {nominal n, blocks,
    p_in, p_out, kind, lambda}."""
    ps = _many_ptrs()
    check(lib().pcd_gen_sbm_batch(seed, count, nmin, nmax, bmin, bmax, pin[0], pin[1], pout[0], pout[1], lam[0],
                                  lam[1], *[C.byref(x) for x in ps]))
    return _unpack_many(count, *ps, pinned)


def gen_requests(seed, count, m_lo=1000, m_hi=50000000, pinned=False):
    """The serving-form workload (pcd_gen_requests): graphs as gen_sbm_batch
    returns them and rows {family, target m, kind, lambda, initial, 0}."""
    ps = _many_ptrs()
    check(lib().pcd_gen_requests(seed, count, int(m_lo), int(m_hi), *[C.byref(x) for x in ps]))
    return _unpack_many(count, *ps, pinned)


class _Pinned:
    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        try:
            lib().pcd_host_free(self.ptr)
        except Exception:
            pass


def pinned_empty(count, dtype=np.int32):
    """A numpy array in page-locked host memory (pcd_host_alloc); the memory
    lives as long as the array or any view of it."""
    dtype = np.dtype(dtype)
    p = C.c_void_p()
    check(lib().pcd_host_alloc(max(1, int(count)) * dtype.itemsize, C.byref(p)))
    holder = _Pinned(p)
    buf = (C.c_char * (max(1, int(count)) * dtype.itemsize)).from_address(p.value)
    buf._holder = holder
    return np.frombuffer(buf, dtype=dtype, count=int(count))


def singleton_partition(n):
    """framework.cpp:14-24 with the process's glibc rand() state."""
    out = np.zeros(n, np.int32)
    check(lib().pcd_singleton_partition(n, _p(out)))
    return out


def kernel_slots():
    L = lib()
    return [L.pcd_kernel_slot_name(i).decode() for i in range(L.pcd_kernel_slot_count())]


def bin_names():
    L = lib()
    return [L.pcd_bin_name(i).decode() for i in range(L.pcd_bin_count())]


# ------------------------------------------------------------ device context
class Context:
    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        check(lib().pcd_ctx_create(device, C.byref(self._ctx)))
        self.n = 0

    def close(self):
        if self._ctx:
            lib().pcd_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_graph(self, n, m, indices, adj):
        self._ind, self._adj = _i32(indices), _i32(adj)
        check(lib().pcd_graph_upload(self._ctx, n, m, _p(self._ind), _p(self._adj)))
        self.n = n

    def upload_labels(self, labels):
        lab = _i32(labels)
        assert lab.size == self.n
        check(lib().pcd_labels_upload(self._ctx, _p(lab)))

    def labels(self):
        out = np.zeros(self.n, np.int32)
        check(lib().pcd_labels_download(self._ctx, _p(out)))
        return out

    def compress(self):
        k = C.c_int()
        check(lib().pcd_compress_labels(self._ctx, C.byref(k)))
        return k.value

    def statistics(self):
        k = C.c_int()
        check(lib().pcd_collect_statistics(self._ctx, C.byref(k)))
        return k.value

    def stats(self, k):
        s, iv, ev = (np.zeros(max(k, 1), np.int32) for _ in range(3))
        check(lib().pcd_stats_download(self._ctx, _p(s), _p(iv), _p(ev)))
        return s[:k], iv[:k], ev[:k]

    def metric(self, kind, lam=0.0):
        out = C.c_double()
        check(lib().pcd_metric(self._ctx, rule(kind, lam), C.byref(out)))
        return out.value

    def clustering(self):
        """The triangle-based initial partition (-i 1) as the context's labels
        (pcd_clustering_labels); returns the number of communities."""
        nc = C.c_int()
        check(lib().pcd_clustering_labels(self._ctx, C.byref(nc)))
        return nc.value

    def propagate(self, kind, lam=0.0, frac=0.5, seed=0, step=0, tie=TIE_RANDOM):
        check(lib().pcd_propagation_step(self._ctx, rule(kind, lam), frac, seed, step, tie))

    def run(self, n, m, indices, adj, labels, job_):
        """The refinement loop of main.cpp:226-266 with -s 0 on one graph:
        returns (labels, last score, iterations, communities)."""
        self.upload_graph(n, m, indices, adj)
        self.upload_labels(labels)
        kind, lam = job_.rule.kind, job_.rule.lam
        prev, it, k, score = 0.0, 0, 0, 0.0
        cap = job_.iter_cap if job_.iter_cap > 0 else 1000
        while True:
            k = self.compress()
            self.statistics()
            score = self.metric(kind, lam)
            if job_.max_iter != -1:
                done = it >= job_.max_iter
            else:
                done = prev > 0 and score > 0 and score / prev - 1 <= job_.threshold
            done = done or it >= cap
            if not done:
                self.propagate(kind, lam, frac=job_.frac, seed=job_.seed, step=it, tie=job_.tie)
            it += 1
            prev = score
            if done:
                return self.labels(), score, it, k

    def refine(self, job_, strategy=0, rounds=1):
        """pcd_refine on the uploaded graph and labels: (score, iterations, k)."""
        sc, it, k = C.c_double(), C.c_int(), C.c_int()
        check(lib().pcd_refine(self._ctx, C.byref(job_), strategy, rounds, C.byref(sc), C.byref(it), C.byref(k)))
        return sc.value, it.value, k.value

    def iterate_many(self, job_, it, count, prev_score, strategy=0, rounds=1):
        """pcd_iterate_many: passes it .. it+count-1 (device-resident once the
        statistics follow the labels, for every rule and schedule): (passes
        run, score, k, done)."""
        p, sc, k, d = C.c_int(), C.c_double(), C.c_int(), C.c_int()
        check(lib().pcd_iterate_many(self._ctx, C.byref(job_), strategy, rounds, int(it), int(count),
                                     float(prev_score), C.byref(p), C.byref(sc), C.byref(k), C.byref(d)))
        return p.value, sc.value, k.value, bool(d.value)

    def iterate(self, job_, it, prev_score, strategy=0, rounds=1):
        """pcd_iterate: one pass of the refinement loop numbered `it`:
        (score, k, done)."""
        sc, k, d = C.c_double(), C.c_int(), C.c_int()
        check(lib().pcd_iterate(self._ctx, C.byref(job_), strategy, rounds, int(it), float(prev_score), C.byref(sc),
                                C.byref(k), C.byref(d)))
        return sc.value, k.value, bool(d.value)

    def merge(self, kind, lam=0.0, rounds=1, seed=0, step=0, tie=TIE_RANDOM):
        check(lib().pcd_merging_step(self._ctx, rule(kind, lam), rounds, seed, step, tie))

    def contract(self):
        U = C.c_int64()
        check(lib().pcd_contract(self._ctx, C.byref(U), None, None, None))
        u = U.value
        l, k, w = (np.zeros(max(u, 1), np.int32) for _ in range(3))
        check(lib().pcd_contract(self._ctx, C.byref(U), _p(l), _p(k), _p(w)))
        return l[:u], k[:u], w[:u]

    def triangles(self):
        tri, td = np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
        cc = C.c_double()
        check(lib().pcd_count_triangles(self._ctx, _p(tri), _p(td), C.byref(cc)))
        return tri, td, cc.value

    def set_cc(self, cc):
        check(lib().pcd_set_clustering_coef(self._ctx, cc))

    def enable_timing(self, on=True):
        check(lib().pcd_enable_timing(self._ctx, 1 if on else 0))

    def phase_ms(self, phase):
        ms = C.c_double()
        check(lib().pcd_phase_time(self._ctx, PHASES.index(phase) if isinstance(phase, str) else phase,
                                   C.byref(ms)))
        return ms.value

    def enable_counters(self, on=True):
        check(lib().pcd_enable_counters(self._ctx, 1 if on else 0))

    def counters(self):
        a, e, d = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().pcd_propagation_counters(self._ctx, C.byref(a), C.byref(e), C.byref(d)))
        return a.value, e.value, d.value

    def synchronize(self):
        check(lib().pcd_synchronize(self._ctx))

    def set_tuning(self, key, value):
        check(lib().pcd_set_tuning(self._ctx, key.encode(), int(value)))

    def copy_probe(self, nbytes=1 << 30, reps=10):
        g = C.c_double()
        check(lib().pcd_copy_probe(self._ctx, int(nbytes), int(reps), C.byref(g)))
        return g.value

    def gather_probe(self, levels=1, reps=10):
        """Random gathers per second over the graph's adjacency (labels, and
        with levels=2 the labels of those labels)."""
        g = C.c_double()
        check(lib().pcd_gather_probe(self._ctx, int(levels), int(reps), C.byref(g)))
        return g.value

    def hub_overflows(self):
        c = C.c_int64()
        check(lib().pcd_hub_overflows(self._ctx, C.byref(c)))
        return c.value

    def sort_fallbacks(self):
        """One-sweep sorts finished by three-launch passes since the last call."""
        c = C.c_int64()
        check(lib().pcd_sort_fallbacks(self._ctx, C.byref(c)))
        return c.value

    def kernel_timer_select(self, names=None):
        """Time only the named kernel slots (None: all of them)."""
        if names is None:
            mask = 0xFFFFFFFF
        else:
            slots = kernel_slots()
            mask = 0
            for nm in names:
                mask |= 1 << slots.index(nm)
        check(lib().pcd_kernel_timer_select(self._ctx, mask))

    def kernel_timer_reset(self):
        check(lib().pcd_kernel_timer_reset(self._ctx))

    def kernel_timers(self):
        """{slot name: (total ms, launches)} since the last reset."""
        out = {}
        for i, name in enumerate(kernel_slots()):
            t, c = C.c_double(), C.c_int64()
            check(lib().pcd_kernel_timer(self._ctx, i, C.byref(t), C.byref(c)))
            if c.value:
                out[name] = (t.value, c.value)
        return out

    def bin_counters_reset(self):
        check(lib().pcd_bin_counters_reset(self._ctx))

    def bin_counters(self):
        """{bin name: (active, entries, distinct)} since the last reset."""
        out = {}
        for i, name in enumerate(bin_names()):
            a, e, d = C.c_int64(), C.c_int64(), C.c_int64()
            check(lib().pcd_bin_counters(self._ctx, i, C.byref(a), C.byref(e), C.byref(d)))
            out[name] = (a.value, e.value, d.value)
        return out

    def bin_sizes(self):
        out = {}
        for i, name in enumerate(bin_names()):
            v, e = C.c_int64(), C.c_int64()
            check(lib().pcd_bin_size(self._ctx, i, C.byref(v), C.byref(e)))
            out[name] = (v.value, e.value)
        return out


# ------------------------------------------------------------ batches
def batch_map_names():
    L = lib()
    return [L.pcd_batch_map_name(i).decode() for i in range(L.pcd_batch_map_count())]


def batch_footprint(n, m):
    b = C.c_int64()
    check(lib().pcd_batch_footprint(n, m, C.byref(b)))
    return b.value


def nmi(a, b):
    """Normalized mutual information of two labelings (pcd_nmi)."""
    a, b = _i32(a), _i32(b)
    if a.size != b.size:
        raise PcdError("labelings differ in length")
    out = C.c_double()
    check(lib().pcd_nmi(a.size, _p(a), _p(b), C.byref(out)))
    return out.value


def nmi_file(labels, path):
    """Normalized mutual information against a ground-truth community file
    (pcd_nmi_file)."""
    lab = _i32(labels)
    out = C.c_double()
    check(lib().pcd_nmi_file(lab.size, _p(lab), os.fsencode(path), C.byref(out)))
    return out.value


class Batch:
    """Many graphs refined together (pcd_batch_*)."""

    def __init__(self, device=0, max_graphs=64, arena_bytes=1 << 30):
        self._b = C.c_void_p()
        check(lib().pcd_batch_create(device, int(max_graphs), int(arena_bytes), C.byref(self._b)))
        self.sizes = {}
        self._held = []  # host arrays of asynchronously admitted graphs

    def close(self):
        if self._b:
            lib().pcd_batch_destroy(self._b)
            self._b = C.c_void_p()
        self._held = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, n, m, indices, adj, labels, job_):
        """Returns the slot, or None when the batch is full (PCD_E_NOMEM)."""
        indices, adj, labels = _i32(indices), _i32(adj), _i32(labels)
        slot = C.c_int()
        rc = lib().pcd_batch_add(self._b, int(n), int(m), _p(indices), _p(adj), _p(labels), C.byref(job_),
                                 C.byref(slot))
        if rc == -3:
            return None
        check(rc)
        self.sizes[slot.value] = int(n)
        return slot.value

    def add_many(self, graphs, labels, jobs, wait=True):
        """Admits graphs[i] = (n, m, indices, adj) with labels[i] and jobs[i]
        in one call (pcd_batch_add_many).  Returns the slots, None where a
        graph was not admitted for lack of room (PCD_E_NOMEM); the codes and
        slots are kept in self.last_codes and self.last_slots.  Raises
        PcdError when a graph was refused for another reason (the admitted
        ones stay admitted).  wait=False admits through
        pcd_batch_add_many_async: the graphs are BATCH_PENDING until a later
        step() joins them, and the batch holds on to their arrays until it is
        closed."""
        count = len(graphs)
        refs = (GraphRef * max(count, 1))()
        keep = []
        for i, (g, lab) in enumerate(zip(graphs, labels)):
            ind, adj, lab = _i32(g[2]), _i32(g[3]), _i32(lab)
            keep.append((ind, adj, lab))
            refs[i] = GraphRef(int(g[0]), int(g[1]), ind.ctypes.data, adj.ctypes.data, lab.ctypes.data)
        jarr = (Job * max(count, 1))(*jobs)
        slots = np.full(max(count, 1), -1, np.int32)
        codes = np.zeros(max(count, 1), np.int32)
        fn = lib().pcd_batch_add_many if wait else lib().pcd_batch_add_many_async
        rc = fn(self._b, count, C.cast(refs, C.c_void_p), C.cast(jarr, C.c_void_p), _p(slots), _p(codes))
        if not wait:
            self._held.append(keep)
        msg = "pcd error %d: %s" % (rc, lib().pcd_last_error().decode(errors="replace")) if rc else ""
        out = []
        for i in range(count):
            if slots[i] >= 0:
                self.sizes[int(slots[i])] = int(graphs[i][0])
                out.append(int(slots[i]))
            else:
                out.append(None)
        self.last_codes = [int(c) for c in codes[:count]]
        self.last_slots = out
        if any(c not in (0, -3) for c in self.last_codes):
            raise PcdError(msg)
        return out

    def labels_many(self, slots):
        """Labels of finished graphs in one download: a list of arrays (views
        of one buffer), in the order of `slots`."""
        slots = np.ascontiguousarray(slots, np.int32)
        sizes = [self.sizes[int(s_)] for s_ in slots]
        out = np.empty(max(sum(sizes), 1), np.int32)
        check(lib().pcd_batch_labels_many(self._b, len(slots), _p(slots), _p(out)))
        return np.split(out[: sum(sizes)], np.cumsum(sizes)[:-1]) if len(sizes) else []

    def step(self, rounds=1):
        r = C.c_int()
        check(lib().pcd_batch_step(self._b, int(rounds), C.byref(r)))
        return r.value

    def status(self, slot):
        st, it, k, sc = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().pcd_batch_status(self._b, slot, C.byref(st), C.byref(it), C.byref(k), C.byref(sc)))
        return st.value, it.value, k.value, sc.value

    def finished(self):
        cnt = C.c_int()
        check(lib().pcd_batch_finished(self._b, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), np.int32)
        check(lib().pcd_batch_finished(self._b, _p(out), len(out), C.byref(cnt)))
        return [int(x) for x in out[: cnt.value]]

    def labels(self, slot):
        out = np.zeros(self.sizes[slot], np.int32)
        check(lib().pcd_batch_labels(self._b, slot, _p(out)))
        return out

    def stats(self, slot):
        k = self.status(slot)[2]
        s, i, e = (np.zeros(max(k, 1), np.int32) for _ in range(3))
        check(lib().pcd_batch_stats(self._b, slot, _p(s), _p(i), _p(e)))
        return s[:k], i[:k], e[:k]

    def release(self, slot):
        check(lib().pcd_batch_release(self._b, slot))
        self.sizes.pop(slot, None)

    def info(self):
        fb, g, r, rd = C.c_int64(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().pcd_batch_info(self._b, C.byref(fb), C.byref(g), C.byref(r), C.byref(rd)))
        return {"free_bytes": fb.value, "graphs": g.value, "running": r.value, "rounds": rd.value}

    def set_map(self, m):
        check(lib().pcd_batch_set_map(self._b, int(m)))

    def autotune(self, reps=3):
        ch = C.c_int()
        ms = (C.c_double * lib().pcd_batch_map_count())()
        check(lib().pcd_batch_autotune(self._b, int(reps), C.byref(ch), ms))
        return ch.value, list(ms)


def clustering_partition(n, indices, adj, tri):
    """framework.cpp:27-59 from triangle counts (pcd_clustering_partition)."""
    out = np.zeros(n, np.int32)
    check(lib().pcd_clustering_partition(n, _p(_i32(indices)), _p(_i32(adj)), _p(_i32(tri)), _p(out)))
    return out


# ------------------------------------------------------------ serving
class Server:
    """Continuous-batching request loop (pcd_server_*).  Keeps the arrays of
    every request alive until its result is taken."""

    def __init__(self, device=0, batch_arena_bytes=1 << 30, batch_slots=256, plain_workers=2, plain_bytes=8 << 30,
                 batch_max_edges=0):
        self._s = C.c_void_p()
        cfg = ServerConfig(device, int(batch_arena_bytes), int(batch_slots), int(plain_workers), int(plain_bytes),
                           int(batch_max_edges))
        check(lib().pcd_server_create(C.byref(cfg), C.byref(self._s)))
        self._keep = {}

    def close(self):
        if self._s:
            lib().pcd_server_destroy(self._s)
            self._s = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, n, m, indices, adj, labels, job_, initial=0):
        ind, adj_ = _i32(indices), _i32(adj)
        lab = None if labels is None else _i32(labels)
        rid = C.c_int64()
        check(lib().pcd_server_submit(self._s, int(n), int(m), _p(ind), _p(adj_), _p(lab), C.byref(job_),
                                      int(initial), C.byref(rid)))
        self._keep[rid.value] = (ind, adj_, int(n))
        return rid.value

    def wait(self, timeout_ms=100, cap=4096):
        ids = np.zeros(cap, np.int64)
        cnt = C.c_int()
        check(lib().pcd_server_wait(self._s, int(timeout_ms), _p(ids), cap, C.byref(cnt)))
        return [int(x) for x in ids[: cnt.value]]

    def result(self, rid):
        """(labels, score, iterations, k, latency_ms, route)."""
        n = self._keep[rid][2]
        out = np.zeros(n, np.int32)
        sc, it, k, lat, route = C.c_double(), C.c_int(), C.c_int(), C.c_double(), C.c_int()
        rc = lib().pcd_server_result(self._s, rid, _p(out), C.byref(sc), C.byref(it), C.byref(k), C.byref(lat),
                                     C.byref(route))
        self._keep.pop(rid, None)
        check(rc)
        return out, sc.value, it.value, k.value, lat.value, route.value

    def info(self):
        q, rb, rp, rd = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().pcd_server_info(self._s, C.byref(q), C.byref(rb), C.byref(rp), C.byref(rd)))
        return {"queued": q.value, "running_batch": rb.value, "running_plain": rp.value, "rounds": rd.value}
