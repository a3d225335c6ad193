"""pcd_compress_labels_noalloc (the allocation-free compress_labels) through
the C ABI: on the GPU its labels and k equal pcd_compress_labels of a
context (framework.cpp:146-170, bit for bit), on two calls in a row with
the second one smaller, so the shared scratch of the first call cannot leak
into the second.  On the CPU: sizes beyond the maximum and aliased arrays
are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

parcd = pytest.importorskip("parcd")


def _path(n):
    ind = np.zeros(n + 1, np.int32)
    deg = np.full(n, 2, np.int32)
    deg[0] = deg[-1] = 1
    ind[1:] = np.cumsum(deg)
    adj = np.empty(ind[-1], np.int32)
    for v in range(n):
        nb = [u for u in (v - 1, v + 1) if 0 <= u < n]
        adj[ind[v]:ind[v + 1]] = nb
    return n, n - 1, ind, adj


def _labels(n, seed, holes):
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, max(1, n // 7), n).astype(np.int32) * 7 % n
    if holes:
        lab[rng.random(n) < 0.15] = parcd.NEW_LABEL if hasattr(parcd, "NEW_LABEL") else -1
    return lab


def test_noalloc_refuses_oversize_and_aliasing():
    with pytest.raises(parcd.PcdError, match="maximum"):
        parcd.compress_labels_noalloc(parcd.NOALLOC_MAX_VERTICES + 1, 1 << 20, 2 << 20, 3 << 20)
    with pytest.raises(parcd.PcdError, match="same array"):
        parcd.compress_labels_noalloc(10, 1 << 20, 1 << 20, 3 << 20)
    with pytest.raises(parcd.PcdError):
        parcd.compress_labels_noalloc(-1, 1 << 20, 2 << 20, 3 << 20)


class Hip:
    """The HIP runtime libparcd.so is linked with (one runtime per process:
    torch bundles another), driven through ctypes for the caller's device
    buffers and stream."""

    def __init__(self):
        parcd.lib()
        path = None
        for line in open("/proc/self/maps"):
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
        self.h = C.CDLL(path or "libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), max(1, nbytes)))
        return p.value

    def put(self, dev, arr):
        self.ok(self.h.hipMemcpy(dev, arr.ctypes.data, arr.nbytes, 1))

    def get(self, dev, count):
        out = np.empty(count, np.int32)
        self.ok(self.h.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2))
        return out

    def stream(self):
        s = C.c_void_p()
        self.ok(self.h.hipStreamCreate(C.byref(s)))
        return s.value


@pytest.mark.gpu
def test_noalloc_equals_context_compress():
    hip = Hip()
    st = hip.stream()
    cases = ((20000, 1, True), (5000, 2, False), (3000, 3, True))
    nmax = max(c[0] for c in cases)
    d_lab, d_out, d_k = hip.alloc(4 * nmax), hip.alloc(4 * nmax), hip.alloc(4)
    with parcd.Context(0) as ctx:
        for n, seed, holes in cases:  # each call smaller than the one before
            print("noalloc case n=%d seed=%d holes=%s" % (n, seed, holes), flush=True)
            g = _path(n)
            lab = _labels(n, seed, holes)
            ctx.upload_graph(*g)
            ctx.upload_labels(lab)
            k_want = ctx.compress()
            want = ctx.labels()
            hip.put(d_lab, lab)
            hip.put(d_out, np.full(nmax, -7, np.int32))
            hip.put(d_k, np.full(1, -7, np.int32))
            parcd.compress_labels_noalloc(n, d_lab, d_out, d_k, st)
            hip.ok(hip.h.hipStreamSynchronize(st))
            assert int(hip.get(d_k, 1)[0]) == k_want
            out = hip.get(d_out, nmax)
            assert np.array_equal(out[:n], want)
            assert (out[n:] == -7).all()  # nothing written past n
    for p in (d_lab, d_out, d_k):
        hip.ok(hip.h.hipFree(p))
    hip.ok(hip.h.hipStreamDestroy(st))


@pytest.mark.gpu
def test_noalloc_flags_labels_out_of_range():
    hip = Hip()
    n = 1000
    d_lab, d_out, d_k = hip.alloc(4 * n), hip.alloc(4 * n), hip.alloc(4)
    lab = np.arange(n, dtype=np.int32)
    lab[17] = n + 3  # outside [-1, n): skipped by every table, reported as k = -1
    hip.put(d_lab, lab)
    parcd.compress_labels_noalloc(n, d_lab, d_out, d_k, None)  # the null stream
    hip.ok(hip.h.hipStreamSynchronize(None))
    assert int(hip.get(d_k, 1)[0]) == -1
    # a valid labeling afterwards: the flag starts cleared
    lab[17] = 17
    hip.put(d_lab, lab)
    parcd.compress_labels_noalloc(n, d_lab, d_out, d_k, None)
    hip.ok(hip.h.hipStreamSynchronize(None))
    assert int(hip.get(d_k, 1)[0]) == n
    assert np.array_equal(hip.get(d_out, n), np.arange(n, dtype=np.int32))
    for p in (d_lab, d_out, d_k):
        hip.ok(hip.h.hipFree(p))
