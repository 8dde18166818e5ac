"""kahip_quality_noalloc (edge cut and block weights of a device-resident partition, no device
allocation) against the allocating entry points kahip_edge_cut / kahip_block_weights, through the
C ABI, on two calls in a row (the second smaller, so stale scratch contents would show).

Device buffers come from the HIP runtime the library itself links (libamdhip64.so.7 of
/opt/rocm, already loaded with libkahip.so), not from torch, whose wheel bundles a second HIP
runtime that cannot initialise in a process where the first one already runs."""
import ctypes as C

import numpy as np
import pytest

import kahip_amd as K
from conftest import random_graph

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    K.lib()  # loads libamdhip64.so.7 as a dependency of libkahip.so
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _noalloc():
    f = K.lib().kahip_quality_noalloc
    f.argtypes = [C.c_int, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


class Dev:
    """Device buffers of one call, allocated with hipMalloc and freed on exit."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _run(hip, f, x, a, vw, ew, part, k, stream):
    with Dev(hip) as d:
        dx, da, dp = d.put(np.asarray(x, np.int32)), d.put(np.asarray(a, np.int32)), d.put(part)
        dvw = d.put(None if vw is None else np.asarray(vw, np.int32))
        dew = d.put(None if ew is None else np.asarray(ew, np.int32))
        cut0 = np.full(1, 7, np.int64)
        w0 = np.full(k, 7, np.int64)
        dcut, dw = d.put(cut0), d.put(w0)
        rc = f(len(x) - 1, int(x[-1]), dx, da, dew, dvw, dp, k, dcut, dw, stream)
        assert rc == 0, K.lib().kahip_last_error().decode()
        assert hip.hipStreamSynchronize(stream) == 0
        return int(d.get(dcut, cut0)[0]), d.get(dw, w0)


@pytest.fixture(scope="module")
def hip_stream():
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    yield hip, s
    hip.hipStreamDestroy(s)


def test_noalloc_quality_matches_allocating_entry_points(hip_stream):
    hip, stream = hip_stream
    f = _noalloc()
    cases = [
        (random_graph(3000, 0.004, 21, weighted=True), 300, 1, stream),
        (random_graph(700, 0.01, 22, weighted=False), 5, 2, None),  # the null stream
    ]
    for (x, a, vw, ew), k, seed, s in cases:
        print(f"noalloc quality: n={len(x) - 1} m2={int(x[-1])} k={k}", flush=True)
        part = np.random.default_rng(seed).integers(0, k, len(x) - 1).astype(np.int32)
        got_cut, got_w = _run(hip, f, x, a, vw, ew, part, k, s)
        assert got_cut == K.edge_cut(x, a, ew, part)
        assert np.array_equal(got_w, K.block_weights(vw, k, part))


def test_noalloc_quality_large_k_and_bad_ids(hip_stream):
    hip, stream = hip_stream
    f = _noalloc()
    x, a, vw, ew = random_graph(2000, 0.003, 23, weighted=True)
    n = len(x) - 1
    # k above the LDS histogram (global atomics), then a smaller call
    for k in (9000, 3):
        print(f"noalloc quality: n={n} k={k}", flush=True)
        part = np.random.default_rng(k).integers(0, k, n).astype(np.int32)
        got_cut, got_w = _run(hip, f, x, a, vw, ew, part, k, stream)
        assert got_cut == K.edge_cut(x, a, ew, part)
        assert np.array_equal(got_w, K.block_weights(vw, k, part))
    part = np.zeros(n, np.int32)
    part[5] = 3  # outside [0, 2)
    got_cut, _ = _run(hip, f, x, a, vw, ew, part, 2, stream)
    assert got_cut == -1
    # k beyond the scratch array is rejected before any GPU work
    rc = f(n, int(x[-1]), None, None, None, None, None, 65537, None, None, None)
    assert rc != 0 and b"65536" in K.lib().kahip_last_error()
