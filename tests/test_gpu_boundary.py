"""The drop-in boundary on the GPU: a C program built against include/kaHIP_interface.h alone
(tests/c_abi/caller.c, linked -lkahip) runs kaffpa, kaffpa_balance and kaffpa_balance_NE; their
cuts and the balance rules of the perfectly balanced and node+edge modes are checked with the
oracle, and an unknown mode behaves as ECO."""
import ctypes as C
import math
import subprocess

import numpy as np
import pytest

import kahip_amd as K
import oracle as O
from conftest import manual_graph
from test_boundary_cpu import build_caller

pytestmark = pytest.mark.gpu


def cycle(m):
    x = (2 * np.arange(m + 1)).astype(np.int32)
    a = np.empty(2 * m, np.int32)
    for v in range(m):
        p, q = (v + m - 1) % m, (v + 1) % m
        a[2 * v], a[2 * v + 1] = min(p, q), max(p, q)
    return x, a


def test_c_caller_results(tmp_path):
    exe = build_caller(str(tmp_path))
    print("running the C caller", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = [ln.split() for ln in out.stdout.strip().splitlines()]
    assert [ln[0] for ln in lines] == ["kaffpa", "kaffpa", "kaffpa_balance", "kaffpa_balance_NE"]
    mx, ma, _, _ = manual_graph()
    cx, ca = cycle(1000)
    for i, ln in enumerate(lines):
        k, cut = int(ln[1]), int(ln[3])
        part = np.array([int(t) for t in ln[4:]], np.int32)
        x, a = (mx, ma) if i == 0 else (cx, ca)
        n = len(x) - 1
        assert len(part) == n and part.min() >= 0 and part.max() < k
        assert cut == O.edge_cut(x, a, None, part)
        w = O.block_weights(None, k, part)
        if i == 0:
            assert w.max() <= O.upper_bound(n, k, 0.03)
        elif i == 1:  # imbalance 0: perfectly balanced
            assert w.max() <= math.ceil(n / k)
        elif i == 2:  # perfectly_balance with 3 %
            assert w.max() <= int(1.03 * math.ceil(n / k))
        else:  # node weight 1 + weighted degree 2, bound on the sum
            assert 3 * w.max() <= int(1.03 * math.ceil(3 * n / k))


def _call(fn, x, a, k, imb, seed, mode):
    L = K.lib()
    n = C.c_int(len(x) - 1)
    nparts = C.c_int(k)
    im = C.c_double(imb)
    cut = C.c_int(0)
    part = np.zeros(len(x) - 1, np.int32)
    ip = C.POINTER(C.c_int)
    xa = np.ascontiguousarray(x, np.int32)
    aa = np.ascontiguousarray(a, np.int32)
    getattr(L, fn)(C.byref(n), None, xa.ctypes.data_as(ip), None, aa.ctypes.data_as(ip),
                   C.byref(nparts), C.byref(im), True, seed, mode, C.byref(cut),
                   part.ctypes.data_as(ip))
    return cut.value, part


def test_unknown_mode_is_eco(rgg15):
    x, a, _, _ = rgg15
    c1, p1 = _call("kaffpa", x, a, 4, 0.03, 5, 77)
    c2, p2 = _call("kaffpa", x, a, 4, 0.03, 5, 1)  # ECO
    assert c1 == c2 and np.array_equal(p1, p2)


def test_perfect_balance_equals_mirror_and_postpass(rgg15):
    """kaffpa with imbalance 0: the GPU pipeline under 1.01 ceil(W/k) (equal to its CPU mirror)
    followed by the CPU post-pass to ceil(W/k), bit for bit."""
    from pipeline_mirror import mirror_partition
    from test_boundary_cpu import postpass_seed
    x, a, _, _ = rgg15
    n = len(x) - 1
    for k, seed in ((2, 3), (7, 1)):
        cut, part = _call("kaffpa", x, a, k, 0.0, seed, 3)
        want, info = mirror_partition(x, a, None, None, k, 0.0, seed)
        want, wcut = K.rebalance(x, a, None, None, k, math.ceil(n / k), want,
                                 seed=postpass_seed(seed))
        print(f"k={k}: perfectly balanced cut {cut}", flush=True)
        assert np.array_equal(part, want) and cut == wcut
        assert O.block_weights(None, k, part).max() <= math.ceil(n / k)
