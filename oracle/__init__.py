"""ctypes bindings of the oracle (liboracle.so) and of the reference build (_ref/libkahip_ref.so).

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this package, and only as the checker. The shipped library never imports it.

Arrays are numpy int32 CSR (xadj[n+1], adjncy[2m]); vwgt / adjwgt may be None (unit weights).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkahip_ref.so")
REFERENCE_ROOT = "/root/reference"

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(with_reference: bool | None = None) -> None:
    """Compile liboracle.so and, where the reference sources exist, the reference build."""
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if with_reference is None:
        with_reference = os.path.isdir(REFERENCE_ROOT)
    if with_reference:
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref"])


def _p(a):
    """Pointer or NULL for an optional int32 array."""
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run oracle.build()")
        L = C.CDLL(ORACLE_SO)
        ip = C.POINTER(C.c_int)
        L.orc_lp_cluster_async.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_uint, C.c_int,
                                           C.c_int, _i32p]
        L.orc_remap_first_appearance.argtypes = [C.c_int, _i32p]
        L.orc_contract.argtypes = [C.c_int, _i32p, _i32p, ip, ip, _i32p, C.c_int, _i32p, _i32p,
                                   _i32p, _i32p]
        L.orc_lp_refine_async.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_uint,
                                          C.c_int, C.c_int, _i32p]
        L.orc_lp_refine_async.restype = None
        L.orc_edge_cut.argtypes = [C.c_int, _i32p, _i32p, ip, _i32p]
        L.orc_edge_cut.restype = C.c_int64
        L.orc_block_weights.argtypes = [C.c_int, ip, C.c_int, _i32p, _i64p]
        L.orc_block_weights.restype = None
        L.orc_balance.argtypes = [C.c_int, ip, C.c_int, _i32p]
        L.orc_balance.restype = C.c_double
        L.orc_upper_bound.argtypes = [C.c_int64, C.c_int, C.c_double]
        L.orc_upper_bound.restype = C.c_uint
        L.orc_cluster_bound.argtypes = [C.c_uint, C.c_int]
        L.orc_cluster_bound.restype = C.c_uint
        L.orc_stop_rule_continue.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_project.argtypes = [C.c_int, _i32p, _i32p, _i32p]
        L.orc_project.restype = None
        L.orc_order_nodes_degree.argtypes = [C.c_int, _i32p, _i32p]
        L.orc_order_nodes_degree.restype = None
        L.orc_hash3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_hash3.restype = C.c_uint32
        L.orc_lp_cluster_sync.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_uint, C.c_int,
                                          C.c_int, C.c_uint32, _i32p, ip]
        L.orc_lp_refine_sync.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_uint,
                                         C.c_int, C.c_int, C.c_uint32, _i32p, ip]
        L.orc_lp_refine_sync.restype = None
        u8p = C.POINTER(C.c_uint8)
        L.orc_lp_cluster_sync_ex.argtypes = [C.c_int, _i32p, _i32p, ip, ip, _i32p, u8p, C.c_int,
                                             C.c_int, C.c_uint32, _i32p, ip]
        L.orc_lp_refine_sync_ex.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, _i32p, u8p,
                                            C.c_int, C.c_int, C.c_uint32, _i32p, ip]
        L.orc_lp_refine_sync_ex.restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_SO} missing: build it where the reference exists")
        L = C.CDLL(REF_SO)
        ip = C.POINTER(C.c_int)
        L.ref_lp_cluster.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_uint, C.c_int, C.c_int,
                                     _i32p]
        L.ref_contract.argtypes = [C.c_int, _i32p, _i32p, ip, ip, _i32p, C.c_int, _i32p, _i32p,
                                   _i32p, _i32p]
        L.ref_lp_refine.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_uint, C.c_int,
                                    C.c_int, _i32p]
        L.ref_lp_refine.restype = None
        L.ref_edge_cut.argtypes = [C.c_int, _i32p, _i32p, ip, _i32p]
        L.ref_balance.argtypes = [C.c_int, _i32p, _i32p, ip, C.c_int, _i32p]
        L.ref_balance.restype = C.c_double
        L.ref_kaffpa.argtypes = [C.c_int, ip, _i32p, ip, _i32p, C.c_int, C.c_double, C.c_int,
                                 C.c_int, _i32p, _i32p]
        L.ref_kaffpa.restype = None
        L.ref_kaffpa_fastsocial_phases.argtypes = [C.c_int, ip, _i32p, ip, _i32p, C.c_int,
                                                   C.c_double, C.c_int, _i32p, _i32p, _f64p,
                                                   _i32p]
        L.ref_kaffpa_fastsocial_phases.restype = None
        L.ref_coarsen_one_level.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_double,
                                            C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.ref_time_lp_level.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_uint, C.c_int,
                                         C.c_int, _f64p]
        L.ref_upper_bound.argtypes = [C.c_int, ip, C.c_int, C.c_double]
        L.ref_upper_bound.restype = C.c_uint
        _ref = L
    return _ref


# ------------------------------------------------------------------ oracle wrappers (numpy)
def _n(xadj):
    return len(xadj) - 1


def lp_cluster_async(xadj, adjncy, vwgt, adjwgt, bound, iterations, seed):
    out = np.empty(_n(xadj), np.int32)
    c = lib().orc_lp_cluster_async(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), bound,
                                   iterations, seed, out)
    return out, c


def remap_first_appearance(cluster):
    c = np.array(cluster, dtype=np.int32)
    nc = lib().orc_remap_first_appearance(len(c), c)
    return c, nc


def contract(xadj, adjncy, vwgt, adjwgt, cluster, nc):
    n = _n(xadj)
    m = int(xadj[-1])
    cx = np.zeros(nc + 1, np.int32)
    ca = np.zeros(max(m, 1), np.int32)
    cw = np.zeros(max(m, 1), np.int32)
    cv = np.zeros(max(nc, 1), np.int32)
    lib().orc_contract(n, xadj, adjncy, _p(vwgt), _p(adjwgt),
                       np.ascontiguousarray(cluster, np.int32), nc, cx, ca, cw, cv)
    mc = int(cx[-1])
    return cx, ca[:mc].copy(), cw[:mc].copy(), cv[:nc].copy()


def lp_refine_async(xadj, adjncy, vwgt, adjwgt, k, bound, rounds, seed, part):
    p = np.array(part, dtype=np.int32)
    lib().orc_lp_refine_async(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), k, bound, rounds,
                              seed, p)
    return p


def edge_cut(xadj, adjncy, adjwgt, part):
    return int(lib().orc_edge_cut(_n(xadj), xadj, adjncy, _p(adjwgt),
                                  np.ascontiguousarray(part, np.int32)))


def block_weights(vwgt, k, part):
    part = np.ascontiguousarray(part, np.int32)
    w = np.zeros(k, np.int64)
    lib().orc_block_weights(len(part), _p(vwgt), k, part, w)
    return w


def balance(vwgt, k, part):
    part = np.ascontiguousarray(part, np.int32)
    return float(lib().orc_balance(len(part), _p(vwgt), k, part))


def upper_bound(total_weight, k, imbalance):
    return int(lib().orc_upper_bound(int(total_weight), k, imbalance))


def cluster_bound(upper, factor=18):
    return int(lib().orc_cluster_bound(upper, factor))


def stop_rule_continue(n_finer, n_coarser, k, stop_factor=5000):
    return bool(lib().orc_stop_rule_continue(n_finer, n_coarser, k, stop_factor))


def project(mapping, coarse_part):
    mapping = np.ascontiguousarray(mapping, np.int32)
    out = np.empty(len(mapping), np.int32)
    lib().orc_project(len(mapping), mapping, np.ascontiguousarray(coarse_part, np.int32), out)
    return out


def order_nodes_degree(xadj):
    out = np.empty(_n(xadj), np.int32)
    lib().orc_order_nodes_degree(_n(xadj), xadj, out)
    return out


def hash3(s, a, b):
    return int(lib().orc_hash3(s & 0xFFFFFFFF, a & 0xFFFFFFFF, b & 0xFFFFFFFF))


def lp_cluster_sync(xadj, adjncy, vwgt, adjwgt, bound, rounds, subrounds, seed):
    out = np.empty(_n(xadj), np.int32)
    moves = np.zeros(max(rounds, 1), np.int32)
    c = lib().orc_lp_cluster_sync(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), bound, rounds,
                                  subrounds, seed & 0xFFFFFFFF, out,
                                  moves.ctypes.data_as(C.POINTER(C.c_int)))
    return out, c, moves


def _u8(mask):
    if mask is None:
        return None
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    return m, m.ctypes.data_as(C.POINTER(C.c_uint8))


def lp_cluster_sync_ex(xadj, adjncy, vwgt, adjwgt, lbound, include, rounds, subrounds, seed):
    """R2 clustering with the bound of label (node) c in lbound[c]; nodes with include[v] == 0
    (include may be None) never propose."""
    out = np.empty(_n(xadj), np.int32)
    moves = np.zeros(max(rounds, 1), np.int32)
    inc = _u8(include)
    c = lib().orc_lp_cluster_sync_ex(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt),
                                     np.ascontiguousarray(lbound, dtype=np.int32),
                                     None if inc is None else inc[1], rounds, subrounds,
                                     seed & 0xFFFFFFFF, out,
                                     moves.ctypes.data_as(C.POINTER(C.c_int)))
    return out, c, moves


def lp_refine_sync_ex(xadj, adjncy, vwgt, adjwgt, k, lbound, include, rounds, subrounds, seed,
                      part):
    p = np.array(part, dtype=np.int32)
    moves = np.zeros(max(rounds, 1), np.int32)
    inc = _u8(include)
    lib().orc_lp_refine_sync_ex(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), k,
                                np.ascontiguousarray(lbound, dtype=np.int32),
                                None if inc is None else inc[1], rounds, subrounds,
                                seed & 0xFFFFFFFF, p, moves.ctypes.data_as(C.POINTER(C.c_int)))
    return p, moves


def lp_refine_sync(xadj, adjncy, vwgt, adjwgt, k, bound, rounds, subrounds, seed, part):
    p = np.array(part, dtype=np.int32)
    moves = np.zeros(max(rounds, 1), np.int32)
    lib().orc_lp_refine_sync(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), k, bound, rounds,
                             subrounds, seed & 0xFFFFFFFF, p,
                             moves.ctypes.data_as(C.POINTER(C.c_int)))
    return p, moves


# ------------------------------------------------------------------ reference wrappers
def ref_lp_cluster(xadj, adjncy, vwgt, adjwgt, bound, iterations, seed):
    out = np.empty(_n(xadj), np.int32)
    c = ref().ref_lp_cluster(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), bound, iterations,
                             seed, out)
    return out, c


def ref_contract(xadj, adjncy, vwgt, adjwgt, cluster, nc):
    n = _n(xadj)
    m = int(xadj[-1])
    cx = np.zeros(nc + 1, np.int32)
    ca = np.zeros(max(m, 1), np.int32)
    cw = np.zeros(max(m, 1), np.int32)
    cv = np.zeros(max(nc, 1), np.int32)
    ref().ref_contract(n, xadj, adjncy, _p(vwgt), _p(adjwgt),
                       np.ascontiguousarray(cluster, np.int32), nc, cx, ca, cw, cv)
    mc = int(cx[-1])
    return cx, ca[:mc].copy(), cw[:mc].copy(), cv[:nc].copy()


def ref_lp_refine(xadj, adjncy, vwgt, adjwgt, k, bound, rounds, seed, part):
    p = np.array(part, dtype=np.int32)
    ref().ref_lp_refine(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), k, bound, rounds, seed, p)
    return p


def ref_edge_cut(xadj, adjncy, adjwgt, part):
    return int(ref().ref_edge_cut(_n(xadj), xadj, adjncy, _p(adjwgt),
                                  np.ascontiguousarray(part, np.int32)))


def ref_balance(xadj, adjncy, vwgt, k, part):
    return float(ref().ref_balance(_n(xadj), xadj, adjncy, _p(vwgt), k,
                                   np.ascontiguousarray(part, np.int32)))


def ref_kaffpa(xadj, adjncy, vwgt, adjwgt, k, imbalance, seed, mode):
    n = _n(xadj)
    part = np.zeros(n, np.int32)
    cut = np.zeros(1, np.int32)
    ref().ref_kaffpa(n, _p(vwgt), xadj, _p(adjwgt), adjncy, k, imbalance, seed, mode, cut, part)
    return int(cut[0]), part


def ref_kaffpa_fastsocial_phases(xadj, adjncy, vwgt, adjwgt, k, imbalance, seed):
    n = _n(xadj)
    part = np.zeros(n, np.int32)
    cut = np.zeros(1, np.int32)
    times = np.zeros(4, np.float64)
    levels = np.zeros(64, np.int32)
    ref().ref_kaffpa_fastsocial_phases(n, _p(vwgt), xadj, _p(adjwgt), adjncy, k, imbalance, seed,
                                       cut, part, times, levels)
    return int(cut[0]), part, times, levels


def ref_coarsen_one_level(xadj, adjncy, vwgt, adjwgt, k, imbalance, seed):
    n = _n(xadj)
    m = int(xadj[-1])
    mapping = np.zeros(n, np.int32)
    cx = np.zeros(n + 1, np.int32)
    ca = np.zeros(max(m, 1), np.int32)
    cw = np.zeros(max(m, 1), np.int32)
    cv = np.zeros(max(n, 1), np.int32)
    nc = ref().ref_coarsen_one_level(n, xadj, adjncy, _p(vwgt), _p(adjwgt), k, imbalance, seed,
                                     mapping, cx, ca, cw, cv)
    mc = int(cx[nc])
    return mapping, (cx[:nc + 1].copy(), ca[:mc].copy(), cw[:mc].copy(), cv[:nc].copy())


def ref_time_lp_level(xadj, adjncy, vwgt, adjwgt, bound, iterations, seed):
    """(clusters, lp_seconds, contraction_seconds) of the reference on one graph."""
    t = np.zeros(2, np.float64)
    nc = ref().ref_time_lp_level(_n(xadj), xadj, adjncy, _p(vwgt), _p(adjwgt), bound, iterations,
                                 seed, t)
    return int(nc), float(t[0]), float(t[1])


def ref_upper_bound(n, vwgt, k, imbalance):
    return int(ref().ref_upper_bound(n, _p(vwgt), k, imbalance))
