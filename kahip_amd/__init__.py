"""kahip_amd: KaHIP's label-propagation coarsening / refinement path on MI355X (gfx950).

The compute lives in ``kahip_amd/lib/libkahip.so`` (HIP kernels + C++ host driver behind a C ABI,
see include/kaHIP_interface.h and include/kahip_amd.h). This module is a thin ctypes layer:

* ``kaffpa(vwgt, xadj, adjcwgt, adjncy, nparts, imbalance, suppress_output, seed, mode)``
  mirrors the reference's Python binding (misc/pymodule/kahip.cpp:4-43,
  python/kahip/__init__.py:14) and returns ``(edgecut, blocks)``.
* ``partition(...)``, ``Graph`` (a device-resident CSR) and the fine-grained steps
  (``lp_cluster``, ``contract``, ``lp_refine``, ``edge_cut``, ...) expose the pipeline's parts.

There is no CPU fallback: if the library is missing or a GPU call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from .graphs import (gen_rgg, gen_rmat, read_metis, read_partition, write_metis,  # noqa: F401
                     write_partition, check_graph)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkahip.so")

FAST, ECO, STRONG, FASTSOCIAL, ECOSOCIAL, STRONGSOCIAL = range(6)
# presets whose reference refinement is FM-based: kaffpa adds the CPU pairwise FM on the finest
# level (kahip_params.fm_rounds) after the LP refinement
FM_MODES = (ECO, STRONG, ECOSOCIAL, STRONGSOCIAL)
PRESET_FM_ROUNDS = 3
MAX_LEVELS = 64
RATING_NONE, RATING_EXPANSION_STAR2, RATING_LEARNED = 0, 1, 2
MLP_FEATURES, MLP_HIDDEN, MLP_PARAMS = 8, 64, 4801

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class KahipError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("k", C.c_int), ("imbalance", C.c_double), ("seed", C.c_int),
                ("cluster_rounds", C.c_int), ("refine_rounds", C.c_int), ("subrounds", C.c_int),
                ("cluster_factor", C.c_int), ("stop_factor", C.c_int),
                ("ip_repetitions", C.c_int), ("ip_threads", C.c_int), ("bins", C.c_int * 6),
                ("profile", C.c_int), ("rating", C.c_int), ("rating_levels", C.c_int),
                ("mlp_weights", C.POINTER(C.c_float)), ("reorder", C.c_int),
                ("ensemble", C.c_int), ("ip_tries", C.c_int), ("fm_rounds", C.c_int),
                ("graph_refine", C.c_int), ("group_items", C.c_int), ("wave_small", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("ms_coarsen", C.c_double), ("ms_initial", C.c_double),
                ("ms_refine", C.c_double), ("ms_cut", C.c_double), ("levels", C.c_int),
                ("edges_cluster", C.c_longlong), ("edges_contract", C.c_longlong),
                ("edges_refine", C.c_longlong), ("edges_cut", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("cut", C.c_longlong), ("balance", C.c_double), ("upper_bound", C.c_uint),
                ("cluster_bound", C.c_uint), ("levels", C.c_int),
                ("level_n", C.c_int * (MAX_LEVELS + 1)),
                ("level_m2", C.c_longlong * (MAX_LEVELS + 1)),
                ("ms_coarsen", C.c_double), ("ms_initial", C.c_double),
                ("ms_refine", C.c_double), ("ms_cut", C.c_double),
                ("edges_cluster", C.c_longlong), ("edges_contract", C.c_longlong),
                ("edges_refine", C.c_longlong), ("edges_cut", C.c_longlong),
                ("max_block_weight", C.c_longlong), ("ms_rating", C.c_double),
                ("reordered", C.c_int), ("ms_fm", C.c_double), ("fm_gain", C.c_longlong)]

    def as_dict(self):
        L = self.levels
        return {
            "cut": int(self.cut), "balance": float(self.balance),
            "upper_bound": int(self.upper_bound), "cluster_bound": int(self.cluster_bound),
            "levels": L, "level_n": [int(x) for x in self.level_n[:L + 1]],
            "level_m2": [int(x) for x in self.level_m2[:L + 1]],
            "ms_coarsen": self.ms_coarsen, "ms_initial": self.ms_initial,
            "ms_refine": self.ms_refine, "ms_cut": self.ms_cut,
            "edges_cluster": int(self.edges_cluster), "edges_contract": int(self.edges_contract),
            "edges_refine": int(self.edges_refine), "edges_cut": int(self.edges_cut),
            "max_block_weight": int(self.max_block_weight), "ms_rating": self.ms_rating,
            "reordered": bool(self.reordered), "ms_fm": self.ms_fm, "fm_gain": int(self.fm_gain),
        }


# Every symbol the headers in include/ declare (checked by the tests).
EXPORTS = [
    "kaffpa", "kaffpa_balance", "kaffpa_balance_NE", "kahip_default_params", "kahip_last_error",
    "kaffpa_gpu_ex", "kahip_graph_from_host", "kahip_graph_rmat", "kahip_graph_info",
    "kahip_graph_to_host", "kahip_graph_free", "kahip_graph_rgg", "kahip_partition",
    "kahip_release_workspaces", "kahip_profile_json",
    "kahip_lp_cluster", "kahip_remap", "kahip_contract", "kahip_lp_refine", "kahip_edge_cut",
    "kahip_block_weights", "kahip_project", "kahip_initial_partition", "kahip_edge_features",
    "kahip_graph_edge_features", "kahip_mlp_forward", "kahip_edge_scores", "kahip_mlp_gradient",
    "kahip_mlp_train", "kahip_batch_from_host", "kahip_batch_generate", "kahip_batch_info",
    "kahip_batch_to_host", "kahip_batch_free", "kahip_batch_partition", "kahip_autotune_bins",
    "kahip_fm_refine", "kahip_quality_noalloc", "kahip_rebalance", "kahip_autotune",
]

_lib = None


def build(verbose=False):
    """Compiles libkahip.so for gfx950 in-tree (see build_lib.py)."""
    from . import build_lib
    return build_lib.build(verbose=verbose)


def lib():
    """Loads libkahip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KahipError(f"{LIB_PATH} is missing: run `python -m kahip_amd.build_lib`")
    L = C.CDLL(LIB_PATH)
    ip = C.POINTER(C.c_int)
    G = C.c_void_p
    L.kahip_default_params.argtypes = [C.POINTER(Params)]
    L.kahip_default_params.restype = None
    L.kahip_last_error.restype = C.c_char_p
    L.kaffpa_gpu_ex.argtypes = [C.c_int, ip, _i32p, ip, _i32p, C.POINTER(Params), ip, _i32p,
                                C.POINTER(Stats)]
    L.kaffpa.argtypes = [ip, ip, ip, ip, ip, ip, C.POINTER(C.c_double), C.c_bool, C.c_int,
                         C.c_int, ip, ip]
    L.kaffpa.restype = None
    L.kaffpa_balance.argtypes = [ip, ip, ip, ip, ip, ip, C.POINTER(C.c_double), C.c_bool,
                                 C.c_bool, C.c_int, C.c_int, ip, ip]
    L.kaffpa_balance.restype = None
    L.kaffpa_balance_NE.argtypes = L.kaffpa.argtypes
    L.kaffpa_balance_NE.restype = None
    L.kahip_graph_from_host.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.POINTER(G)]
    L.kahip_graph_rmat.argtypes = [C.c_int, C.c_longlong, C.c_ulonglong, C.c_double,
                                   C.c_double, C.c_double, C.POINTER(G)]
    L.kahip_graph_rgg.argtypes = [C.c_int, C.c_ulonglong, C.c_double, C.POINTER(G)]
    L.kahip_graph_info.argtypes = [G, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip]
    L.kahip_graph_to_host.argtypes = [G, ip, ip, ip, ip]
    L.kahip_graph_free.argtypes = [G]
    L.kahip_graph_free.restype = None
    L.kahip_release_workspaces.argtypes = []
    L.kahip_partition.argtypes = [G, C.POINTER(Params), ip, C.POINTER(Stats)]
    L.kahip_profile_json.argtypes = [C.c_char_p, C.c_int]
    L.kahip_lp_cluster.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_int, C.c_int,
                                   C.c_uint, _i32p, ip, ip]
    L.kahip_remap.argtypes = [C.c_int, _i32p, _i32p, ip]
    L.kahip_contract.argtypes = [C.c_int, _i32p, _i32p, ip, ip, _i32p, C.c_int, _i32p, _i32p,
                                 _i32p, _i32p]
    L.kahip_lp_refine.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_uint, _i32p, ip]
    L.kahip_edge_cut.argtypes = [C.c_int, _i32p, _i32p, ip, _i32p, C.POINTER(C.c_longlong)]
    L.kahip_block_weights.argtypes = [C.c_int, ip, C.c_int, _i32p, _i64p]
    L.kahip_project.argtypes = [C.c_int, _i32p, C.c_int, _i32p, _i32p]
    L.kahip_initial_partition.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_longlong,
                                          C.c_int, C.c_int, C.c_ulonglong, _i32p]
    L.kahip_rebalance.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_longlong,
                                  C.c_ulonglong, _i32p, C.POINTER(C.c_longlong)]
    L.kahip_fm_refine.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_longlong, C.c_int,
                                  C.c_ulonglong, _i32p, C.POINTER(C.c_longlong)]
    L.kahip_edge_features.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int, C.c_uint, _f32p]
    L.kahip_graph_edge_features.argtypes = [G, C.c_int, C.c_uint, _f32p]
    L.kahip_mlp_forward.argtypes = [C.c_longlong, _f32p, _f32p, _f32p]
    L.kahip_edge_scores.argtypes = [C.c_int, _i32p, _i32p, ip, ip, C.c_int,
                                    C.POINTER(C.c_float), C.c_int, C.c_uint, _i32p]
    L.kahip_mlp_gradient.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p,
                                     C.POINTER(C.c_double)]
    L.kahip_mlp_train.argtypes = [C.c_longlong, _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                  _f32p, _f64p]
    L.kahip_batch_from_host.argtypes = [C.c_int, _i64p, _i64p, _i32p, _i32p, ip, ip,
                                        C.POINTER(G)]
    L.kahip_batch_generate.argtypes = [C.c_int, C.c_double, C.c_double, C.c_ulonglong,
                                       C.POINTER(G)]
    L.kahip_batch_info.argtypes = [G, ip, ip, C.POINTER(C.c_longlong), ip]
    L.kahip_batch_to_host.argtypes = [G, ip, ip, ip, ip]
    L.kahip_batch_free.argtypes = [G]
    L.kahip_batch_free.restype = None
    L.kahip_batch_partition.argtypes = [G, C.POINTER(Params), C.c_int, ip,
                                        C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                        C.POINTER(BatchStats)]
    L.kahip_autotune_bins.argtypes = [G, G, C.POINTER(Params), ip, C.POINTER(C.c_double)]
    L.kahip_autotune.argtypes = [G, G, C.POINTER(Params), C.POINTER(Params),
                                 C.POINTER(C.c_double)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise KahipError(lib().kahip_last_error().decode())


def _p(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_params(**kw) -> Params:
    """kahip_default_params() with overrides; ``mlp_weights`` takes an array of MLP_PARAMS
    floats (kept alive by the returned object)."""
    p = Params()
    lib().kahip_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "bins":
            for i, b in enumerate(v):
                p.bins[i] = b
        elif k == "mlp_weights":
            if v is None:
                p.mlp_weights = None
                continue
            w = np.ascontiguousarray(v, dtype=np.float32)
            if w.size != MLP_PARAMS:
                raise KahipError(f"mlp_weights must hold {MLP_PARAMS} floats")
            p._mlp_keep = w
            p.mlp_weights = w.ctypes.data_as(C.POINTER(C.c_float))
        else:
            setattr(p, k, v)
    return p


def ensemble_size(k):
    """LP clusterings per level of the STRONGSOCIAL preset (app/configuration.h:494-509)."""
    return 18 if k <= 3 else 17 if k <= 7 else 15 if k <= 15 else 7 if k <= 31 else 3


# ----------------------------------------------------------------------------- drop-in API
def kaffpa(vwgt, xadj, adjcwgt, adjncy, nparts, imbalance, suppress_output, seed, mode):
    """Same arguments and return value as the reference's ``kahip.kaffpa``: ``(edgecut, blocks)``.

    Raises KahipError on failure (the reference has no error channel)."""
    xadj = _i32(xadj)
    adjncy = _i32(adjncy)
    n = len(xadj) - 1
    vw = None if vwgt is None or len(vwgt) == 0 else _i32(vwgt)
    ew = None if adjcwgt is None or len(adjcwgt) == 0 else _i32(adjcwgt)
    part = np.zeros(max(n, 1), np.int32)
    cut = C.c_int(0)
    p = default_params(k=int(nparts), imbalance=float(imbalance), seed=int(seed),
                       ensemble=ensemble_size(int(nparts)) if mode == STRONGSOCIAL else 1,
                       fm_rounds=PRESET_FM_ROUNDS if mode in FM_MODES else 0)
    _keep = (vw, ew)
    _check(lib().kaffpa_gpu_ex(n, _p(vw), xadj, _p(ew), adjncy, C.byref(p), C.byref(cut), part,
                               None))
    del _keep
    return int(cut.value), part[:n].tolist()


def partition(xadj, adjncy, vwgt=None, adjwgt=None, *, k, imbalance=0.03, seed=0, **params):
    """kaffpa's pipeline with explicit parameters; returns (part ndarray, stats dict)."""
    xadj = _i32(xadj)
    adjncy = _i32(adjncy)
    n = len(xadj) - 1
    vw = None if vwgt is None else _i32(vwgt)
    ew = None if adjwgt is None else _i32(adjwgt)
    part = np.zeros(max(n, 1), np.int32)
    cut = C.c_int(0)
    st = Stats()
    p = default_params(k=k, imbalance=imbalance, seed=seed, **params)
    _check(lib().kaffpa_gpu_ex(n, _p(vw), xadj, _p(ew), adjncy, C.byref(p), C.byref(cut), part,
                               C.byref(st)))
    return part[:n], st.as_dict()


class Graph:
    """A CSR graph resident in GPU memory."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def from_host(cls, xadj, adjncy, vwgt=None, adjwgt=None):
        h = C.c_void_p()
        xadj, adjncy = _i32(xadj), _i32(adjncy)
        vw = None if vwgt is None else _i32(vwgt)
        ew = None if adjwgt is None else _i32(adjwgt)
        _check(lib().kahip_graph_from_host(len(xadj) - 1, xadj, adjncy, _p(vw), _p(ew),
                                           C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, samples, seed, a=0.57, b=0.19, c=0.19):
        h = C.c_void_p()
        _check(lib().kahip_graph_rmat(scale, samples, seed, a, b, c, C.byref(h)))
        return cls(h)

    @classmethod
    def rgg(cls, n, seed, radius=0.0):
        """Random geometric graph generated on the GPU (equal to graphs.gen_rgg(n, seed))."""
        h = C.c_void_p()
        _check(lib().kahip_graph_rgg(int(n), int(seed), float(radius), C.byref(h)))
        return cls(h)

    def info(self):
        n, md = C.c_int(), C.c_int()
        m2, tw = C.c_longlong(), C.c_longlong()
        _check(lib().kahip_graph_info(self._h, C.byref(n), C.byref(m2), C.byref(tw),
                                      C.byref(md)))
        return {"n": n.value, "m2": m2.value, "total_vwgt": tw.value, "max_degree": md.value}

    def to_host(self):
        inf = self.info()
        xadj = np.zeros(inf["n"] + 1, np.int32)
        adj = np.zeros(max(inf["m2"], 1), np.int32)
        ew = np.zeros(max(inf["m2"], 1), np.int32)
        vw = np.zeros(max(inf["n"], 1), np.int32)
        ip = C.POINTER(C.c_int)
        _check(lib().kahip_graph_to_host(self._h, xadj.ctypes.data_as(ip), adj.ctypes.data_as(ip),
                                         vw.ctypes.data_as(ip), ew.ctypes.data_as(ip)))
        return xadj, adj[:inf["m2"]], vw[:inf["n"]], ew[:inf["m2"]]

    def partition(self, k, imbalance=0.03, seed=0, want_part=True, **params):
        n = self.info()["n"]
        part = np.zeros(max(n, 1), np.int32) if want_part else None
        st = Stats()
        p = default_params(k=k, imbalance=imbalance, seed=seed, **params)
        _check(lib().kahip_partition(self._h, C.byref(p),
                                     part.ctypes.data_as(C.POINTER(C.c_int)) if want_part else None,
                                     C.byref(st)))
        return (part[:n] if want_part else None), st.as_dict()

    def free(self):
        if self._h:
            lib().kahip_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def release_workspaces():
    """Frees the device workspaces that the partition entry points keep per device between
    calls (those of engines no call is using)."""
    _check(lib().kahip_release_workspaces())


def profile():
    """Per-launch-name timing summary of the last profiled call on this thread."""
    need = lib().kahip_profile_json(None, 0)
    buf = C.create_string_buffer(need + 16)
    lib().kahip_profile_json(buf, len(buf))
    return json.loads(buf.value.decode() or "{}")


# ----------------------------------------------------------------------------- single steps
def lp_cluster(xadj, adjncy, vwgt, adjwgt, bound, rounds, subrounds, seed):
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    out = np.zeros(max(n, 1), np.int32)
    nc = C.c_int(0)
    moves = np.zeros(max(rounds, 1), np.int32)
    _check(lib().kahip_lp_cluster(n, xadj, adjncy, _p(vwgt), _p(adjwgt), int(bound), rounds,
                                  subrounds, seed & 0xFFFFFFFF, out, C.byref(nc),
                                  moves.ctypes.data_as(C.POINTER(C.c_int))))
    return out[:n], nc.value, moves[:rounds]


def remap(label):
    label = _i32(label)
    n = len(label)
    out = np.zeros(max(n, 1), np.int32)
    nc = C.c_int(0)
    _check(lib().kahip_remap(n, label if n else np.zeros(1, np.int32), out, C.byref(nc)))
    return out[:n], nc.value


def contract(xadj, adjncy, vwgt, adjwgt, cluster, nc):
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    cx = np.zeros(nc + 1, np.int32)
    ca = np.zeros(max(m2, 1), np.int32)
    cw = np.zeros(max(m2, 1), np.int32)
    cv = np.zeros(max(nc, 1), np.int32)
    cl = _i32(cluster) if n else np.zeros(1, np.int32)
    _check(lib().kahip_contract(n, xadj, adjncy, _p(vwgt), _p(adjwgt), cl, nc, cx, ca, cw, cv))
    mc = int(cx[-1])
    return cx, ca[:mc], cw[:mc], cv[:nc]


def lp_refine(xadj, adjncy, vwgt, adjwgt, k, bound, rounds, subrounds, seed, part):
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    p = np.array(part, dtype=np.int32) if n else np.zeros(1, np.int32)
    moves = np.zeros(max(rounds, 1), np.int32)
    _check(lib().kahip_lp_refine(n, xadj, adjncy, _p(vwgt), _p(adjwgt), k, int(bound), rounds,
                                 subrounds, seed & 0xFFFFFFFF, p,
                                 moves.ctypes.data_as(C.POINTER(C.c_int))))
    return p[:n], moves[:rounds]


def edge_cut(xadj, adjncy, adjwgt, part):
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    cut = C.c_longlong(0)
    pa = _i32(part) if n else np.zeros(1, np.int32)
    _check(lib().kahip_edge_cut(n, xadj, adjncy if len(adjncy) else np.zeros(1, np.int32),
                                _p(adjwgt), pa, C.byref(cut)))
    return int(cut.value)


def block_weights(vwgt, k, part):
    part = _i32(part)
    n = len(part)
    w = np.zeros(k, np.int64)
    _check(lib().kahip_block_weights(n, _p(vwgt), k, part if n else np.zeros(1, np.int32), w))
    return w


def project(mapping, coarse_part):
    mapping, coarse_part = _i32(mapping), _i32(coarse_part)
    out = np.zeros(max(len(mapping), 1), np.int32)
    _check(lib().kahip_project(len(mapping), mapping if len(mapping) else np.zeros(1, np.int32),
                               len(coarse_part),
                               coarse_part if len(coarse_part) else np.zeros(1, np.int32), out))
    return out[:len(mapping)]


def initial_partition(xadj, adjncy, vwgt, adjwgt, k, block_bound, repetitions=-1, seed=0,
                      tries=-1):
    """The CPU initial partitioner of the coarsest graph (runs without a GPU); tries = BFS-grown
    starts per bisection (-1: the default, kahip_params.ip_tries)."""
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    out = np.zeros(max(n, 1), np.int32)
    _check(lib().kahip_initial_partition(n, xadj, adjncy if len(adjncy) else np.zeros(1, np.int32),
                                         _p(vwgt), _p(adjwgt), k, int(block_bound), repetitions,
                                         tries, seed, out))
    return out[:n]


def fm_refine(xadj, adjncy, vwgt, adjwgt, k, block_bound, part, rounds=3, seed=0):
    """CPU 2-way FM over the pairs of adjacent blocks (no GPU): (refined part, cut reduction)."""
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    out = np.array(part, np.int32, copy=True)
    if n == 0:
        return out, 0
    gain = C.c_longlong(0)
    _check(lib().kahip_fm_refine(n, xadj, adjncy if len(adjncy) else np.zeros(1, np.int32),
                                 _p(vwgt), _p(adjwgt), int(k), int(block_bound), int(rounds),
                                 int(seed) & 0xFFFFFFFFFFFFFFFF, out, C.byref(gain)))
    return out, int(gain.value)


def rebalance(xadj, adjncy, vwgt, adjwgt, k, bound, part, seed=0):
    """CPU post-pass of the perfectly balanced modes (rebalancing to `bound`, then 2-way FM under
    it); returns (part, cut). No GPU."""
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    p = np.array(part, dtype=np.int32) if n else np.zeros(1, np.int32)
    cut = C.c_longlong(0)
    _check(lib().kahip_rebalance(n, xadj, adjncy if len(adjncy) else np.zeros(1, np.int32),
                                 _p(vwgt), _p(adjwgt), k, int(bound), seed, p, C.byref(cut)))
    return p[:n], int(cut.value)


def fm_seed(seed):
    """Seed of the finest-level pairwise FM of a partition with this seed."""
    return (((int(seed) & 0xFFFFFFFF) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF) ^ 0xF3F3F3F3


# ----------------------------------------------------------------------------- edge ratings
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def edge_features(xadj, adjncy, vwgt, adjwgt, bound, seed):
    """Per-directed-edge features (xadj[n] x 8 float32) of the learned rating."""
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    out = np.zeros((max(m2, 1), MLP_FEATURES), np.float32)
    _check(lib().kahip_edge_features(n, xadj, adjncy if m2 else np.zeros(1, np.int32), _p(vwgt),
                                     _p(adjwgt), int(bound), seed & 0xFFFFFFFF, out))
    return out[:m2]


def graph_edge_features(graph, bound, seed):
    m2 = graph.info()["m2"]
    out = np.zeros((max(m2, 1), MLP_FEATURES), np.float32)
    _check(lib().kahip_graph_edge_features(graph._h, int(bound), seed & 0xFFFFFFFF, out))
    return out[:m2]


def mlp_forward(features, params):
    """MLP logits of feature rows (BF16 on MFMA, FP32 accumulate)."""
    f = _f32(features).reshape(-1, MLP_FEATURES)
    m = f.shape[0]
    out = np.zeros(max(m, 1), np.float32)
    _check(lib().kahip_mlp_forward(m, f if m else np.zeros((1, 8), np.float32), _f32(params),
                                   out))
    return out[:m]


def edge_scores(xadj, adjncy, vwgt, adjwgt, rating, params=None, bound=0, seed=0):
    xadj, adjncy = _i32(xadj), _i32(adjncy)
    n = len(xadj) - 1
    m2 = int(xadj[-1])
    out = np.zeros(max(m2, 1), np.int32)
    w = None if params is None else _f32(params)
    _check(lib().kahip_edge_scores(n, xadj, adjncy if m2 else np.zeros(1, np.int32), _p(vwgt),
                                   _p(adjwgt), rating,
                                   None if w is None else w.ctypes.data_as(C.POINTER(C.c_float)),
                                   int(bound), seed & 0xFFFFFFFF, out))
    return out[:m2]


def mlp_gradient(features, labels, params):
    """(mean BCE loss, mean gradient) of the MLP on the GPU."""
    f = _f32(features).reshape(-1, MLP_FEATURES)
    y = _f32(labels)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_double(0)
    _check(lib().kahip_mlp_gradient(f.shape[0], f, y, _f32(params), g, C.byref(loss)))
    return loss.value, g


def mlp_train(features, labels, params, batch=4096, epochs=1, lr=1e-3):
    """Adam on the GPU over consecutive mini-batches; returns (params, loss per epoch)."""
    f = _f32(features).reshape(-1, MLP_FEATURES)
    y = _f32(labels)
    w = _f32(params).copy()
    losses = np.zeros(max(epochs, 1), np.float64)
    _check(lib().kahip_mlp_train(f.shape[0], f, y, batch, epochs, lr, w, losses))
    return w, losses[:epochs]


def load_rating_weights(path=None):
    """The committed trained MLP weights (float32[MLP_PARAMS])."""
    path = path or os.path.join(HERE, "data", "rating_mlp.npy")
    w = np.load(path, allow_pickle=False).astype(np.float32)
    if w.size != MLP_PARAMS:
        raise KahipError(f"{path}: expected {MLP_PARAMS} floats")
    return w


# ----------------------------------------------------------------------------- batched mode
class Batch:
    """G independent graphs resident in GPU memory as one disjoint union."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_graphs(cls, graphs):
        """graphs: list of (xadj, adjncy, vwgt, adjwgt) with local ids (weights may be None)."""
        G = len(graphs)
        node_off = np.zeros(G + 1, np.int64)
        edge_off = np.zeros(G + 1, np.int64)
        for i, (x, a, _, _) in enumerate(graphs):
            node_off[i + 1] = node_off[i] + len(x) - 1
            edge_off[i + 1] = edge_off[i] + int(x[-1])
        xs = np.concatenate([_i32(x) for x, _, _, _ in graphs])
        adj = np.concatenate([_i32(a) for _, a, _, _ in graphs] + [np.zeros(1, np.int32)])
        any_vw = any(v is not None for _, _, v, _ in graphs)
        any_ew = any(w is not None for _, _, _, w in graphs)
        vw = np.concatenate([_i32(v) if v is not None else np.ones(len(x) - 1, np.int32)
                             for x, _, v, _ in graphs]) if any_vw else None
        ew = np.concatenate([_i32(w) if w is not None else np.ones(int(x[-1]), np.int32)
                             for x, _, _, w in graphs]) if any_ew else None
        h = C.c_void_p()
        _check(lib().kahip_batch_from_host(G, node_off, edge_off, xs, adj, _p(vw), _p(ew),
                                           C.byref(h)))
        return cls(h)

    @classmethod
    def generate(cls, num_graphs, log2_min=10.0, log2_max=16.0, seed=0):
        h = C.c_void_p()
        _check(lib().kahip_batch_generate(num_graphs, log2_min, log2_max, seed, C.byref(h)))
        return cls(h)

    def info(self):
        G, n = C.c_int(), C.c_int()
        m2 = C.c_longlong()
        _check(lib().kahip_batch_info(self._h, C.byref(G), C.byref(n), C.byref(m2), None))
        off = np.zeros(G.value + 1, np.int32)
        _check(lib().kahip_batch_info(self._h, None, None, None,
                                      off.ctypes.data_as(C.POINTER(C.c_int))))
        return {"graphs": G.value, "n": n.value, "m2": m2.value, "node_off": off}

    def to_host(self):
        inf = self.info()
        ip = C.POINTER(C.c_int)
        x = np.zeros(inf["n"] + 1, np.int32)
        a = np.zeros(max(inf["m2"], 1), np.int32)
        w = np.zeros(max(inf["m2"], 1), np.int32)
        v = np.zeros(max(inf["n"], 1), np.int32)
        _check(lib().kahip_batch_to_host(self._h, x.ctypes.data_as(ip), a.ctypes.data_as(ip),
                                         v.ctypes.data_as(ip), w.ctypes.data_as(ip)))
        return x, a[:inf["m2"]], v[:inf["n"]], w[:inf["m2"]]

    def partition(self, k, imbalance=0.03, seed=0, use_graph=False, **params):
        """Returns (local part ndarray, cuts[G], max block weights[G], stats dict)."""
        inf = self.info()
        G = inf["graphs"]
        part = np.zeros(max(inf["n"], 1), np.int32)
        cuts = np.zeros(G, np.int64)
        mbw = np.zeros(G, np.int64)
        st = BatchStats()
        p = default_params(k=k, imbalance=imbalance, seed=seed, **params)
        LL = C.POINTER(C.c_longlong)
        _check(lib().kahip_batch_partition(self._h, C.byref(p), 1 if use_graph else 0,
                                           part.ctypes.data_as(C.POINTER(C.c_int)),
                                           cuts.ctypes.data_as(LL), mbw.ctypes.data_as(LL),
                                           C.byref(st)))
        return part[:inf["n"]], cuts, mbw, st.as_dict()

    def free(self):
        if self._h:
            lib().kahip_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def autotune_bins(target, k=8, **params):
    """Autotuned LP bins for a Graph or Batch: (bins list, ms per candidate)."""
    bins = np.zeros(6, np.int32)
    ms = np.zeros(6, np.float64)
    p = default_params(k=k, **params)
    g = target._h if isinstance(target, Graph) else None
    b = target._h if isinstance(target, Batch) else None
    _check(lib().kahip_autotune_bins(g, b, C.byref(p), bins.ctypes.data_as(C.POINTER(C.c_int)),
                                     ms.ctypes.data_as(C.POINTER(C.c_double))))
    return bins.tolist(), ms.tolist()


AUTOTUNE_EVALS = 12


def autotune(target, k=8, **params):
    """Autotuned LP degree bins and tile shapes for a Graph or Batch: (dict with bins,
    group_items, wave_small; ms per evaluation). Pass the dict's entries to partition()."""
    ms = np.zeros(AUTOTUNE_EVALS, np.float64)
    p = default_params(k=k, **params)
    out = Params()
    g = target._h if isinstance(target, Graph) else None
    b = target._h if isinstance(target, Batch) else None
    _check(lib().kahip_autotune(g, b, C.byref(p), C.byref(out),
                                ms.ctypes.data_as(C.POINTER(C.c_double))))
    return ({"bins": list(out.bins), "group_items": out.group_items,
             "wave_small": out.wave_small}, ms.tolist())
