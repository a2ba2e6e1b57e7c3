"""Python binding of the MI355X kNN-graph builder (ctypes over include/knng.h).

The product is the C-ABI library ``lib/libknng.so`` (HIP kernels) and the CLI
``bin/knng_main``; this module is the host-side mirror used by bench.py and
the tests.  It never falls back to a CPU path: if the HIP library is missing
or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
CLI = os.path.join(PKG, "bin", "knng_main")

# every symbol include/knng.h and include/knng_h5.h declare
KNNG_SYMBOLS = [
    "knng_default_params", "knng_create", "knng_destroy", "knng_last_error",
    "knng_ingest_chunk", "knng_ingest_device", "knng_attach_fp32_device",
    "knng_host_alloc", "knng_host_free",
    "knng_build", "knng_build_forest", "knng_build_init", "knng_join_iteration",
    "knng_export", "knng_export_device", "knng_get_stats", "knng_synchronize",
    "knng_copy_quantized", "knng_copy_forest", "knng_copy_graph", "knng_copy_cells",
    "knng_load_graph",
    "knng_quantize_device", "knng_score_pairs_device", "knng_join_iteration_noalloc",
    "knng_server_default_params", "knng_server_create", "knng_server_destroy",
    "knng_server_last_error", "knng_server_submit", "knng_server_poll", "knng_server_wait",
    "knng_server_size", "knng_server_search", "knng_server_wait_info", "knng_server_info",
    "knng_server_mask_error",
    "knng_build_batch",
]
H5_SYMBOLS = [
    "knng_h5_open", "knng_h5_shape_of", "knng_h5_read_rows", "knng_h5_close",
    "knng_h5_shape", "knng_h5_read_f32", "knng_h5_read_i32", "knng_h5_write_f32",
    "knng_h5_write_i32", "knng_h5_create_f32", "knng_h5_write_rows", "knng_h5_finish",
    "knng_h5_write_result", "knng_h5_read_attr_str", "knng_h5_read_attr_f32",
]


STOP_REFERENCE, STOP_INSERTED, STOP_NONE = 0, 1, 2


class KnngError(RuntimeError):
    pass


class Params(C.Structure):
    """knng_params; defaults are src/main.cpp:19-29."""
    _fields_ = [
        ("dim", C.c_int32), ("k_eval", C.c_int32), ("k_build", C.c_int32),
        ("num_trees", C.c_int32), ("rerank_mult", C.c_int32), ("leaf_size", C.c_int32),
        ("max_iters", C.c_int32), ("early_stop_rate", C.c_float), ("rate_diff_tol", C.c_float),
        ("seed", C.c_uint64), ("keep_fp32", C.c_int32), ("rerank_width", C.c_int32),
        ("stop_rule", C.c_int32), ("insert_stop_rate", C.c_float),
        ("incremental", C.c_int32), ("cells", C.c_int32), ("kmeans_iters", C.c_int32),
        ("cell_partitions", C.c_int32), ("cell_oversample", C.c_int32),
        ("cell_sample", C.c_int32), ("cell_merge", C.c_float), ("cell_full_iters", C.c_int32),
        ("cell_prune", C.c_float), ("reverse_width", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("dim", C.c_int32), ("iterations", C.c_int32),
        ("forest_levels", C.c_int32), ("pad0", C.c_int32), ("num_leaves", C.c_int64),
        ("updates", C.c_int64 * 32), ("rates", C.c_float * 32),
        ("ms_ingest", C.c_float), ("ms_forest", C.c_float), ("ms_init", C.c_float),
        ("ms_join", C.c_float), ("ms_join_iter", C.c_float * 32), ("ms_export", C.c_float),
        ("ms_build", C.c_float), ("ms_kmeans", C.c_float), ("cells", C.c_int32),
        ("join_bytes", C.c_double * 32),
        ("join_candidates", C.c_int64 * 32), ("join_rescored", C.c_int64 * 32),
        ("join_inserted", C.c_int64 * 32), ("insert_rates", C.c_float * 32),
        ("join_variant", C.c_int32), ("ms_tune", C.c_float), ("join_passed_on", C.c_int64 * 32),
        ("join_leafmates", C.c_int64 * 32), ("init_tile_ops", C.c_double), ("assign_ops", C.c_double),
        ("leaf_variant", C.c_int32), ("ms_tune_leaf", C.c_float),
    ]

    def as_dict(self) -> dict:
        it = self.iterations
        return {
            "n": self.n, "dim": self.dim, "iterations": it, "forest_levels": self.forest_levels,
            "num_leaves": self.num_leaves, "updates": list(self.updates[:it]),
            "rates": [float(r) for r in self.rates[:it]],
            "ms_ingest": self.ms_ingest, "ms_forest": self.ms_forest, "ms_init": self.ms_init,
            "ms_join": self.ms_join, "ms_join_iter": [float(v) for v in self.ms_join_iter[:it]],
            "ms_export": self.ms_export, "ms_build": self.ms_build,
            "ms_kmeans": self.ms_kmeans, "cells": self.cells,
            "join_bytes": list(self.join_bytes[:it]),
            "join_candidates": list(self.join_candidates[:it]),
            "join_rescored": list(self.join_rescored[:it]),
            "join_inserted": list(self.join_inserted[:it]),
            "insert_rates": [float(r) for r in self.insert_rates[:it]],
            "join_variant": self.join_variant, "ms_tune": self.ms_tune,
            "join_passed_on": list(self.join_passed_on[:it]),
            "join_leafmates": list(self.join_leafmates[:it]),
            "init_tile_ops": self.init_tile_ops, "assign_ops": self.assign_ops,
            "leaf_variant": self.leaf_variant, "ms_tune_leaf": self.ms_tune_leaf,
        }


_lib = None
_h5 = None


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise KnngError(f"{path} is not built: run `python sisap-2026-task1_amd/build.py`")
    return C.CDLL(path)


def lib() -> C.CDLL:
    """The HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        L = _load("libknng.so")
        P, V = C.POINTER, C.c_void_p
        i32, i64 = C.c_int32, C.c_int64
        sig = {
            "knng_default_params": ([P(Params)], i32),
            "knng_create": ([P(V), i32, P(Params), i64], i32),
            "knng_destroy": ([V], i32),
            "knng_last_error": ([V], C.c_char_p),
            "knng_ingest_chunk": ([V, V, i64, i64], i32),
            "knng_ingest_device": ([V, V, i64, i64], i32),
            "knng_host_alloc": ([V, P(V), i64], i32),
            "knng_host_free": ([V, V], i32),
            "knng_attach_fp32_device": ([V, V], i32),
            "knng_build": ([V], i32),
            "knng_build_forest": ([V], i32),
            "knng_build_init": ([V], i32),
            "knng_join_iteration": ([V, P(i64)], i32),
            "knng_export": ([V, V, V, i32, i32], i32),
            "knng_export_device": ([V, V, V, i32, i32], i32),
            "knng_get_stats": ([V, P(Stats)], i32),
            "knng_synchronize": ([V], i32),
            "knng_copy_quantized": ([V, V, V], i32),
            "knng_copy_forest": ([V, V, V, i64, P(i64)], i32),
            "knng_copy_graph": ([V, V, V], i32),
            "knng_copy_cells": ([V, i32, P(i32), V, V, V], i32),
            "knng_load_graph": ([V, V, V], i32),
            "knng_quantize_device": ([V, i64, i32, V, V], i32),
            "knng_score_pairs_device": ([V, V, i32, V, V, i64, V, V], i32),
            "knng_join_iteration_noalloc": ([V, V, i32, i64, i32, i32, V, V, V, V, V, V, V], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _lib = L
    return _lib


def h5() -> C.CDLL:
    """The HDF5 boundary library."""
    global _h5
    if _h5 is None:
        L = _load("libknng_h5.so")
        V, i32, i64, s = C.c_void_p, C.c_int32, C.c_int64, C.c_char_p
        sig = {
            "knng_h5_open": ([s, s, C.POINTER(V)], i32),
            "knng_h5_shape_of": ([V, C.POINTER(i64), C.POINTER(i64)], i32),
            "knng_h5_read_rows": ([V, i64, i64, i64, V], i32),
            "knng_h5_close": ([V], i32),
            "knng_h5_shape": ([s, s, C.POINTER(i64), C.POINTER(i64)], i32),
            "knng_h5_read_f32": ([s, s, i64, i64, i64, V], i32),
            "knng_h5_read_i32": ([s, s, i64, i64, i64, V], i32),
            "knng_h5_write_f32": ([s, s, V, i64, i64, i64, i32], i32),
            "knng_h5_write_i32": ([s, s, V, i64, i64, i64, i32], i32),
            "knng_h5_create_f32": ([s, s, i64, i64, i64, C.POINTER(V)], i32),
            "knng_h5_write_rows": ([V, i64, i64, V], i32),
            "knng_h5_finish": ([V], i32),
            "knng_h5_write_result": ([s, V, V, i64, i32, s, s, C.c_float, C.c_float, s, s, s], i32),
            "knng_h5_read_attr_str": ([s, s, C.c_char_p, i64], i32),
            "knng_h5_read_attr_f32": ([s, s, C.POINTER(C.c_float)], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _h5 = L
    return _h5


def default_params(**over) -> Params:
    p = Params()
    lib().knng_default_params(C.byref(p))
    for k, v in over.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


class Builder:
    """One kNN-graph build on one GPU (a knng_ctx)."""

    def __init__(self, n: int, params: Params | None = None, device: int = 0, **over):
        self.L = lib()
        self.p = params if params is not None else default_params(**over)
        self.n = int(n)
        self.ctx = C.c_void_p()
        rc = self.L.knng_create(C.byref(self.ctx), device, C.byref(self.p), self.n)
        if rc:
            raise KnngError(f"knng_create failed ({rc})")

    # -- plumbing
    def _ck(self, rc: int, what: str) -> None:
        if rc:
            msg = self.L.knng_last_error(self.ctx).decode(errors="replace")
            raise KnngError(f"{what} failed ({rc}): {msg}")

    def close(self) -> None:
        if self.ctx:
            self.L.knng_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ingest
    def ingest(self, rows: np.ndarray, row0: int = 0) -> None:
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        assert rows.ndim == 2 and rows.shape[1] == self.p.dim
        self._ck(self.L.knng_ingest_chunk(self.ctx, _ptr(rows), row0, rows.shape[0]), "ingest")

    def ingest_device(self, ptr: int, row0: int, nrows: int) -> None:
        self._ck(self.L.knng_ingest_device(self.ctx, C.c_void_p(ptr), row0, nrows), "ingest_device")

    def attach_fp32(self, ptr: int) -> None:
        self._ck(self.L.knng_attach_fp32_device(self.ctx, C.c_void_p(ptr)), "attach")

    # -- build
    def build(self) -> None:
        self._ck(self.L.knng_build(self.ctx), "build")

    def build_forest(self) -> None:
        self._ck(self.L.knng_build_forest(self.ctx), "build_forest")

    def build_init(self) -> None:
        self._ck(self.L.knng_build_init(self.ctx), "build_init")

    def join_iteration(self) -> int:
        u = C.c_int64()
        self._ck(self.L.knng_join_iteration(self.ctx, C.byref(u)), "join_iteration")
        return int(u.value)

    def synchronize(self) -> None:
        self._ck(self.L.knng_synchronize(self.ctx), "synchronize")

    # -- export
    def export(self, k: int | None = None, rerank: bool = False):
        k = self.p.k_eval if k is None else k
        knns = np.empty((self.n, k), np.int32)
        dists = np.empty((self.n, k), np.float32)
        self._ck(self.L.knng_export(self.ctx, _ptr(knns), _ptr(dists), k, int(rerank)), "export")
        return knns, dists

    def export_device(self, knns_ptr: int, dists_ptr: int, k: int, rerank: bool) -> None:
        self._ck(self.L.knng_export_device(self.ctx, C.c_void_p(knns_ptr), C.c_void_p(dists_ptr),
                                           k, int(rerank)), "export_device")

    def stats(self) -> dict:
        s = Stats()
        self._ck(self.L.knng_get_stats(self.ctx, C.byref(s)), "stats")
        return s.as_dict()

    # -- introspection
    def quantized(self):
        q = np.empty((self.n, self.p.dim), np.int8)
        sk = np.empty((self.n, self.p.dim // 64), np.uint64)
        self._ck(self.L.knng_copy_quantized(self.ctx, _ptr(q), _ptr(sk)), "copy_quantized")
        return q, sk

    def forest(self):
        T = self.p.num_trees
        perm = np.empty((T, self.n), np.int32)
        nl = C.c_int64()
        self._ck(self.L.knng_copy_forest(self.ctx, _ptr(perm), None, 0, C.byref(nl)), "copy_forest")
        leaves = np.empty((max(nl.value, 1), 3), np.int64)
        self._ck(self.L.knng_copy_forest(self.ctx, None, _ptr(leaves), leaves.shape[0], C.byref(nl)),
                 "copy_forest")
        return perm, leaves[: nl.value]

    def cells(self, part: int = 0):
        """(z [C, d] int8, w [C] f32, assign [n] int32) of a cell partition, or None."""
        cnt = C.c_int32()
        self._ck(self.L.knng_copy_cells(self.ctx, part, C.byref(cnt), None, None, None), "copy_cells")
        if cnt.value == 0:
            return None
        z = np.empty((cnt.value, self.p.dim), np.int8)
        w = np.empty(cnt.value, np.float32)
        a = np.empty(self.n, np.int32)
        self._ck(self.L.knng_copy_cells(self.ctx, part, C.byref(cnt), _ptr(a), _ptr(z), _ptr(w)),
                 "copy_cells")
        return z, w, a

    def graph(self):
        K = self.p.k_build
        ids = np.empty((self.n, K), np.int32)
        sc = np.empty((self.n, K), np.int32)
        self._ck(self.L.knng_copy_graph(self.ctx, _ptr(ids), _ptr(sc)), "copy_graph")
        return ids, sc

    def load_graph(self, ids: np.ndarray, scores: np.ndarray) -> None:
        ids = np.ascontiguousarray(ids, np.int32)
        scores = np.ascontiguousarray(scores, np.int32)
        self._ck(self.L.knng_load_graph(self.ctx, _ptr(ids), _ptr(scores)), "load_graph")


def quantize_device(x_ptr: int, n: int, d: int, q_ptr: int, sk_ptr: int) -> None:
    rc = lib().knng_quantize_device(C.c_void_p(x_ptr), n, d, C.c_void_p(q_ptr), C.c_void_p(sk_ptr))
    if rc:
        raise KnngError(f"knng_quantize_device failed ({rc})")


def score_pairs_device(q_ptr, sk_ptr, d, a_ptr, b_ptr, npairs, dot_ptr, sim_ptr) -> None:
    rc = lib().knng_score_pairs_device(C.c_void_p(q_ptr), C.c_void_p(sk_ptr), d, C.c_void_p(a_ptr),
                                       C.c_void_p(b_ptr), npairs, C.c_void_p(dot_ptr),
                                       C.c_void_p(sim_ptr))
    if rc:
        raise KnngError(f"knng_score_pairs_device failed ({rc})")


def join_iteration_noalloc(q_ptr, sk_ptr, d, n, k_build, rerank_mult, ids_in_ptr, sc_in_ptr,
                           ids_out_ptr, sc_out_ptr, counts_ptr, order_ptr=None, stream=None) -> None:
    """knng_join_iteration_noalloc on device buffers (stream-ordered, no sync)."""
    rc = lib().knng_join_iteration_noalloc(
        C.c_void_p(q_ptr), C.c_void_p(sk_ptr), d, n, k_build, rerank_mult, C.c_void_p(ids_in_ptr),
        C.c_void_p(sc_in_ptr), C.c_void_p(ids_out_ptr), C.c_void_p(sc_out_ptr),
        C.c_void_p(order_ptr) if order_ptr else None, C.c_void_p(counts_ptr),
        C.c_void_p(stream) if stream else None)
    if rc:
        raise KnngError(f"knng_join_iteration_noalloc failed ({rc})")


# -- HDF5 helpers -----------------------------------------------------------

def h5_shape(path: str, dset: str = "/train"):
    r, c = C.c_int64(), C.c_int64()
    rc = h5().knng_h5_shape(path.encode(), dset.encode(), C.byref(r), C.byref(c))
    if rc:
        raise KnngError(f"h5 shape of {path}:{dset} failed ({rc})")
    return int(r.value), int(c.value)


def h5_read(path: str, dset: str, dtype=np.float32) -> np.ndarray:
    rows, cols = h5_shape(path, dset)
    out = np.empty((rows, cols), dtype)
    f = h5().knng_h5_read_f32 if dtype == np.float32 else h5().knng_h5_read_i32
    rc = f(path.encode(), dset.encode(), 0, rows, cols, _ptr(out))
    if rc:
        raise KnngError(f"h5 read of {path}:{dset} failed ({rc})")
    return out


def h5_write_train(path: str, x: np.ndarray, chunk_rows: int = 0) -> None:
    x = np.ascontiguousarray(x, np.float32)
    rc = h5().knng_h5_write_f32(path.encode(), b"train", _ptr(x), x.shape[0], x.shape[1],
                                chunk_rows, 0)
    if rc:
        raise KnngError(f"h5 write of {path} failed ({rc})")


class H5Writer:
    """Chunked float32 /train writer (knng_h5_create_f32 / write_rows / finish)."""

    def __init__(self, path: str, rows: int, cols: int, chunk_rows: int, dset: str = "train"):
        self.w = C.c_void_p()
        rc = h5().knng_h5_create_f32(path.encode(), dset.encode(), rows, cols, chunk_rows, C.byref(self.w))
        if rc:
            raise KnngError(f"h5 create of {path} failed ({rc})")

    def write(self, row0: int, rows: np.ndarray) -> None:
        rows = np.ascontiguousarray(rows, np.float32)
        rc = h5().knng_h5_write_rows(self.w, row0, rows.shape[0], _ptr(rows))
        if rc:
            raise KnngError(f"h5 write of rows {row0}+{rows.shape[0]} failed ({rc})")

    def close(self) -> None:
        if self.w:
            rc = h5().knng_h5_finish(self.w)
            self.w = C.c_void_p()
            if rc:
                raise KnngError(f"h5 finish failed ({rc})")


def h5_attr_str(path: str, name: str) -> str:
    buf = C.create_string_buffer(4096)
    rc = h5().knng_h5_read_attr_str(path.encode(), name.encode(), buf, 4096)
    if rc:
        raise KnngError(f"h5 attribute {name} of {path}: {rc}")
    return buf.value.decode()


def h5_attr_f32(path: str, name: str) -> float:
    v = C.c_float()
    rc = h5().knng_h5_read_attr_f32(path.encode(), name.encode(), C.byref(v))
    if rc:
        raise KnngError(f"h5 attribute {name} of {path}: {rc}")
    return float(v.value)


# -- serving ------------------------------------------------------------------

class ServerParams(C.Structure):
    _fields_ = [
        ("beam", C.c_int32), ("k", C.c_int32), ("entry_trees", C.c_int32), ("rerank", C.c_int32),
        ("max_steps", C.c_int32), ("max_batch", C.c_int32), ("max_slots", C.c_int32),
        ("pad0", C.c_int32), ("insert_capacity", C.c_int64), ("reserve_bytes", C.c_int64),
        ("express_rows", C.c_int32), ("express_slots", C.c_int32), ("express_cus", C.c_int32),
        ("strict_order", C.c_int32), ("bulk_yield", C.c_int32),
    ]


def _server_lib() -> C.CDLL:
    L = lib()
    if not getattr(L, "_server_sigs", False):
        V, i32, i64, P = C.c_void_p, C.c_int32, C.c_int64, C.POINTER
        sig = {
            "knng_server_default_params": ([P(ServerParams)], i32),
            "knng_server_create": ([P(V), V, P(ServerParams)], i32),
            "knng_server_destroy": ([V], i32),
            "knng_server_last_error": ([V], C.c_char_p),
            "knng_server_submit": ([V, i32, V, i32, P(i64)], i32),
            "knng_server_poll": ([V, i64, P(i32)], i32),
            "knng_server_wait": ([V, i64, V, V, P(C.c_double)], i32),
            "knng_server_wait_info": ([V, i64, V, V, P(C.c_double), P(i64)], i32),
            "knng_server_info": ([V, P(i32), P(i32), P(i32)], i32),
            "knng_server_mask_error": ([V], C.c_char_p),
            "knng_server_size": ([V], i64),
            "knng_server_search": ([V, V, i32, V, V], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        L._server_sigs = True
    return L


def server_params(**over) -> ServerParams:
    p = ServerParams()
    _server_lib().knng_server_default_params(C.byref(p))
    for k, v in over.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Server:
    """Device-resident kNN serving over a built Builder (knng_server)."""

    def __init__(self, builder: Builder, params: ServerParams | None = None, **over):
        self.L = _server_lib()
        self.builder = builder          # keeps the forest alive
        self.p = params if params is not None else server_params(**over)
        self.d = builder.p.dim
        self.srv = C.c_void_p()
        rc = self.L.knng_server_create(C.byref(self.srv), builder.ctx, C.byref(self.p))
        if rc:
            raise KnngError(f"knng_server_create failed ({rc})")

    def _ck(self, rc, what):
        if rc:
            msg = self.L.knng_server_last_error(self.srv).decode(errors="replace")
            raise KnngError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.srv:
            self.L.knng_server_destroy(self.srv)
            self.srv = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, rows: np.ndarray, insert: bool = False) -> int:
        rows = np.ascontiguousarray(rows, np.float32)
        assert rows.ndim == 2 and rows.shape[1] == self.d
        t = C.c_int64()
        self._ck(self.L.knng_server_submit(self.srv, int(insert), _ptr(rows), rows.shape[0],
                                           C.byref(t)), "submit")
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int32()
        self._ck(self.L.knng_server_poll(self.srv, ticket, C.byref(st)), "poll")
        return st.value

    def wait(self, ticket: int, nrows: int, insert: bool = False):
        """-> (ids, scores, latency_ms); inserts return (new ids, None, latency)."""
        lat = C.c_double()
        if insert:
            ids = np.empty(nrows, np.int32)
            self._ck(self.L.knng_server_wait(self.srv, ticket, _ptr(ids), None, C.byref(lat)), "wait")
            return ids, None, lat.value
        ids = np.empty((nrows, self.p.k), np.int32)
        sc = np.empty((nrows, self.p.k), np.float32)
        self._ck(self.L.knng_server_wait(self.srv, ticket, _ptr(ids), _ptr(sc), C.byref(lat)), "wait")
        return ids, sc, lat.value

    def wait_info(self, ticket: int, nrows: int, insert: bool = False):
        """-> (ids, scores, latency_ms, index_rows): index_rows is the index size a
        query was dispatched against (an insert: the size just after it)."""
        lat, rows = C.c_double(), C.c_int64()
        if insert:
            ids = np.empty(nrows, np.int32)
            self._ck(self.L.knng_server_wait_info(self.srv, ticket, _ptr(ids), None, C.byref(lat),
                                                  C.byref(rows)), "wait")
            return ids, None, lat.value, rows.value
        ids = np.empty((nrows, self.p.k), np.int32)
        sc = np.empty((nrows, self.p.k), np.float32)
        self._ck(self.L.knng_server_wait_info(self.srv, ticket, _ptr(ids), _ptr(sc), C.byref(lat),
                                              C.byref(rows)), "wait")
        return ids, sc, lat.value, rows.value

    def search(self, queries: np.ndarray):
        queries = np.ascontiguousarray(queries, np.float32)
        nq = queries.shape[0]
        ids = np.empty((nq, self.p.k), np.int32)
        sc = np.empty((nq, self.p.k), np.float32)
        self._ck(self.L.knng_server_search(self.srv, _ptr(queries), nq, _ptr(ids), _ptr(sc)), "search")
        return ids, sc

    def size(self) -> int:
        return int(self.L.knng_server_size(self.srv))

    def info(self) -> dict:
        m, b, e = C.c_int32(), C.c_int32(), C.c_int32()
        self._ck(self.L.knng_server_info(self.srv, C.byref(m), C.byref(b), C.byref(e)), "info")
        return {"cu_masked": bool(m.value), "bulk_slots": b.value, "express_slots": e.value,
                "mask_error": self.L.knng_server_mask_error(self.srv).decode(errors="replace")}


# -- batched builds ------------------------------------------------------------

class Job(C.Structure):
    _fields_ = [
        ("rows", C.c_void_p), ("rows_on_device", C.c_int32), ("rerank", C.c_int32),
        ("n", C.c_int64), ("params", Params), ("knns", C.c_void_p), ("dists", C.c_void_p),
        ("status", C.c_int32), ("iterations", C.c_int32), ("build_ms", C.c_double),
    ]


def build_batch(jobs, device: int = 0, workers: int = 8) -> int:
    """jobs: list of dicts {rows: device ptr or numpy array, n, params, rerank};
    returns the batch status; fills job['knns'], ['dists'], ['status'],
    ['iterations'], ['build_ms']."""
    L = lib()
    if not getattr(L, "_batch_sig", False):
        L.knng_build_batch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.knng_build_batch.restype = C.c_int32
        L._batch_sig = True
    arr = (Job * len(jobs))()
    keep = []
    for j, job in enumerate(jobs):
        p = job["params"]
        k = p.k_eval
        job["knns"] = np.empty((job["n"], k), np.int32)
        job["dists"] = np.empty((job["n"], k), np.float32)
        rows = job["rows"]
        if isinstance(rows, np.ndarray):
            rows = np.ascontiguousarray(rows, np.float32)
            keep.append(rows)
            arr[j].rows = rows.ctypes.data
            arr[j].rows_on_device = 0
        else:
            arr[j].rows = int(rows)
            arr[j].rows_on_device = 1
        arr[j].rerank = int(job.get("rerank", 0))
        arr[j].n = job["n"]
        arr[j].params = p
        arr[j].knns = job["knns"].ctypes.data
        arr[j].dists = job["dists"].ctypes.data
    rc = L.knng_build_batch(C.cast(arr, C.c_void_p), len(jobs), device, workers)
    for j, job in enumerate(jobs):
        job["status"] = arr[j].status
        job["iterations"] = arr[j].iterations
        job["build_ms"] = arr[j].build_ms
    return rc
