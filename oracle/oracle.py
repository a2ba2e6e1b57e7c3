"""TEST INFRASTRUCTURE ONLY: ctypes binding of oracle/liboracle.so, the CPU
restatement of the reference engine (src/main.cpp).  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.

See knng_oracle.cpp for what each function restates and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "apex_main")
REF_BIN_NATIVE = os.path.join(HERE, "_ref", "apex_main_native")


class OrcParams(C.Structure):
    _fields_ = [("k_build", C.c_int32), ("num_trees", C.c_int32), ("rerank_mult", C.c_int32),
                ("leaf_size", C.c_int32), ("max_iters", C.c_int32),
                ("early_stop_rate", C.c_float), ("rate_diff_tol", C.c_float)]


class OrcStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("num_leaves", C.c_int32),
                ("updates", C.c_int64 * 64), ("rates", C.c_float * 64)]


_L = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        V, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
        sig = {
            "orc_quantize": ([V, i64, i32, V, V], i32),
            "orc_dot_i8": ([V, V, i32], i32),
            "orc_sketch_sim": ([V, V, i32], i32),
            "orc_pivots": ([u64, i32, i32, i64, i64, C.POINTER(i64), C.POINTER(i64)], None),
            "orc_build_reference": ([V, V, i64, i32, C.POINTER(OrcParams), V, V, C.POINTER(OrcStats)], i32),
            "orc_export": ([V, V, i64, i32, i32, V, V], i32),
            "orc_forest_device": ([V, i64, i32, i32, i32, u64, V, V, i64], i64),
            "orc_init_device": ([V, i64, i32, i32, i32, V, V, i64, V, V], i32),
            "orc_sketch_threshold": ([V, i64, i32, i32], i32),
            "orc_join_device": ([V, V, i64, i32, i32, i32, V, V, V, V, C.POINTER(i64), V, V, V, i32, i32], i64),
            "orc_rerank_device": ([V, i64, i32, i32, i32, i32, V, V, V], i32),
            "orc_search_device": ([V, i64, i32, V, i32, i32, i32, u64, i32, i32, i32, i32, V, i64,
                                   V, V, i32, V, V, V], i32),
            "orc_assign_device": ([V, i64, i32, V, V, i32, V], i32),
            "orc_kmeans_device": ([V, i64, i32, i32, i32, i32, i32, i32, C.c_float, i32,
                                   C.c_float, V, V, V], i32),
            "orc_forest_cells_device": ([V, i64, i32, i32, i32, u64, V, V, i32, V, V, i64], i64),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _L = L
    return _L


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def ref_params(**over) -> OrcParams:
    """src/main.cpp:22-29 constants."""
    p = OrcParams(50, 16, 8, 64, 30, 0.04, 0.01)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def quantize(x: np.ndarray):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    q = np.empty((n, d), np.int8)
    sk = np.empty((n, d // 64), np.uint64)
    assert lib().orc_quantize(_p(x), n, d, _p(q), _p(sk)) == 0
    return q, sk


def dot_i8(a: np.ndarray, b: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    return int(lib().orc_dot_i8(_p(a), _p(b), a.shape[0]))


def sketch_sim(a: np.ndarray, b: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    return int(lib().orc_sketch_sim(_p(a), _p(b), a.shape[0]))


def pivots(seed: int, tree: int, depth: int, start: int, size: int):
    a, b = C.c_int64(), C.c_int64()
    lib().orc_pivots(seed, tree, depth, start, size, C.byref(a), C.byref(b))
    return int(a.value), int(b.value)


def build_reference(q: np.ndarray, sk: np.ndarray, params: OrcParams | None = None):
    """Reference-mode build (bit-exact to src/main.cpp at one thread)."""
    p = params or ref_params()
    n, d = q.shape
    ids = np.empty((n, p.k_build), np.int32)
    sc = np.empty((n, p.k_build), np.int32)
    st = OrcStats()
    rc = lib().orc_build_reference(_p(q), _p(sk), n, d, C.byref(p), _p(ids), _p(sc), C.byref(st))
    assert rc == 0
    return ids, sc, st


def export(ids: np.ndarray, sc: np.ndarray, k_eval: int = 15):
    n, K = ids.shape
    knns = np.empty((n, k_eval), np.int32)
    dists = np.empty((n, k_eval), np.float32)
    assert lib().orc_export(_p(np.ascontiguousarray(ids)), _p(np.ascontiguousarray(sc)), n, K,
                            k_eval, _p(knns), _p(dists)) == 0
    return knns, dists


def forest_device(q: np.ndarray, trees: int, leaf: int, seed: int = 1337):
    n, d = q.shape
    perm = np.empty((trees, n), np.int32)
    cap = trees * n + 16
    leaves = np.empty((cap, 3), np.int64)
    nl = lib().orc_forest_device(_p(q), n, d, trees, leaf, seed, _p(perm), _p(leaves), cap)
    assert nl >= 0
    return perm, leaves[:nl].copy()


def cells_for(n: int, cells: int = -1) -> int:
    """Cell count the library picks: `cells` if >= 0, else n // 1536 when that is >= 2."""
    if cells >= 0:
        return cells
    c = n // 1536
    return c if c >= 2 else 0


def assign_device(q: np.ndarray, z: np.ndarray, w: np.ndarray) -> np.ndarray:
    n, d = q.shape
    C = z.shape[0]
    out = np.empty(n, np.int32)
    assert lib().orc_assign_device(_p(np.ascontiguousarray(q)), n, d, _p(np.ascontiguousarray(z)),
                                   _p(np.ascontiguousarray(w, np.float32)), C, _p(out)) == 0
    return out


def kmeans_device(q: np.ndarray, Ct: int, g: int = 0, iters: int = 2, over: int = 4,
                  sample: int = 8, merge: float = 0.6, full_iters: int = 0, prune: float = 0.25):
    """Cells of partition g (over-seeded spherical int8 k-means with merging)
    -> (z [C, d] int8, w [C] f32, assign [n] int32)."""
    n, d = q.shape
    off = g % sample if g % sample < n else 0
    ns = (n - off + sample - 1) // sample
    cap = max(1, min(Ct * over, ns))
    z = np.empty((cap, d), np.int8)
    w = np.empty(cap, np.float32)
    a = np.empty(n, np.int32)
    C = lib().orc_kmeans_device(_p(np.ascontiguousarray(q)), n, d, Ct, g, iters, over, sample,
                                merge, full_iters, prune, _p(z), _p(w), _p(a))
    assert C >= 1
    return z[:C].copy(), w[:C].copy(), a


def cells_device(q: np.ndarray, trees: int, leaf: int, cells: int = -1, iters: int = 2,
                 partitions: int = 2, over: int = 4, sample: int = 8, merge: float = 0.6,
                 full_iters: int = 0, prune: float = 0.25):
    """The library's cell plan for a build: None when no cells are used, else
    (counts [G], G, [z_g], [w_g], assign [G, n])."""
    n = q.shape[0]
    Ct = min(cells_for(n, cells), n)
    if Ct < 2 or n <= leaf:
        return None
    G = max(1, min(partitions, trees))
    zs, ws, As = zip(*(kmeans_device(q, Ct, g, iters, over, sample, merge, full_iters, prune)
                       for g in range(G)))
    counts = np.array([len(w) for w in ws], np.int32)
    return counts, G, list(zs), list(ws), np.stack(As)


def forest_cells_device(q: np.ndarray, trees: int, leaf: int, plan, seed: int = 1337):
    """Forest seeded by a cell plan (from cells_device); plan None = plain forest."""
    if plan is None:
        return forest_device(q, trees, leaf, seed)
    counts, G, _, _, A = plan
    n, d = q.shape
    perm = np.empty((trees, n), np.int32)
    cap = trees * n + 16
    leaves = np.empty((cap, 3), np.int64)
    nl = lib().orc_forest_cells_device(_p(q), n, d, trees, leaf, seed,
                                       _p(np.ascontiguousarray(A, np.int32)),
                                       _p(np.ascontiguousarray(counts, np.int32)), G, _p(perm),
                                       _p(leaves), cap)
    assert nl >= 0
    return perm, leaves[:nl].copy()


def init_device(q: np.ndarray, perm: np.ndarray, leaves: np.ndarray, K: int):
    n, d = q.shape
    T = perm.shape[0]
    ids = np.empty((n, K), np.int32)
    sc = np.empty((n, K), np.int32)
    leaves = np.ascontiguousarray(leaves, np.int64)
    assert lib().orc_init_device(_p(q), n, d, T, K, _p(np.ascontiguousarray(perm)), _p(leaves),
                                 leaves.shape[0], _p(ids), _p(sc)) == 0
    return ids, sc


def sketch_threshold(h: np.ndarray, M: int, bits: int) -> int:
    h = np.ascontiguousarray(h, np.int32)
    return int(lib().orc_sketch_threshold(_p(h), h.shape[0], M, bits))


def leaf_keys(perm: np.ndarray, leaves: np.ndarray) -> np.ndarray:
    """[n][T] int32: row r's leaf in tree t, as the leaf's first position in perm[t]."""
    T, n = perm.shape
    key = np.full((n, T), -1, np.int32)
    for t, s, z in np.asarray(leaves, np.int64):
        key[perm[t, s:s + z], t] = s
    assert (key >= 0).all(), "every row lies in one leaf of every tree"
    return key


def join_device(q, sk, ids, sc, M: int, with_inserted: bool = False, fresh=None, leafkey=None,
                reverse: int = 0):
    """One Jacobi join iteration -> (ids, scores, updates[, inserted][, fresh_out]).
    fresh (u64 per row) selects the incremental join and adds fresh_out;
    leafkey ([n][T], from leaf_keys) applies the leaf-mate rule; reverse = R > 0
    adds the reverse-neighbour candidates of R-entry reverse lists."""
    n, d = q.shape
    K = ids.shape[1]
    oi = np.empty_like(ids)
    os_ = np.empty_like(sc)
    ins = C.c_int64()
    fo = np.empty(n, np.uint64)
    fi = None if fresh is None else _p(np.ascontiguousarray(fresh, np.uint64))
    lk = None if leafkey is None else np.ascontiguousarray(leafkey, np.int32)
    upd = lib().orc_join_device(_p(q), _p(sk), n, d, K, M, _p(np.ascontiguousarray(ids)),
                                _p(np.ascontiguousarray(sc)), _p(oi), _p(os_), C.byref(ins),
                                fi, _p(fo), None if lk is None else _p(lk),
                                0 if lk is None else lk.shape[1], int(reverse))
    assert upd >= 0
    out = [oi, os_, int(upd)]
    if with_inserted:
        out.append(int(ins.value))
    if fresh is not None:
        out.append(fo)
    return tuple(out)


def rerank_device(x, ids, R: int, k_eval: int = 15):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    K = ids.shape[1]
    knns = np.empty((n, k_eval), np.int32)
    dists = np.empty((n, k_eval), np.float64)
    assert lib().orc_rerank_device(_p(x), n, d, K, R, k_eval, _p(np.ascontiguousarray(ids)),
                                   _p(knns), _p(dists)) == 0
    return knns, dists


STOP_REFERENCE, STOP_INSERTED, STOP_NONE = 0, 1, 2


def build_device_semantics(q, sk, trees=16, leaf=64, K=50, mult=8, max_iters=30,
                           stop=0.04, tol=0.01, seed=1337, stop_rule=STOP_INSERTED,
                           insert_stop=0.01, incremental=False, cells=-1, kmeans_iters=2,
                           partitions=2, over=4, sample=8, merge=0.6, full_iters=0,
                           prune=0.25, reverse=0):
    """Full device-mode build.  Stop rules: STOP_REFERENCE is src/main.cpp:432-439
    on the rate of slots whose id changed; STOP_INSERTED stops once fewer than
    insert_stop of all slots received a new id; STOP_NONE runs max_iters.
    cells / kmeans_iters / partitions select the cell seeding of the forest
    with the library's rule (cells=-1: n // 1536 when >= 2, 0: none); reverse
    is the reverse-list width of the join iterations (0: none)."""
    plan = cells_device(q, trees, leaf, cells, kmeans_iters, partitions, over, sample, merge,
                        full_iters, prune)
    perm, leaves = forest_cells_device(q, trees, leaf, plan, seed)
    ids, sc = init_device(q, perm, leaves, K)
    lk = leaf_keys(perm, leaves)   # the lists descend from this forest: leaf-mate rule
    n = q.shape[0]
    prev = np.float32(1.0)
    rates = []
    fresh = np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF)) if incremental else None
    for it in range(max_iters):
        if incremental:
            ids, sc, upd, ins, fresh = join_device(q, sk, ids, sc, K * mult, with_inserted=True,
                                                   fresh=fresh, leafkey=lk, reverse=reverse)
        else:
            ids, sc, upd, ins = join_device(q, sk, ids, sc, K * mult, with_inserted=True, leafkey=lk,
                                            reverse=reverse)
        rate = np.float32(upd) / np.float32(n * K)
        rates.append(float(rate))
        if stop_rule == STOP_INSERTED:
            if np.float32(ins) / np.float32(n * K) < np.float32(insert_stop):
                break
        elif stop_rule == STOP_REFERENCE:
            if rate < np.float32(stop) or (it > 0 and (prev - rate) < np.float32(tol)):
                break
        prev = rate
    return ids, sc, rates


def run_reference_binary(h5_path: str, workdir: str, threads: int = 1, native: bool = False,
                         task: str = "task1", timeout: float = 3600.0) -> dict:
    """Run the compiled reference engine on an HDF5 /train file; returns the
    path of its output file and its stdout."""
    exe = REF_BIN_NATIVE if native else REF_BIN
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([exe, "--task", task, "--dataset", os.path.abspath(h5_path)], cwd=workdir,
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    return {"returncode": r.returncode, "stdout": r.stdout,
            "out": os.path.join(workdir, "results", task, "apex_simd.h5")}


def search_device(q, gid, queries_q, T=16, leaf=64, seed=1337, entry_trees=4, beam=64, k=15,
                  max_steps=128, plan=None):
    """Beam search over graph `gid` (device semantics of the serving kernel);
    `plan` (from cells_device) routes each tree's entry through the query's cell."""
    n, d = q.shape
    K = gid.shape[1]
    queries_q = np.ascontiguousarray(queries_q)
    nq = queries_q.shape[0]
    ids = np.empty((nq, k), np.int32)
    sc = np.empty((nq, k), np.int32)
    A = qc = counts = None
    G = 0
    if plan is not None:
        counts, G, Z, W, A = plan
        qc = np.ascontiguousarray(np.stack([assign_device(queries_q, Z[g], W[g]) for g in range(G)]))
        A = np.ascontiguousarray(A, np.int32)
        counts = np.ascontiguousarray(counts, np.int32)
    rc = lib().orc_search_device(_p(np.ascontiguousarray(q)), n, d, _p(np.ascontiguousarray(gid)), K,
                                 T, leaf, seed, entry_trees, beam, k, max_steps,
                                 _p(queries_q), nq, None if A is None else _p(A),
                                 None if counts is None else _p(counts), G,
                                 None if qc is None else _p(qc), _p(ids), _p(sc))
    assert rc == 0
    return ids, sc
