"""Synthetic inputs for the kNN-graph workloads (no dataset download exists).

gaussian_mixture / gaussian_mixture_torch
    The workload generator: C cluster centres drawn as random unit vectors,
    point = centre + N(0, 0.35^2 * (k+1)^-0.5) noise on dim k, L2-normalised,
    float32; similarity is the dot product.  The numpy and torch versions draw
    from the same distribution but not the same stream.

fixture_rows
    Integer-exact clustered rows for golden fixtures: every value is an
    integer multiple of 2^-13 computed with uint64 arithmetic only, so the
    bytes are identical on every platform (a SHA-256 pins them).

recall_at_k
    recall = sum_i |pred_i & (gt_i \\ {i+1})[:k]| / (N k), 1-based ids, as the
    SISAP evaluator computes it (scripts/evaluation/evaluate_submission.py:15-26).
"""
from __future__ import annotations

import hashlib

import numpy as np

NOISE_SIGMA = 0.35


def noise_scale(d: int) -> np.ndarray:
    return (NOISE_SIGMA * np.arange(1, d + 1, dtype=np.float64) ** -0.25).astype(np.float64)


def gaussian_mixture(seed: int, n: int, d: int, clusters: int, chunk: int = 65536) -> np.ndarray:
    rng = np.random.default_rng(seed)
    cent = rng.standard_normal((clusters, d))
    cent /= np.linalg.norm(cent, axis=1, keepdims=True)
    assign = rng.integers(0, clusters, n)
    scale = noise_scale(d)
    out = np.empty((n, d), np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        y = cent[assign[a:b]] + rng.standard_normal((b - a, d)) * scale
        y /= np.linalg.norm(y, axis=1, keepdims=True)
        out[a:b] = y
    return out


def gaussian_mixture_torch(seed: int, n: int, d: int, clusters: int, device="cuda",
                           chunk: int = 262144, out=None, draw_seed: int | None = None):
    """Same distribution generated directly in device memory (float32).
    draw_seed: held-out rows of the mixture `seed` describes -- the same cluster
    centres, with assignments and noise drawn from a stream seeded by draw_seed."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    cent = torch.randn((clusters, d), generator=g, device=device, dtype=torch.float32)
    cent /= cent.norm(dim=1, keepdim=True)
    if draw_seed is not None:
        g = torch.Generator(device=device)
        g.manual_seed(draw_seed)
    assign = torch.randint(0, clusters, (n,), generator=g, device=device)
    scale = torch.tensor(noise_scale(d), device=device, dtype=torch.float32)
    if out is None:
        out = torch.empty((n, d), device=device, dtype=torch.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        y = torch.randn((b - a, d), generator=g, device=device, dtype=torch.float32)
        y.mul_(scale).add_(cent[assign[a:b]])
        y.div_(y.norm(dim=1, keepdim=True))
        out[a:b] = y
    return out


def _splitmix(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def fixture_rows(seed: int, n: int, d: int, clusters: int) -> np.ndarray:
    """Integer-exact clustered rows (values k / 8192, |k| < 2^13)."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) * np.uint64(0x100000001B3)
        kk = np.arange(d, dtype=np.uint64)
        cid = np.arange(clusters, dtype=np.uint64)[:, None]
        cent = (_splitmix(s ^ (cid << np.uint64(20)) ^ kk) % np.uint64(1201)).astype(np.int64) - 600
        pid = np.arange(n, dtype=np.uint64)
        assign = (_splitmix(s ^ np.uint64(0xA5A5) ^ (pid << np.uint64(1))) % np.uint64(clusters)).astype(np.int64)
        # noise amplitude falls with the dimension index: 1 + 896 // (k+1)^(1/2) (integer)
        amp = 1 + 896 // np.maximum(1, np.floor(np.sqrt(np.arange(1, d + 1))).astype(np.int64))
        out = np.empty((n, d), np.float32)
        for a in range(0, n, 4096):
            b = min(n, a + 4096)
            p = pid[a:b, None]
            h = _splitmix(s ^ np.uint64(0x5EED) ^ (p << np.uint64(16)) ^ kk)
            u = (h % np.uint64(2001)).astype(np.int64) - 1000          # [-1000, 1000]
            v = cent[assign[a:b]] + (u * amp) // 1000
            out[a:b] = (v.astype(np.float64) / 8192.0).astype(np.float32)
    return out


def strided_sha256_torch(x, stride: int = 997) -> str:
    """SHA-256 of rows 0, stride, 2 stride, ... of a device tensor (float32
    bytes): a fingerprint of a generated workload without copying it whole."""
    return sha256(x[::stride].cpu().numpy())


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def exact_knn(x: np.ndarray, queries: np.ndarray, k: int, block: int = 1024) -> np.ndarray:
    """Exact FP32 dot-product top-k (0-based ids) of rows `queries`, self excluded."""
    out = np.empty((len(queries), k), np.int64)
    for a in range(0, len(queries), block):
        qi = queries[a:a + block]
        s = x[qi] @ x.T
        s[np.arange(len(qi)), qi] = -np.inf
        top = np.argpartition(-s, k, axis=1)[:, :k]
        o = np.argsort(-np.take_along_axis(s, top, 1), axis=1, kind="stable")
        out[a:a + block] = np.take_along_axis(top, o, 1)
    return out


def recall_at_k(knns_1based: np.ndarray, gt_0based: np.ndarray, rows: np.ndarray, k: int) -> float:
    """Evaluator rule: hits of pred (1-based) against gt (self already removed)."""
    hits = 0
    for r, g in zip(rows, gt_0based):
        hits += len(np.intersect1d(knns_1based[r] - 1, g[:k]))
    return hits / (len(rows) * k)


def exact_knn_torch(x, rows, k: int, chunk: int = 262144):
    """Exact FP32 dot-product top-k (0-based ids, best first) of rows x[rows]
    over all rows of the device tensor x, self excluded.  The product is taken
    in row chunks: one unchunked 500 x 6.35M float32 product in torch returned
    a different top-15 for a quarter of the entries."""
    import torch
    q = x[rows]
    R = q.shape[0]
    best_s = torch.full((R, k + 1), float("-inf"), device=x.device)
    best_i = torch.zeros((R, k + 1), dtype=torch.int64, device=x.device)
    ar = torch.arange(R, device=x.device)
    for a in range(0, x.shape[0], chunk):
        b = min(x.shape[0], a + chunk)
        s = q @ x[a:b].T
        inside = (rows >= a) & (rows < b)
        s[ar[inside], rows[inside] - a] = float("-inf")
        ts, ti = s.topk(min(k + 1, b - a), dim=1)
        cs = torch.cat([best_s, ts], 1)
        ci = torch.cat([best_i, ti + a], 1)
        best_s, o = cs.topk(k + 1, dim=1)
        best_i = ci.gather(1, o)
    return best_i[:, :k]


def recall_rows(pred_1based: np.ndarray, gt_0based: np.ndarray, k: int) -> float:
    """Evaluator rule on aligned rows: pred_1based[j] (1-based ids) against gt_0based[j]."""
    hits = 0
    for p, g in zip(pred_1based, gt_0based):
        hits += len(np.intersect1d(np.asarray(p) - 1, g[:k]))
    return hits / (len(gt_0based) * k)


def exact_knn_prefix_torch(parts, qv, limits, k: int, chunk: int = 262144):
    """Exact FP32 top-k (0-based ids, best first) of query vectors qv over the
    rows of the device tensors `parts`, taken as one concatenated row space,
    each query j seeing only rows [0, limits[j]) (an index that grew by inserts)."""
    import torch
    R = qv.shape[0]
    best_s = torch.full((R, k), float("-inf"), device=qv.device)
    best_i = torch.zeros((R, k), dtype=torch.int64, device=qv.device)
    lim = limits.view(-1, 1)
    top = int(limits.max().item())
    base = 0
    for part in parts:
        for a in range(0, part.shape[0], chunk):
            if base + a >= top:
                break
            b = min(part.shape[0], a + chunk)
            s = qv @ part[a:b].T
            cols = torch.arange(base + a, base + b, device=qv.device).view(1, -1)
            s.masked_fill_(cols >= lim, float("-inf"))
            ts, ti = s.topk(min(k, b - a), dim=1)
            cs = torch.cat([best_s, ts], 1)
            ci = torch.cat([best_i, ti + base + a], 1)
            best_s, o = cs.topk(k, dim=1)
            best_i = ci.gather(1, o)
        base += part.shape[0]
    return best_i
