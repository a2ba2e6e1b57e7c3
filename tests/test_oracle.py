"""CPU tests that pin the oracle (oracle/knng_oracle.cpp).

Reference mode is pinned byte-for-byte by the compiled reference's golden
outputs (tests/golden/, written by make_golden.py at OMP_NUM_THREADS=1).
The primitives are pinned by closed forms and known-answer rows, the
device-mode stages by brute-force restatements of their definitions on small
inputs and by invariants.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import make_golden
import oracle
import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))


# ---------------------------------------------------------------- primitives

def test_quantize_known_answers():
    # src/main.cpp:184-190: multiply by 127, clamp to [-128,127], truncate; bit = clamped > 0
    vals = np.array([0.0, -0.0, 1e-40, -1e-40, 0.5 / 127, -0.5 / 127, 1.0, -1.0, 1.5, -1.5,
                     126.9 / 127, -128.5 / 127, 127.0 / 127, -128.0 / 127, np.inf, -np.inf,
                     0.99 / 127, 3.3 / 127, -3.3 / 127, 2.0 / 127] + [0.0] * 44, np.float32)
    q, sk = oracle.quantize(vals[None, :])
    exp_q = [0, 0, 0, 0, 0, 0, 127, -127, 127, -128, 126, -128, 127, -128, 127, -128, 0, 3, -3, 2]
    assert q[0, :20].tolist() == exp_q
    bits = [(int(sk[0, 0]) >> b) & 1 for b in range(20)]
    # 1e-40 is denormal: DAZ reads it as 0 -> bit 0; 0.5/127 -> q 0 but bit 1 (clamped > 0)
    assert bits == [0, 0, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 1, 0, 1]


def test_fast_math_sets_daz():
    """The reference is linked with -ffast-math (CMakeLists.txt:6); GCC then
    sets FTZ/DAZ at start-up, which is why the oracle reads denormals as 0."""
    src = ("#include <xmmintrin.h>\n#include <stdio.h>\n"
           "int main(){unsigned c=_mm_getcsr();printf(\"%d %d\\n\",(c>>6)&1,(c>>15)&1);return 0;}\n")
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "p")
        subprocess.run(["gcc", "-O3", "-ffast-math", c, "-o", exe], check=True)
        daz, ftz = subprocess.run([exe], stdout=subprocess.PIPE, text=True).stdout.split()
    assert (daz, ftz) == ("1", "1")


def test_quantize_matches_closed_form_random():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((37, 256)) * 0.3).astype(np.float32)
    x[0, :8] = [1e-39, -1e-39, 5e-41, 0.0, -0.0, 2.0, -2.0, 1e-45]
    q, sk = oracle.quantize(x)
    xd = np.where(np.abs(x) < np.finfo(np.float32).tiny, np.float32(0), x)
    c = np.clip(xd * np.float32(127), -128, 127)
    assert np.array_equal(q, np.trunc(c).astype(np.int8))
    bits = np.packbits((c > 0).reshape(37, 4, 64)[:, :, ::-1], axis=2).view(">u8")[:, :, 0]
    assert np.array_equal(sk, bits.astype(np.uint64))


def test_dot_and_popcount_closed_form():
    rng = np.random.default_rng(1)
    a = rng.integers(-128, 128, 1024).astype(np.int8)
    b = rng.integers(-128, 128, 1024).astype(np.int8)
    assert oracle.dot_i8(a, b) == int(a.astype(np.int64) @ b.astype(np.int64))
    extreme = np.full(1024, -128, np.int8)
    assert oracle.dot_i8(extreme, extreme) == 1024 * 128 * 128
    sa = rng.integers(0, 2 ** 63, 16, dtype=np.uint64)
    sb = rng.integers(0, 2 ** 63, 16, dtype=np.uint64)
    pc = sum(bin(int(u) ^ int(v)).count("1") for u, v in zip(sa, sb))
    assert oracle.sketch_sim(sa, sb) == 1024 - pc
    assert oracle.sketch_sim(sa, sa) == 1024


# ---------------------------------------------------------- reference mode

@pytest.mark.parametrize("case", make_golden.CASES, ids=[c[0] for c in make_golden.CASES])
def test_reference_mode_matches_golden(case):
    name, seed, n, clusters, edge, keep = case
    x = make_golden.case_input(seed, n, clusters, edge)
    assert synth.sha256(x) == META[name]["input_sha256"]   # the generator is pinned
    q, sk = oracle.quantize(x)
    ids, sc, st = oracle.build_reference(q, sk)
    knns, dists = oracle.export(ids, sc)
    f = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    rows = f["rows"] if "rows" in f.files else np.arange(n)
    assert np.array_equal(knns[rows], f["knns"])
    assert np.array_equal(dists[rows], f["dists"])
    # the whole output, also where only sampled rows are committed (n = 20000)
    assert synth.sha256(knns) == META[name]["knns_sha256"]
    assert synth.sha256(dists) == META[name]["dists_sha256"]
    rates = [round(float(st.rates[i]) * 100, 4) for i in range(st.iterations)]
    assert rates == pytest.approx(META[name]["rates_percent"], abs=1e-3)


def test_golden_dists_are_int8_dots():
    """Every emitted dist is the int8 dot of the quantised rows (closed form)."""
    x = make_golden.case_input(12, 2000, 32, False)
    q, _ = oracle.quantize(x)
    f = np.load(os.path.join(GOLDEN, "ref_n2000.npz"))
    Q = q.astype(np.int64)
    i = np.repeat(np.arange(2000), 15)
    j = f["knns"].reshape(-1) - 1
    assert np.array_equal((Q[i] * Q[j]).sum(1).astype(np.float32), f["dists"].reshape(-1))
    # rows ascending by score, best neighbour last (src/main.cpp:455-456)
    assert np.all(np.diff(f["dists"], axis=1) >= 0)


# -------------------------------------------------------------- device mode

M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _pivots(seed, t, depth, start, size):
    h1 = _mix64(seed ^ _mix64((t << 48) ^ (depth << 40) ^ start))
    h2 = _mix64(h1)
    a, b = h1 % size, h2 % (size - 1)
    return a, b + (b >= a)


def test_pivot_hash():
    for args in [(1337, 0, 0, 0, 1000), (1337, 5, 3, 12345, 77), (7, 63, 30, 2 ** 30, 65)]:
        assert oracle.pivots(*args) == _pivots(*args)


def _forest_bruteforce(q, T, leaf, seed, cell_assign=None):
    """Recursive two-pivot forest; with cell_assign ([G, n]) tree t starts
    from partition t % G's cells (rows by (cell, id)) at depth 1."""
    Q = q.astype(np.int64)
    n = len(q)
    perms, leaves = [], []
    for t in range(T):
        if cell_assign is None:
            perm = list(range(n))
        else:
            a = cell_assign[t % len(cell_assign)]
            perm = sorted(range(n), key=lambda i: (a[i], i))

        def rec(start, size, depth):
            if size <= leaf or depth >= 64:   # depth cap: cut into leaves of <= leaf
                for s0 in range(0, size, leaf):
                    leaves.append((t, start + s0, min(leaf, size - s0)))
                return
            a, b = _pivots(seed, t, depth, start, size)
            seg = perm[start:start + size]
            diff = Q[seg] @ (Q[seg[a]] - Q[seg[b]])
            lo = [seg[j] for j in range(size) if j == a or (j != b and diff[j] > 0)]
            hi = [seg[j] for j in range(size) if not (j == a or (j != b and diff[j] > 0))]
            perm[start:start + size] = lo + hi
            rec(start, len(lo), depth + 1)
            rec(start + len(lo), size - len(lo), depth + 1)

        if cell_assign is None:
            rec(0, n, 0)
        else:
            sizes = np.bincount(a, minlength=int(a.max()) + 1)
            start = 0
            for z in sizes:
                if z:
                    rec(start, int(z), 1)
                start += int(z)
        perms.append(perm)
    return np.array(perms, np.int32), np.array(sorted(leaves), np.int64)


def _small(seed=3, n=300, d=128, clusters=6):
    x = synth.gaussian_mixture(seed, n, d, clusters)
    return x, *oracle.quantize(x)


def test_forest_device_bruteforce():
    x, q, sk = _small()
    perm, leaves = oracle.forest_device(q, 3, 8, seed=1337)
    bp, bl = _forest_bruteforce(q, 3, 8, 1337)
    assert np.array_equal(perm, bp)
    assert np.array_equal(leaves, bl)
    # invariants: permutations, leaves tile each tree, leaf bound
    for t in range(3):
        assert sorted(perm[t].tolist()) == list(range(300))
        lt = leaves[leaves[:, 0] == t]
        assert lt[0, 1] == 0 and np.all(lt[1:, 1] == lt[:-1, 1] + lt[:-1, 2])
        assert lt[:, 2].sum() == 300 and lt[:, 2].max() <= 8 and lt[:, 2].min() >= 1


def test_forest_device_duplicate_rows_terminate():
    """Identical rows always tie (ties go right), so a node of duplicates only
    sheds its pivot each level; past depth 64 the node is cut into leaves."""
    x, q, sk = _small(n=300)
    q[50:260] = q[50]
    sys_limit = __import__("sys").getrecursionlimit()
    __import__("sys").setrecursionlimit(max(sys_limit, 5000))
    perm, leaves = oracle.forest_device(q, 2, 8, seed=5)
    bp, bl = _forest_bruteforce(q, 2, 8, 5)
    assert np.array_equal(perm, bp) and np.array_equal(leaves, bl)
    assert leaves[:, 2].max() <= 8 and leaves[:, 2].sum() == 2 * 300


# ------------------------------------------------------------------ cells

def _assign_np(q, z, w):
    """argmax_c float32(<q, z_c>) * w_c with the first maximum, in numpy."""
    dots = (q.astype(np.int64) @ z.astype(np.int64).T).astype(np.float32)
    return np.argmax(dots * w[None, :].astype(np.float32), axis=1).astype(np.int32)


def test_assign_device_bruteforce():
    rng = np.random.default_rng(5)
    q = rng.integers(-128, 128, size=(500, 192), dtype=np.int8)
    z = rng.integers(-127, 128, size=(37, 192), dtype=np.int8)
    w = (1.0 / np.sqrt((z.astype(np.float64) ** 2).sum(1))).astype(np.float32)
    assert np.array_equal(oracle.assign_device(q, z, w), _assign_np(q, z, w))
    # exact ties (duplicate centroids) go to the smaller cell
    z[20] = z[3]
    w[20] = w[3]
    a = oracle.assign_device(q, z, w)
    assert not np.any(a == 20)
    assert np.array_equal(a, _assign_np(q, z, w))
    # all-zero centroids score 0 and never beat a positive score
    z[:10] = 0
    w[:10] = 0
    assert np.array_equal(oracle.assign_device(q, z, w), _assign_np(q, z, w))


def _quant_centroid_np(S):
    """int8 direction of an exact integer sum: trunc(S * 127 / max|S|), w = 1/|z|."""
    S = S.astype(np.int64)
    M = np.abs(S).max()
    if M == 0:
        return np.zeros(S.shape, np.int8), np.float32(0)
    z = (np.sign(S) * ((np.abs(S) * 127) // M)).astype(np.int8)
    return z, np.float32(1.0 / np.sqrt((z.astype(np.float64) ** 2).sum()))


def _kmeans_np(q, Ct, g, iters, over, sample, merge, full_iters, prune):
    """numpy restatement of the cell k-means: seeds, sample Lloyd, mutual-best
    merging, pruning, full Lloyd, final assignment."""
    n, d = q.shape
    off = g % sample if g % sample < n else 0
    rows = np.arange(off, n, sample)
    Cs = max(1, min(Ct * over, len(rows)))
    seeds = ((2 * np.arange(Cs, dtype=np.int64) + 1) * n // (2 * Cs) + 7919 * g) % n
    Z = np.zeros((Cs, d), np.int8)
    W = np.zeros(Cs, np.float32)
    for c in range(Cs):
        Z[c], W[c] = _quant_centroid_np(q[seeds[c]])

    def sums(idx, a, C):
        S = np.zeros((C, d), np.int64)
        np.add.at(S, a, q[idx].astype(np.int64))
        return S, np.bincount(a, minlength=C)

    for it in range(iters + 1):
        a = _assign_np(q[rows], Z, W)
        S, cnt = sums(rows, a, Cs)
        if it == iters:
            break
        for c in np.nonzero(cnt)[0]:
            Z[c], W[c] = _quant_centroid_np(S[c])
    live = [c for c in range(Cs) if cnt[c]]
    for c in live:
        Z[c], W[c] = _quant_centroid_np(S[c])
    for _ in range(64):
        if len(live) < 2:
            break
        zl, wl = Z[live], W[live]
        dots = (zl.astype(np.int64) @ zl.astype(np.int64).T).astype(np.float32) * wl[None, :]
        np.fill_diagonal(dots, -np.inf)
        best = np.argmax(dots, 1)
        val = dots[np.arange(len(live)), best]
        gone = set()
        for j in range(len(live)):
            k = best[j]
            if k > j and best[k] == j and np.float32(val[j] * wl[j]) > np.float32(merge):
                S[live[j]] += S[live[k]]
                cnt[live[j]] += cnt[live[k]]
                Z[live[j]], W[live[j]] = _quant_centroid_np(S[live[j]])
                gone.add(k)
        if not gone:
            break
        live = [c for j, c in enumerate(live) if j not in gone]
    med = np.sort(cnt[live])[len(live) // 2]
    live = [c for c in live if not np.float32(cnt[c]) < np.float32(prune) * np.float32(med)]
    Zf, Wf = Z[live].copy(), W[live].copy()
    allr = np.arange(n)
    for it in range(full_iters + 1):
        a = _assign_np(q, Zf, Wf)
        if it == full_iters:
            break
        S, cnt = sums(allr, a, len(live))
        for c in np.nonzero(cnt)[0]:
            Zf[c], Wf[c] = _quant_centroid_np(S[c])
    return Zf, Wf, a


@pytest.mark.parametrize("Ct,g,iters,over,sample,merge,full,prune", [
    (6, 0, 2, 4, 3, 0.6, 1, 0.25), (6, 1, 0, 1, 1, 2.0, 0, 0.0), (9, 1, 1, 3, 2, 0.3, 2, 0.5),
    (40, 0, 2, 4, 8, 0.6, 1, 0.25), (30, 1, 2, 4, 2, 0.6, 1, 1.0),
])
def test_kmeans_device_numpy_restatement(Ct, g, iters, over, sample, merge, full, prune):
    x, q, sk = _small(seed=8, n=600, d=128, clusters=10)
    z, w, a = oracle.kmeans_device(q, Ct, g, iters, over, sample, merge, full, prune)
    ez, ew, ea = _kmeans_np(q, Ct, g, iters, over, sample, merge, full, prune)
    assert np.array_equal(z, ez) and np.array_equal(w, ew) and np.array_equal(a, ea)


def test_kmeans_recovers_separated_clusters():
    """Eight sign-pattern clusters with small noise: four seeds per cluster
    are merged back, and the cells are exactly the clusters."""
    rng = np.random.default_rng(11)
    C, per, d = 8, 60, 128
    centers = rng.choice([-1.0, 1.0], size=(C, d)).astype(np.float32) * 0.5
    lab = rng.integers(0, C, C * per)
    x = centers[lab] + rng.normal(0, 0.05, size=(C * per, d)).astype(np.float32)
    q, sk = oracle.quantize(x.astype(np.float32))
    z, w, a = oracle.kmeans_device(q, C, 0, iters=2, over=4, sample=2, merge=0.6, full_iters=1)
    assert len(w) == C
    # a bijection between cells and clusters
    pairs = set(zip(a.tolist(), lab.tolist()))
    assert len(pairs) == C and len({p[0] for p in pairs}) == C
    # without merging the pieces stay apart
    z2, w2, a2 = oracle.kmeans_device(q, C, 0, iters=2, over=4, sample=2, merge=2.0, full_iters=1)
    assert len(w2) > C


def test_forest_cells_bruteforce():
    x, q, sk = _small(seed=4, n=500, d=128, clusters=8)
    for Ct, G, T, leaf in [(6, 2, 3, 16), (40, 3, 4, 8), (125, 1, 2, 4), (4, 2, 2, 300)]:
        plan = oracle.cells_device(q, T, leaf, cells=Ct, partitions=G, sample=2)
        counts, G, _, _, A = plan
        perm, leaves = oracle.forest_cells_device(q, T, leaf, plan, seed=1337)
        bp, bl = _forest_bruteforce(q, T, leaf, 1337, cell_assign=A)
        assert np.array_equal(perm, bp) and np.array_equal(leaves, bl)
        for t in range(T):
            # each cell is one contiguous run of the tree's order, cells ascending
            cell_of = A[t % G][perm[t]]
            assert np.all(np.diff(cell_of) >= 0) and cell_of.max() < counts[t % G]
            lt = leaves[leaves[:, 0] == t]
            assert lt[:, 2].sum() == 500 and lt[:, 2].max() <= leaf
            for _, s0, m in lt:   # no leaf straddles two cells
                assert len(set(cell_of[s0:s0 + m].tolist())) == 1


def test_cells_plan_rule():
    assert oracle.cells_for(3071) == 0 and oracle.cells_for(3072) == 2
    assert oracle.cells_for(6_350_000) == 4134 and oracle.cells_for(10, 5) == 5
    x, q, sk = _small(n=300)
    assert oracle.cells_device(q, 4, 64) is None          # auto: too few rows
    counts, G, z, w, A = oracle.cells_device(q, 4, 64, cells=5, partitions=3, sample=2)
    assert G == 3 and A.shape == (3, 300) and len(z) == 3 and z[0].shape[1] == 128
    assert all(1 <= c <= 20 for c in counts)
    assert oracle.cells_device(q, 4, 400, cells=5) is None   # the whole set is one leaf
    assert oracle.cells_device(q, 1, 64, cells=5, partitions=2)[1] == 1   # G <= trees


def test_init_big_leaves_bruteforce():
    """Leaves above 64 rows: the init equals the brute-force top-K over all
    leaf mates (the device keeps each member's 64 best per tree)."""
    x, q, sk = _small(seed=6, n=700, d=128, clusters=5)
    plan = oracle.cells_device(q, 2, 300, cells=4, partitions=2, sample=2)
    perm, leaves = oracle.forest_cells_device(q, 2, 300, plan)
    assert leaves[:, 2].max() > 64
    ids, sc = oracle.init_device(q, perm, leaves, 50)
    Q = q.astype(np.int64)
    for i in range(0, 700, 37):
        mates = set()
        for t, s0, m in leaves:
            mem = perm[t, s0:s0 + m]
            if i in mem:
                mates |= set(mem.tolist())
        mates.discard(i)
        cand = sorted(((int(Q[i] @ Q[j]), j) for j in mates), key=lambda e: (-e[0], e[1]))[:50]
        assert ids[i].tolist() == [j for _, j in cand]
        assert sc[i].tolist() == [s for s, _ in cand]


def _topk_unique(cands, K):
    cands = sorted(cands, key=lambda e: (-e[1], e[0]))
    out, prev = [], None
    for e in cands:
        if e[0] != prev:
            out.append(e)
            prev = e[0]
    return out[:K]


def test_init_device_bruteforce():
    x, q, sk = _small(n=300)
    T, K = 3, 12
    perm, leaves = oracle.forest_device(q, T, 8, seed=9)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    Q = q.astype(np.int64)
    mates = [set() for _ in range(300)]
    for t, s, z in leaves:
        mem = perm[t, s:s + z]
        for m in mem:
            mates[m].update(int(v) for v in mem if v != m)
    for i in range(300):
        cands = [(c, int(Q[i] @ Q[c])) for c in mates[i]]
        exp = _topk_unique(cands, K)
        if len(exp) == K:
            assert ids[i].tolist() == [e[0] for e in exp]
            assert sc[i].tolist() == [e[1] for e in exp]
        else:   # padding rows: the real candidates still come first
            assert set(e[0] for e in exp) <= set(ids[i].tolist())


def _xorshift_ids(i, n, count):
    """src/main.cpp:271-286: xorshift32 from 1337+i, id = state % n, self skipped."""
    s, out = (1337 + i) & 0xFFFFFFFF, []
    while len(out) < count:
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        if s % n != i:
            out.append(s % n)
    return out


def test_init_padding_follows_reference_rule():
    """n < K: every other row is a candidate; the pool is padded to K with
    xorshift32 ids (duplicates), de-duplicated, and the free slots filled by
    the same xorshift32 stream at INT32_MIN (src/main.cpp:271-312)."""
    n, K = 40, 50
    x = make_golden.case_input(15, n, 4, False)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 16, 64)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    Q = q.astype(np.int64)
    for i in range(n):
        others = [c for c in range(n) if c != i]
        exp = sorted(((c, int(Q[i] @ Q[c])) for c in others), key=lambda e: (-e[1], e[0]))
        assert list(zip(ids[i, :n - 1].tolist(), sc[i, :n - 1].tolist())) == exp
        stream = _xorshift_ids(i, n, K - (n - 1) + K)
        assert ids[i, n - 1:].tolist() == stream[K - (n - 1):][:K - (n - 1)]
        assert np.all(sc[i, n - 1:] == np.iinfo(np.int32).min)


def test_sketch_threshold_bruteforce():
    rng = np.random.default_rng(5)
    for _ in range(50):
        cnt = int(rng.integers(1, 3000))
        h = rng.integers(400, 700, cnt).astype(np.int32)
        M = int(rng.integers(1, 600))
        t = oracle.sketch_threshold(h, M, 1024)
        if cnt <= M:
            assert t == 0
        else:
            assert (h >= t).sum() >= M and (h >= t + 1).sum() < M


def _leaf_sets(perm, leaves, n):
    """per row, the set of (tree, leaf start) it belongs to (numpy, independent of the oracle)"""
    mine = [set() for _ in range(n)]
    for t, s0, z in leaves:
        for r in perm[t, s0:s0 + z]:
            mine[int(r)].add((int(t), int(s0)))
    return mine


@pytest.mark.parametrize("leaf_rule", [False, True])
def test_join_device_bruteforce(leaf_rule):
    """One join iteration against a numpy restatement: 2-hop candidates (a path
    i -> n1 -> c is dropped under the leaf-mate rule when i, n1 and c lie in one
    leaf of one tree), sketch threshold, exact int8 scores, (score desc, id asc,
    slot) merge, change count."""
    x, q, sk = _small(n=250, d=128)
    K, M = 10, 30
    perm, leaves = oracle.forest_device(q, 2, 16)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    lk = oracle.leaf_keys(perm, leaves) if leaf_rule else None
    oi, osc, upd = oracle.join_device(q, sk, ids, sc, M, leafkey=lk)
    mine = _leaf_sets(perm, leaves, 250)
    Q = q.astype(np.int64)
    bits = np.unpackbits(sk.view(np.uint8), axis=1, bitorder="little")
    changes = 0
    for i in range(250):
        excl = set(ids[i].tolist()) | {i}
        cands = []
        for n1 in ids[i]:
            for c in ids[n1]:
                c = int(c)
                # leaf-mate rule: i, n1 and c in one leaf of one tree
                if leaf_rule and mine[i] & mine[int(n1)] & mine[c]:
                    continue
                if c not in excl and c not in cands:
                    cands.append(c)
        h = np.array([128 - int((bits[i] ^ bits[c]).sum()) for c in cands], np.int32)
        if len(cands) > M:
            thr = max(t for t in range(129) if (h >= t).sum() >= M)
            sel = [c for c, hv in zip(cands, h) if hv >= thr]
        else:
            sel = cands
        pool = [(int(ids[i, k]), int(sc[i, k]), k) for k in range(K)]
        pool += [(c, int(Q[i] @ Q[c]), K) for c in sel]
        pool.sort(key=lambda e: (-e[1], e[0], e[2]))
        assert oi[i].tolist() == [e[0] for e in pool[:K]]
        assert osc[i].tolist() == [e[1] for e in pool[:K]]
        changes += sum(int(a != b) for a, b in zip(oi[i], ids[i]))
    assert upd == changes


@pytest.mark.parametrize("leaf", [24, 150])   # 150: leaves above 64 rows keep 64 mates per member
def test_leaf_mates_never_enter_the_list(leaf):
    """What makes the leaf-mate rule safe: after a leaf init, a join without the
    rule never inserts a candidate sharing a leaf with the row (the init scored
    it exactly and kept it unless 64 mates beat it), over three iterations."""
    x, q, sk = _small(n=400, d=128)
    K, M = 10, 30
    perm, leaves = oracle.forest_device(q, 3, leaf)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    mine = _leaf_sets(perm, leaves, 400)
    entered = 0
    for _ in range(3):
        oi, osc, _ = oracle.join_device(q, sk, ids, sc, M)
        for i in range(400):
            for c in set(oi[i].tolist()) - set(ids[i].tolist()):
                entered += 1
                assert not (mine[i] & mine[c]), (i, c)
        ids, sc = oi, osc
    assert entered > 50


def test_incremental_join_bruteforce():
    """Incremental rule: (i -> n1 -> c) is a candidate iff n1 is fresh in i's
    list or c is fresh in n1's list; fresh_out marks entries that came from
    candidates."""
    x, q, sk = _small(n=250, d=128)
    K, M = 10, 30
    perm, leaves = oracle.forest_device(q, 2, 16)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    rng = np.random.default_rng(3)
    fresh = rng.integers(0, 2 ** 63, 250, dtype=np.uint64) & np.uint64((1 << K) - 1)
    oi, osc, upd, fo = oracle.join_device(q, sk, ids, sc, M, fresh=fresh)
    Q = q.astype(np.int64)
    bits = np.unpackbits(sk.view(np.uint8), axis=1, bitorder="little")
    for i in range(250):
        excl = set(ids[i].tolist()) | {i}
        cands = []
        for k, n1 in enumerate(ids[i]):
            for k2, c in enumerate(ids[n1]):
                c = int(c)
                if not ((int(fresh[i]) >> k) & 1 or (int(fresh[n1]) >> k2) & 1):
                    continue
                if c not in excl and c not in cands:
                    cands.append(c)
        h = np.array([128 - int((bits[i] ^ bits[c]).sum()) for c in cands], np.int32)
        sel = cands
        if len(cands) > M:
            thr = max(t for t in range(129) if (h >= t).sum() >= M)
            sel = [c for c, hv in zip(cands, h) if hv >= thr]
        pool = [(int(ids[i, k]), int(sc[i, k]), k) for k in range(K)]
        pool += [(c, int(Q[i] @ Q[c]), K) for c in sel]
        pool.sort(key=lambda e: (-e[1], e[0], e[2]))
        assert oi[i].tolist() == [e[0] for e in pool[:K]]
        assert int(fo[i]) == sum(1 << r for r, e in enumerate(pool[:K]) if e[2] == K)


@pytest.mark.parametrize("R,leaf_rule,incremental", [(4, True, False), (2, True, True), (3, True, True),
                                                    (8, False, True)])
def test_reverse_join_bruteforce(R, leaf_rule, incremental):
    """Reverse-neighbour candidates against a numpy restatement.  RV(v) = the R
    best rows u by (score desc, u asc) whose lists hold v as an entry the
    iteration treats as new (all entries, or the fresh ones); row i adds u in
    RV(i), w in RV(n1) for its neighbours n1 and the list of every u in RV(i).
    Under the leaf-mate rule a path is dropped when all of its rows lie in one
    leaf of one tree (i <- u: i and u share a leaf)."""
    x, q, sk = _small(n=250, d=128)
    K, M = 10, 30
    perm, leaves = oracle.forest_device(q, 2, 16)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    lk = oracle.leaf_keys(perm, leaves) if leaf_rule else None
    fresh = None
    if incremental:
        rng = np.random.default_rng(R)
        fresh = rng.integers(0, 2 ** 63, 250, dtype=np.uint64) & np.uint64((1 << K) - 1)
    out = oracle.join_device(q, sk, ids, sc, M, fresh=fresh, leafkey=lk, reverse=R)
    oi, osc = out[0], out[1]
    mine = _leaf_sets(perm, leaves, 250)
    Q = q.astype(np.int64)
    bits = np.unpackbits(sk.view(np.uint8), axis=1, bitorder="little")
    new = lambda u, k: fresh is None or (int(fresh[u]) >> k) & 1
    inc = [[] for _ in range(250)]
    for u in range(250):
        for k in range(K):
            if new(u, k):
                inc[int(ids[u, k])].append((-int(sc[u, k]), u))
    rv = [[u for _, u in sorted(e)[:R]] for e in inc]
    shared = lambda *rows: leaf_rule and bool(set.intersection(*(mine[r] for r in rows)))
    n_rev = 0
    for i in range(250):
        excl = set(ids[i].tolist()) | {i}
        cands = []

        def add(c):
            if c not in excl and c not in cands:
                cands.append(c)
        for k, n1 in enumerate(ids[i]):
            for k2, c in enumerate(ids[n1]):
                if fresh is not None and not (new(i, k) or new(int(n1), k2)):
                    continue
                if not shared(i, int(n1), int(c)):
                    add(int(c))
        before = len(cands)
        for u in rv[i]:
            if not shared(i, u):
                add(u)
        for n1 in ids[i]:
            for w in rv[int(n1)]:
                if not shared(i, int(n1), w):
                    add(w)
        for u in rv[i]:
            for w in ids[u]:
                if not shared(i, u, int(w)):
                    add(int(w))
        n_rev += len(cands) - before
        h = np.array([128 - int((bits[i] ^ bits[c]).sum()) for c in cands], np.int32)
        sel = cands
        if len(cands) > M:
            thr = max(t for t in range(129) if (h >= t).sum() >= M)
            sel = [c for c, hv in zip(cands, h) if hv >= thr]
        pool = [(int(ids[i, k]), int(sc[i, k]), k) for k in range(K)]
        pool += [(c, int(Q[i] @ Q[c]), K) for c in sel]
        pool.sort(key=lambda e: (-e[1], e[0], e[2]))
        assert oi[i].tolist() == [e[0] for e in pool[:K]], i
        assert osc[i].tolist() == [e[1] for e in pool[:K]], i
    # the reverse lists add candidates the 2-hop paths miss (without the new/old
    # rule almost every reverse candidate is also a 2-hop one at this size)
    assert n_rev > (250 if incremental else 0)


def test_rerank_device_fp64():
    x, q, sk = _small(n=200, d=128)
    perm, leaves = oracle.forest_device(q, 4, 16)
    ids, sc = oracle.init_device(q, perm, leaves, 20)
    knns, dists = oracle.rerank_device(x, ids, 20, 15)
    X = x.astype(np.float64)
    for i in range(200):
        s = [(float(X[i] @ X[c]), int(c)) for c in ids[i]]
        s.sort(key=lambda e: (-e[0], e[1]))
        assert knns[i].tolist() == [c + 1 for _, c in s[:15]][::-1]
        assert np.allclose(dists[i], [v for v, _ in s[:15]][::-1], rtol=0, atol=1e-12)


def test_device_semantics_recall_tracks_reference():
    """Device mode (hash pivots, Jacobi, dedup) reaches the reference's recall
    on the same input."""
    x = synth.gaussian_mixture(21, 4000, 1024, 16)
    q, sk = oracle.quantize(x)
    gt = synth.exact_knn(x, np.arange(4000), 15)
    r_ids, r_sc, _ = oracle.build_reference(q, sk)
    d_ids, d_sc, _ = oracle.build_device_semantics(q, sk)
    rk, _ = oracle.export(r_ids, r_sc)
    dk, _ = oracle.export(d_ids, d_sc)
    rec_ref = synth.recall_at_k(rk, gt, np.arange(4000), 15)
    rec_dev = synth.recall_at_k(dk, gt, np.arange(4000), 15)
    assert rec_dev >= rec_ref - 0.01, (rec_dev, rec_ref)


@pytest.mark.parametrize("beam,leaf,T", [(64, 8, 2), (20, 8, 2), (32, 4, 3)])
def test_search_device_complete_graph_is_bruteforce(beam, leaf, T):
    """orc_search_device on a complete graph (every row lists all others):
    the entry leaf holds a few rows, and the first expansion scores every row,
    so whatever the beam width (>= k) the answer is the brute-force int8
    top-k by (score desc, id asc).  A wrong sort key, a dropped candidate or a
    beam cut before the sort fails it."""
    n, d, k = 60, 128, 15
    x = synth.gaussian_mixture(61, n, d, 4)
    q, _ = oracle.quantize(x)
    rng = np.random.default_rng(beam)
    gid = np.stack([rng.permutation(np.delete(np.arange(n), i)) for i in range(n)]).astype(np.int32)
    queries = synth.gaussian_mixture(62, 25, d, 4)
    qq, _ = oracle.quantize(queries)
    ids, sc = oracle.search_device(q, gid, qq, T=T, leaf=leaf, entry_trees=T, beam=beam, k=k,
                                   max_steps=128)
    S = qq.astype(np.int64) @ q.astype(np.int64).T
    for j in range(len(qq)):
        order = np.lexsort((np.arange(n), -S[j]))[:k]
        assert np.array_equal(ids[j], order), j
        assert np.array_equal(sc[j], S[j, order]), j


def test_search_device_without_expansion_stays_in_entry_leaves():
    """max_steps = 0: no expansion, so every answer comes from the entry leaves
    (each at most `leaf` rows), best first by (score desc, id asc)."""
    n, d, k, leaf = 300, 128, 5, 16
    x = synth.gaussian_mixture(63, n, d, 6)
    q, _ = oracle.quantize(x)
    gid = np.zeros((n, 8), np.int32)
    queries = synth.gaussian_mixture(64, 40, d, 6)
    qq, _ = oracle.quantize(queries)
    ids, sc = oracle.search_device(q, gid, qq, T=1, leaf=leaf, entry_trees=1, beam=16, k=k,
                                   max_steps=0)
    perm, leaves = oracle.forest_device(q, 1, leaf, 1337)
    S = qq.astype(np.int64) @ q.astype(np.int64).T
    for j in range(len(qq)):
        got = ids[j][ids[j] >= 0]
        # all answers lie in one leaf of tree 0, and are that leaf's best rows
        owners = [l for l in range(len(leaves)) if np.isin(got, perm[0, leaves[l, 1]:leaves[l, 1] + leaves[l, 2]]).all()]
        assert len(owners) == 1, j
        s, z = leaves[owners[0], 1], leaves[owners[0], 2]
        mem = perm[0, s:s + z]
        order = mem[np.lexsort((mem, -S[j, mem]))][:k]
        assert np.array_equal(got, order), j
