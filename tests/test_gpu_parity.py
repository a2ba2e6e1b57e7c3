"""GPU parity tests: every stage of the HIP path against the oracle on the
same inputs, through the C-ABI (libknng.so).

Integer stages (quantise, sketch, int8 dots, forest, init, join, plain
export) must be bit-exact against the oracle's device mode; the FP32 rerank
must be within 1e-5 relative of FP64 dot products.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import knng
import make_golden
import oracle
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _params(**kw):
    return knng.default_params(**kw)


def _builder(x, **kw):
    n, d = x.shape
    b = knng.Builder(n, _params(dim=d, **kw))
    b.ingest(x)
    return b


# ------------------------------------------------------------------ K1

@pytest.mark.parametrize("d", [1024, 768, 384, 64])
def test_quantize_bitexact(torch_cuda, d):
    say("quantize", d)
    torch = torch_cuda
    rng = np.random.default_rng(d)
    x = (rng.standard_normal((2000, d)) * 0.2).astype(np.float32)
    edge = make_golden.EDGE_VALUES
    for r in range(16):
        x[r] = np.tile(np.roll(edge, r), d // len(edge))
    x[16, :4] = [np.inf, -np.inf, 1e-45, -1e-45]
    q_ref, sk_ref = oracle.quantize(x)
    dx = _dev(torch, x)
    dq = torch.empty((2000, d), dtype=torch.int8, device="cuda")
    dsk = torch.empty((2000, d // 64), dtype=torch.int64, device="cuda")
    knng.quantize_device(dx.data_ptr(), 2000, d, dq.data_ptr(), dsk.data_ptr())
    assert np.array_equal(dq.cpu().numpy(), q_ref)
    assert np.array_equal(dsk.cpu().numpy().view(np.uint64), sk_ref)


def test_ingest_host_and_device_paths(torch_cuda):
    say("ingest paths")
    torch = torch_cuda
    x = synth.gaussian_mixture(5, 20000, 1024, 64)     # 80 MB: two pinned slices
    q_ref, sk_ref = oracle.quantize(x)
    b = knng.Builder(20000, _params())
    b.ingest(x[:7000], 0)
    b.ingest(x[7000:], 7000)
    q, sk = b.quantized()
    assert np.array_equal(q, q_ref) and np.array_equal(sk, sk_ref)
    b.close()
    b = knng.Builder(20000, _params(keep_fp32=0))
    dx = _dev(torch, x)
    b.ingest_device(dx.data_ptr(), 0, 20000)
    q, sk = b.quantized()
    assert np.array_equal(q, q_ref) and np.array_equal(sk, sk_ref)


def test_score_pairs_bitexact(torch_cuda):
    say("score pairs")
    torch = torch_cuda
    x = synth.gaussian_mixture(6, 3000, 1024, 16)
    q, sk = oracle.quantize(x)
    rng = np.random.default_rng(0)
    a = rng.integers(0, 3000, 5000).astype(np.int32)
    b = rng.integers(0, 3000, 5000).astype(np.int32)
    dq, dsk, da, db = _dev(torch, q), _dev(torch, sk.view(np.int64)), _dev(torch, a), _dev(torch, b)
    dot = torch.empty(5000, dtype=torch.int32, device="cuda")
    sim = torch.empty(5000, dtype=torch.int32, device="cuda")
    knng.score_pairs_device(dq.data_ptr(), dsk.data_ptr(), 1024, da.data_ptr(), db.data_ptr(), 5000,
                            dot.data_ptr(), sim.data_ptr())
    Q = q.astype(np.int64)
    assert np.array_equal(dot.cpu().numpy(), (Q[a] * Q[b]).sum(1))
    exp = np.array([oracle.sketch_sim(sk[i], sk[j]) for i, j in zip(a, b)])
    assert np.array_equal(sim.cpu().numpy(), exp)


# ------------------------------------------------------------- K2-K4

@pytest.mark.parametrize("n,d,T,leaf,cells,parts", [
    (3000, 1024, 16, 64, -1, 2),    # auto: n < 3072, no cells
    (5000, 128, 3, 8, -1, 2),       # auto: 3 cells, 2 partitions
    (700, 384, 5, 33, -1, 2),
    (3000, 1024, 16, 64, 12, 2),
    (2000, 256, 5, 16, 40, 3),      # many small cells: leaves straight from cells
    (1500, 64, 2, 64, 700, 1),      # cells of 0-3 rows, empty cells
    (6000, 256, 2, 512, -1, 2),     # big leaves (k_leaf_big) under RP splits of the cells
    (5000, 128, 3, 2048, 3, 2),     # whole cells as big leaves
    (6000, 1024, 3, 64, 150, 2),    # d = 1024 assignment over 19 centroid tiles (last one partial)
    (4000, 1024, 2, 512, 300, 1),   # 38 centroid tiles, one partition
    (5000, 1024, 2, 2048, 3, 2),    # d = 1024 big leaves: whole cells of ~1700 rows
    (6000, 1024, 2, 512, -1, 2),    # d = 1024 big leaves under RP splits
])
def test_forest_and_init_bitexact(torch_cuda, n, d, T, leaf, cells, parts):
    say("forest+init", n, d, T, leaf, cells, parts)
    x = synth.gaussian_mixture(n + d, n, d, 24)
    q, sk = oracle.quantize(x)
    K = 50 if n >= 1000 else 20
    b = _builder(x, num_trees=T, leaf_size=leaf, k_build=K, k_eval=15, cells=cells,
                 cell_partitions=parts)
    b.build_forest()
    perm, leaves = b.forest()
    plan = oracle.cells_device(q, T, leaf, cells, partitions=parts)
    assert b.stats()["cells"] == (0 if plan is None else int(plan[0].sum()))
    o_perm, o_leaves = oracle.forest_cells_device(q, T, leaf, plan, 1337)
    assert np.array_equal(perm, o_perm)
    assert np.array_equal(leaves, o_leaves)
    b.build_init()
    ids, sc = b.graph()
    o_ids, o_sc = oracle.init_device(q, o_perm, o_leaves, K)
    assert np.array_equal(ids, o_ids)
    assert np.array_equal(sc, o_sc)


def test_leaf_select_ties_bitexact(torch_cuda):
    """1100 copies of one row in leaves of about 2000 rows: at least 550 of
    them share a leaf, where each copy's scores tie by the hundred at its top,
    so the selection's cut keeps more than 512 entries and those rows go to
    the slow kernel (the tie is broken by id)."""
    say("leaf select ties")
    n, d = 4000, 256
    x = synth.gaussian_mixture(77, n, d, 8)
    x[1000:2100] = x[1000]
    q, _ = oracle.quantize(x)
    b = _builder(x, num_trees=2, leaf_size=2048, k_build=50, k_eval=15, cells=0)
    b.build_forest()
    perm, leaves = b.forest()
    assert leaves[:, 2].max() > 512
    b.build_init()
    ids, sc = b.graph()
    o_ids, o_sc = oracle.init_device(q, perm, leaves, 50)
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    b.close()


# both leaf-tile shapes for leaves of at most 64 rows (block per leaf, wave
# per leaf): leaves of at most 32 rows (one tile) and of 33-64 rows (the
# symmetric three-tile path), d from one 32-dim step group to 1024
@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n,d,T,leaf", [(3000, 1024, 8, 64), (2500, 384, 4, 32), (2000, 64, 3, 48),
                                        (1200, 192, 5, 16)])
def test_leaf_tile_shapes_bitexact(torch_cuda, monkeypatch, variant, n, d, T, leaf):
    say("leaf tiles", variant, n, d, T, leaf)
    monkeypatch.setenv("KNNG_LEAF_VARIANT", str(variant))
    x = synth.gaussian_mixture(n + leaf, n, d, 24)
    q, _ = oracle.quantize(x)
    b = _builder(x, num_trees=T, leaf_size=leaf, k_build=20, k_eval=15, cells=0)
    b.build_forest()
    perm, leaves = b.forest()
    b.build_init()
    assert b.stats()["leaf_variant"] == variant
    ids, sc = b.graph()
    o_ids, o_sc = oracle.init_device(q, perm, leaves, 20)
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    b.close()


def test_leaf_tile_autotuner(torch_cuda):
    """Enough leaves to time: the autotuner times both shapes on a prefix of the
    leaf table, keeps one (cached for the shape), and the init equals the
    init of either shape forced."""
    say("leaf tile autotuner")
    n, d = 120000, 128
    x = synth.gaussian_mixture(91, n, d, 64)
    graphs = []
    for env in (None, "0", "1"):
        if env is None:
            os.environ.pop("KNNG_LEAF_VARIANT", None)
        else:
            os.environ["KNNG_LEAF_VARIANT"] = env
        try:
            b = _builder(x, num_trees=2, leaf_size=16, k_build=20, k_eval=15, cells=0)
            b.build_forest()
            b.build_init()
            st = b.stats()
            graphs.append(b.graph())
            if env is None:
                assert st["leaf_variant"] in (0, 1)
                assert st["ms_tune_leaf"] > 0 or st["leaf_variant"] >= 0
            b.close()
        finally:
            os.environ.pop("KNNG_LEAF_VARIANT", None)
    for ids, sc in graphs[1:]:
        assert np.array_equal(ids, graphs[0][0]) and np.array_equal(sc, graphs[0][1])


def test_large_assignment_forest_bitexact(torch_cuda):
    """70,000 rows at d = 1024 (45 target cells, 2 partitions): the full-row
    cell assignment spans 547 row tiles, the last one partial, over 69 / 68
    cells (the last centroid tile partial); forest bit-exact."""
    n, d, T, leaf = 70000, 1024, 2, 512
    say("large assignment forest", n)
    x = synth.gaussian_mixture(77, n, d, 64)
    q, sk = oracle.quantize(x)
    b = _builder(x, num_trees=T, leaf_size=leaf, cells=-1, cell_partitions=2)
    b.build_forest()
    perm, leaves = b.forest()
    plan = oracle.cells_device(q, T, leaf, -1, partitions=2)
    o_perm, o_leaves = oracle.forest_cells_device(q, T, leaf, plan, 1337)
    assert np.array_equal(perm, o_perm)
    assert np.array_equal(leaves, o_leaves)
    b.close()


def test_duplicate_rows_build_bitexact(torch_cuda):
    """Many identical rows: the depth cap cuts them into leaves; whole build
    still bit-exact."""
    say("duplicates")
    x = synth.gaussian_mixture(31, 3000, 256, 8)
    x[100:2100] = x[100]
    q, sk = oracle.quantize(x)
    b = _builder(x, num_trees=4, leaf_size=16, k_build=20, k_eval=15)
    b.build_forest()
    perm, leaves = b.forest()
    o_perm, o_leaves = oracle.forest_device(q, 4, 16, 1337)
    assert np.array_equal(perm, o_perm) and np.array_equal(leaves, o_leaves)
    b.build_init()
    o_ids, o_sc = oracle.init_device(q, o_perm, o_leaves, 20)
    ids, sc = b.graph()
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)


# ------------------------------------------------------------- K5-K7

def test_join_iterations_bitexact(torch_cuda):
    say("join")
    x = synth.gaussian_mixture(7, 6000, 1024, 32)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 4, 64, 1337)   # a weak forest leaves work to the join
    ids, sc = oracle.init_device(q, perm, leaves, 50)
    b = _builder(x)
    b.load_graph(ids, sc)
    for it in range(3):
        upd = b.join_iteration()
        ids, sc, o_upd = oracle.join_device(q, sk, ids, sc, 400)
        g_ids, g_sc = b.graph()
        assert upd == o_upd, it
        assert np.array_equal(g_ids, ids), it
        assert np.array_equal(g_sc, sc), it


@pytest.mark.parametrize("fused", [False, True])
def test_incremental_join_bitexact(torch_cuda, monkeypatch, fused):
    """New/old rule: only pairs with a fresh edge are candidates.  Default
    path: the wave-per-row phases with the rows that hold too many candidates
    passed to the block-per-row kernel; fused: the block-per-row kernel, and
    after the first iteration its small-row variant passing rows on."""
    say("incremental join", "fused" if fused else "phased")
    if fused:
        monkeypatch.setenv("KNNG_JOIN_FUSED", "1")
    x = synth.gaussian_mixture(17, 5000, 1024, 32)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 4, 64, 1337)
    ids, sc = oracle.init_device(q, perm, leaves, 50)
    b = _builder(x, incremental=1)
    b.load_graph(ids, sc)
    fresh = np.full(5000, np.uint64(0xFFFFFFFFFFFFFFFF))
    for it in range(4):
        upd = b.join_iteration()
        ids, sc, o_upd, fresh = oracle.join_device(q, sk, ids, sc, 400, fresh=fresh)
        g_ids, g_sc = b.graph()
        assert upd == o_upd, it
        assert np.array_equal(g_ids, ids) and np.array_equal(g_sc, sc), it
    # rows were passed on to the block-per-row kernel in some iteration while
    # others stayed: both kernels wrote rows
    passed = b.stats()["join_passed_on"]
    if fused:
        assert passed[0] == -1 and all(p >= 0 for p in passed[1:]), passed
    else:
        assert all(0 <= p <= 5000 for p in passed), passed
    assert any(0 < p < 5000 for p in passed), passed


def test_incremental_full_build_bitexact(torch_cuda):
    say("incremental build")
    x = synth.gaussian_mixture(18, 4000, 1024, 16)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, incremental=True)
    b = _builder(x, incremental=1)
    b.build()
    ids, sc = b.graph()
    assert b.stats()["cells"] >= 2   # n // 1536 target, merged
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    assert b.stats()["iterations"] == len(o_rates)


# reverse lists (reverse_width R): the phased join (rows over 512 candidates
# passed to the block-per-row kernel), the fused join (its small-row variant
# under the new/old rule), with and without the new/old rule
@pytest.mark.parametrize("n,d,T,leaf,incremental,R,fused", [
    (4000, 1024, 16, 64, 1, 4, False),
    (3000, 512, 4, 32, 0, 8, False),
    (3000, 384, 8, 64, 1, 3, True),
    (2500, 1024, 4, 2048, 1, 2, True)])
def test_reverse_lists_full_build_bitexact(torch_cuda, monkeypatch, n, d, T, leaf, incremental, R, fused):
    say("reverse lists", n, d, T, leaf, incremental, R, fused)
    if fused:
        monkeypatch.setenv("KNNG_JOIN_FUSED", "1")
    x = synth.gaussian_mixture(70 + R, n, d, 16)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, trees=T, leaf=leaf, incremental=bool(incremental),
                                                         reverse=R)
    b = _builder(x, num_trees=T, leaf_size=leaf, incremental=incremental, reverse_width=R)
    b.build()
    ids, sc = b.graph()
    st = b.stats()
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    assert st["iterations"] == len(o_rates)
    if not fused and T == 16:
        assert max(st["join_passed_on"][:st["iterations"]]) > 0   # k_join ran some rows
    # the reverse candidates change the build
    p_ids, _, _ = oracle.build_device_semantics(q, sk, trees=T, leaf=leaf, incremental=bool(incremental))
    assert not np.array_equal(p_ids, o_ids)
    b.close()


@pytest.mark.parametrize("cells,parts,T", [(24, 2, 8), (9, 3, 3)])
def test_cells_full_build_bitexact(torch_cuda, cells, parts, T):
    """k-means cells seed the forest; the whole build stays bit-exact."""
    say("cells build", cells, parts, T)
    x = synth.gaussian_mixture(19 + cells, 3500, 512, 30)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, trees=T, cells=cells,
                                                         partitions=parts)
    b = _builder(x, num_trees=T, cells=cells, cell_partitions=parts)
    b.build()
    ids, sc = b.graph()
    st = b.stats()
    plan = oracle.cells_device(q, T, 64, cells, partitions=parts)
    assert st["cells"] == int(plan[0].sum()) and st["forest_levels"] >= 2
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    assert st["iterations"] == len(o_rates)


@pytest.mark.parametrize("K,mult", [(64, 8), (10, 3)])
def test_join_other_shapes_bitexact(torch_cuda, K, mult):
    say("join shapes", K, mult)
    x = synth.gaussian_mixture(8, 3000, 768, 16)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 3, 64, 1337)
    ids, sc = oracle.init_device(q, perm, leaves, K)
    b = _builder(x, k_build=K, rerank_mult=mult, k_eval=min(15, K))
    b.load_graph(ids, sc)
    upd = b.join_iteration()
    o_ids, o_sc, o_upd = oracle.join_device(q, sk, ids, sc, K * mult)
    g_ids, g_sc = b.graph()
    assert upd == o_upd
    assert np.array_equal(g_ids, o_ids) and np.array_equal(g_sc, o_sc)


def test_big_leaf_full_build_bitexact(torch_cuda):
    """Cells as leaves of up to 2048 rows (two partitions, two trees): the
    whole build, then the plain export, bit-exact."""
    say("big-leaf build")
    x = synth.gaussian_mixture(23, 6000, 512, 6)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, trees=2, leaf=2048)
    b = _builder(x, num_trees=2, leaf_size=2048)
    b.build()
    ids, sc = b.graph()
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    assert b.stats()["iterations"] == len(o_rates)


# ---------------------------------------------------- whole build + export

@pytest.mark.parametrize("n,rule", [(3000, knng.STOP_INSERTED), (3000, knng.STOP_REFERENCE),
                                    (40, knng.STOP_INSERTED), (2, knng.STOP_INSERTED)])
def test_full_build_bitexact(torch_cuda, n, rule):
    say("full build", n, rule)
    x = synth.gaussian_mixture(9 + n, n, 1024, 8) if n > 100 else make_golden.case_input(15, n, 2, False)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, stop_rule=rule)
    b = _builder(x, stop_rule=rule)
    b.build()
    ids, sc = b.graph()
    st = b.stats()
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    assert st["iterations"] == len(o_rates)
    knns, dists = b.export(15, rerank=False)
    e_knns, e_dists = oracle.export(o_ids, o_sc)
    assert np.array_equal(knns, e_knns) and np.array_equal(dists, e_dists)


# slices of 64 dims (16 / 12 / 5 / 1 passes at d = 1024 / 768 / 320 / 64),
# 128, 256 (the default) and 512 dims (320: a partial last slice; 64, and 320
# in 512-dim slices: a single pass)
@pytest.mark.parametrize("d,slice_dims", [(1024, 64), (768, 64), (320, 64), (64, 64), (1024, 128),
                                          (320, 128), (1024, 256), (768, 256), (320, 256), (64, 256),
                                          (1024, 512), (768, 512), (320, 512)])
def test_rerank_within_fp64_tolerance(torch_cuda, monkeypatch, d, slice_dims):
    say("rerank", d, slice_dims)
    monkeypatch.setenv("KNNG_RERANK_SLICE", str(slice_dims))
    x = synth.gaussian_mixture(10, 4000, d, 16)
    b = _builder(x)
    b.build()
    ids, _ = b.graph()
    knns, dists = b.export(15, rerank=True)
    o_knns, o_dists = oracle.rerank_device(x, ids, 50, 15)
    # |fp32 - fp64| / |fp64| <= 1e-5 for every emitted dist
    X = x.astype(np.float64)
    i = np.repeat(np.arange(4000), 15)
    exact = (X[i] * X[knns.reshape(-1) - 1]).sum(1).reshape(4000, 15)
    rel = np.abs(dists - exact) / np.maximum(np.abs(exact), 1e-30)
    assert rel.max() <= 1e-5, rel.max()
    # same selection as the FP64 oracle except at FP32-indistinguishable ties
    same = (knns == o_knns).all(1)
    if not same.all():
        bad = np.where(~same)[0]
        for r in bad:
            assert np.allclose(np.sort(dists[r]), np.sort(o_dists[r]), rtol=1e-5, atol=0)
    assert same.mean() > 0.99


def test_recall_plain_and_rerank(torch_cuda):
    say("recall n=20000")
    n = 20000
    x = synth.gaussian_mixture(1, n, 1024, 256)
    b = _builder(x, num_trees=8)
    b.build()
    kp, _ = b.export(15, rerank=False)
    kr, _ = b.export(15, rerank=True)
    rows = np.random.default_rng(0).choice(n, 2000, replace=False)
    gt = synth.exact_knn(x, rows, 15)
    rp = synth.recall_at_k(kp, gt, rows, 15)
    rr = synth.recall_at_k(kr, gt, rows, 15)
    say(f"recall plain {rp:.4f} rerank {rr:.4f} stats {b.stats()['iterations']}")
    assert rp > 0.70
    assert rr > 0.95
    assert rr >= rp + 0.01


def test_cli_end_to_end(torch_cuda):
    say("cli")
    x = synth.gaussian_mixture(11, 3000, 1024, 8)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, _ = oracle.build_device_semantics(q, sk)
    e_knns, e_dists = oracle.export(o_ids, o_sc)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.h5")
        knng.h5_write_train(p, x, chunk_rows=1000)
        r = subprocess.run([knng.CLI, "--task", "task1", "--dataset", p, "--chunk-rows", "1100"],
                           cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout
        out = os.path.join(td, "results", "task1", "apex_simd.h5")
        knns = knng.h5_read(out, "knns", np.int32)
        dists = knng.h5_read(out, "dists", np.float32)
        assert np.array_equal(knns, e_knns) and np.array_equal(dists, e_dists)
        assert knng.h5_attr_str(out, "task") == "task1"
        assert knng.h5_attr_str(out, "dists_kind") == "int8_dot"
        assert knng.h5_attr_f32(out, "buildtime") > 0
        assert knng.h5_attr_f32(out, "querytime") == 0.0
        r = subprocess.run([knng.CLI, p, "--rerank", "--task", "t2"], cwd=td,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        out2 = os.path.join(td, "results", "t2", "apex_simd.h5")
        assert knng.h5_attr_str(out2, "dists_kind") == "fp32_dot"
        k2 = knng.h5_read(out2, "knns", np.int32)
        assert k2.shape == (3000, 15) and k2.min() >= 1 and k2.max() <= 3000


def test_large_build_properties(torch_cuda):
    """n = 200,000 x 1024 generated in HBM: size-independent properties."""
    torch = torch_cuda
    n = 200000
    say("large", n)
    xt = synth.gaussian_mixture_torch(2, n, 1024, 1024)
    torch.cuda.synchronize()
    b = knng.Builder(n, _params(keep_fp32=0))
    b.ingest_device(xt.data_ptr(), 0, n)
    b.attach_fp32(xt.data_ptr())
    b.build()
    ids, sc = b.graph()
    assert ids.min() >= 0 and ids.max() < n
    assert not (ids == np.arange(n)[:, None]).any()                    # no self loops
    assert np.all(np.diff(sc, axis=1) <= 0)                            # sorted by score
    srt = np.sort(ids, axis=1)
    assert not (srt[:, 1:] == srt[:, :-1]).any()                       # no duplicate ids
    rows = np.random.default_rng(1).choice(n, 500, replace=False)
    q, _ = b.quantized()
    Q = q.astype(np.int64)
    assert np.array_equal((Q[rows][:, None, :] * Q[ids[rows]]).sum(-1), sc[rows])
    kr, dr = b.export(15, rerank=True)
    X = xt[torch.from_numpy(rows).cuda()].double()
    nb = xt[torch.from_numpy(kr[rows].reshape(-1) - 1).cuda()].double().view(500, 15, 1024)
    exact = (X[:, None, :] * nb).sum(-1).cpu().numpy()
    assert np.abs(dr[rows] - exact).max() / np.abs(exact).min() <= 1e-5


@pytest.mark.parametrize("d,K,mult,T,leaf", [(1024, 64, 8, 4, 64), (768, 10, 3, 3, 200), (512, 50, 8, 16, 64)])
def test_leaf_rule_join_bitexact(torch_cuda, d, K, mult, T, leaf):
    """Join iterations after a leaf init drop candidates that share a leaf with
    the row (leaf-mate rule), in every join shape: the K > 50 kernel, a small
    K with leaves above 64 rows, 16 trees; two iterations bit-exact against
    the oracle's join with the same leaf table."""
    say("leaf rule join", d, K, mult, T, leaf)
    x = synth.gaussian_mixture(90 + K, 4000, d, 40)
    q, sk = oracle.quantize(x)
    b = _builder(x, num_trees=T, leaf_size=leaf, k_build=K, k_eval=min(15, K), rerank_mult=mult, cells=0)
    b.build_forest()
    b.build_init()
    perm, leaves = b.forest()
    ids, sc = b.graph()
    lk = oracle.leaf_keys(perm, leaves)
    for it in range(2):
        upd = b.join_iteration()
        ids, sc, o_upd = oracle.join_device(q, sk, ids, sc, K * mult, leafkey=lk)
        g_ids, g_sc = b.graph()
        assert upd == o_upd, it
        assert np.array_equal(g_ids, ids) and np.array_equal(g_sc, sc), it
    assert b.stats()["join_leafmates"][0] > 0
