"""GPU tests at the benchmark workloads: the exact build form bench.py times,
and configurations c1-c5 at their own sizes (100k, 1M and 6.35M rows x 1024,
512 uneven jobs, serving on the 6.35M graph).

Integer results are checked against the oracle (bit-exact) where the oracle
finishes in seconds, and otherwise through properties the size does not
change: valid, distinct, non-self ids; lists sorted by score; every sampled
score equal to the int8 dot product of the quantised rows (src/main.cpp:403-409,
exact int32); recall@15 against exact FP32 brute force as the SISAP evaluator
computes it (scripts/evaluation/evaluate_submission.py:15-26); and FP32
rerank dists within 1e-5 relative of FP64 dot products.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import knng
import oracle
import synth

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5   # FP32-rerank dists against FP64, relative


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


# The form bench.py times at 6.35M rows: 4 trees seeded by 4 k-means cell
# partitions with 2 seeds per target cell, leaves of up to 2048 rows, one full
# join iteration and one under the new/old rule with at most 3 * k_build
# candidates re-scored per row, FP32 rerank of the 28 best.
BENCH_FORM = dict(num_trees=4, leaf_size=2048, cells=-1, cell_partitions=4, cell_oversample=2,
                  incremental=1, max_iters=2, rerank_width=28, rerank_mult=3)


def _fp64_rel_err(torch, x, rows, knns_1b, dists):
    """max |dist - FP64 dot| / |FP64 dot| over the given rows (device tensors)."""
    k = knns_1b.shape[1]
    xs = x[rows].double()
    nb = x[(knns_1b[rows].long().reshape(-1) - 1)].double().view(len(rows), k, -1)
    exact = (xs[:, None, :] * nb).sum(-1)
    return float(((dists[rows].double() - exact).abs() / exact.abs().clamp_min(1e-30)).max())


def _graph_properties(torch, ids, sc, n):
    """valid, distinct, non-self ids; each list sorted by score (device check)."""
    gi = torch.from_numpy(ids).cuda()
    gs = torch.from_numpy(sc).cuda()
    assert int(gi.min()) >= 0 and int(gi.max()) < n
    ar = torch.arange(n, device="cuda", dtype=gi.dtype)[:, None]
    assert not bool((gi == ar).any()), "self loop"
    assert bool((gs[:, 1:] <= gs[:, :-1]).all()), "list not sorted by score"
    srt, _ = gi.sort(dim=1)
    assert not bool((srt[:, 1:] == srt[:, :-1]).any()), "duplicate id in a list"


def _scores_are_int8_dots(q, ids, sc, rows):
    """graph scores of the sampled rows = exact int8 dots (closed form)."""
    Q = q.astype(np.int64)
    got = sc[rows]
    exp = np.einsum("rd,rkd->rk", Q[rows], Q[ids[rows]])
    assert np.array_equal(got, exp)


# ------------------------------------------------------------ bench form

@pytest.fixture(scope="module")
def bench_form_case():
    """n = 8000 x 1024 with 2000 clusters: the init from 5 target cells leaves
    enough to fix that both join iterations run (the second under the new/old
    rule, through the small-row kernel)."""
    n = 8000
    x = synth.gaussian_mixture(77, n, 1024, 2000)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, trees=4, leaf=2048, partitions=4,
                                                         over=2, incremental=True, max_iters=2, mult=3)
    return x, q, sk, o_ids, o_sc, o_rates


@pytest.mark.timeout(600)
def test_bench_form_bitexact(torch_cuda, bench_form_case):
    """The exact form of the headline measurement, whole build bit-exact against
    the oracle's device mode; plain export bit-exact; the width-28 FP32 rerank
    within 1e-5 of FP64 and selecting the oracle's FP64 top 15 of the 28."""
    torch = torch_cuda
    x, q, sk, o_ids, o_sc, o_rates = bench_form_case
    n = x.shape[0]
    say("bench form", n)
    b = knng.Builder(n, knng.default_params(dim=1024, **BENCH_FORM))
    b.ingest(x)
    b.build()
    ids, sc = b.graph()
    st = b.stats()
    assert st["cells"] >= 2 and st["iterations"] == 2 == len(o_rates)
    assert all(0 <= p < n for p in st["join_passed_on"])
    assert np.array_equal(ids, o_ids) and np.array_equal(sc, o_sc)
    knns, dists = b.export(15, rerank=False)
    e_knns, e_dists = oracle.export(o_ids, o_sc)
    assert np.array_equal(knns, e_knns) and np.array_equal(dists, e_dists)
    rk, rd = b.export(15, rerank=True)
    o_rk, o_rd = oracle.rerank_device(x, o_ids, 28, 15)
    X = x.astype(np.float64)
    i = np.repeat(np.arange(n), 15)
    exact = (X[i] * X[rk.reshape(-1) - 1]).sum(1).reshape(n, 15)
    assert (np.abs(rd - exact) / np.abs(exact)).max() <= REL_TOL
    same = (rk == o_rk).all(1)
    for r in np.where(~same)[0]:   # only where FP32 cannot tell FP64 neighbours apart
        assert np.allclose(np.sort(rd[r]), np.sort(o_rd[r]), rtol=REL_TOL, atol=0), r
    assert same.mean() > 0.99
    # the width-28 rerank only ever picks from the list's first 28 entries
    first = np.sort(o_ids[:, :28], axis=1)
    for r in range(0, n, 97):
        assert np.isin(rk[r] - 1, first[r]).all()
    b.close()


@pytest.mark.timeout(600)
def test_cli_bench_form_bitexact(torch_cuda, bench_form_case):
    """The drop-in CLI reaches the measured form (--form chosen), reads /train in
    hyperslabs through its pinned double buffer, and writes the oracle's lists
    (--plain: the int8 export) or FP32-reranked ones (the preset's default)."""
    x, q, sk, o_ids, o_sc, _ = bench_form_case
    n = x.shape[0]
    say("cli bench form", n)
    e_knns, e_dists = oracle.export(o_ids, o_sc)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.h5")
        knng.h5_write_train(p, x, chunk_rows=2000)
        r = subprocess.run([knng.CLI, "--task", "task1", "--dataset", p, "--form", "chosen", "--plain",
                            "--chunk-rows", "2200"], cwd=td, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        out = os.path.join(td, "results", "task1", "apex_simd.h5")
        assert np.array_equal(knng.h5_read(out, "knns", np.int32), e_knns)
        assert np.array_equal(knng.h5_read(out, "dists", np.float32), e_dists)
        params = knng.h5_attr_str(out, "params")
        assert "TREES=4" in params and "LEAF=2048" in params and "PARTITIONS=4" in params
        assert "STOP_NEW_IDS=0.010" in params and "ITERATIONS=2" in params
        assert knng.h5_attr_str(out, "dists_kind") == "int8_dot"
        r = subprocess.run([knng.CLI, "--dataset", p, "--task", "t2", "--form", "chosen"], cwd=td,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        out2 = os.path.join(td, "results", "t2", "apex_simd.h5")
        assert knng.h5_attr_str(out2, "dists_kind") == "fp32_dot"
        rk = knng.h5_read(out2, "knns", np.int32)
        rd = knng.h5_read(out2, "dists", np.float32)
        X = x.astype(np.float64)
        i = np.repeat(np.arange(n), 15)
        exact = (X[i] * X[rk.reshape(-1) - 1]).sum(1).reshape(n, 15)
        assert (np.abs(rd - exact) / np.abs(exact)).max() <= REL_TOL
        assert np.all(np.diff(rd, axis=1) >= 0)   # ascending, best neighbour last


def test_cli_reference_form_stop_rule(torch_cuda):
    """--form reference: plain RP trees and the reference stop rule, bit-exact
    against the oracle; the params attribute names the rule that ran."""
    say("cli reference form")
    x = synth.gaussian_mixture(79, 3000, 1024, 8)
    q, sk = oracle.quantize(x)
    o_ids, o_sc, o_rates = oracle.build_device_semantics(q, sk, cells=0, stop_rule=oracle.STOP_REFERENCE)
    e_knns, e_dists = oracle.export(o_ids, o_sc)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "in.h5")
        knng.h5_write_train(p, x)
        r = subprocess.run([knng.CLI, p, "--form", "reference"], cwd=td, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        out = os.path.join(td, "results", "task1", "apex_simd.h5")
        assert np.array_equal(knng.h5_read(out, "knns", np.int32), e_knns)
        assert np.array_equal(knng.h5_read(out, "dists", np.float32), e_dists)
        params = knng.h5_attr_str(out, "params")
        assert "STOP=0.04, DIFF=0.01" in params and "CELLS=0" in params
        assert f"ITERATIONS={len(o_rates)}" in params
        assert r.stdout.count("-> Iteration") == len(o_rates)


@pytest.mark.parametrize("fused", [False, True])
def test_small_row_join_other_dims_bitexact(torch_cuda, monkeypatch, fused):
    """Join iterations at d != 1024 with a small K (M = K * mult = 30), over four
    iterations under the new/old rule: the wave-per-row phases with rows passed
    to the block-per-row kernel, and (fused) the small-row kernel with rows
    passed to the full one; both kernels write rows, every list bit-exact."""
    say("small-row join d=768 K=10", "fused" if fused else "phased")
    if fused:
        monkeypatch.setenv("KNNG_JOIN_FUSED", "1")
    n = 4000
    x = synth.gaussian_mixture(27, n, 768, 40)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 3, 32, 1337)
    ids, sc = oracle.init_device(q, perm, leaves, 10)
    b = knng.Builder(n, knng.default_params(dim=768, k_build=10, k_eval=10, rerank_mult=3,
                                            incremental=1, cells=0, num_trees=3, leaf_size=32))
    b.ingest(x)
    b.load_graph(ids, sc)
    fresh = np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF))
    for it in range(4):
        upd = b.join_iteration()
        ids, sc, o_upd, fresh = oracle.join_device(q, sk, ids, sc, 30, fresh=fresh)
        g_ids, g_sc = b.graph()
        assert upd == o_upd, it
        assert np.array_equal(g_ids, ids) and np.array_equal(g_sc, sc), it
    passed = b.stats()["join_passed_on"]
    if fused:
        assert passed[0] == -1 and all(0 <= p <= n for p in passed[1:]), passed
        assert any(0 < p < n for p in passed[1:]), passed
    else:   # at most K * K = 100 candidates: every row stays in the phases
        assert all(p == 0 for p in passed), passed


# ------------------------------------------------------------ c1: 100k rows

# Fingerprints of the generated workloads (rows 0, 997, 1994, ... of the float32
# rows, generated on the device by torch's seeded generator): c1, c2 and c3 as
# bench.py and these tests draw them.
WORKLOAD_SHA = {
    (1, 100_000, 256): "9801fe668de6cc4b929676ef1b81384498befaa70f8b6e4cd36e938406d84ffd",
    (2, 1_000_000, 1024): "b762e79a4e65da34acd3aa15018a3d3e46f7b06bebfc7e3c72b9959194881873",
    (3, 6_350_000, 4096): "69a1e1e681ff787482d5224b546c7198cc67953d74cbb763a4dae33e1b5f9eeb",
}


@pytest.mark.parametrize("seed,n,clusters", sorted(WORKLOAD_SHA))
def test_workload_inputs_pinned(torch_cuda, seed, n, clusters):
    """The c1-c3 inputs are the same bytes on every run (strided SHA-256)."""
    torch = torch_cuda
    say("workload sha", seed, n, clusters)
    x = synth.gaussian_mixture_torch(seed, n, 1024, clusters)
    torch.cuda.synchronize()
    assert synth.strided_sha256_torch(x) == WORKLOAD_SHA[(seed, n, clusters)]
    del x
    torch.cuda.empty_cache()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("trees,plain_min,rerank_min", [(8, 0.80, 0.95), (16, 0.82, 0.97)])
def test_c1_plain_forest(torch_cuda, trees, plain_min, rerank_min):
    """c1: 100,000 x 1024 float32, 256 clusters (seed 1), plain RP forest (no
    cells) of 8 and 16 trees with leaf 64, NN-Descent to K = 50, k = 15.
    Quantised rows and sketches bit-exact against the oracle; sampled list
    scores equal the int8 dots; recall@15 on every row against brute force."""
    torch = torch_cuda
    n = 100000
    say("c1", n, trees)
    x = synth.gaussian_mixture_torch(1, n, 1024, 256)
    torch.cuda.synchronize()
    b = knng.Builder(n, knng.default_params(dim=1024, num_trees=trees, leaf_size=64, cells=0,
                                            keep_fp32=0))
    b.ingest_device(x.data_ptr(), 0, n)
    b.attach_fp32(x.data_ptr())
    b.build()
    xh = x.cpu().numpy()
    q, sk = b.quantized()
    q_ref, sk_ref = oracle.quantize(xh)
    assert np.array_equal(q, q_ref) and np.array_equal(sk, sk_ref)
    ids, sc = b.graph()
    _graph_properties(torch, ids, sc, n)
    rows = np.sort(np.random.default_rng(11).choice(n, 2000, replace=False))
    _scores_are_int8_dots(q, ids, sc, rows)
    kp, dp = b.export(15, rerank=False)
    kr, dr = b.export(15, rerank=True)
    assert np.array_equal(dp[rows], sc[rows, :15][:, ::-1].astype(np.float32))
    gt = torch.empty((n, 15), dtype=torch.int64)
    for a in range(0, n, 10000):
        r = torch.arange(a, min(n, a + 10000), device="cuda")
        gt[a:a + len(r)] = synth.exact_knn_torch(x, r, 15).cpu()
    gt = gt.numpy()
    rp = synth.recall_rows(kp, gt, 15)
    rr = synth.recall_rows(kr, gt, 15)
    st = b.stats()
    say(f"c1 trees {trees}: recall plain {rp:.4f} rerank {rr:.4f}, {st['iterations']} iterations, "
        f"build {st['ms_build']:.1f} ms")
    assert rp >= plain_min and rr >= rerank_min
    err = _fp64_rel_err(torch, x, torch.from_numpy(rows).cuda(), torch.from_numpy(kr).cuda(),
                        torch.from_numpy(dr).cuda())
    assert err <= REL_TOL
    b.close()


# ------------------------------------------------------------ c2: 1M rows

@pytest.mark.timeout(600)
def test_c2_bench_form(torch_cuda):
    """c2: 1,000,000 x 1024, 1024 clusters (seed 2), the bench form: graph
    properties, sampled int8 scores, recall@15 on 10,000 sampled rows, FP32
    rerank dists against FP64."""
    torch = torch_cuda
    n = 1_000_000
    say("c2", n)
    x = synth.gaussian_mixture_torch(2, n, 1024, 1024)
    torch.cuda.synchronize()
    b = knng.Builder(n, knng.default_params(dim=1024, keep_fp32=0, **BENCH_FORM))
    b.ingest_device(x.data_ptr(), 0, n)
    b.attach_fp32(x.data_ptr())
    b.build()
    ids, sc = b.graph()
    _graph_properties(torch, ids, sc, n)
    q, _ = b.quantized()
    _scores_are_int8_dots(q, ids, sc, np.random.default_rng(12).choice(n, 500, replace=False))
    kp, _ = b.export(15, rerank=False)
    kr, dr = b.export(15, rerank=True)
    rows_np = np.sort(np.random.default_rng(12345).choice(n, 10000, replace=False))
    rows = torch.from_numpy(rows_np).cuda()
    gt = synth.exact_knn_torch(x, rows, 15).cpu().numpy()
    rp = synth.recall_rows(kp[rows_np], gt, 15)
    rr = synth.recall_rows(kr[rows_np], gt, 15)
    err = _fp64_rel_err(torch, x, rows, torch.from_numpy(kr).cuda(), torch.from_numpy(dr).cuda())
    st = b.stats()
    say(f"c2: recall plain {rp:.4f} rerank {rr:.4f}, dist err {err:.2e}, build {st['ms_build']:.1f} ms")
    assert rp >= 0.78 and rr >= 0.92 and err <= REL_TOL
    b.close()


# ------------------------------------------------------------ c4: 512 jobs

@pytest.mark.timeout(900)
def test_c4_batched_uneven_builds(torch_cuda):
    """c4: 512 independent builds, n log-uniform in [2000, 200000], d from
    {384, 768, 1024}, 16-512 clusters, k from {10, 15, 32}, rows in HBM,
    packed per dimension.  Per-job recall@k on 200 sampled rows; the three
    smallest jobs export plain int8 lists that must equal the oracle's
    device-mode build bit for bit."""
    torch = torch_cuda
    rng = np.random.default_rng(4)
    jobs, data = [], []
    for j in range(512):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(200000))))
        d = int(rng.choice([384, 768, 1024]))
        cl = int(np.exp(rng.uniform(np.log(16), np.log(512))))
        k = int(rng.choice([10, 15, 32]))
        x = synth.gaussian_mixture_torch(4 * (1 << 20) + j, n, d, cl)
        data.append(x)
        p = knng.default_params(dim=d, k_eval=k, k_build=max(50, k), num_trees=8, keep_fp32=0, cells=0)
        jobs.append({"rows": x.data_ptr(), "n": n, "params": p, "rerank": 1, "k": k, "d": d})
    oracle_jobs = sorted(range(512), key=lambda j: jobs[j]["n"])[:3]
    for j in oracle_jobs:
        jobs[j]["rerank"] = 0
    torch.cuda.synchronize()
    say("c4", sum(j["n"] for j in jobs), "rows")
    assert knng.build_batch(jobs, 0, 8) == 0
    recs = []
    for j, job in enumerate(jobs):
        assert job["status"] == 0, j
        x = data[j]
        rows = np.sort(np.random.default_rng(j).choice(job["n"], min(200, job["n"]), replace=False))
        gt = synth.exact_knn_torch(x, torch.from_numpy(rows).cuda(), job["k"]).cpu().numpy()
        recs.append(synth.recall_rows(job["knns"][rows], gt, job["k"]))
    recs = np.array(recs)
    say(f"c4 recall@k mean {recs.mean():.4f} min {recs.min():.4f}")
    assert recs.mean() >= 0.9 and recs.min() >= 0.6
    for j in oracle_jobs:
        job = jobs[j]
        xh = data[j].cpu().numpy()
        q, sk = oracle.quantize(xh)
        o_ids, o_sc, _ = oracle.build_device_semantics(q, sk, trees=8, leaf=64, K=max(50, job["k"]),
                                                       cells=0)
        e_knns, e_dists = oracle.export(o_ids, o_sc, job["k"])
        assert np.array_equal(job["knns"], e_knns), j
        assert np.array_equal(job["dists"], e_dists), j
    del data
    torch.cuda.empty_cache()


# ------------------------------------------------- c3 (6.35M) and c5 (serving)

C3_N = 6_350_000


@pytest.fixture(scope="module")
def c3_build(torch_cuda):
    """c3: 6,350,000 x 1024 float32 (26 GB), 4096 clusters (seed 3), generated
    in HBM and built in the bench form; shared by the c3 and c5 tests."""
    torch = torch_cuda
    say("c3 generate + build", C3_N)
    x = synth.gaussian_mixture_torch(3, C3_N, 1024, 4096)
    torch.cuda.synchronize()
    b = knng.Builder(C3_N, knng.default_params(dim=1024, keep_fp32=0, **BENCH_FORM))
    b.ingest_device(x.data_ptr(), 0, C3_N)
    b.attach_fp32(x.data_ptr())
    b.build()
    b.synchronize()
    yield x, b
    b.close()
    del x
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_c3_bench_form(torch_cuda, c3_build):
    """c3 at full size: graph properties over all 6.35M lists, sampled int8
    scores, recall@15 >= 0.85 of the FP32-rerank export on 10,000 sampled rows
    (the headline's quality bar), plain-export recall reported, FP32 dists
    within 1e-5 of FP64."""
    torch = torch_cuda
    x, b = c3_build
    n = C3_N
    say("c3 checks")
    ids, sc = b.graph()
    _graph_properties(torch, ids, sc, n)
    q, _ = b.quantized()
    _scores_are_int8_dots(q, ids, sc, np.random.default_rng(13).choice(n, 300, replace=False))
    del q
    kp = torch.empty((n, 15), dtype=torch.int32, device="cuda")
    dp = torch.empty((n, 15), dtype=torch.float32, device="cuda")
    kr = torch.empty_like(kp)
    dr = torch.empty_like(dp)
    b.export_device(kp.data_ptr(), dp.data_ptr(), 15, False)
    b.export_device(kr.data_ptr(), dr.data_ptr(), 15, True)
    b.synchronize()
    rows_np = np.sort(np.random.default_rng(12345).choice(n, 10000, replace=False))
    rows = torch.from_numpy(rows_np).cuda()
    gt = synth.exact_knn_torch(x, rows, 15).cpu().numpy()
    rp = synth.recall_rows(kp[rows].cpu().numpy(), gt, 15)
    rr = synth.recall_rows(kr[rows].cpu().numpy(), gt, 15)
    err = _fp64_rel_err(torch, x, rows, kr, dr)
    st = b.stats()
    say(f"c3: recall plain {rp:.4f} rerank {rr:.4f}, dist err {err:.2e}, build {st['ms_build']:.1f} ms, "
        f"{st['iterations']} iterations")
    # plain: bounded by the int8 quantiser's own top-15 (about 0.84 on this
    # cluster density even for an exact int8 graph)
    assert rp >= 0.74
    assert rr >= 0.85 and err <= REL_TOL
    # the plain export is the graph's first 15 entries, reversed (src/main.cpp:449-457)
    r0 = rows_np[:200]
    assert np.array_equal(kp[torch.from_numpy(r0).cuda()].cpu().numpy(), ids[r0, :15][:, ::-1] + 1)


@pytest.mark.timeout(900)
def test_c5_serving_on_c3_graph(torch_cuda, c3_build):
    """c5: the server over the 6.35M graph under a Poisson stream at 2,000
    req/s (query batches log-uniform in [1, 4096], beam 64, k 15, FP32
    rerank).  Phase 1, queries only: every sampled request's answer in the mix
    equals its answer alone.  Phase 2, with 10% inserts of 1-1024 rows:
    inserts get consecutive ids in submission order and find themselves, and
    query recall@15 against brute force over the rows present when the query
    ran stays high."""
    import threading
    import time
    torch = torch_cuda
    x, b = c3_build
    n, d = C3_N, 1024
    say("c5 serving")
    rng = np.random.default_rng(7)

    def stream(seconds, insert_frac):
        reqs, t = [], 0.0
        while True:
            t += rng.exponential(1.0 / 2000.0)
            if t > seconds:
                return reqs
            if rng.random() < insert_frac:
                reqs.append((t, 1, int(rng.integers(1, 1025))))
            else:
                reqs.append((t, 0, int(np.exp(rng.uniform(0.0, np.log(4096.0))))))

    ph1 = stream(0.25, 0.0)
    ph2 = stream(0.25, 0.1)
    nq = sum(r[2] for r in ph1 + ph2 if r[1] == 0)
    n_ins = sum(r[2] for r in ph2 if r[1] == 1)
    # held-out rows of the same mixture: the c3 centres, another draw stream
    extra = synth.gaussian_mixture_torch(3, nq + n_ins, d, 4096, draw_seed=5)
    eh = extra.cpu().numpy()
    qrows, irows = eh[:nq], eh[nq:]
    sp = dict(beam=64, k=15, rerank=1, entry_trees=4, max_batch=2048, max_slots=4, express_rows=64,
              express_slots=2, insert_capacity=n_ins + 16)

    def replay(srv, plan):
        tickets = [None] * len(plan)
        start = time.time()

        def submitter():
            for j, (tt, kind, rows) in enumerate(plan):
                delay = start + tt - time.time()
                if delay > 0:
                    time.sleep(delay)
                tickets[j] = srv.submit(rows, insert=bool(kind))

        th = threading.Thread(target=submitter)
        th.start()
        out = []
        for j, (tt, kind, rows) in enumerate(plan):
            while tickets[j] is None:
                time.sleep(0.0002)
            out.append(srv.wait_info(tickets[j], len(rows), insert=bool(kind)))
        th.join()
        return out

    with knng.Server(b, **sp) as srv:
        srv.search(qrows[:64])   # warm-up
        qo = 0
        plan1 = []
        for tt, kind, sz in ph1:
            plan1.append((tt, 0, qrows[qo:qo + sz]))
            qo += sz
        res1 = replay(srv, plan1)
        pick = np.random.default_rng(1).choice(len(plan1), min(40, len(plan1)), replace=False)
        for j in pick:
            ids_alone, sc_alone = srv.search(plan1[j][2])
            assert np.array_equal(res1[j][0], ids_alone), j
            assert np.array_equal(res1[j][1], sc_alone), j
        small = [r[2] for (tt, kind, rows), r in zip(plan1, res1) if len(rows) <= 64]
        say(f"c5 phase 1: {len(plan1)} requests, p99 batches <= 64 {np.percentile(small, 99):.2f} ms")
        plan2, io = [], 0
        for tt, kind, sz in ph2:
            if kind:
                plan2.append((tt, 1, irows[io:io + sz]))
                io += sz
            else:
                plan2.append((tt, 0, qrows[qo:qo + sz]))
                qo += sz
        res2 = replay(srv, plan2)
        new_ids = np.concatenate([r[0] for (tt, kind, rows), r in zip(plan2, res2) if kind])
        assert np.array_equal(new_ids, np.arange(n, n + n_ins))
        assert srv.size() == n + n_ins
        got, _ = srv.search(irows[:512])
        assert (got[:, 0] == np.arange(n, n + min(512, n_ins))).mean() > 0.95
    # recall of sampled queries of phase 2 against the rows present when they ran
    samp = [(rows[0], r[3], r[0][0]) for (tt, kind, rows), r in zip(plan2, res2) if not kind]
    samp = samp[:: max(1, len(samp) // 300)]
    qv = torch.from_numpy(np.stack([s[0] for s in samp])).cuda()
    lim = torch.tensor([s[1] for s in samp], device="cuda")
    gt = synth.exact_knn_prefix_torch([x, extra[nq:]], qv, lim, 15).cpu().numpy()
    rec = np.mean([len(np.intersect1d(s[2], gt[j])) / 15 for j, s in enumerate(samp)])
    say(f"c5 phase 2: {len(plan2)} requests, {n_ins} inserted rows, query recall@15 {rec:.4f}")
    assert rec >= 0.8
