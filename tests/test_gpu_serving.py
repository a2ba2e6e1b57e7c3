"""GPU tests of the serving path (knng_server_*): the int8 beam search is
bit-exact against the oracle's restatement, FP32-reranked answers reach high
recall against brute force, every result is the same alone and in a
concurrent mix, and inserts make new rows findable."""
import sys

import numpy as np
import pytest

import knng
import oracle
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _built(x, **kw):
    n, d = x.shape
    b = knng.Builder(n, knng.default_params(dim=d, **kw))
    b.ingest(x)
    b.build()
    return b


@pytest.mark.parametrize("T,leaf", [(4, 32), (2, 2048)])
def test_search_int8_bitexact_vs_oracle(torch_cuda, T, leaf):
    say("search bit-exact", T, leaf)
    x = synth.gaussian_mixture(41, 4000, 256, 16)
    b = _built(x, num_trees=T, leaf_size=leaf, k_build=20, k_eval=15)
    ids, _ = b.graph()
    q, _ = oracle.quantize(x)
    queries = synth.gaussian_mixture(42, 300, 256, 16)
    qq, _ = oracle.quantize(queries)
    with knng.Server(b, beam=32, k=10, rerank=0, entry_trees=4, max_batch=64, max_slots=3) as s:
        g_ids, g_sc = s.search(queries)
    plan = oracle.cells_device(q, T, leaf)   # n = 4000: a 2-cell target, two partitions
    assert plan is not None and b.stats()["cells"] == int(plan[0].sum())
    o_ids, o_sc = oracle.search_device(q, ids, qq, T=T, leaf=leaf, entry_trees=4, beam=32, k=10,
                                       max_steps=128, plan=plan)
    assert np.array_equal(g_ids, o_ids)
    assert np.array_equal(g_sc, o_sc.astype(np.float32))


@pytest.mark.parametrize("d", [1024, 320])
def test_search_many_cells_routing_bitexact(torch_cuda, d):
    """Queries routed through hundreds of cells per partition (several
    8-tile slices of the small-batch routing kernel, partitions of different
    sizes): launches of at most 256 queries (the one-launch routing kernel)
    and one launch of 600 queries (the per-partition assignment kernel) both
    return the oracle's search results over the builder's own cell plan."""
    say("search many cells", d)
    n = 20000
    x = synth.gaussian_mixture(61, n, d, 64)
    b = _built(x, num_trees=2, leaf_size=256, cells=320, cell_partitions=2, k_build=20, k_eval=15)
    ids, _ = b.graph()
    q, _ = oracle.quantize(x)
    queries = synth.gaussian_mixture(62, 600, d, 64)
    qq, _ = oracle.quantize(queries)
    parts = [b.cells(g) for g in range(2)]
    counts = np.array([len(p[1]) for p in parts], np.int32)
    assert counts.min() > 128, counts   # more than 8 tiles of 16 cells: several slices
    plan = (counts, 2, [p[0] for p in parts], [p[1] for p in parts], np.stack([p[2] for p in parts]))
    o_ids, o_sc = oracle.search_device(q, ids, qq, T=2, leaf=256, entry_trees=4, beam=32, k=10,
                                       max_steps=128, plan=plan)
    for max_batch in (600, 100, 7):
        with knng.Server(b, beam=32, k=10, rerank=0, entry_trees=4, max_batch=max_batch, max_slots=2,
                         express_rows=0) as s:
            g_ids, g_sc = s.search(queries)
        assert np.array_equal(g_ids, o_ids), max_batch
        assert np.array_equal(g_sc, o_sc.astype(np.float32)), max_batch
    b.close()


def test_search_recall_rerank(torch_cuda):
    say("search recall")
    n = 20000
    x = synth.gaussian_mixture(43, n + 500, 1024, 64)
    base, queries = x[:n], x[n:]
    b = _built(base)
    with knng.Server(b, beam=64, k=15, rerank=1) as s:
        ids, sc = s.search(queries)
    full = np.vstack([base, queries])
    gt = synth.exact_knn(full, np.arange(n, n + 500), 15)
    rec = np.mean([len(np.intersect1d(ids[j], gt[j])) / 15 for j in range(500)])
    say(f"query recall@15 {rec:.4f}")
    assert rec > 0.9
    # FP32 scores are the exact dot products, best first
    exact = np.einsum("qd,qkd->qk", queries.astype(np.float64), base[ids].astype(np.float64))
    assert np.abs(sc - exact).max() < 1e-5
    assert np.all(np.diff(sc, axis=1) <= 0)


def test_alone_equals_in_the_mix(torch_cuda):
    """Continuous batching: results of requests that ran concurrently with
    others (split into pieces on several streams) equal their results alone."""
    say("alone vs mix")
    x = synth.gaussian_mixture(44, 12000, 512, 32)
    b = _built(x)
    rng = np.random.default_rng(0)
    reqs = [synth.gaussian_mixture(100 + j, int(rng.integers(1, 300)), 512, 32) for j in range(40)]
    with knng.Server(b, beam=48, k=15, rerank=1, max_batch=64, max_slots=4) as s:
        alone = [s.search(r) for r in reqs]
        tickets = [s.submit(r) for r in reqs]         # all in flight together
        states = [s.poll(t) for t in tickets]
        mixed = [s.wait(t, len(r)) for t, r in zip(tickets, reqs)]
    assert all(st in (0, 1, 2) for st in states)
    for (ai, asc), (mi, msc, lat) in zip(alone, mixed):
        assert np.array_equal(ai, mi)
        assert np.array_equal(asc, msc)
        assert lat >= 0


def test_insert_then_find(torch_cuda):
    say("insert")
    x = synth.gaussian_mixture(45, 10000, 1024, 32)
    new = synth.gaussian_mixture(46, 300, 1024, 32)
    b = _built(x)
    with knng.Server(b, beam=64, k=15, rerank=1, insert_capacity=400) as s:
        t = s.submit(new, insert=True)
        ids, _, _ = s.wait(t, 300, insert=True)
        assert ids.tolist() == list(range(10000, 10300))
        assert s.size() == 10300
        got, sc = s.search(new)
        assert (got[:, 0] == np.arange(10000, 10300)).mean() > 0.98   # each finds itself first
        with pytest.raises(knng.KnngError):
            s.wait(s.submit(new, insert=True), 300, insert=True)   # beyond insert_capacity
        assert s.size() == 10300


def test_admission_with_tight_memory_reserve(torch_cuda):
    """A reserve larger than free HBM still admits the first slot, so
    requests complete (one at a time)."""
    say("admission")
    x = synth.gaussian_mixture(47, 5000, 256, 8)
    b = _built(x)
    queries = synth.gaussian_mixture(48, 700, 256, 8)
    with knng.Server(b, beam=32, k=10, max_batch=100, max_slots=8, reserve_bytes=1 << 50) as s:
        ts = [s.submit(queries[j:j + 100]) for j in range(0, 700, 100)]
        outs = [s.wait(t, 100)[0] for t in ts]
        ref, _ = s.search(queries)
    assert np.array_equal(np.vstack(outs), ref)


def test_express_lane_order_and_equality(torch_cuda):
    """Small requests take the express lane (reserved CUs) and return the same
    answers as the bulk lane.  strict_order = 1: a query submitted after an
    insert sees the inserted rows, one submitted before it does not, without
    waiting in between."""
    say("express lane, strict order")
    x = synth.gaussian_mixture(49, 8000, 256, 16)
    new = synth.gaussian_mixture(50, 40, 256, 16)
    b = _built(x)
    with knng.Server(b, beam=32, k=10, rerank=1, max_batch=256, max_slots=2, insert_capacity=64,
                     express_rows=64, express_slots=2, express_cus=16, strict_order=1) as s:
        big, big_sc = s.search(np.vstack([new] * 4))          # 160 rows: bulk lane
        t_before = s.submit(new)                               # 40 rows: express lane
        t_ins = s.submit(new, insert=True)
        t_after = s.submit(new)
        ids_b, sc_b, _, rows_b = s.wait_info(t_before, 40)
        ins_ids, _, _, rows_i = s.wait_info(t_ins, 40, insert=True)
        ids_a, _, _, rows_a = s.wait_info(t_after, 40)
    assert np.array_equal(ids_b, big[:40]) and np.array_equal(sc_b, big_sc[:40])
    assert ids_b.max() < 8000 and rows_b == 8000
    assert ins_ids.tolist() == list(range(8000, 8040)) and rows_i == 8040
    assert rows_a == 8040
    assert (ids_a[:, 0] == np.arange(8000, 8040)).mean() > 0.95   # each finds itself first


def test_concurrent_inserts_beside_queries(torch_cuda):
    """Default mode: inserts run on their own thread and stream while queries
    keep flowing; inserts are applied in submission order, a query reports
    the index size it was dispatched against, and every inserted row becomes
    findable."""
    say("concurrent inserts")
    x = synth.gaussian_mixture(51, 9000, 256, 16)
    new = synth.gaussian_mixture(52, 600, 256, 16)
    qs = synth.gaussian_mixture(53, 500, 256, 16)
    b = _built(x)
    with knng.Server(b, beam=32, k=10, rerank=1, max_batch=128, max_slots=3,
                     insert_capacity=600) as s:
        tickets = []
        for j in range(6):   # interleave inserts of 100 rows with query batches
            tickets.append(("i", s.submit(new[100 * j:100 * (j + 1)], insert=True), 100))
            tickets.append(("q", s.submit(qs[80 * j:80 * j + 80]), 80))
        firsts = []
        for kind, t, m in tickets:
            if kind == "i":
                ids, _, _, rows = s.wait_info(t, m, insert=True)
                assert rows == ids[-1] + 1
                firsts.append(int(ids[0]))
            else:
                _, _, _, rows = s.wait_info(t, m)
                assert 9000 <= rows <= 9600
        assert firsts == list(range(9000, 9600, 100))   # applied in submission order
        assert s.size() == 9600
        got, _ = s.search(new)
        assert (got[:, 0] == np.arange(9000, 9600)).mean() > 0.95


def test_bulk_yield_answers_unchanged(torch_cuda):
    """bulk_yield = 1 holds new bulk launches back while express requests wait
    or run; every request still gets the answer it gets served alone."""
    say("bulk yield")
    x = synth.gaussian_mixture(51, 8000, 256, 16)
    qs = synth.gaussian_mixture(52, 1600, 256, 16)
    b = _built(x)
    with knng.Server(b, beam=32, k=10, rerank=1, max_batch=128, max_slots=3, express_rows=64,
                     express_slots=2, bulk_yield=1) as s:
        reqs = []
        for j in range(0, 1600, 200):   # a 160-row bulk request, then five 8-row express ones
            reqs.append((j, 160, s.submit(qs[j:j + 160])))
            for u in range(5):
                a = j + 160 + 8 * u
                reqs.append((a, 8, s.submit(qs[a:a + 8])))
        got = [(a, m, s.wait(t, m)[:2]) for a, m, t in reqs]
        for a, m, (ids, sc) in got:
            ids1, sc1 = s.search(qs[a:a + m])
            assert np.array_equal(ids, ids1) and np.array_equal(sc, sc1), (a, m)
