"""GPU test of knng_join_iteration_noalloc, the join iteration that allocates
nothing (its counters live in a static device array): through the C-ABI it
must give exactly the lists and change count of knng_join_iteration on the
same graph, on two calls in a row, the second on a smaller input."""
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


def _case(seed, n, d, K, mult, clusters):
    x = synth.gaussian_mixture(seed, n, d, clusters)
    q, sk = oracle.quantize(x)
    perm, leaves = oracle.forest_device(q, 3, 64, 1337)   # a weak forest leaves work to the join
    ids, sc = oracle.init_device(q, perm, leaves, K)
    return x, ids, sc


def _with_context(x, ids, sc, K, mult):
    n, d = x.shape
    with knng.Builder(n, knng.default_params(dim=d, k_build=K, rerank_mult=mult, k_eval=min(15, K),
                                             cells=0)) as b:
        b.ingest(x)
        b.load_graph(ids, sc)
        upd = b.join_iteration()
        g_ids, g_sc = b.graph()
    return g_ids, g_sc, upd


def _noalloc(torch, x, ids, sc, K, mult, order=None):
    n, d = x.shape
    dx = torch.from_numpy(x).cuda()
    dq = torch.empty((n, d), dtype=torch.int8, device="cuda")
    dsk = torch.empty((n, d // 64), dtype=torch.int64, device="cuda")
    knng.quantize_device(dx.data_ptr(), n, d, dq.data_ptr(), dsk.data_ptr())
    di = torch.from_numpy(ids).cuda()
    ds = torch.from_numpy(sc).cuda()
    oi = torch.full_like(di, -7)
    os_ = torch.full_like(ds, -7)
    counts = torch.zeros(6, dtype=torch.int64, device="cuda")
    dord = torch.from_numpy(order.astype(np.int32)).cuda() if order is not None else None
    torch.cuda.synchronize()
    knng.join_iteration_noalloc(dq.data_ptr(), dsk.data_ptr(), d, n, K, mult, di.data_ptr(), ds.data_ptr(),
                                oi.data_ptr(), os_.data_ptr(), counts.data_ptr(),
                                dord.data_ptr() if dord is not None else None)
    torch.cuda.synchronize()
    return oi.cpu().numpy(), os_.cpu().numpy(), counts.cpu().numpy()


def test_join_noalloc_equals_join_iteration(torch_cuda):
    torch = torch_cuda
    cases = [(7, 6000, 1024, 50, 8, 32), (8, 2500, 512, 20, 4, 16)]   # the second call is smaller
    for j, (seed, n, d, K, mult, cl) in enumerate(cases):
        say("join noalloc", n, d, K, mult)
        x, ids, sc = _case(seed, n, d, K, mult, cl)
        e_ids, e_sc, e_upd = _with_context(x, ids, sc, K, mult)
        order = np.random.default_rng(seed).permutation(n) if j == 0 else None
        g_ids, g_sc, counts = _noalloc(torch, x, ids, sc, K, mult, order)
        assert np.array_equal(g_ids, e_ids), j
        assert np.array_equal(g_sc, e_sc), j
        assert int(counts[0]) == e_upd, j
        assert counts[1] > 0 and 0 < counts[2] <= counts[1] and counts[5] == 0, j
        # and the oracle agrees
        o_ids, o_sc, o_upd = oracle.join_device(*oracle.quantize(x), ids, sc, K * mult)
        assert np.array_equal(g_ids, o_ids) and np.array_equal(g_sc, o_sc) and o_upd == e_upd


def test_join_noalloc_rejects_sizes_beyond_its_limits(torch_cuda):
    torch = torch_cuda
    buf = torch.zeros(64, dtype=torch.int64, device="cuda")
    p = buf.data_ptr()
    with pytest.raises(knng.KnngError):   # d not a multiple of 64
        knng.join_iteration_noalloc(p, p, 1000, 4, 2, 2, p, p, p + 8, p + 16, p + 24)
    with pytest.raises(knng.KnngError):   # k_build above 64
        knng.join_iteration_noalloc(p, p, 64, 4, 65, 2, p, p, p + 8, p + 16, p + 24)
    with pytest.raises(knng.KnngError):   # outputs alias the inputs
        knng.join_iteration_noalloc(p, p, 64, 4, 2, 2, p, p + 8, p, p + 16, p + 24)
