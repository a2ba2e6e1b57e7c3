"""GPU test of knng_build_batch: every job of a concurrent batch of uneven
builds gives exactly the result of the same build run alone."""
import sys

import numpy as np
import pytest

import knng
import synth

pytestmark = pytest.mark.gpu


def test_batch_equals_single_builds(torch_cuda):
    print("[case] batch", flush=True)
    sys.stdout.flush()
    torch = torch_cuda
    specs = [(5000, 1024, 15, 32), (700, 384, 10, 8), (2500, 768, 32, 16), (64, 128, 15, 2),
             (3000, 256, 10, 64), (1200, 1024, 32, 8)]
    jobs, data = [], []
    for j, (n, d, k, cl) in enumerate(specs):
        x = synth.gaussian_mixture(200 + j, n, d, cl)
        data.append(x)
        p = knng.default_params(dim=d, k_eval=k, k_build=max(50, k), num_trees=8, keep_fp32=1)
        if j % 2:   # half the jobs read device rows
            xt = torch.from_numpy(x).cuda()
            data[-1] = (x, xt)
            rows = xt.data_ptr()
        else:
            rows = x
        jobs.append({"rows": rows, "n": n, "params": p, "rerank": 1})
    torch.cuda.synchronize()
    assert knng.build_batch(jobs, 0, 3) == 0
    for j, (n, d, k, cl) in enumerate(specs):
        x = data[j][0] if isinstance(data[j], tuple) else data[j]
        p = knng.default_params(dim=d, k_eval=k, k_build=max(50, k), num_trees=8, keep_fp32=1)
        with knng.Builder(n, p) as b:
            b.ingest(x)
            b.build()
            kn, di = b.export(k, rerank=True)
        assert jobs[j]["status"] == 0
        assert np.array_equal(jobs[j]["knns"], kn), j
        assert np.array_equal(jobs[j]["dists"], di), j
        assert jobs[j]["iterations"] >= 1 and jobs[j]["build_ms"] > 0


def _single(x, p, k, rerank):
    with knng.Builder(x.shape[0], p) as b:
        b.ingest(x)
        b.build()
        kn, di = b.export(k, rerank=rerank)
        return kn, di, b.stats()["iterations"]


@pytest.mark.parametrize("graph", [True, False])
def test_packed_batch_equals_single_builds(torch_cuda, monkeypatch, graph):
    """Plain-tree jobs with equal build parameters are packed into one context
    (one forest, one init, join iterations over a device work list, each job
    with its own stop rule, hipGraph-replayed iterations): every job's lists
    and iteration count equal its build alone.  Covers tiny jobs (a single
    leaf), k_build 64 (its own group), the reference stop rule and jobs that
    need cells (run whole beside the groups)."""
    print(f"[case] packed batch graph={graph}", flush=True)
    if not graph:
        monkeypatch.setenv("KNNG_BATCH_NO_GRAPH", "1")
    torch = torch_cuda
    sizes = [2, 40, 64, 65, 700, 1500, 2047, 2900, 900, 5000, 1200, 3100]
    ks = [10, 15, 32, 15, 10, 32, 15, 10, 15, 15, 64, 15]
    jobs, data, refs = [], [], []
    for j, (n, k) in enumerate(zip(sizes, ks)):
        x = synth.gaussian_mixture(300 + j, n, 256, 4 + j)
        over = dict(dim=256, k_eval=min(k, 50) if k != 64 else 15, k_build=max(50, k), num_trees=8,
                    keep_fp32=1)
        if j == 8:
            over["stop_rule"] = knng.STOP_REFERENCE
        if j == 11:
            over["cells"] = 2          # needs cells: built alone on a worker
        p = knng.default_params(**over)
        rerank = bool(j % 2)
        refs.append(_single(x, p, p.k_eval, rerank))
        if j % 3 == 0:
            xt = torch.from_numpy(x).cuda()
            data.append(xt)
            rows = xt.data_ptr()
        else:
            data.append(x)
            rows = x
        jobs.append({"rows": rows, "n": n, "params": p, "rerank": int(rerank)})
    torch.cuda.synchronize()
    assert knng.build_batch(jobs, 0, 3) == 0
    for j, (kn, di, its) in enumerate(refs):
        assert jobs[j]["status"] == 0, j
        assert np.array_equal(jobs[j]["knns"], kn), j
        assert np.array_equal(jobs[j]["dists"], di), j
        assert jobs[j]["iterations"] == its, (j, jobs[j]["iterations"], its)
