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
