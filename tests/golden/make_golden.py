"""Write the golden fixtures of the reference engine (run in the build
container, where the reference sources are mounted).

For each case the input rows come from synth.fixture_rows (integer-exact,
SHA-256 pinned), are written to an HDF5 /train file, and the compiled
reference (oracle/_ref/apex_main_native, built by oracle/Makefile from
src/main.cpp with the reference's own flags) is run at OMP_NUM_THREADS=1,
where it is deterministic.  Its knns/dists and attributes are stored as
fixtures; the inputs are regenerated from the seed by the tests.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "sisap-2026-task1_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import knng  # noqa: E402
import oracle  # noqa: E402
import synth  # noqa: E402

# (name, seed, n, clusters, edge rows, rows kept in full (None = all))
CASES = [
    ("ref_n257", 11, 257, 8, False, None),
    ("ref_n2000", 12, 2000, 32, False, None),
    ("ref_edge_n300", 13, 300, 8, True, None),
    ("ref_n20000", 14, 20000, 128, False, 200),
    ("ref_n40", 15, 40, 4, False, None),      # fewer than K_BUILD rows: xorshift padding
    ("ref_n2", 16, 2, 1, False, None),
]

EDGE_VALUES = np.array([1.5, -1.5, 1.0, -1.0, 0.5 / 127, -0.5 / 127, 126.9 / 127, -128.5 / 127,
                        0.0, -0.0, 1e-40, -1e-40, 127.0 / 127, -128.0 / 127, 0.99 / 127, 3.3 / 127],
                       dtype=np.float32)


def case_input(seed, n, clusters, edge):
    x = synth.fixture_rows(seed, n, 1024, clusters)
    if edge:
        for r in range(8):
            pat = np.roll(EDGE_VALUES, r)
            x[r] = np.tile(pat, 1024 // len(pat))
            x[r, r::17] = x[(r + 8) % n, r::17]     # keep some overlap with normal rows
    return x


def main():
    oracle.build()
    out = {}
    for name, seed, n, clusters, edge, keep in CASES:
        x = case_input(seed, n, clusters, edge)
        with tempfile.TemporaryDirectory() as td:
            h5p = os.path.join(td, "in.h5")
            knng.h5_write_train(h5p, x)
            r = oracle.run_reference_binary(h5p, td, threads=1, native=True)
            assert r["returncode"] == 0, r["stdout"]
            knns = knng.h5_read(r["out"], "knns", np.int32)
            dists = knng.h5_read(r["out"], "dists", np.float32)
            attrs = {a: knng.h5_attr_str(r["out"], a) for a in ("algo", "task", "params")}
            attrs["querytime"] = knng.h5_attr_f32(r["out"], "querytime")
        rates = [float(v) for v in re.findall(r"\(Rate: ([0-9.eE+-]+)%\)", r["stdout"])]
        meta = {"seed": seed, "n": n, "d": 1024, "clusters": clusters, "edge_rows": edge,
                "input_sha256": synth.sha256(x), "knns_sha256": synth.sha256(knns),
                "dists_sha256": synth.sha256(dists), "rates_percent": rates, "attrs": attrs}
        if keep is None:
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), knns=knns, dists=dists)
        else:
            rows = np.sort(np.random.default_rng(seed).choice(n, keep, replace=False))
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), rows=rows, knns=knns[rows],
                                dists=dists[rows])
            meta["rows_kept"] = keep
        out[name] = meta
        print(name, meta["input_sha256"][:16], "iterations", len(rates), flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
