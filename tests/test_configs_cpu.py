"""The full-size configuration fixtures (tests/golden/configs.npz, .json) on
the CPU: their layout, that the inputs rebuild to the recorded sinograms, and
that the smallest case (the 256^2 uneven slice, 959 angles, 25 SIRT
iterations) is what the oracle computes today, bit for bit."""
import json
import os
import sys

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_config_fixtures as fx  # noqa: E402


def _load():
    meta = json.load(open(os.path.join(GOLD, "configs.json")))
    return meta, np.load(os.path.join(GOLD, "configs.npz"))


def test_fixture_layout():
    meta, arr = _load()
    keys = list(fx.CASES) + [f"cfg3_{i}" for i, _ in fx.cfg3_picked()]
    assert sorted(meta["cases"]) == sorted(keys)
    for k in keys:
        m = meta["cases"][k]
        alg, N, P, S, iters, prm = fx.case_spec(k)
        assert (m["alg"], m["N"], m["P"], m["slices"], m["iters"]) == (alg, N, P, S, iters)
        assert arr[f"{k}_ref"].shape == (S, meta["nsample"]) and arr[f"{k}_ref"].dtype == np.float32
        assert arr[f"{k}_f64"].shape == (S, meta["nsample"])
        assert arr[f"{k}_rows_ref"].shape == (S, N)
        # the reference's own fp32 error: small, and PML's the largest
        assert 0 < m["reference_vs_fp64"] < 1e-3
        assert np.isfinite(arr[f"{k}_ref"]).all()


@pytest.mark.parametrize("key", ["cfg1", "cfg3_1", "cfg3_7"])
def test_inputs_rebuild(key):
    meta, _ = _load()
    th, sino, init = fx.case_inputs(key)
    assert fx.sino_sha(sino) == meta["cases"][key]["sino_sha256"]


def test_smallest_case_is_the_oracle(oracle):
    meta, arr = _load()
    key = next(k for k in meta["cases"] if meta["cases"][k]["N"] == 256)
    alg, N, P, S, iters, _ = fx.case_spec(key)
    th, sino, init = fx.case_inputs(key)
    x = oracle.reconstruct(alg, th, sino, init, N / 2.0, iters)
    samp, rows = fx.reduce_image(np.asarray(x))
    assert np.array_equal(samp.view(np.uint32), arr[f"{key}_ref"].view(np.uint32))
    assert np.array_equal(rows, arr[f"{key}_rows_ref"])
