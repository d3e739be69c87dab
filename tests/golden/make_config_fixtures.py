"""Writes tests/golden/configs.npz and configs.json: the one-thread oracle
(bit-exact to the reference at one thread) and its fp64 twin on the full-size
configurations, reduced to fixtures small enough to commit.

  python tests/golden/make_config_fixtures.py [--workers 6] [--cases cfg1,cfg2p,...]

Cases (inputs: seeded ellipse phantoms and their Poisson sinograms, generated
on the CPU by trace_amd.phantom; a SHA-256 of every sinogram is recorded so a
test can tell that it rebuilt the same input):

  cfg1    MLEM, 1024^2, 720 angles, 50 iterations, slice 0 of the 16
          (mlem_main.cc:171-182; slices are independent)
  cfg2p   PML, beta 0.1, 2048^2, 1500 angles, 30 iterations, slice 0 of 64
          (pml_main.cc:178-193)
  cfg2a   APMLR, beta0 0.1, delta0 0.01, beta1 = delta1 = regw = 1, compat
          slice offset, 2048^2, 1500 angles, 30 iterations, a 2-slice stack
          (apmlr_main.cc:209-244; APMLR couples slices)
  cfg2a_fixed
          the same with the slice offset corrected (the oracle's restatement
          of a mode the reference does not have)
  cfg3_i  slice i of the 200 uneven slices (trace_amd.batch.uneven_specs,
          seed 3): the first of width 256, 1024, 1536 and 2048, 25 iterations
  stack_mlem, stack_sirt
          MLEM and SIRT, 10 iterations, at the benchmark's geometry (2048^2,
          1500 angles) on its slice-0 phantom (seed 2), noise drawn on the CPU

With --cases the named cases are recomputed and merged into the existing
files.

Per slice the fixture keeps the fp32 and fp64 values at 2048 pixels drawn
by PCG64(20261203 + slice index) and the float64 sum of every image row (a
check that covers every pixel), plus the full-image norms that give the
reference's own error against fp64.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NSAMPLE = 2048
APMLR_PRM = dict(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=1)

# name: (alg, N, P, S, iters, seed, I0, K, params)
CASES = {
    "cfg1": ("mlem", 1024, 720, 1, 50, 1, 1e5, 20, None),
    "cfg2p": ("pml", 2048, 1500, 1, 30, 2, 1e4, 20, dict(beta=0.1)),
    "cfg2a": ("apmlr", 2048, 1500, 2, 30, 2, 1e4, 20, APMLR_PRM),
    # APMLR with the ghost offset corrected (no reference counterpart)
    "cfg2a_fixed": ("apmlr", 2048, 1500, 2, 30, 2, 1e4, 20, dict(APMLR_PRM, apmlr_compat=0)),
    # the benchmark's geometry and algorithms (its slice-0 phantom)
    "stack_mlem": ("mlem", 2048, 1500, 1, 10, 2, 1e4, 20, None),
    "stack_sirt": ("sirt", 2048, 1500, 1, 10, 2, 1e4, 20, None),
}
CFG3_WIDTHS = (256, 1024, 1536, 2048)
CFG3_ITERS = 25


def cfg3_picked():
    """[(index, (C, P, alg))] of the cfg3 slices in the fixture."""
    from trace_amd.batch import uneven_specs
    specs = uneven_specs(200, seed=3)
    out = []
    for w in CFG3_WIDTHS:
        i = next(k for k, s in enumerate(specs) if s[0] == w)
        out.append((i, specs[i]))
    return out


def case_spec(key):
    """(alg, N, P, S, iters, params dict or None) of a fixture key."""
    if key.startswith("cfg3_"):
        i = int(key.split("_")[1])
        C, P, alg = dict(cfg3_picked())[i]
        return alg, C, P, 1, CFG3_ITERS, None
    alg, N, P, S, iters, _, _, _, prm = CASES[key]
    return alg, N, P, S, iters, prm


def case_inputs(key):
    """(theta [P] radians, sinogram [P][S][C] as the algorithm takes it,
    initial image [S][N][N]), built on the CPU.  The preprocessing is the
    mains': MLEM/PML/APMLR take |sinogram| (a no-op here: the clamped
    sinograms are non-negative), PML takes theta as given."""
    from trace_amd.phantom import angles_rad, noisy_sinogram
    if key.startswith("cfg3_"):
        i = int(key.split("_")[1])
        C, P, alg = dict(cfg3_picked())[i]
        th = angles_rad(P)
        sino = noisy_sinogram(th, C, 1, seed=3, I0=1e4, K=20, device="cpu",
                              first_slice=i).numpy()
        init = np.full((1, C, C), 0.0 if alg == "sirt" else 1.0, np.float32)
        return th, sino, init
    alg, N, P, S, iters, seed, I0, K, _ = CASES[key]
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=seed, I0=I0, K=K, device="cpu").numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    return th, sino, init


def sino_sha(sino) -> str:
    return hashlib.sha256(np.ascontiguousarray(sino, np.float32).tobytes()).hexdigest()


def sample_index(N: int, s: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(20261203 + s))
    return np.sort(rng.choice(N * N, NSAMPLE, replace=False)).astype(np.int64)


def reduce_image(x):
    """(samples [S][NSAMPLE], row sums [S][N] float64) of x [S][N][N]."""
    S, N, _ = x.shape
    samp = np.stack([x[s].reshape(-1)[sample_index(N, s)] for s in range(S)])
    rows = np.asarray(x, np.float64).sum(axis=2)
    return samp, rows


def _job(job):
    key, kind = job
    from oracle import Params, load_oracle
    o = load_oracle()
    alg, N, P, S, iters, prm = case_spec(key)
    th, sino, init = case_inputs(key)
    p = Params(**prm) if prm else None
    t0 = time.time()
    if kind == "f32":
        x = o.reconstruct(alg, th, sino, init, N / 2.0, iters, p)
    else:
        x = o.reconstruct_f64(alg, th, sino, init, N / 2.0, iters, p)
    return key, kind, np.asarray(x), sino_sha(sino), time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=6)
    ap.add_argument("--cases", default="")
    a = ap.parse_args()
    from concurrent.futures import ProcessPoolExecutor
    keys = list(CASES) + [f"cfg3_{i}" for i, _ in cfg3_picked()]
    if a.cases:
        keys = [k for k in keys if k in a.cases.split(",")]
    cost = {k: case_spec(k)[1] ** 2 * case_spec(k)[2] * case_spec(k)[3] * case_spec(k)[4]
            for k in keys}
    jobs = sorted([(k, kind) for k in keys for kind in ("f32", "f64")],
                  key=lambda j: -cost[j[0]])
    res = {}
    with ProcessPoolExecutor(a.workers) as ex:
        for key, kind, x, sha, dt in ex.map(_job, jobs):
            print(f"[fixtures] {key} {kind} in {dt:.0f} s", flush=True)
            res[(key, kind)] = (x, sha)
    arrays, meta = {}, {}
    if a.cases and os.path.exists(os.path.join(HERE, "configs.json")):
        old = np.load(os.path.join(HERE, "configs.npz"))
        arrays = {k: old[k] for k in old.files}
        meta = json.load(open(os.path.join(HERE, "configs.json")))["cases"]
    for k in keys:
        x32, sha = res[(k, "f32")]
        x64, _ = res[(k, "f64")]
        alg, N, P, S, iters, prm = case_spec(k)
        s32, r32 = reduce_image(x32)
        s64, r64 = reduce_image(x64)
        arrays[f"{k}_ref"] = s32.astype(np.float32)
        arrays[f"{k}_f64"] = s64.astype(np.float64)
        arrays[f"{k}_rows_ref"] = r32
        arrays[f"{k}_rows_f64"] = r64
        d = np.asarray(x32, np.float64) - np.asarray(x64, np.float64)
        meta[k] = {
            "alg": alg, "N": N, "P": P, "slices": S, "iters": iters, "params": prm,
            "sino_sha256": sha,
            "reference_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(x64)),
            "norm_reference": float(np.linalg.norm(np.asarray(x32, np.float64))),
        }
        print(k, meta[k], flush=True)
    np.savez_compressed(os.path.join(HERE, "configs.npz"), **arrays)
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump({"nsample": NSAMPLE, "cases": meta}, f, indent=1)


if __name__ == "__main__":
    main()
