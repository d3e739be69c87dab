"""Error of the GPU reconstructions against the fp64 host oracle, per
configuration (not collected by pytest; run by hand).

  python tests/fp64_report.py gpu    --out gpurun_out/fp64 [--cases cfg0,cfg1,...]  (GPU box)
  python tests/fp64_report.py oracle --gpu gpurun_out/fp64 --out tests/golden/fp64_errors.json

Inputs are generated on the CPU (seeded ellipse phantoms, CPU Poisson
stream), so both phases see identical sinograms.  Cases (iterations as in
the configurations; slices are independent for SIRT/MLEM/PML, so a slice
alone equals the same slice inside its stack - tested in
test_gpu_recon.py::test_slice_independence):

  cfg0   SIRT, 256^2, 180 angles, 20 iterations, 1 slice (the whole case)
  cfg1   MLEM, 1024^2, 720 angles, 50 iterations, slices 0 and 1 of 16
  cfg2p  PML (beta 0.1), 2048^2, 1500 angles, 30 iterations, slice 0 of 64
  cfg2a  APMLR (beta 0.1, delta 0.01), 2048^2, 1500 angles, 30 iterations,
         a 2-slice stack (APMLR couples slices; both are boundary slices)
  cfg3   4 of the 200 uneven slices (widths 256, 512, 1024, 1536), 25 iterations,
         through the batched graph path

The oracle phase runs the fp32 restatement (bit-exact to the reference at
one thread) and the fp64 twin for every case, in parallel processes, and
reports relative L2 errors: GPU vs fp64, reference vs fp64, GPU vs
reference.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# name: (alg, N, P, S, iters, seed, I0, K, params)
CASES = {
    "cfg0": ("sirt", 256, 180, 1, 20, 0, 1e4, 10, None),
    "cfg1": ("mlem", 1024, 720, 2, 50, 1, 1e5, 20, None),
    "cfg2p": ("pml", 2048, 1500, 1, 30, 2, 1e4, 20, dict(beta=0.1)),
    "cfg2a": ("apmlr", 2048, 1500, 2, 30, 2, 1e4, 20,
              dict(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=1)),
}


def cfg3_specs():
    from trace_amd.batch import uneven_specs
    specs = uneven_specs(200, seed=3)
    picked = []
    for w in (256, 512, 1024, 1536):
        i = next(k for k, s in enumerate(specs) if s[0] == w)
        picked.append((i, specs[i]))
    return picked


def case_inputs(name):
    from trace_amd.phantom import angles_rad, noisy_sinogram
    alg, N, P, S, iters, seed, I0, K, prm = CASES[name]
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=seed, I0=I0, K=K, device="cpu").numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    return th, sino, init


def cfg3_inputs(i, spec):
    from trace_amd.phantom import angles_rad, noisy_sinogram
    C, P, alg = spec
    th = angles_rad(P)
    sino = noisy_sinogram(th, C, 1, seed=3, I0=1e4, K=20, device="cpu", first_slice=i).numpy()
    init = np.full((1, C, C), 0.0 if alg == "sirt" else 1.0, np.float32)
    return th, sino, init


def run_gpu(out, cases):
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.batch import Batch, BatchSlice
    os.makedirs(out, exist_ok=True)
    for name, (alg, N, P, S, iters, seed, I0, K, prm) in CASES.items():
        if name not in cases:
            continue
        print(f"[fp64] gpu {name}: {alg} {S}x{N}^2, {P} angles, {iters} iterations", flush=True)
        th, sino, init = case_inputs(name)
        rp = RegParams(**prm) if prm else None
        with Projector(th, S, N, N / 2.0) as pj:
            ds = torch.from_numpy(sino).cuda()
            dx = torch.from_numpy(init).cuda()
            pj.reconstruct(alg, iters, ds, dx, rp)
            torch.cuda.synchronize()
            np.save(os.path.join(out, f"{name}.npy"), dx.cpu().numpy())
    if "cfg3" not in cases:
        return
    slices = []
    for i, spec in cfg3_specs():
        th, sino, init = cfg3_inputs(i, spec)
        C = spec[0]
        slices.append(BatchSlice(th, torch.from_numpy(sino[:, 0, :].copy()).cuda(),
                                 torch.from_numpy(init[0].copy()).cuda(), spec[2], 25, C / 2.0))
    print("[fp64] gpu cfg3: 4 uneven slices through the batched path", flush=True)
    with Batch(slices) as b:
        b.run()
        torch.cuda.synchronize()
    for (i, spec), s in zip(cfg3_specs(), slices):
        np.save(os.path.join(out, f"cfg3_{i}.npy"), s.recon.cpu().numpy()[None])
    print("[fp64] gpu done", flush=True)


def _oracle_job(job):
    """(key, alg, kind, inputs-spec) -> (key, kind, image); runs in a worker."""
    key, kind = job
    from oracle import Params, load_oracle
    o = load_oracle()
    if key.startswith("cfg3_"):
        i = int(key.split("_")[1])
        spec = dict(cfg3_specs())[i]
        th, sino, init = cfg3_inputs(i, spec)
        alg, iters, prm, N = spec[2], 25, None, spec[0]
    else:
        alg, N, P, S, iters, seed, I0, K, prm = CASES[key]
        th, sino, init = case_inputs(key)
    p = Params(**prm) if prm else None
    t0 = time.time()
    if kind == "f32":
        x = o.reconstruct(alg, th, sino, init, N / 2.0, iters, p)
    else:
        x = o.reconstruct_f64(alg, th, sino, init, N / 2.0, iters, p)
    return key, kind, np.asarray(x, np.float64), time.time() - t0


def run_oracle(gpu_dir, out, workers):
    from concurrent.futures import ProcessPoolExecutor
    keys = list(CASES) + [f"cfg3_{i}" for i, _ in cfg3_specs()]
    jobs = [(k, kind) for k in keys for kind in ("f64", "f32")]
    # longest first
    cost = {"cfg2a": 6, "cfg2p": 4, "cfg1": 3}
    jobs.sort(key=lambda j: -cost.get(j[0], 1))
    res = {}
    with ProcessPoolExecutor(workers) as ex:
        for key, kind, x, dt in ex.map(_oracle_job, jobs):
            print(f"[fp64] oracle {key} {kind} in {dt:.0f} s", flush=True)
            res[(key, kind)] = x

    def rel(a, b):
        a = np.asarray(a, np.float64).ravel()
        b = np.asarray(b, np.float64).ravel()
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    report = {}
    for k in keys:
        g = np.load(os.path.join(gpu_dir, f"{k}.npy")).astype(np.float64)
        f64, f32 = res[(k, "f64")], res[(k, "f32")]
        report[k] = {"gpu_vs_fp64": rel(g, f64), "reference_vs_fp64": rel(f32, f64),
                     "gpu_vs_reference": rel(g, f32)}
        if k.startswith("cfg3_"):
            report[k]["slice"] = dict(cfg3_specs())[int(k.split("_")[1])]
        else:
            alg, N, P, S, iters = CASES[k][:5]
            report[k]["case"] = {"alg": alg, "N": N, "P": P, "slices": S, "iters": iters}
        print(k, report[k], flush=True)
    with open(out, "w") as f:
        json.dump(report, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gpu")
    g.add_argument("--out", required=True)
    g.add_argument("--cases", default="cfg0,cfg1,cfg2p,cfg2a,cfg3")
    o = sub.add_parser("oracle")
    o.add_argument("--gpu", required=True)
    o.add_argument("--out", required=True)
    o.add_argument("--workers", type=int, default=8)
    a = ap.parse_args()
    if a.cmd == "gpu":
        run_gpu(a.out, set(a.cases.split(",")))
    else:
        run_oracle(a.gpu, a.out, a.workers)


if __name__ == "__main__":
    main()
