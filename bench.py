#!/usr/bin/env python3
"""Throughput of the iterative-reconstruction hot path on one MI355X.

Workload (default): MLEM on a stack of 64 slices of 2048 x 2048 fp32 seeded
ellipse phantoms (20 ellipses each, seed 2), 1500 angles evenly spaced in
[0, pi), 2048 detector columns, Poisson noise at 1e4 incident counts --
the 2048^2 / 1500-angle geometry the metric is quoted on.  Inputs are
resident in HBM before timing starts.

A step is one full iteration over the whole stack (zero accumulators, the
ray pass A x -> weights -> A^T w, image update).  `value` is slice-iterations
per second (= iterations per second per 2048^2 slice, summed over slices).

`roofline` reports the dominant kernel of the step against the HBM roofline
(8 TB/s).  The ray kernels regenerate the system matrix on the fly, so no
byte model of a stored operator bounds them; what they move is measured:
  achieved  = memory-side bytes per launch (rocprofv3's size-resolved
              L2-to-memory request counters, TCC_EA0_RDREQ_32B/64B/128B and
              TCC_EA0_WRREQ/_64B, exact on a copy kernel and in agreement
              with FETCH_SIZE and WRITE_SIZE times the copy's factors;
              profiles/traffic.json) / the kernel's average launch time,
              timed live with HIP events on the launch stream;
  frac      = achieved / 8 TB/s;
  traffic   = those bytes per launch, an upper bound on HBM bytes (the
              Infinity Cache serves part of them);
  copy_GBs  = what a 1 GiB device copy reaches on the same GPU;
  compulsory_bytes_per_launch = what the kernel cannot avoid moving: its
              image, sinogram and weight arrays once, and the ray table once
              per slice batch; traffic_over_compulsory is the re-read factor;
  limiter   = the binding resource found from counters (profiles/
              counters.json) with its utilisation;
  valu_issue_frac = that kernel's wave64 VALU instructions per SIMD clock
              over the 0.25 a SIMD retires: the ray kernels are bound by
              VALU issue, not by HBM;
  stored_operator_equiv_GBs = the byte model of streaming a stored A once
              per product (8 B per nonzero: utils/inters.py:11-15 counts the
              nonzeros) over the same time, reported for comparison only.

`cpu_baseline` times the reference's own threaded CPU engine (compiled from
its sources into oracle/_ref) on this host: one MLEM iteration of one
2048^2 slice over an evenly spaced subset of the angles, scaled linearly to
the full angle count, with as many threads as the host gives this process
(the reference's -t 0 takes hardware_concurrency, disp_engine_base.h:89-114).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SIRT/MLEM iter/s per 2048² slice; batched slices/s vs plain at "
          "≤1e-4 rel err")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus():
    """(threads to use, CPUs in the affinity set, cgroup CPU quota or None,
    online CPUs): the CPUs this process may run on, capped by a cgroup CPU
    quota where one is set."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    for path, parse in (("/sys/fs/cgroup/cpu.max", lambda t: t.split()),
                        ("/sys/fs/cgroup/cpu/cpu.cfs_quota_us", None)):
        try:
            if parse:
                q, per = parse(open(path).read())
                if q != "max":
                    quota = int(q) / int(per)
            else:
                q = int(open(path).read())
                per = int(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
                if q > 0:
                    quota = q / per
            break
        except Exception:
            continue
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return threads, aff, quota, os.cpu_count()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(N, P_full, threads, target_s=12.0):
    """Reference CPU engine (oracle/_ref) on a bounded sample; returns a
    dict or None when the compiled reference is absent."""
    try:
        from oracle import load_ref, ref_available
    except Exception:
        return None
    if not ref_available():
        return None
    import numpy as np
    from trace_amd.phantom import angles_deg
    R = load_ref()
    rng = np.random.default_rng(0)

    def run(P, iters):
        stride = max(1, P_full // P)
        deg = angles_deg(P_full)[::stride][:P]
        sino = rng.uniform(0.5, 50.0, (deg.size, 1, N)).astype(np.float32)
        t0 = time.perf_counter()
        R.run("mlem", deg, sino, N / 2.0, iters, threads)
        return time.perf_counter() - t0, deg.size

    # size the angle subset so that one iteration takes about target_s
    t1, p1 = run(max(8, P_full // 100), 1)
    t2, _ = run(p1, 2)
    per_angle = max((t2 - t1) / p1, 1e-6)
    P = int(min(P_full, max(p1, target_s / per_angle)))
    ta, pa = run(P, 1)
    tb, _ = run(pa, 2)
    per_iter_sample = tb - ta  # one iteration, setup cancelled
    per_iter_full = per_iter_sample * P_full / pa
    _, aff, quota, online = host_cpus()
    return {
        "value": 1.0 / per_iter_full,
        "unit": "slice-iter/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference MLEM engine (oracle/_ref, {threads} threads), one "
                   f"{N}x{N} slice, {pa} of {P_full} evenly spaced angles, "
                   f"iteration time = t(2 iters) - t(1 iter) = {per_iter_sample:.2f} s, "
                   f"scaled x{P_full / pa:.2f} to all angles"),
        "seconds_per_slice_iter": per_iter_full,
        "host": {"cpu_model": cpu_model(), "online_cpus": online, "affinity_cpus": aff,
                 "cgroup_cpu_quota": quota, "threads": threads,
                 "compiler": R.compiler(),
                 "flags": "-O3 -std=c++11 -ffp-contract=off (no fast-math, no -march)"},
    }


def counter_record(kernel, workload, shape=None):
    """Utilisation counters of a kernel (profiles/counters.json) measured on
    the current kernel sources, or None."""
    p = os.path.join(ROOT, "profiles", "counters.json")
    try:
        with open(p) as f:
            recs = json.load(f)["records"]
    except Exception:
        return None
    src = kernel_source_hash()
    for r in recs:
        if (r.get("kernel") == kernel and r.get("workload") == workload
                and r.get("source_sha") == src
                and (shape is None or r.get("shape") == list(shape))):
            return r
    return None


def _code_only(text: str) -> str:
    """C/C++ source without comments and with whitespace runs collapsed, so
    that editing a comment does not change the hash of the code."""
    import re
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return " ".join(text.split())


# The device code: the kernels and what they include.  Host files (the
# C-ABI, the batch set-up) choose launch shapes, which every counter record
# names separately.
KERNEL_SOURCES = ("internal.h", "kernels.hip", "ordered.hip", "staged.h", "staged.hip",
                  "traverse.h")


def kernel_source_hash(d: str | None = None) -> str:
    """sha256 over the code (comments stripped) of the kernel sources:
    counter records apply only to the kernels they were measured on."""
    import hashlib
    d = d or os.path.join(ROOT, "trace_amd", "csrc")
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update(f.encode())
        with open(os.path.join(d, f), encoding="utf-8") as fh:
            h.update(_code_only(fh.read()).encode())
    return h.hexdigest()[:16]


def fwd_staged(pj) -> bool:
    """True when the projector's forward launches run the slab-staged form
    (32-slice batches and TRACE_OPT_FWD_STAGED on)."""
    from trace_amd import _capi
    return pj.launch_shape()[0] == 32 and pj.get_option(_capi.TRACE_OPT_FWD_STAGED) == 1


def fwd_shape(pj) -> list:
    """Launch shape of the forward kernel as the profile records name it:
    [32, angles, columns, slab] of the staged form's bundles, else [slice
    batch, lanes per ray, bundle angles]."""
    if fwd_staged(pj):
        return [32, 16, 64, 4]
    return [pj.launch_shape()[0], pj.fwd_lanes(), pj.fwd_bundle()]


def traffic_calibration() -> dict:
    """The copy-kernel calibration of the traffic counters (profiles/
    traffic.json), or {}."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get("calibration", {})
    except Exception:
        return {}


def traffic_record(kernel, workload, shape=None):
    """Counter-derived memory-side bytes per launch (size-resolved L2
    request counters, checked on a copy kernel) committed under profiles/,
    or None unless a record matches this kernel, workload, launch shape and
    kernel sources."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            recs = json.load(f)["records"]
    except Exception:
        return None
    src = kernel_source_hash()
    for r in recs:
        if (r.get("kernel") == kernel and r.get("workload") == workload
                and r.get("source_sha") == src
                and (shape is None or r.get("shape") == list(shape))):
            return r.get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--alg", default="mlem", choices=["mlem", "sirt"])
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--angles", type=int, default=1500)
    ap.add_argument("--fast-sqrt", action="store_true",
                    help="hardware v_sqrt_f32 segment lengths (default: exact)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shape")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--shape", default="",
                    help="fwd_batch,tile_w,tile_h,tile_batch,lanes,bundle: use this launch "
                         "shape instead of autotuning (stack, pml, apmlr)")
    ap.add_argument("--workload", default="stack",
                    choices=["stack", "pml", "apmlr", "autograd", "batched"],
                    help="stack: the headline SIRT/MLEM stack (default); pml/apmlr: "
                         "regularised reconstruction, stencil-update GB/s; autograd: "
                         "TV+Adam and unrolled SIRT through torch.autograd; batched: "
                         "uneven slices in one graph launch vs per-slice launches")
    ap.add_argument("--batch-slices", type=int, default=200,
                    help="batched: number of uneven slices")
    ap.add_argument("--batched-iters", type=int, default=5,
                    help="stack: iterations of the batched-vs-plain comparison "
                         "appended to the line (200 uneven slices)")
    ap.add_argument("--no-batched", action="store_true",
                    help="stack: skip the batched-vs-plain comparison")
    ap.add_argument("--iters", type=int, default=0,
                    help="pml/apmlr: iterations per step (default 30); autograd: "
                         "Adam steps (default 100)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(slice sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("pml", "apmlr"):
        return run_regularised(args)
    if args.workload == "autograd":
        return run_autograd(args)
    if args.workload == "batched":
        return run_batched(args)
    return run_stack(args)


def apply_shape(pj, tune):
    from trace_amd.autotune import apply_shape as _apply
    _apply(pj, tune)


def forced_shape(args):
    """--shape fb,tw,th,tb,lanes,bundle as an autotuner result, or None."""
    if not args.shape:
        return None
    v = [int(x) for x in args.shape.split(",")]
    if len(v) != 6:
        raise SystemExit("--shape takes fwd_batch,tile_w,tile_h,tile_batch,lanes,bundle")
    return {"shape": v[:4], "lanes": v[4], "bundle": v[5], "forced": True}


def tune_shape(theta, sino, N, S, alg, fast_sqrt, params=None, init=None):
    """Autotunes launch shapes on a sub-stack of at most 32 slices of the same
    geometry (the shapes depend on the geometry and the slice batches, not on
    the stack depth); returns the autotuner's result."""
    import torch
    from trace_amd import Projector
    from trace_amd.autotune import autotune
    st = min(S, 32)
    if init is None:
        init = 0.0 if alg == "sirt" else 1.0
    with Projector(theta, st, N, center=N / 2.0, device=0, exact_sqrt=not fast_sqrt) as tp:
        sub_s = sino[:, :st, :].contiguous()
        sub_x = torch.full((st, N, N), init, dtype=torch.float32, device=sino.device)
        res = autotune(tp, sub_s, sub_x, alg, params)
    return res


def run_stack(args):
    import numpy as np
    import torch

    from trace_amd import Projector
    from trace_amd.phantom import angles_rad, noisy_sinogram

    N, P, S = args.size, args.angles, args.slices
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    workload = (f"{args.alg.upper()}, {S} slices of {N}x{N} fp32, {P} angles in [0,pi), "
                f"{N} columns, 1 iteration per step")
    log(f"workload: {workload}")

    theta = angles_rad(P)
    t0 = time.perf_counter()
    sino = noisy_sinogram(theta, N, S, seed=2, I0=1e4, K=20, device=dev)
    torch.cuda.synchronize()
    log(f"synthetic sinogram {tuple(sino.shape)} in {time.perf_counter() - t0:.1f} s")
    init = 0.0 if args.alg == "sirt" else 1.0
    recon = torch.full((S, N, N), init, dtype=torch.float32, device=dev)

    pj = Projector(theta, S, N, center=N / 2.0, device=0, exact_sqrt=not args.fast_sqrt)
    nnz, oor = pj.count_segments()
    log(f"segments per slice: {nnz} (outside image: {oor})")
    tune = forced_shape(args)
    if tune is not None:
        apply_shape(pj, tune)
        log(f"launch shape {tune['shape']} lanes {tune['lanes']} bundle {tune['bundle']} (given)")
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tune = tune_shape(theta, sino, N, S, args.alg, args.fast_sqrt)
        apply_shape(pj, tune)
        log(f"autotuned shape {tune['shape']} bundle {tune['bundle']} "
            f"in {time.perf_counter() - t0:.1f} s")

    for w in range(args.warmup):
        pj.reconstruct(args.alg, 1, sino, recon)
        torch.cuda.synchronize()
        log(f"warmup step {w + 1}/{args.warmup} done")

    pj.set_phase_timing(True)
    pj.phase_times(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        pj.reconstruct(args.alg, 1, sino, recon)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = pj.phase_times(reset=True)
    pj.set_phase_timing(False)
    elapsed = t1 - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s; phases {phases}")

    finite = bool(torch.isfinite(recon).all().item())
    nonneg = bool((recon >= 0).all().item()) if args.alg == "mlem" else None
    invariants = stack_invariants(pj, sino, recon, args.alg, dev)
    log(f"invariants: {invariants}")

    R = P * N
    fb, tw, th, tb = pj.launch_shape()
    lanes, bundle = pj.fwd_lanes(), pj.fwd_bundle()
    nb_f, nb_b = -(-S // fb), -(-S // tb)
    staged = fwd_staged(pj)
    # slab plan of the staged form (read once per slice batch): per bundle of
    # 16 angles x 64 columns its axis and an 8-byte window per 4-pixel slab
    nwg = -(-P // 16) * -(-N // 64)
    plan_b = nwg * (4.0 + 8.0 * -(-N // 4)) if staged else 0.0
    # compulsory bytes per launch: each array once, the 32-byte ray table once
    # per slice batch, den once per update batch
    model = {
        "forward": ("k_forward_staged" if staged else "k_forward",
                    S * (4.0 * N * N + 8.0 * R) + nb_f * (32.0 * R + plan_b),
                    S * (8.0 * nnz + 8.0 * R + 4.0 * N * N)),
        "back": ("k_bp_tile", S * (4.0 * R + 4.0 * N * N) + nb_b * 32.0 * R,
                 S * (8.0 * nnz + 4.0 * R + 4.0 * N * N)),
        "update": ("k_update_em", S * 16.0 * N * N + nb_f * 4.0 * N * N, S * 16.0 * N * N),
    }
    shapes = {"forward": fwd_shape(pj), "back": [tw, th, tb], "update": [fb]}
    per_kernel = {}
    for ph, (kname, comp, stored) in model.items():
        ms, n = phases[ph]
        t = ms / max(n, 1) / 1e3
        traffic = traffic_record(kname, workload, shapes[ph])
        per_kernel[ph] = {
            "kernel": kname, "avg_launch_ms": t * 1e3, "launches": n,
            "traffic_bytes_per_launch": traffic,
            "hbm_GBs": traffic / t / 1e9 if (traffic and t > 0) else None,
            "compulsory_bytes_per_launch": comp,
            "traffic_over_compulsory": traffic / comp if traffic else None,
            "stored_operator_bytes_per_launch": stored,
            "stored_operator_equiv_GBs": stored / t / 1e9 if t > 0 else None,
        }
    dom = max(("forward", "back"), key=lambda k: per_kernel[k]["avg_launch_ms"])
    kd = per_kernel[dom]
    kernel = kd["kernel"]
    t_dom = kd["avg_launch_ms"] / 1e3
    traffic = kd["traffic_bytes_per_launch"]
    achieved = kd["hbm_GBs"]
    crec = counter_record(kernel, workload, shapes[dom])
    limiter = None
    if crec:
        limiter = {"resource": crec.get("limiter"), "utilisation": crec.get("derived"),
                   "source": crec.get("source")}
    b_stored_step = S * (16.0 * nnz + 12.0 * R + 24.0 * N * N)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or host_cpus()[0]
        log(f"cpu baseline with {threads} threads ...")
        cpu = cpu_baseline(N, P, threads)
        if cpu:
            log(f"cpu baseline: {cpu['sample']}")

    batched = None
    if not args.no_batched:
        batched = batched_compare(args.batch_slices, args.batched_iters,
                                  autotune=not args.no_autotune)

    value = S * args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "slice-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic: 20 seeded ellipses per slice (seed 2), exact line "
                 "integrals, Poisson noise at 1e4 counts"),
        "config": {
            "workload": workload,
            "alg": args.alg,
            "slices": S,
            "image": N,
            "angles": P,
            "columns": N,
            "iters_per_step": 1,
            "segment_lengths": "v_sqrt_f32" if args.fast_sqrt else "correctly rounded",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "achieved_from": "counter memory-side bytes per launch / HIP-event launch time",
            "copy_GBs": traffic_calibration().get("copy_1GiB_GBs"),
            "avg_launch_ms": t_dom * 1e3,
            "compulsory_bytes_per_launch": kd["compulsory_bytes_per_launch"],
            "traffic_over_compulsory": kd["traffic_over_compulsory"],
            "limiter": limiter,
            # the binding resource of the ray kernels: wave64 VALU instructions
            # per SIMD clock over the 0.25 a SIMD retires (counters)
            "valu_issue_frac": ((crec or {}).get("derived") or {}).get("valu_issue_share"),
            "stored_operator_bytes_per_launch": kd["stored_operator_bytes_per_launch"],
            "stored_operator_model": ("per slice 8*nnz + 8*R + 4*N^2 (forward) / 8*nnz + 4*R "
                                      "+ 4*N^2 (back): an int32 index and an fp32 length per "
                                      "nonzero of a stored A, for comparison only"),
            "stored_operator_equiv_GBs": kd["stored_operator_equiv_GBs"],
            "stored_operator_step_equiv_GBs": b_stored_step / (elapsed / args.steps) / 1e9,
            "nnz_per_slice": nnz,
        },
        "kernels": per_kernel,
        "launch_shape": {"fwd_batch_tile_w_h_batch": list(pj.launch_shape()),
                         "fwd_bundle_angles": pj.fwd_bundle(),
                         "fwd_lanes_per_ray": pj.fwd_lanes(),
                         "fwd_staged": fwd_staged(pj),
                         "autotuned": tune is not None and not tune.get("forced")},
        "phases_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in phases.items()},
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "batched": batched,
        "checks": {"finite": finite, "nonnegative": nonneg, **invariants},
    }
    print(json.dumps(out), flush=True)


def stack_invariants(pj, sino, recon, alg, dev):
    """Checks of the timed stack that need no reference run, at the launch
    shape that was timed:
      dot_product_rel  |<A u, v> - <u, A^T v>| / |<A u, v>| for random u, v
                       on the whole stack (fp64 sums);
      mlem_conservation_rel  after an MLEM iteration sum_j den_j x'_j =
                       sum_r p_r per slice (x' = x A^T(p/Ax)/den, so
                       sum_j den_j x'_j = sum_r p_r (Ax)_r/(Ax)_r): the largest
                       relative violation over the slices."""
    import torch
    S, N = recon.shape[0], recon.shape[1]
    g = torch.Generator(device=dev).manual_seed(5)
    u = torch.rand(recon.shape, device=dev, generator=g)
    v = torch.randn(sino.shape, device=dev, generator=g)
    au = pj.forward(u)
    atv = pj.adjoint(v)
    lhs = (au.double() * v.double()).sum().item()
    rhs = (u.double() * atv.double()).sum().item()
    out = {"dot_product_rel": abs(lhs - rhs) / max(abs(lhs), 1e-300)}
    del u, v, au, atv
    if alg == "mlem":
        den = pj.column_sums().double()
        lhs = (recon.double() * den).sum(dim=(1, 2))
        rhs = sino.double().sum(dim=(0, 2))
        out["mlem_conservation_rel"] = float(((lhs - rhs).abs() / rhs.abs()).max().item())
    return out


def run_regularised(args):
    """PML or APMLR on the stack (default 64 x 2048², 1500 angles, 30
    iterations per step; beta = 0.1, delta = 0.01, beta1 = delta1 = regw = 1).
    The update kernel fuses the stencil, the closed-form root and the
    batched-copy refresh; its algorithmic traffic is 16 B per pixel (read x,
    A^T w, den; write x)."""
    import torch

    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram

    alg = args.workload
    N, P, S = args.size, args.angles, args.slices
    iters = args.iters or 30
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    prm = (RegParams(beta=0.1, delta=0.01) if alg == "pml"
           else RegParams(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=True))
    workload = (f"{alg.upper()}, {S} slices of {N}x{N} fp32, {P} angles in [0,pi), {N} columns, "
                f"{iters} iterations per step")
    log(f"workload: {workload}")
    theta = angles_rad(P)
    sino = noisy_sinogram(theta, N, S, seed=2, I0=1e4, K=20, device=dev)
    recon = torch.ones((S, N, N), dtype=torch.float32, device=dev)
    pj = Projector(theta, S, N, center=N / 2.0, device=0, exact_sqrt=not args.fast_sqrt)
    nnz, _ = pj.count_segments()
    tune = forced_shape(args)
    if tune is not None:
        apply_shape(pj, tune)
    elif not args.no_autotune:
        tune = tune_shape(theta, sino, N, S, alg, args.fast_sqrt, prm, init=1.0)
        apply_shape(pj, tune)
        log(f"autotuned shape {tune['shape']} bundle {tune['bundle']}")
    for w in range(args.warmup):
        pj.reconstruct(alg, 1, sino, recon.clone(), prm)
        torch.cuda.synchronize()
    pj.set_phase_timing(True)
    pj.phase_times(reset=True)
    x = recon.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x.copy_(recon)
        pj.reconstruct(alg, iters, sino, x, prm)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    phases = pj.phase_times(reset=True)
    pj.set_phase_timing(False)
    log(f"timed {args.steps} steps in {elapsed:.3f} s; phases {phases}")
    upd_ms, upd_n = phases["update"]
    t_upd = upd_ms / max(upd_n, 1) / 1e3
    upd_bytes = 16.0 * S * N * N
    R = P * N
    slice_iters = S * iters * args.steps
    out = {
        "metric": f"{alg.upper()} update+stencil kernel HBM GB/s (fused, 16 B/pixel model)",
        "value": upd_bytes / t_upd / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic: 20 seeded ellipses per slice (seed 2), exact line "
                 "integrals, Poisson noise at 1e4 counts"),
        "config": {"workload": workload, "alg": alg, "slices": S, "image": N, "angles": P,
                   "iters_per_step": iters,
                   "params": {"beta": prm.beta, "beta1": prm.beta1, "delta": prm.delta,
                              "delta1": prm.delta1, "regw": prm.regw}},
        "roofline": {"bound": "hbm", "kernel": f"k_update_{alg}",
                     "achieved": upd_bytes / t_upd / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": upd_bytes / t_upd / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic_record(f"k_update_{alg}", workload,
                                               [pj.launch_shape()[0]]),
                     "algorithmic_bytes_per_launch": upd_bytes,
                     "avg_launch_ms": t_upd * 1e3},
        "slice_iter_per_s": slice_iters / elapsed,
        "phases_ms_per_iter": {k: v[0] / max(v[1], 1) for k, v in phases.items()},
        "back_kernel": back_kernel_record(phases, workload),
        "stored_operator_frac": (S * (16.0 * nnz + 12.0 * R + 24.0 * N * N))
                                / (elapsed / (iters * args.steps)) / 1e9 / HBM_PEAK_GBS,
        "launch_shape": {"fwd_batch_tile_w_h_batch": list(pj.launch_shape()),
                         "fwd_bundle_angles": pj.fwd_bundle(),
                         "fwd_lanes_per_ray": pj.fwd_lanes(),
                         "fwd_staged": fwd_staged(pj),
                         "autotuned": tune is not None and not tune.get("forced")},
        "checks": {"finite": bool(torch.isfinite(x).all().item()),
                   "nonnegative": bool((x >= 0).all().item())},
    }
    print(json.dumps(out), flush=True)


def back_kernel_record(phases, workload):
    """The backprojection of a PML/APMLR iteration (the per-pixel sums in
    the reference's order, k_bp_ordered_tile): launch time and counter
    traffic."""
    ms, n = phases["back"]
    t = ms / max(n, 1) / 1e3
    traffic = traffic_record("k_bp_ordered_tile", workload, [32])
    return {"kernel": "k_bp_ordered_tile", "avg_launch_ms": t * 1e3,
            "traffic_bytes_per_launch": traffic,
            "memory_GBs": traffic / t / 1e9 if (traffic and t > 0) else None}


def run_autograd(args):
    """Differentiable reconstruction on one 2048² slice, 1500 angles: Adam
    (lr 1e-2) on 1/2||Ax - y||^2 + 1e-3 TV(x) through torch.autograd (the
    forward projector and the hand-written adjoint kernel), then a
    10-iteration unrolled SIRT with per-iteration checkpoints, forward and
    backward."""
    import torch

    from trace_amd.autograd import RayTransform, ls_tv_loss, sirt_unrolled, tv_grad
    from trace_amd.phantom import angles_rad, noisy_sinogram, phantom_slices

    N, P = args.size, args.angles
    steps = args.iters or 100
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    workload = (f"TV+Adam, one {N}x{N} fp32 image, {P} angles in [0,pi), {N} columns, "
                f"{steps} Adam steps (lr 1e-2, TV weight 1e-3)")
    log(f"workload: {workload}")
    theta = angles_rad(P)
    y3 = noisy_sinogram(theta, N, 1, seed=4, I0=1e4, K=20, device=dev)
    y = y3[:, 0, :].contiguous()
    x_true = phantom_slices(N, 1, seed=4, K=20, device=dev)[0].float()
    op = RayTransform(theta, N, 1, N / 2.0, exact_sqrt=not args.fast_sqrt)
    tune = None
    if not args.no_autotune:
        tune = tune_shape(theta, y3, N, 1, "sirt", args.fast_sqrt)
        apply_shape(op.pj, tune)
        log(f"autotuned shape {tune['shape']} bundle {tune['bundle']}")
    # adjoint dot-product test on this geometry
    g = torch.Generator(device="cuda").manual_seed(0)
    xr = torch.rand((N, N), device=dev, generator=g)
    yr = torch.rand((P, N), device=dev, generator=g)
    lhs = (op.apply_A(xr).double() * yr.double()).sum().item()
    rhs = (xr.double() * op.apply_AT(yr).double()).sum().item()
    dot_rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))

    x = torch.zeros((N, N), device=dev, requires_grad=True)
    opt = torch.optim.Adam([x], lr=1e-2)

    def adam_step():
        opt.zero_grad(set_to_none=True)
        loss = ls_tv_loss(op, x, y, tv_weight=1e-3)
        loss.backward()
        opt.step()
        return loss

    for _ in range(max(args.warmup, 1)):
        adam_step()
    with torch.no_grad():
        x.zero_()
    opt = torch.optim.Adam([x], lr=1e-2)
    op.pj.set_phase_timing(True)
    op.pj.phase_times(reset=True)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        losses.append(adam_step().detach())
    torch.cuda.synchronize()
    t_adam = time.perf_counter() - t0
    phases = op.pj.phase_times(reset=True)
    losses = torch.stack(losses).cpu().tolist()
    err_adam = ((x.detach() - x_true).norm() / x_true.norm()).item()
    log(f"{steps} Adam steps in {t_adam:.3f} s; loss {losses[0]:.4e} -> {losses[-1]:.4e}")

    # unrolled SIRT, 10 iterations, learnable per-iteration TV step
    K = 10
    tau = torch.full((K,), 1e-3, device=dev, requires_grad=True)
    p = y.clone().requires_grad_(True)

    def unrolled():
        xk = sirt_unrolled(op, p, K, prox=lambda z, k: z - tau[k] * tv_grad(z),
                           use_checkpoint=True)
        return ((xk - x_true) ** 2).sum()

    unrolled().backward()  # warm-up
    p.grad = None
    tau.grad = None
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    loss_u = unrolled()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    kept = torch.cuda.memory_allocated() - base
    loss_u.backward()
    torch.cuda.synchronize()
    t_unrolled = time.perf_counter() - t0
    peak = torch.cuda.max_memory_allocated() - base
    op.pj.set_phase_timing(False)
    out = {
        "metric": "TV+Adam steps/s through torch.autograd (forward projector + hand-written adjoint)",
        "value": steps / t_adam,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": t_adam / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 20 seeded ellipses (seed 4), exact line integrals, Poisson 1e4",
        "config": {"workload": workload, "image": N, "angles": P, "columns": N},
        "phases_ms_per_step": {k: v[0] / steps for k, v in phases.items()},
        "adam": {"loss_first": losses[0], "loss_last": losses[-1],
                 "rel_error_vs_phantom": err_adam},
        "dot_product_rel": dot_rel,
        "unrolled_sirt": {"iterations": K, "checkpointed": True,
                          "forward_ms": t_fwd * 1e3, "forward_backward_ms": t_unrolled * 1e3,
                          "kept_after_forward_MiB": kept / 2**20,
                          "peak_MiB": peak / 2**20,
                          "grad_finite": bool(torch.isfinite(p.grad).all().item())},
        "launch_shape": {"fwd_batch_tile_w_h_batch": list(op.pj.launch_shape()),
                         "fwd_bundle_angles": op.pj.fwd_bundle(),
                         "fwd_lanes_per_ray": op.pj.fwd_lanes(),
                         "autotuned": tune is not None and not tune.get("forced")},
    }
    print(json.dumps(out), flush=True)

def batched_compare(n, iters, autotune=True):
    """Uneven slices (widths {256..2048}, 180-1800 angles, SIRT or MLEM,
    `iters` iterations) reconstructed (a) slice by slice with the plain
    per-slice entry point, each slice with the shape the autotuner picks for
    its width, and (b) all together by one graph launch of the batched
    persistent kernels with the shape the batch autotuner picks.  Without
    `autotune` both keep the library's default shapes (and a kernel trace
    of the run holds only the timed work, one warm-up iteration per plain
    slice and the set-up).  Setup (tables, graph instantiation) is outside
    both timings; every slice's two results are compared bit for bit.
    Returns a dict."""
    import torch

    from trace_amd import Projector
    from trace_amd.autotune import autotune as autotune_shape
    from trace_amd.batch import Batch, autotune_batch, uneven_slices, uneven_specs

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    specs = uneven_specs(n, seed=3)
    workload = (f"{n} uneven slices (width 256-2048, 180-1800 angles, SIRT/MLEM, seed 3), "
                f"{iters} iterations each")
    log(f"batched: {workload}")
    t0 = time.perf_counter()
    slices = uneven_slices(specs, seed=3, iters=iters, device=0)
    torch.cuda.synchronize()
    log(f"synthetic sinograms in {time.perf_counter() - t0:.1f} s")
    init = [s.recon.clone() for s in slices]
    nnz_total = 0

    # (a) plain: one handle per slice, shapes tuned per width
    t0 = time.perf_counter()
    shapes = {}
    for C in sorted({c for c, _, _ in specs}):
        i = next(k for k, sp in enumerate(specs) if sp[0] == C)
        s = slices[i]
        with Projector(s.theta, 1, C, s.center) as tp:
            if autotune:
                x = s.recon.clone().reshape(1, C, C)
                shapes[C] = autotune_shape(tp, s.sino.reshape(-1, 1, C).contiguous(), x, s.alg)
            else:
                shapes[C] = {"shape": list(tp.launch_shape()), "bundle": tp.fwd_bundle(),
                             "lanes": tp.fwd_lanes(), "forced": True}
    handles = []
    for s in slices:
        C = s.recon.shape[0]
        pj = Projector(s.theta, 1, C, s.center)
        apply_shape(pj, shapes[C])
        x = s.recon.clone().reshape(1, C, C)
        pj.reconstruct(s.alg, 1, s.sino.reshape(-1, 1, C).contiguous(), x)  # tables, warm-up
        nnz_total += pj.count_segments()[0]
        handles.append(pj)
    torch.cuda.synchronize()
    log(f"plain set-up ({'per-width autotune + ' if autotune else ''}{n} handles) in "
        f"{time.perf_counter() - t0:.1f} s")
    sinos = [s.sino.reshape(-1, 1, s.recon.shape[0]).contiguous() for s in slices]
    outs = [x.clone().reshape(1, *x.shape) for x in init]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pj, s, sn, x in zip(handles, slices, sinos, outs):
        pj.reconstruct(s.alg, iters, sn, x)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    log(f"plain: {n} slices x {iters} iterations in {t_plain:.3f} s")
    for pj in handles:
        pj.close()
    del handles

    # (b) batched
    t0 = time.perf_counter()
    if autotune:
        tune = autotune_batch(slices, bundles=(2, 4, 8, 16), tiles=((32, 32), (64, 64)),
                              sample=min(24, n), iters=1)
    for s, x0 in zip(slices, init):
        s.recon.copy_(x0)
    b = Batch(slices)
    if autotune:
        b.set_launch_shape(tune["bundle"], *tune["tile"])
    else:
        bundle, tw, th = b.launch_shape()
        tune = {"bundle": bundle, "tile": (tw, th), "ms": None}
    b.prepare()
    torch.cuda.synchronize()
    log(f"batch set-up ({'autotune ' if autotune else 'shape '}{tune['bundle']}/{tune['tile']}, "
        f"tables, graph) in "
        f"{time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"batched: {n} slices x {iters} iterations in {t_batch:.3f} s")
    identical = 0
    worst = 0.0
    for s, x in zip(slices, outs):
        xb = s.recon
        x2 = x.reshape(xb.shape)
        if torch.equal(xb, x2):
            identical += 1
        else:
            d = ((xb.double() - x2.double()).norm() / x2.double().norm().clamp_min(1e-30)).item()
            worst = max(worst, d)
    b.close()
    return {
        "workload": workload, "slices": n, "iters": iters,
        "widths": {str(c): sum(1 for sp in specs if sp[0] == c)
                   for c in sorted({sp[0] for sp in specs})},
        "sirt": sum(1 for sp in specs if sp[2] == "sirt"),
        "plain_s": t_plain, "batched_s": t_batch, "ratio": t_plain / t_batch,
        "plain_shapes": {str(k): {"shape": v["shape"], "bundle": v["bundle"],
                                  "lanes": v.get("lanes", 1)} for k, v in shapes.items()},
        "batched_shape": {"bundle": tune["bundle"], "tile": list(tune["tile"])},
        "batched_autotune_ms": tune["ms"],
        "slice_iter_per_s": n * iters / t_batch,
        "segments_per_s": nnz_total * iters * 2 / t_batch,
        "parity": {"bit_identical_slices": identical, "slices": n,
                   "worst_rel_l2_of_others": worst},
    }


def run_batched(args):
    """Config 3 on its own line: 200 uneven slices, 25 iterations each
    (batched_compare)."""
    r = batched_compare(args.batch_slices, args.iters or 25, autotune=not args.no_autotune)
    out = {
        "metric": "batched uneven slices/s (one graph launch) vs plain per-slice launches",
        "value": r["slices"] / r["batched_s"],
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": r["batched_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded ellipse phantoms (seed 3, 20 ellipses), Poisson 1e4",
        "config": {"workload": r["workload"], "slices": r["slices"], "iters": r["iters"],
                   "widths": r["widths"], "sirt": r["sirt"]},
        "plain": {"seconds": r["plain_s"], "slices_per_s": r["slices"] / r["plain_s"],
                  "shapes": r["plain_shapes"]},
        "batched": {"seconds": r["batched_s"], **r["batched_shape"],
                    "autotune_ms": r["batched_autotune_ms"]},
        "speedup_vs_plain": r["ratio"],
        "slice_iter_per_s": r["slice_iter_per_s"],
        "segments_per_s": r["segments_per_s"],
        "parity": r["parity"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
