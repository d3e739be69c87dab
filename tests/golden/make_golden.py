"""Writes the golden fixtures of tests/golden/ from the compiled reference.

The reference repository has no tests, golden vectors or fixtures of its own,
so these vectors are produced here by running the reference's own hot-path
code (compiled from its sources by oracle/Makefile into oracle/_ref/, with
the reference SIRT build's FP flags) on seeded inputs.  Each .npz holds the
inputs and the reference's outputs; tests/test_oracle_golden.py checks that
the oracle restatement reproduces them bit for bit, and the GPU tests use
them as known answers.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Params, load_ref  # noqa: E402

KPI = np.float32(3.14159265358979)


def ray_cases(rng):
    cases = []
    adv = np.array([0.0, KPI / 4, KPI / 2, 3 * KPI / 4, KPI, 1.25 * KPI,
                    1.5 * KPI, 2 * KPI, -KPI / 4, 2.5 * KPI, 1e-7], np.float32)
    adv = np.concatenate([adv, np.nextafter(adv, np.float32(9)),
                          np.nextafter(adv, np.float32(-9))])
    for N in (4, 7, 8, 16, 33, 64, 256):
        centers = [N / 2.0, 0.0, N / 2.0 + 0.3]
        for th in adv:
            for ctr in centers:
                for col in sorted(set([0, N // 2, N - 1, int(rng.integers(0, N))])):
                    cases.append((N, ctr, np.float32(th), col))
        for _ in range(40):
            th = np.float32(rng.uniform(-7.0, 7.0))
            cases.append((N, float(rng.choice(centers)), th, int(rng.integers(0, N))))
    return cases


def make_rays(R, rng):
    cases = ray_cases(rng)
    Ns, ctrs, ths, cols, offs = [], [], [], [], [0]
    idx_all, len_all, len2_all = [], [], []
    for N, ctr, th, col in cases:
        idx, ln, l2 = R.trace_ray(N, ctr, th, col)
        Ns.append(N); ctrs.append(ctr); ths.append(th); cols.append(col)
        idx_all.append(idx); len_all.append(ln); len2_all.append(l2)
        offs.append(offs[-1] + idx.size)
    np.savez_compressed(
        os.path.join(HERE, "rays.npz"),
        N=np.array(Ns, np.int32), center=np.array(ctrs, np.float32),
        theta=np.array(ths, np.float32), col=np.array(cols, np.int32),
        offsets=np.array(offs, np.int64),
        idx=np.concatenate(idx_all).astype(np.int32),
        len=np.concatenate(len_all).astype(np.float32),
        len2=np.concatenate(len2_all).astype(np.float32))
    return len(cases)


def make_scalar(R, rng):
    deg = np.concatenate([np.arange(0, 360, 0.7), rng.uniform(-720, 720, 500)]).astype(np.float32)
    q = np.concatenate([rng.uniform(-10, 10, 500),
                        np.array([0, KPI / 2, KPI, 1.5 * KPI, 2 * KPI])]).astype(np.float32)
    q = np.concatenate([q, np.nextafter(q, np.float32(99)), np.nextafter(q, np.float32(-99))])
    Cs = np.array([4, 7, 8, 33, 256, 1024, 2048], np.int32)
    ctr = np.array([-1.0, 0.0, 1.5, 2.0, 3.7, 128.0, 1024.0], np.float32)
    mov = np.array([[R.mov(int(C), float(c)) for c in ctr] for C in Cs], np.float32)
    np.savez_compressed(os.path.join(HERE, "scalars.npz"), deg=deg, rad=R.deg2rad(deg),
                        q=q, quadrant=np.array([R.quadrant(v) for v in q], np.int32),
                        C=Cs, center=ctr, mov=mov)


def make_recon(R, rng):
    """Small full runs of the four apps (all stages of the last iteration)."""
    out = {}
    cases = [
        ("sirt", 16, 12, 2, 4), ("sirt", 33, 20, 1, 3), ("mlem", 16, 12, 2, 4),
        ("mlem", 24, 16, 1, 3), ("pml", 16, 12, 2, 3), ("pml", 20, 10, 1, 2),
        ("apmlr", 16, 12, 1, 2), ("apmlr", 16, 12, 2, 2), ("apmlr", 12, 10, 4, 3),
    ]
    for k, (alg, N, P, S, iters) in enumerate(cases):
        sino = rng.uniform(0.05, 3.0, (P, S, N)).astype(np.float32)
        if alg == "sirt":
            sino[rng.uniform(size=sino.shape) < 0.1] *= -1  # SIRT keeps signs
        deg = (np.arange(P) * 180.0 / P).astype(np.float32)
        theta_file = R.deg2rad(deg) if alg == "pml" else deg
        prm = Params(beta=0.1 if alg != "apmlr" else 0.5, beta1=0.7, delta=0.05,
                     delta1=0.8, regw=0.6)
        center = N / 2.0 if k % 3 else N / 2.0 + 0.25
        x, st = R.run(alg, theta_file, sino, center, iters, 1, prm, want_stages=True)
        key = f"{alg}_{N}_{P}_{S}_{iters}"
        out[key + "_meta"] = np.array([N, P, S, iters], np.int32)
        out[key + "_center"] = np.array([center], np.float32)
        out[key + "_params"] = np.array([prm.beta, prm.beta1, prm.delta, prm.delta1,
                                         prm.regw], np.float32)
        out[key + "_theta_file"] = theta_file
        out[key + "_sino"] = sino
        out[key + "_recon"] = x
        for s, v in st.items():
            out[key + "_" + s] = v
    np.savez_compressed(os.path.join(HERE, "recon.npz"), **out)
    return len(cases)


def main():
    R = load_ref()
    rng = np.random.default_rng(20240601)
    n = make_rays(R, rng)
    make_scalar(R, rng)
    m = make_recon(R, rng)
    print(f"wrote {n} rays and {m} reconstructions")


if __name__ == "__main__":
    main()
