"""Full-size configurations on the GPU against the one-thread oracle (bit-exact
to the reference at one thread), through fixtures written by
tests/golden/make_config_fixtures.py: sampled pixels and per-row sums of the
oracle's fp32 result and of its fp64 twin.

  cfg1   MLEM 1024^2, 720 angles, 50 iterations      (mlem_main.cc:171-182)
  cfg2p  PML 2048^2, 1500 angles, 30 iterations      (pml_main.cc:178-193)
  cfg2a  APMLR 2 x 2048^2, 1500 angles, 30 iterations (apmlr_main.cc:209-244)
  cfg2a_fixed  the same with the slice offset corrected (the oracle's
         restatement of a mode the reference does not have)
  cfg3_i four of 200 uneven slices (widths 256, 1024, 1536, 2048),
         25 iterations, through the batched graph path
  stack_mlem, stack_sirt
         MLEM and SIRT at the benchmark's geometry (2048^2, 1500 angles),
         10 iterations; these and cfg2p also as stacks of 64 (32) copies of
         the slice, which run the 32-slice kernels of the benchmark

Bar: relative L2 <= 1e-4 against the reference on the sampled pixels and on
the row sums.  PML and APMLR sum den and num in the reference's order, so
there the sampled pixels must be bit-identical.  The error against fp64 is
printed for the record."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_config_fixtures as fx  # noqa: E402

TOL = 1e-4


def _fixture():
    meta = json.load(open(os.path.join(GOLD, "configs.json")))
    return meta["cases"], np.load(os.path.join(GOLD, "configs.npz"))


def _inputs(key, meta):
    th, sino, init = fx.case_inputs(key)
    sha = fx.sino_sha(sino)
    assert sha == meta["sino_sha256"], (
        f"{key}: the rebuilt sinogram differs from the fixture's input "
        f"({sha[:12]} vs {meta['sino_sha256'][:12]})")
    return th, sino, init


def _check(key, x, meta, arr, exact):
    samp, rows = fx.reduce_image(x)
    ref, f64 = arr[f"{key}_ref"], arr[f"{key}_f64"]
    e_ref = rel_l2(samp, ref)
    e_rows = rel_l2(rows, arr[f"{key}_rows_ref"])
    e64 = rel_l2(samp, f64)
    bad = int(np.count_nonzero(samp.astype(np.float32).view(np.uint32) != ref.view(np.uint32)))
    print(f"{key} {meta['alg']} {meta['slices']}x{meta['N']}^2 P={meta['P']} "
          f"it={meta['iters']}: vs reference {e_ref:.2e} (rows {e_rows:.2e}, "
          f"{bad}/{samp.size} sampled pixels differ); vs fp64 {e64:.2e}; "
          f"reference vs fp64 {meta['reference_vs_fp64']:.2e}")
    assert e_ref <= TOL and e_rows <= TOL
    if exact:
        assert bad == 0


@pytest.mark.parametrize("key", ["cfg1", "cfg2p", "cfg2a", "cfg2a_fixed", "stack_mlem", "stack_sirt"])
def test_config_vs_reference(gpu, key):
    import torch
    from trace_amd import Projector, RegParams
    cases, arr = _fixture()
    meta = cases[key]
    print(f"case {key}: building inputs", flush=True)
    th, sino, init = _inputs(key, meta)
    P, S, C = sino.shape
    prm = meta["params"]
    rp = RegParams(**prm) if prm else None
    print(f"case {key}: {meta['alg']} {S}x{C}^2, {P} angles, {meta['iters']} iterations",
          flush=True)
    with Projector(th, S, C, C / 2.0) as pj:
        ds = torch.from_numpy(sino).cuda()
        dx = torch.from_numpy(init).cuda()
        pj.reconstruct(meta["alg"], meta["iters"], ds, dx, rp)
        torch.cuda.synchronize()
        x = dx.cpu().numpy()
    _check(key, x, meta, arr, exact=meta["alg"] in ("pml", "apmlr"))


@pytest.mark.parametrize("key,S", [("stack_mlem", 64), ("stack_sirt", 64), ("cfg2p", 32)])
def test_stack_form_vs_reference(gpu, key, S):
    """The benchmark's stack form on S copies of the case's slice: 32-slice
    batches, so the slab-staged forward projector and 16x16x32 backprojector
    tiles run (a one-slice stack runs the one-slice kernels).  Every slice
    must equal the first bit for bit, and the first meets the case's bar
    against the reference (PML: bit-identical sampled pixels)."""
    import torch
    from trace_amd import Projector, RegParams
    cases, arr = _fixture()
    meta = cases[key]
    print(f"case {key} as a stack of {S}: building inputs", flush=True)
    th, sino, init = _inputs(key, meta)
    P, _, C = sino.shape
    prm = meta["params"]
    rp = RegParams(**prm) if prm else None
    with Projector(th, S, C, C / 2.0) as pj:
        ds = torch.from_numpy(sino).cuda().expand(P, S, C).contiguous()
        dx = torch.from_numpy(init).cuda().expand(S, C, C).contiguous()
        assert pj.launch_shape()[0] == 32
        pj.reconstruct(meta["alg"], meta["iters"], ds, dx, rp)
        torch.cuda.synchronize()
        bits = dx.view(torch.int32)
        same = [bool(torch.equal(bits[k], bits[0])) for k in range(S)]
        x = dx[:1].cpu().numpy()
    print(f"{key}: {sum(same)} of {S} slices equal the first", flush=True)
    assert all(same)
    _check(key, x, meta, arr, exact=meta["alg"] in ("pml", "apmlr"))


def test_config3_batched_vs_reference(gpu):
    import torch
    from trace_amd.batch import Batch, BatchSlice
    cases, arr = _fixture()
    keys = [f"cfg3_{i}" for i, _ in fx.cfg3_picked()]
    slices = []
    for key in keys:
        meta = cases[key]
        print(f"case {key}: building inputs", flush=True)
        th, sino, init = _inputs(key, meta)
        C = sino.shape[2]
        slices.append(BatchSlice(th, torch.from_numpy(sino[:, 0, :].copy()).cuda(),
                                 torch.from_numpy(init[0].copy()).cuda(), meta["alg"],
                                 meta["iters"], C / 2.0))
    print("four uneven slices through the batched graph", flush=True)
    with Batch(slices) as b:
        b.run()
        torch.cuda.synchronize()
    for key, s in zip(keys, slices):
        _check(key, s.recon.cpu().numpy()[None], cases[key], arr, exact=False)
