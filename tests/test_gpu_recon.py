"""Reconstructions on the GPU against the oracle (fp32 restatement, which is
bit-exact to the reference at one thread) and the golden reference runs.
Bar: relative L2 <= 1e-4 after the given iterations; the error against the
fp64 twin is printed for the record."""
import os

import numpy as np
import pytest

from conftest import rel_l2
from oracle import Params

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-4


def _run_gpu(alg, th, sino, init, center, iters, prm=None, exact=True):
    import torch
    from trace_amd import Projector, RegParams
    P, S, C = sino.shape
    rp = None
    if prm is not None:
        rp = RegParams(prm.beta, prm.beta1, prm.delta, prm.delta1, prm.regw,
                       bool(prm.apmlr_compat))
    with Projector(th, S, C, center, exact_sqrt=exact) as pj:
        ds = torch.from_numpy(np.ascontiguousarray(sino, np.float32)).cuda()
        dx = torch.from_numpy(np.ascontiguousarray(init, np.float32)).cuda()
        pj.reconstruct(alg, iters, ds, dx, rp)
        torch.cuda.synchronize()
        return dx.cpu().numpy()


def _recon_keys():
    g = np.load(os.path.join(GOLD, "recon.npz"))
    return sorted({k.rsplit("_meta", 1)[0] for k in g.files if k.endswith("_meta")})


@pytest.mark.parametrize("key", _recon_keys())
def test_golden_recon_gpu(gpu, oracle, key):
    g = np.load(os.path.join(GOLD, "recon.npz"))
    alg = key.split("_")[0]
    N, P, S, iters = (int(v) for v in g[key + "_meta"])
    center = float(g[key + "_center"][0])
    b, b1, d, d1, rw = (float(v) for v in g[key + "_params"])
    prm = Params(beta=b, beta1=b1, delta=d, delta1=d1, regw=rw, apmlr_compat=1)
    sino = g[key + "_sino"]
    thf = g[key + "_theta_file"]
    th = thf if alg == "pml" else oracle.deg2rad(thf)
    s_in = sino if alg == "sirt" else oracle.absolute(sino)
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    print(f"golden {key}", flush=True)
    x = _run_gpu(alg, th, s_in, init, center, iters, prm)
    e = rel_l2(x, g[key + "_recon"])
    print(f"{key}: rel L2 vs reference {e:.2e}")
    assert e < TOL


CASES = [
    ("sirt", 64, 48, 3, 10, 1), ("mlem", 64, 48, 3, 10, 1), ("pml", 64, 48, 3, 6, 1),
    ("apmlr", 48, 36, 1, 5, 1), ("apmlr", 48, 36, 2, 5, 1), ("apmlr", 48, 36, 4, 5, 1),
    ("apmlr", 48, 36, 4, 5, 0), ("mlem", 127, 60, 2, 8, 1),
]


@pytest.mark.parametrize("alg,N,P,S,iters,compat", CASES)
def test_recon_vs_oracle(gpu, oracle, alg, N, P, S, iters, compat):
    from trace_amd.phantom import angles_rad, noisy_sinogram
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=N + S, I0=1e4, K=8).numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    prm = Params(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=compat)
    print(f"recon {alg} N={N} P={P} S={S} iters={iters} compat={compat}", flush=True)
    x = _run_gpu(alg, th, sino, init, N / 2, iters, prm)
    xo = oracle.reconstruct(alg, th, sino, init, N / 2, iters, prm)
    x64 = oracle.reconstruct_f64(alg, th, sino, init, N / 2, iters, prm)
    e, e64, eo64 = rel_l2(x, xo), rel_l2(x, x64), rel_l2(xo, x64)
    print(f"{alg} N={N} S={S}: gpu-vs-oracle {e:.2e}  gpu-vs-fp64 {e64:.2e}  "
          f"oracle-vs-fp64 {eo64:.2e}")
    assert e < TOL


def test_fast_sqrt_recon(gpu, oracle):
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 96, 64, 2, 10
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=3, I0=1e4, K=8).numpy()
    init = np.ones((S, N, N), np.float32)
    x = _run_gpu("mlem", th, sino, init, N / 2, iters, exact=False)
    xo = oracle.reconstruct("mlem", th, sino, init, N / 2, iters)
    assert rel_l2(x, xo) < TOL


def test_config0_sirt(gpu, oracle):
    """SIRT, one 256x256 slice of 10 seeded ellipses, 180 angles, Poisson
    1e4, 20 iterations (the reference's CPU-runnable configuration)."""
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 256, 180, 1, 20
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=0, I0=1e4, K=10).numpy()
    init = np.zeros((S, N, N), np.float32)
    x = _run_gpu("sirt", th, sino, init, N / 2, iters)
    xo = oracle.reconstruct("sirt", th, sino, init, N / 2, iters)
    x64 = oracle.reconstruct_f64("sirt", th, sino, init, N / 2, iters)
    e, e64 = rel_l2(x, xo), rel_l2(x, x64)
    print(f"config0 SIRT: gpu-vs-oracle {e:.2e} gpu-vs-fp64 {e64:.2e} "
          f"oracle-vs-fp64 {rel_l2(xo, x64):.2e}")
    assert e < TOL


def test_host_entry_matches_device(gpu, oracle):
    import trace_amd
    from trace_amd.phantom import angles_deg, noisy_sinogram
    N, P, S = 48, 40, 2
    deg = angles_deg(P)
    sino = noisy_sinogram(trace_amd.degree_to_radian(deg), N, S, seed=5, I0=1e4, K=6).numpy()
    xh = trace_amd.mlem(deg, sino, 4, center=N / 2)
    th = oracle.deg2rad(deg)
    xo = oracle.reconstruct("mlem", th, oracle.absolute(sino), np.ones((S, N, N), np.float32),
                            N / 2, 4)
    assert rel_l2(xh, xo) < TOL


def test_nonfinite_sinogram_rejected(gpu):
    import torch
    from trace_amd import Projector, TraceNonFinite
    N, P = 16, 8
    th = np.linspace(0, 3, P).astype(np.float32)
    sino = torch.ones((P, 1, N), device="cuda")
    sino[3, 0, 5] = float("nan")
    x = torch.ones((1, N, N), device="cuda")
    with Projector(th, 1, N, N / 2) as pj:
        with pytest.raises(TraceNonFinite):
            pj.reconstruct("mlem", 1, sino, x)


@pytest.mark.parametrize("shape", [(4, 64, 64, 2), (8, 32, 32, 8), (32, 16, 16, 16),
                                   (16, 16, 16, 8), (16, 24, 24, 8), (1, 32, 32, 1),
                                   (2, 64, 64, 1), (2, 16, 16, 2), (16, 16, 16, 32, 2),
                                   (16, 32, 32, 16, 4), (8, 32, 16, 16, 2)])
@pytest.mark.parametrize("alg", ["sirt", "mlem"])
def test_launch_shapes_recon(gpu, oracle, alg, shape):
    """Every compiled launch shape gives the same reconstruction."""
    import torch
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 72, 50, 9, 4
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=7, I0=1e4, K=6).numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_launch_shape(*shape[:4])
        if len(shape) > 4:
            pj.set_fwd_lanes(shape[4])  # lanes per ray
        dx = torch.from_numpy(init).cuda()
        pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx)
        torch.cuda.synchronize()
        x = dx.cpu().numpy()
    xo = oracle.reconstruct(alg, th, sino, init, N / 2, iters)
    assert rel_l2(x, xo) < TOL, shape


@pytest.mark.parametrize("fb", [1, 2, 4, 8])
@pytest.mark.parametrize("alg,compat", [("pml", 1), ("apmlr", 1), ("apmlr", 0)])
def test_regularised_slice_batches(gpu, oracle, alg, compat, fb):
    """The fused stencil update refreshes the slice-batched image copy for
    every forward slice batch; 8 slices fill the last batch exactly, so the
    compat APMLR slot of the last slice lies past the batched copy."""
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 40, 30, 8, 4
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=11, I0=1e4, K=6).numpy()
    init = np.ones((S, N, N), np.float32)
    prm = Params(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=compat)
    rp = RegParams(0.1, 1.0, 0.01, 1.0, 1.0, bool(compat))
    print(f"{alg} compat={compat} forward batch {fb}", flush=True)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_launch_shape(fb, 16, 16, 4)
        dx = torch.from_numpy(init).cuda()
        pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx, rp)
        x = dx.cpu().numpy()
    xo = oracle.reconstruct(alg, th, sino, init, N / 2, iters, prm)
    assert rel_l2(x, xo) < TOL


@pytest.mark.parametrize("alg,compat", [("pml", 1), ("apmlr", 1), ("apmlr", 0)])
def test_stencil_recon_independent_of_forward_batch(gpu, alg, compat):
    """PML and APMLR reconstructions are bit-identical for forward slice
    batches of 32 and 8 over 40 slices: the last batch is partial at 32
    and full at 8, and the fused stencil update refreshes the slice-batched
    copy of every slot either way."""
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 64, 40, 40, 3
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=5, I0=1e4, K=6).numpy()
    init = np.ones((S, N, N), np.float32)
    rp = RegParams(0.1, 1.0, 0.01, 1.0, 1.0, bool(compat))
    outs = []
    for fb in (32, 8):
        print(f"{alg} compat={compat} forward batch {fb}", flush=True)
        with Projector(th, S, N, N / 2) as pj:
            pj.set_launch_shape(fb, 16, 16, 8)
            dx = torch.from_numpy(init).cuda()
            pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx, rp)
            outs.append(dx.cpu().numpy())
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0].view(np.int32), outs[1].view(np.int32))


@pytest.mark.parametrize("alg", ["mlem", "sirt", "pml"])
def test_slice_independence(gpu, alg):
    """A slice reconstructed alone equals the same slice inside a stack, bit
    for bit (per-slice forward sums, per-slice fixed-point scale, per-pixel
    update), for the stack's default launch shape."""
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 96, 64, 37, 4
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=21, I0=1e4, K=6).numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    rp = RegParams(beta=0.1) if alg == "pml" else None
    with Projector(th, S, N, N / 2) as pj:
        dx = torch.from_numpy(init).cuda()
        pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx, rp)
        stack = dx.cpu().numpy()
    for s in (0, 17, 36):
        with Projector(th, 1, N, N / 2) as pj:
            dx = torch.from_numpy(init[s:s + 1].copy()).cuda()
            pj.reconstruct(alg, iters, torch.from_numpy(sino[:, s:s + 1, :].copy()).cuda(), dx, rp)
            alone = dx.cpu().numpy()[0]
        assert np.array_equal(alone.view(np.int32), stack[s].view(np.int32)), s


@pytest.mark.parametrize("alg", ["sirt", "mlem", "pml"])
def test_default_centre_is_reproducible(gpu, oracle, alg):
    """With the reference's default centre (center <= 0 selects C/2 + 1,
    trace_data.h:97) the rays of one edge column miss the image: y = 0 and
    a2 = 0 give them inf/NaN weights (as in the reference), which add
    nothing to A^T w.  They must not set the slice's fixed-point scale, or
    the slice falls back to float accumulation: the result then varies
    from run to run.  Checks that such rays exist, that two runs agree bit
    for bit, and parity with the oracle."""
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 64, 60, 3, 3
    th = angles_rad(P)
    sino = np.abs(noisy_sinogram(th, N, S, seed=5, I0=1e4, K=8).numpy())
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    rp = RegParams(beta=0.5) if alg == "pml" else None
    outs = []
    with Projector(th, S, N, 0.0) as pj:
        cnt, _, _, _ = pj.dump_rays(np.tile(np.arange(P), 2), np.repeat([0, N - 1], P))
        assert (np.asarray(cnt) == 0).any(), "no ray misses the image"
        for _ in range(2):
            dx = torch.from_numpy(init).cuda()
            pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx, rp)
            outs.append(dx.cpu().numpy())
    assert np.array_equal(outs[0].view(np.int32), outs[1].view(np.int32))
    xo = oracle.reconstruct(alg, th, sino, init, 0.0, iters, Params(beta=0.5) if rp else None)
    assert np.isfinite(outs[0]).all()
    assert rel_l2(outs[0], xo) <= 1e-4
