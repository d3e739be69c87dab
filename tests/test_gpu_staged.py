"""Slab-staged forward projector (TRACE_OPT_FWD_STAGED) against the
whole-ray form and the oracle.

Each lane of the staged form walks its own ray in the reference's order
(sirt.cc:4-19) and only the source of each pixel line changes (LDS or HBM),
so its ray sums, weights and reconstructions must equal the whole-ray
form's bit for bit, on geometries that exercise the fallbacks: rays on grid
lines, row wraps, N != C, edge columns outside the image, partial bundles
and slice batches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# N, C, P, centre, S
GEOMS = [
    (64, 64, 48, 32.0, 32),      # centred, one slice batch
    (63, 63, 50, 31.5, 40),      # odd side, half-pixel centre; 32 + 8 slices
    (64, 64, 40, 32.5, 32),      # off-centre axis: segments that wrap rows
    (96, 64, 37, 48.0, 33),      # image wider than the detector
    (64, 96, 37, 48.0, 32),      # detector wider than the image
    (64, 64, 30, 0.0, 32),       # the reference's default centre
    (300, 300, 181, 150.0, 64),  # many slabs and bundles, two batches
    (8, 8, 5, 4.0, 32),          # tiny: two slabs, a bundle of 5 angles
    (5, 7, 9, 3.5, 32),          # odd side (a partial last slab), detector wider
    (130, 70, 17, 35.0, 32),     # bundles cut at 16 angles and 64 columns
]


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


def _angles(kind, P, seed):
    from trace_amd.phantom import angles_rad
    if kind == "even":
        return angles_rad(P)
    return np.random.default_rng(seed).uniform(-0.5, 3.5, P).astype(np.float32)


def _forward(th, S, C, center, N, x, staged):
    import torch
    from trace_amd import Projector
    with Projector(th, S, C, center, num_grids=N) as pj:
        pj.set_fwd_staged(staged)
        y = pj.forward(torch.from_numpy(x).cuda())
        torch.cuda.synchronize()
        return y.cpu().numpy()


@pytest.mark.parametrize("kind", ["even", "random"])
@pytest.mark.parametrize("N,C,P,center,S", GEOMS)
def test_forward_staged_bit_identical(gpu, N, C, P, center, S, kind):
    th = _angles(kind, P, N + P)
    x = np.random.default_rng(N * S).uniform(0, 1, (S, N, N)).astype(np.float32)
    print(f"staged forward N={N} C={C} P={P} centre={center} S={S} angles={kind}", flush=True)
    ys = _forward(th, S, C, center, N, x, True)
    yw = _forward(th, S, C, center, N, x, False)
    bad = int(np.count_nonzero(_bits(ys) != _bits(yw)))
    print(f"{bad} of {ys.size} ray sums differ")
    assert bad == 0


def test_forward_staged_vs_oracle(gpu, oracle):
    N, P, S = 64, 45, 32
    th = _angles("random", P, 7)
    x = np.random.default_rng(3).uniform(0, 1, (S, N, N)).astype(np.float32)
    print(f"staged forward vs oracle N={N} P={P} S={S}", flush=True)
    ys = _forward(th, S, N, N / 2, N, x, True)
    assert np.array_equal(_bits(ys), _bits(oracle.forward(th, x, N / 2)))


@pytest.mark.parametrize("alg", ["sirt", "mlem", "pml", "apmlr"])
def test_recon_staged_bit_identical(gpu, alg):
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S, iters = 96, 64, 34, 3
    th = angles_rad(P)
    sino = noisy_sinogram(th, N, S, seed=11, I0=1e4, K=8).numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    print(f"staged {alg} N={N} P={P} S={S} iters={iters}", flush=True)
    out = []
    for staged in (True, False):
        with Projector(th, S, N, N / 2.0) as pj:
            pj.set_fwd_staged(staged)
            ds = torch.from_numpy(sino).cuda()
            dx = torch.from_numpy(init).cuda()
            pj.reconstruct(alg, iters, ds, dx, RegParams(beta=0.1))
            torch.cuda.synchronize()
            out.append(dx.cpu().numpy())
    bad = int(np.count_nonzero(_bits(out[0]) != _bits(out[1])))
    print(f"{alg}: {bad} of {out[0].size} pixels differ")
    assert bad == 0


def test_forced_32_slice_batch_with_fewer_slices(gpu):
    """A 32-slice forward batch chosen for a 5-slice stack: the staged form
    reads whole 32-slot lines and writes only the live slices."""
    import torch
    from trace_amd import Projector
    N, P, S = 64, 40, 5
    th = _angles("random", P, 11)
    x = np.random.default_rng(5).uniform(0, 1, (S, N, N)).astype(np.float32)
    print(f"staged forward, forced 32-slice batch, S={S}", flush=True)
    ys = []
    for staged in (True, False):
        with Projector(th, S, N, N / 2.0) as pj:
            pj.set_launch_shape(32, 16, 16, 32)
            pj.set_fwd_staged(staged)
            y = pj.forward(torch.from_numpy(x).cuda())
            torch.cuda.synchronize()
            ys.append(y.cpu().numpy())
    assert np.array_equal(_bits(ys[0]), _bits(ys[1]))


def test_staged_option(gpu):
    """TRACE_OPT_FWD_STAGED: on by default, 0 or 1 only."""
    from trace_amd import Projector, _capi
    from trace_amd.phantom import angles_rad
    with Projector(angles_rad(8), 32, 16, 8.0) as pj:
        assert pj.get_option(_capi.TRACE_OPT_FWD_STAGED) == 1
        pj.set_fwd_staged(False)
        assert pj.get_option(_capi.TRACE_OPT_FWD_STAGED) == 0
        rc = _capi.lib().trace_set_option(pj.handle, _capi.TRACE_OPT_FWD_STAGED, 2)
        assert rc == _capi.TRACE_E_INVALID
        assert pj.get_option(_capi.TRACE_OPT_FWD_STAGED) == 0


def test_forward_staged_fast_sqrt(gpu):
    """Hardware square roots (TRACE_OPT_EXACT_SQRT = 0): the two forms still
    agree bit for bit."""
    import torch
    from trace_amd import Projector
    N, P, S = 96, 60, 32
    th = _angles("random", P, 13)
    x = np.random.default_rng(9).uniform(0, 1, (S, N, N)).astype(np.float32)
    print(f"staged forward, hardware sqrt, N={N} P={P} S={S}", flush=True)
    ys = []
    for staged in (True, False):
        with Projector(th, S, N, N / 2.0, exact_sqrt=False) as pj:
            pj.set_fwd_staged(staged)
            y = pj.forward(torch.from_numpy(x).cuda())
            torch.cuda.synchronize()
            ys.append(y.cpu().numpy())
    assert np.array_equal(_bits(ys[0]), _bits(ys[1]))
