"""Backprojected sums in the reference's order (TRACE_OPT_ORDERED_SUMS).

The reference sums den = A^T 1 and num in float, segment by segment, in the
order of its one-thread engine (projections, then columns, then segments:
sirt.cc:61-62, mlem.cc:56-57, pml.cc:243-244, apmlr.cc:1024-1025).  The
oracle restates that order and is bit-exact to the compiled reference at one
thread, so here the GPU's ordered sums, and the reconstructions built on
them, must equal the oracle's bit for bit."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import Params

pytestmark = pytest.mark.gpu

# (N, C, P, centre): centred; odd side; rays on grid lines at a half-pixel
# centre, odd C - N and an off-centre axis (segments that wrap to the next
# row, window-unsafe rays); image side differing from the detector width;
# the reference's default centre (an edge column outside the image)
GEOMS = [
    (64, 64, 48, 32.0), (63, 63, 50, 31.5), (64, 64, 40, 32.5), (64, 64, 36, 31.0),
    (40, 40, 30, 19.5), (64, 48, 40, 24.0), (48, 64, 40, 32.0), (64, 64, 30, 0.0),
]


def _den_oracle(oracle, th, N, C, center):
    sino = np.zeros((th.size, 1, C), np.float32)
    _, st = oracle.reconstruct("sirt", th, sino, np.zeros((1, N, N), np.float32), center,
                               1, want_stages=True)
    return st["den"][0]


@pytest.mark.parametrize("tile", [True, False])
@pytest.mark.parametrize("N,C,P,center", GEOMS)
def test_column_sums_bit_exact(gpu, oracle, N, C, P, center, tile):
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    th = angles_rad(P)
    print(f"ordered den N={N} C={C} P={P} centre={center} tile={tile}", flush=True)
    with Projector(th, 1, C, center, num_grids=N) as pj:
        pj.set_ordered_tile(tile)
        d = pj.column_sums().cpu().numpy()
    do = _den_oracle(oracle, th, N, C, center)
    bad = int(np.count_nonzero(d.view(np.uint32) != do.view(np.uint32)))
    print(f"N={N} C={C} P={P} centre={center}: {bad} of {d.size} pixels differ")
    assert bad == 0


def _recon(alg, th, sino, init, center, iters, prm, level, N, tile=True):
    import torch
    from trace_amd import Projector, RegParams
    P, S, C = sino.shape
    rp = RegParams(prm.beta, prm.beta1, prm.delta, prm.delta1, prm.regw,
                   bool(prm.apmlr_compat))
    with Projector(th, S, C, center, num_grids=N) as pj:
        pj.set_ordered_sums(level)
        pj.set_ordered_tile(tile)
        ds = torch.from_numpy(np.ascontiguousarray(sino, np.float32)).cuda()
        dx = torch.from_numpy(np.ascontiguousarray(init, np.float32)).cuda()
        pj.reconstruct(alg, iters, ds, dx, rp)
        torch.cuda.synchronize()
        return dx.cpu().numpy()


CASES = [
    # alg, N, C, P, S, iters, centre, compat, level
    ("pml", 64, 64, 48, 3, 6, 32.0, 1, 1),
    ("pml", 63, 63, 40, 2, 5, 31.5, 1, 1),
    ("pml", 64, 64, 40, 2, 4, 32.5, 1, 1),
    ("apmlr", 48, 48, 36, 4, 5, 24.0, 1, 1),
    ("apmlr", 48, 48, 36, 4, 5, 24.0, 0, 1),
    ("apmlr", 40, 40, 30, 3, 4, 19.5, 1, 1),
    ("sirt", 64, 64, 48, 3, 6, 32.0, 1, 2),
    ("mlem", 64, 64, 48, 3, 6, 32.0, 1, 2),
    ("mlem", 64, 48, 40, 2, 4, 24.0, 1, 2),
    ("pml", 256, 256, 300, 1, 5, 128.0, 1, 1),
    ("pml", 96, 96, 64, 40, 3, 48.0, 1, 1),   # 32-slice batches of the tile form
    ("mlem", 72, 72, 50, 33, 3, 36.0, 1, 2),  # 32 + 1
]


@pytest.mark.parametrize("tile", [True, False])
@pytest.mark.parametrize("alg,N,C,P,S,iters,center,compat,level", CASES)
def test_recon_bit_exact(gpu, oracle, alg, N, C, P, S, iters, center, compat, level, tile):
    from trace_amd.phantom import angles_rad, noisy_sinogram
    th = angles_rad(P)
    sino = noisy_sinogram(th, C, S, seed=N + S + P, I0=1e4, K=8).numpy()
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    prm = Params(beta=0.1, beta1=1.0, delta=0.01, delta1=1.0, regw=1.0, apmlr_compat=compat)
    print(f"ordered {alg} N={N} C={C} P={P} S={S} it={iters} c={center} compat={compat}",
          flush=True)
    x = _recon(alg, th, sino, init, center, iters, prm, level, N, tile)
    xo = oracle.reconstruct(alg, th, sino, init, center, iters, prm)
    bad = int(np.count_nonzero(x.view(np.uint32) != xo.view(np.uint32)))
    print(f"{alg}: {bad} of {x.size} pixels differ, rel L2 {rel_l2(x, xo):.2e}")
    assert bad == 0


def test_fixed_point_sums_still_selectable(gpu, oracle):
    """Level 0 keeps the fixed-point sums: same bar as before (1e-4)."""
    from trace_amd.phantom import angles_rad, noisy_sinogram
    th = angles_rad(48)
    sino = noisy_sinogram(th, 64, 2, seed=5, I0=1e4, K=8).numpy()
    init = np.ones((2, 64, 64), np.float32)
    prm = Params(beta=0.1)
    x = _recon("pml", th, sino, init, 32.0, 4, prm, 0, 64)
    xo = oracle.reconstruct("pml", th, sino, init, 32.0, 4, prm)
    assert rel_l2(x, xo) < 1e-4
