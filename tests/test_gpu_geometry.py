"""Image side N different from the detector column count C, on the GPU
against the oracle.  The reference's mains pass N = C (sirt.cc:107-108),
but its TraceMetadata and traversal functions take them separately; the
oracle's N != C traversal is pinned bit for bit against those functions
(test_oracle_vs_ref.py::test_traversal_bitexact_image_side_differs)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import Params

pytestmark = pytest.mark.gpu

# (N, C): image wider and narrower than the detector, sides that are not
# multiples of the tile widths
CASES = [(48, 64), (80, 56), (100, 64), (37, 50)]


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


def _setup(N, C, P=40, S=5, seed=0):
    from trace_amd.phantom import angles_rad
    rng = np.random.default_rng(seed + N * 100 + C)
    th = angles_rad(P)
    x = rng.uniform(0.0, 1.0, (S, N, N)).astype(np.float32)
    r = rng.standard_normal((P, S, C)).astype(np.float32)
    return th, x, r


@pytest.mark.parametrize("N,C", CASES)
@pytest.mark.parametrize("center", ["half", "default"])
@pytest.mark.parametrize("S", [5, 1])  # 1: the one-slice tiles' two-axis windows
def test_operators(gpu, oracle, N, C, center, S):
    import torch
    from trace_amd import Projector
    ctr = C / 2.0 if center == "half" else 0.0
    th, x, r = _setup(N, C, S=S)
    P = th.size
    with Projector(th, S, C, ctr, num_grids=N) as pj:
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
        z = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        den = pj.column_sums().cpu().numpy()
        p = np.array([0, 7, P // 2, P - 1] * 3, np.int32)
        c = np.array([0] * 4 + [C // 2] * 4 + [C - 1] * 4, np.int32)
        cnt, idx, ln, _ = pj.dump_rays(p, c)
    assert y.shape == (P, S, C) and z.shape == (S, N, N) and den.shape == (N, N)
    # forward: the reference's segment order, bit for bit
    assert np.array_equal(_bits(y), _bits(oracle.forward(th, x, ctr, C=C)))
    # adjoint and column sums: fixed-point accumulation vs fp64 sums
    assert rel_l2(z, oracle.adjoint(th, r, ctr, N=N)) < 1e-6
    ones = np.ones((P, 1, C), np.float32)
    assert rel_l2(den, oracle.adjoint(th, ones, ctr, N=N)[0]) < 1e-6
    for i in range(p.size):
        oi, ol, _ = oracle.trace_ray(N, C, ctr, th[p[i]], int(c[i]))
        assert cnt[i] == oi.size
        assert np.array_equal(idx[i, :cnt[i]], oi)
        assert np.array_equal(_bits(ln[i, :cnt[i]]), _bits(ol))


@pytest.mark.parametrize("N,C", CASES[:3])
@pytest.mark.parametrize("alg", ["sirt", "mlem", "pml", "apmlr"])
def test_reconstruct(gpu, oracle, N, C, alg):
    import torch
    from trace_amd import Projector, RegParams
    from trace_amd.phantom import angles_rad
    P, S, iters = 40, 4, 3
    th = angles_rad(P)
    rng = np.random.default_rng(N + C)
    truth = rng.uniform(0.1, 1.0, (S, N, N)).astype(np.float32)
    sino = np.abs(oracle.forward(th, truth, C / 2.0, C=C))
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    prm = {"pml": (RegParams(beta=0.5), Params(beta=0.5)),
           "apmlr": (RegParams(0.5, 0.25, 0.1, 0.2, 0.8, True),
                     Params(beta=0.5, beta1=0.25, delta=0.1, delta1=0.2, regw=0.8))}
    rp, op = prm.get(alg, (None, None))
    with Projector(th, S, C, C / 2.0, num_grids=N) as pj:
        dx = torch.from_numpy(init).cuda()
        pj.reconstruct(alg, iters, torch.from_numpy(sino).cuda(), dx, rp)
        got = dx.cpu().numpy()
    want = oracle.reconstruct(alg, th, sino, init, C / 2.0, iters, op)
    assert np.isfinite(got).all()
    assert rel_l2(got, want) <= 1e-4, rel_l2(got, want)


def test_batch_slice_with_other_image_side(gpu):
    """A batched slice with N != C equals its own single-slice
    reconstruction bit for bit, beside N = C slices."""
    import torch
    from trace_amd import Projector
    from trace_amd.batch import Batch, BatchSlice
    from trace_amd.phantom import angles_rad
    rng = np.random.default_rng(9)
    specs = [(64, 64, 48, "mlem"), (48, 64, 60, "sirt"), (90, 70, 36, "mlem")]  # (N, C, P, alg)
    slices, want = [], []
    for N, C, P, alg in specs:
        th = angles_rad(P)
        sino = rng.uniform(0.5, 2.0, (P, C)).astype(np.float32)
        init = np.full((N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
        with Projector(th, 1, C, C / 2.0, num_grids=N) as pj:
            dx = torch.from_numpy(init[None].copy()).cuda()
            pj.reconstruct(alg, 4, torch.from_numpy(sino[:, None, :].copy()).cuda(), dx)
            want.append(dx.cpu().numpy()[0])
        slices.append(BatchSlice(th, torch.from_numpy(sino).cuda(),
                                 torch.from_numpy(init.copy()).cuda(), alg, 4, C / 2.0))
    with Batch(slices) as b:
        b.run()
        torch.cuda.synchronize()
    for s, w, spec in zip(slices, want, specs):
        assert np.array_equal(_bits(s.recon.cpu().numpy()), _bits(w)), spec
