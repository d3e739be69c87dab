"""Batched reconstruction of uneven slices (one HIP-graph launch) on the GPU.

Bar: every slice equals its own single-slice reconstruction bit for bit
(same kernels per slice, deterministic fixed-point backprojection), and
matches the oracle to 1e-4 relative L2.  Graph replays, interleaved with
eager work, reproduce the first run exactly."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _specs():
    rng = np.random.default_rng(31)
    out = []
    for i in range(10):
        C = int(rng.choice([48, 64, 97, 128, 160]))
        P = int(rng.integers(30, 121))
        alg = "sirt" if rng.random() < 0.5 else "mlem"
        out.append((C, P, alg, int(rng.integers(2, 7))))
    return out


def _make(specs, seed=5):
    from trace_amd.batch import uneven_slices
    sl = uneven_slices([(C, P, a) for C, P, a, _ in specs], seed=seed, I0=1e4, K=8)
    for s, (_, _, _, it) in zip(sl, specs):
        s.iters = it
    return sl


def _single(s):
    import torch
    from trace_amd import Projector
    C = s.recon.shape[0]
    x = s.recon.clone().reshape(1, C, C)
    with Projector(s.theta, 1, C, s.center) as pj:
        pj.set_launch_shape(1, 32, 32, 1)
        pj.reconstruct(s.alg, s.iters, s.sino.reshape(-1, 1, C).contiguous(), x)
    torch.cuda.synchronize()
    return x.reshape(C, C)


def test_batch_equals_single_slices(gpu, oracle):
    torch = gpu
    from trace_amd.batch import Batch
    specs = _specs()
    sl = _make(specs)
    init = [s.recon.clone() for s in sl]
    expect = []
    for s, x0 in zip(sl, init):
        s.recon.copy_(x0)
        expect.append(_single(s))
        s.recon.copy_(x0)
    print("batch of", [(C, P, a, it) for C, P, a, it in specs], flush=True)
    with Batch(sl) as b:
        b.run()
        torch.cuda.synchronize()
    for i, (s, e) in enumerate(zip(sl, expect)):
        assert torch.equal(s.recon, e), f"slice {i} {specs[i]}"
    for i in (0, 3, 7):
        C, P, alg, it = specs[i]
        sino = sl[i].sino.cpu().numpy().reshape(P, 1, C)
        xo = oracle.reconstruct(alg, sl[i].theta, sino, init[i].cpu().numpy().reshape(1, C, C),
                                C / 2.0, it)
        e = rel_l2(sl[i].recon.cpu().numpy(), xo[0])
        print(f"slice {i} {specs[i]}: batch vs oracle rel L2 {e:.2e}")
        assert e < 1e-4


def test_batch_graph_replays(gpu):
    """Replays of the captured graph, with eager work and synchronisation
    between them, give the first run's result every time."""
    torch = gpu
    from trace_amd.batch import Batch
    sl = _make(_specs()[:6], seed=9)
    init = [s.recon.clone() for s in sl]
    first = None
    with Batch(sl) as b:
        for rep in range(4):
            for s, x0 in zip(sl, init):
                s.recon.copy_(x0)
            junk = torch.randn(1 << 20, device="cuda").sum()  # eager work in between
            b.run()
            torch.cuda.synchronize()
            out = [s.recon.clone() for s in sl]
            assert torch.isfinite(junk)
            if first is None:
                first = out
            else:
                for a, c in zip(first, out):
                    assert torch.equal(a, c), f"replay {rep}"


@pytest.mark.parametrize("bundle", [1, 4, 16])
@pytest.mark.parametrize("tile", [(16, 16), (32, 32), (64, 64)])
def test_batch_launch_shapes(gpu, bundle, tile):
    torch = gpu
    from trace_amd.batch import Batch
    sl = _make(_specs()[:5], seed=13)
    init = [s.recon.clone() for s in sl]
    ref = []
    with Batch(sl) as b:
        b.run()
        torch.cuda.synchronize()
        ref = [s.recon.clone() for s in sl]
        for s, x0 in zip(sl, init):
            s.recon.copy_(x0)
        b.set_launch_shape(bundle, *tile)
        assert b.launch_shape() == (bundle, *tile)
        b.run()
        torch.cuda.synchronize()
    for a, s in zip(ref, sl):
        assert torch.equal(a, s.recon)


def test_batch_max_iters(gpu):
    torch = gpu
    from trace_amd.batch import Batch
    specs = [(C, P, a, 5) for C, P, a, _ in _specs()[:4]]
    sl = _make(specs, seed=17)
    init = [s.recon.clone() for s in sl]
    with Batch(sl) as b:
        b.run(max_iters=2)
        torch.cuda.synchronize()
    for s, x0 in zip(sl, init):
        got = s.recon.clone()
        s.recon.copy_(x0)
        s.iters = 2
        assert torch.equal(got, _single(s))


def test_batch_nonfinite(gpu):
    torch = gpu
    from trace_amd._capi import TraceNonFinite
    from trace_amd.batch import Batch
    sl = _make(_specs()[:3], seed=19)
    sl[1].sino[3, 4] = float("nan")
    with Batch(sl) as b:
        with pytest.raises(TraceNonFinite):
            b.run()


@pytest.mark.parametrize("N,side", [(256, 16), (767, 16), (768, 32), (1536, 32), (1792, 64),
                                    (2048, 64)])
def test_one_slice_default_tile_follows_image_side(gpu, N, side):
    """Without autotuning a one-slice handle takes 16x16 tiles up to 767
    pixels, 32x32 up to 1791 and 64x64 above (the autotuner's picks at 1500
    angles); the batch defaults to 8-angle bundles and 64x64 tiles."""
    from trace_amd import Projector
    from trace_amd.batch import Batch
    from trace_amd.phantom import angles_rad
    th = angles_rad(30)
    with Projector(th, 1, N, N / 2.0) as pj:
        assert pj.launch_shape() == (1, side, side, 1)
    if N == 256:
        sl = _make(_specs()[:2], seed=3)
        with Batch(sl) as b:
            assert b.launch_shape() == (8, 64, 64)
