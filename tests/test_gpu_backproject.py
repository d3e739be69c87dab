"""The LDS-tiled backprojector against the whole-ray traversal and the
oracle.  The per-pixel segment counts are integers: the tile kernel (which
walks only a window of each ray around each tile) must produce exactly the
same counts as the whole-ray traversal and the oracle, for every launch
shape, including angles outside [0, pi), odd grids and off-centre rotation
axes.  Weighted backprojection agrees with the fp64 oracle adjoint to fp32
summation-order level."""
import numpy as np
import pytest

from conftest import adversarial_angles, rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16, 8), (4, 16, 16, 16), (8, 16, 16, 4), (32, 24, 24, 8), (4, 32, 16, 8),
          (4, 16, 32, 8), (16, 32, 32, 4), (8, 32, 32, 8), (4, 8, 8, 16), (4, 64, 64, 2),
          (1, 32, 32, 1), (2, 64, 64, 1), (1, 32, 32, 2), (2, 16, 16, 2), (1, 16, 16, 1),
          (16, 16, 16, 32), (16, 32, 32, 16), (8, 32, 16, 16), (4, 24, 24, 16)]


def _oracle_counts(oracle, th, N, center):
    cnt = np.zeros(N * N, np.int64)
    for t in th:
        for c in range(N):
            idx, _, _ = oracle.trace_ray(N, N, center, t, c)
            idx = idx[(idx >= 0) & (idx < N * N)]
            np.add.at(cnt, idx, 1)
    return cnt.reshape(N, N)


@pytest.mark.parametrize("N,center,kind", [
    (64, 32.0, "even"), (63, 31.5, "odd"), (96, 0.0, "default-centre"),
    (128, 60.3, "off-centre"),
    # rays on grid lines: vertical rays along the right edge put every
    # segment in pixel 0 of the next row (ix == N, trace_utils.cc:177-185)
    (64, 32.5, "on-lines"), (63, 31.0, "on-lines-odd"), (40, 19.5, "on-lines-small")])
def test_segment_counts_exact(gpu, oracle, N, center, kind):
    from trace_amd import Projector
    rng = np.random.default_rng(N)
    th = np.concatenate([adversarial_angles(), rng.uniform(-7, 7, 20).astype(np.float32),
                         np.float32(np.pi / 2) + rng.normal(0, 1e-3, 6).astype(np.float32)])
    expect = _oracle_counts(oracle, th, N, center)
    print(f"segment counts N={N} {kind}: {expect.sum()} segments", flush=True)
    with Projector(th, 1, N, center) as pj:
        pj.set_backprojector("atomic")
        atomic = pj.segment_counts().cpu().numpy()
        assert np.array_equal(atomic, expect)
        pj.set_backprojector("tile")
        for fb, tw, th_, tb in SHAPES:
            pj.set_launch_shape(fb, tw, th_, tb)
            tile = pj.segment_counts().cpu().numpy()
            assert np.array_equal(tile, expect), (fb, tw, th_, tb)


def test_segment_counts_large(gpu):
    """2048 x 2048, 1500 angles: tile windows vs whole rays, integer-exact."""
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    N, P = 2048, 1500
    with Projector(angles_rad(P), 1, N, N / 2) as pj:
        pj.set_backprojector("atomic")
        a = pj.segment_counts().cpu().numpy()
        pj.set_backprojector("tile")
        t = pj.segment_counts().cpu().numpy()
        nnz, _ = pj.count_segments()
    assert int(a.astype(np.int64).sum()) == nnz
    assert np.array_equal(a, t)


@pytest.mark.parametrize("shape", SHAPES)
def test_adjoint_vs_oracle(gpu, oracle, shape):
    torch = gpu
    from trace_amd import Projector
    N, P, S = 96, 40, 5
    rng = np.random.default_rng(sum(shape))
    th = np.concatenate([rng.uniform(0, np.pi, P - 4), [0.0, 1.5707964, 3.1415925, 4.0]]).astype(np.float32)
    r = rng.normal(size=(P, S, N)).astype(np.float32)
    r[rng.uniform(size=r.shape) < 0.2] = 0.0
    z64 = oracle.adjoint(th, r, N / 2)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_launch_shape(*shape)
        z = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        pj.set_backprojector("atomic")
        za = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
    assert rel_l2(z, z64) < 1e-6, shape
    assert rel_l2(za, z64) < 1e-6


@pytest.mark.parametrize("fb", [1, 2, 4, 8, 16, 32])
def test_forward_batches_bitexact(gpu, oracle, fb):
    torch = gpu
    from trace_amd import Projector
    N, P, S = 80, 30, 19
    rng = np.random.default_rng(fb)
    th = rng.uniform(-1, 4, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_launch_shape(fb, 32, 32, 4)
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y.view(np.int32), oracle.forward(th, x, N / 2).view(np.int32))


def test_column_sums_match(gpu, oracle):
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    N, P = 200, 120
    th = angles_rad(P)
    den64 = oracle.adjoint(th, np.ones((P, 1, N), np.float32), N / 2)[0]
    with Projector(th, 1, N, N / 2) as pj:
        d_tile = pj.column_sums().cpu().numpy()
        pj.set_backprojector("atomic")
        d_atomic = pj.column_sums().cpu().numpy()
    assert rel_l2(d_tile, den64) < 1e-6
    assert rel_l2(d_atomic, den64) < 1e-6


def test_backprojection_deterministic(gpu):
    """Fixed-point LDS accumulation: repeated backprojections are
    bit-identical, whatever the order in which waves arrive."""
    torch = gpu
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    N, P, S = 256, 180, 6
    rng = np.random.default_rng(0)
    r = torch.from_numpy(rng.normal(size=(P, S, N)).astype(np.float32)).cuda()
    with Projector(angles_rad(P), S, N, N / 2) as pj:
        z1 = pj.adjoint(r).cpu().numpy()
        for _ in range(3):
            z2 = pj.adjoint(r).cpu().numpy()
            assert np.array_equal(z1.view(np.int32), z2.view(np.int32))


def test_nonfinite_weights_propagate(gpu):
    """A slice whose weights hold inf (the reference's y = 0 rays) is summed
    in float so that inf/NaN reach the same pixels as a float scatter;
    the other slices keep the fixed-point path."""
    torch = gpu
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    N, P, S = 64, 40, 3
    rng = np.random.default_rng(1)
    r = rng.uniform(0.5, 1.5, (P, S, N)).astype(np.float32)
    r[7, 1, 30] = np.inf
    with Projector(angles_rad(P), S, N, N / 2) as pj:
        z = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        pj.set_backprojector("atomic")
        za = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
    assert np.array_equal(np.isfinite(z), np.isfinite(za))
    assert not np.isfinite(z[1]).all()
    assert np.isfinite(z[0]).all() and np.isfinite(z[2]).all()
    fin = np.isfinite(za)
    assert rel_l2(z[fin], za[fin]) < 1e-6


@pytest.mark.parametrize("ba", [1, 2, 4, 8, 16])
def test_forward_bundles_bitexact(gpu, oracle, ba):
    """Ray bundles across adjacent angles change only the lane mapping."""
    torch = gpu
    from trace_amd import Projector
    N, P, S = 70, 37, 5
    rng = np.random.default_rng(ba)
    th = rng.uniform(-1, 4, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_fwd_bundle(ba)
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y.view(np.int32), oracle.forward(th, x, N / 2).view(np.int32))


@pytest.mark.parametrize("fb,lanes,ba", [(8, 2, 8), (16, 2, 4), (32, 2, 16), (16, 4, 8),
                                         (32, 4, 16), (16, 4, 1)])
def test_forward_lanes_per_ray_bitexact(gpu, oracle, fb, lanes, ba):
    """Splitting a ray's slice batch over 2 or 4 lanes leaves every ray sum
    bit-identical (same segment order per slice)."""
    torch = gpu
    from trace_amd import Projector
    N, P, S = 80, 30, 37
    rng = np.random.default_rng(fb * lanes + ba)
    th = rng.uniform(-1, 4, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    with Projector(th, S, N, N / 2) as pj:
        pj.set_launch_shape(fb, 16, 16, 8)
        pj.set_fwd_lanes(lanes)
        pj.set_fwd_bundle(ba)
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y.view(np.int32), oracle.forward(th, x, N / 2).view(np.int32))


def test_forward_lanes_rejects_uncompiled(gpu):
    from trace_amd import Projector, TraceUnsupported
    th = np.linspace(0, 3, 10).astype(np.float32)
    with Projector(th, 4, 32, 16.0) as pj:
        pj.set_launch_shape(4, 16, 16, 4)
        with pytest.raises(TraceUnsupported):
            pj.set_fwd_lanes(2)  # batch 4 has no 2-lane form


def test_bench_launch_shape_full_batches(gpu, oracle):
    """The launch shape bench.py runs (forward: 32 slices per ray batch, 4
    lanes per ray, 16-angle bundles; tiles 16x16 with 32 slices) on a stack
    that fills two whole 32-slice batches: forward bit-exact, adjoint within
    1e-6 of the oracle's fp64 adjoint, and two MLEM iterations within 1e-4
    of the oracle."""
    torch = gpu
    from trace_amd import Projector
    from trace_amd.autotune import apply_shape
    from trace_amd.phantom import angles_rad, noisy_sinogram
    N, P, S = 128, 90, 64
    th = angles_rad(P)
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    r = rng.normal(size=(P, S, N)).astype(np.float32)
    sino = noisy_sinogram(th, N, S, seed=9, I0=1e4, K=6).numpy()
    print(f"bench shape N={N} P={P} S={S}", flush=True)
    with Projector(th, S, N, N / 2.0) as pj:
        apply_shape(pj, {"shape": [32, 16, 16, 32], "lanes": 4, "bundle": 16})
        assert pj.launch_shape() == (32, 16, 16, 32)
        assert pj.fwd_lanes() == 4 and pj.fwd_bundle() == 16
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
        z = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        xr = torch.ones((S, N, N), device="cuda")
        pj.reconstruct("mlem", 2, torch.from_numpy(sino).cuda(), xr)
        torch.cuda.synchronize()
        xr = xr.cpu().numpy()
    assert np.array_equal(y.view(np.int32), oracle.forward(th, x, N / 2.0).view(np.int32))
    assert rel_l2(z, oracle.adjoint(th, r, N / 2.0)) < 1e-6
    xo = oracle.reconstruct("mlem", th, sino, np.ones((S, N, N), np.float32), N / 2.0, 2)
    assert rel_l2(xr, xo) < 1e-4
