"""The oracle restatement against the compiled reference (oracle/_ref), on
fresh random inputs: bit-exact at one thread; the reference's threaded
engine (replicas + tree merge) agrees to fp32 summation-order level."""
import numpy as np
import pytest

from conftest import adversarial_angles, rel_l2
from oracle import Params


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


@pytest.mark.parametrize("N", [4, 5, 16, 63, 128, 512])
def test_traversal_bitexact(oracle, ref, N):
    rng = np.random.default_rng(N)
    angles = np.concatenate([adversarial_angles(), rng.uniform(-8, 8, 30).astype(np.float32)])
    for th in angles:
        for ctr in (N / 2.0, -1.0, N / 2.0 - 0.37):
            for col in sorted({0, 1, N // 2, N - 2, N - 1, int(rng.integers(0, N))}):
                a = oracle.trace_ray(N, N, ctr, th, col)
                b = ref.trace_ray(N, ctr, th, col)
                assert a[0].size == b[0].size, (N, th, ctr, col)
                assert np.array_equal(a[0], b[0])
                assert np.array_equal(_bits(a[1]), _bits(b[1]))
                assert np.array_equal(_bits(a[2]), _bits(b[2]))


def test_forward_bitexact(oracle, ref):
    rng = np.random.default_rng(7)
    for N, P, S in [(16, 9, 2), (40, 31, 1), (64, 17, 3)]:
        x = rng.uniform(-1, 2, (S, N, N)).astype(np.float32)
        th = rng.uniform(-1, 4, P).astype(np.float32)
        assert np.array_equal(_bits(oracle.forward(th, x, N / 2)), _bits(ref.forward(th, x, N / 2)))


CASES = [
    ("sirt", 24, 15, 1, 3), ("sirt", 17, 9, 3, 2),
    ("mlem", 24, 15, 2, 3), ("mlem", 31, 12, 1, 4),
    ("pml", 20, 14, 2, 3), ("pml", 16, 9, 1, 2),
    ("apmlr", 16, 10, 1, 2), ("apmlr", 16, 10, 2, 2), ("apmlr", 12, 8, 3, 2),
    ("apmlr", 14, 8, 5, 3),
]


@pytest.mark.parametrize("alg,N,P,S,iters", CASES)
def test_reconstruct_bitexact(oracle, ref, alg, N, P, S, iters):
    rng = np.random.default_rng(N * 100 + P + S)
    sino = rng.uniform(0.05, 4.0, (P, S, N)).astype(np.float32)
    if alg == "sirt":
        sino *= np.where(rng.uniform(size=sino.shape) < 0.2, -1, 1).astype(np.float32)
    deg = (np.arange(P) * 180.0 / P).astype(np.float32) + np.float32(rng.uniform(0, 3))
    prm = Params(beta=float(rng.uniform(0.05, 2)), beta1=float(rng.uniform(0.1, 2)),
                 delta=float(rng.uniform(0.01, 1)), delta1=float(rng.uniform(0.1, 2)),
                 regw=float(rng.uniform(0.1, 2)), apmlr_compat=1)
    center = N / 2.0
    th_file = oracle.deg2rad(deg) if alg == "pml" else deg
    xr, sr = ref.run(alg, th_file, sino, center, iters, 1, prm, want_stages=True)
    th = th_file if alg == "pml" else oracle.deg2rad(th_file)
    s_in = sino if alg == "sirt" else oracle.absolute(sino)
    init = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    xo, so = oracle.reconstruct(alg, th, s_in, init, center, iters, prm, want_stages=True)
    assert np.array_equal(_bits(xo), _bits(xr))
    for k in sr:
        assert np.array_equal(_bits(so[k]), _bits(sr[k])), k


@pytest.mark.parametrize("alg", ["sirt", "mlem", "pml", "apmlr"])
def test_threaded_reference_close(oracle, ref, alg):
    """The reference's multi-threaded engine (per-thread replicas, pairwise
    merge) only reorders fp32 sums; PML's root amplifies that a little, so
    the bound is the parity bar, 1e-4 (measured: <= 1.6e-5)."""
    N, P, S, iters = 32, 24, 3, 3
    rng = np.random.default_rng(11)
    sino = rng.uniform(0.1, 3.0, (P, S, N)).astype(np.float32)
    deg = (np.arange(P) * 180.0 / P).astype(np.float32)
    prm = Params(beta=0.2, beta1=0.5, delta=0.1, delta1=0.5, regw=0.5)
    th_file = oracle.deg2rad(deg) if alg == "pml" else deg
    x1 = ref.run(alg, th_file, sino, N / 2, iters, 1, prm)
    x4 = ref.run(alg, th_file, sino, N / 2, iters, 4, prm)
    assert rel_l2(x4, x1) < 1e-4


@pytest.mark.parametrize("N,C", [(16, 24), (24, 16), (63, 40), (40, 63), (128, 97)])
def test_traversal_bitexact_image_side_differs(oracle, ref, N, C):
    """Image side N != detector columns C: the oracle against the
    reference's own TraceMetadata and traversal functions on such a grid."""
    rng = np.random.default_rng(N * 1000 + C)
    angles = np.concatenate([adversarial_angles(), rng.uniform(-8, 8, 20).astype(np.float32)])
    for th in angles:
        for ctr in (C / 2.0, -1.0, C / 2.0 - 0.37):
            for col in sorted({0, 1, C // 2, C - 2, C - 1, int(rng.integers(0, C))}):
                a = oracle.trace_ray(N, C, ctr, th, col)
                b = ref.trace_ray_n(N, C, ctr, th, col)
                assert a[0].size == b[0].size, (N, C, th, ctr, col)
                assert np.array_equal(a[0], b[0])
                assert np.array_equal(_bits(a[1]), _bits(b[1]))
                assert np.array_equal(_bits(a[2]), _bits(b[2]))
