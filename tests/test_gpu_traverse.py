"""Traversal on the GPU against the oracle and the golden vectors: the
per-ray (index, length, length^2) lists of the kernels' traversal are
bit-identical to the reference's (exact-sqrt mode), including angles
outside [0, pi), non-default centres and odd grid sizes."""
import os

import numpy as np
import pytest

from conftest import adversarial_angles

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


def _check_group(Projector, N, center, thetas, cols, expect, exact=True):
    """thetas/cols: ray list; expect(i) -> (idx, len, len2) reference lists."""
    uniq, inv = np.unique(np.asarray(thetas, np.float32), return_inverse=True)
    with Projector(uniq, 1, N, center) as pj:
        if not exact:
            pj.set_exact_sqrt(False)
        counts, idx, ln, l2 = pj.dump_rays(inv.astype(np.int32), np.asarray(cols, np.int32))
    for i in range(len(cols)):
        e_idx, e_len, e_l2 = expect(i)
        n = int(counts[i])
        assert n == e_idx.size, (N, center, thetas[i], cols[i])
        assert np.array_equal(idx[i, :n], e_idx), (N, center, thetas[i], cols[i])
        if exact:
            assert np.array_equal(_bits(ln[i, :n]), _bits(e_len)), (N, thetas[i], cols[i])
        else:
            assert np.all(np.abs(ln[i, :n] - e_len) <= 2.5e-7 * np.maximum(e_len, 1e-30) + 1e-38)
        assert np.array_equal(_bits(l2[i, :n]), _bits(e_l2)), (N, thetas[i], cols[i])


def test_golden_rays_gpu(gpu):
    from trace_amd import Projector
    g = np.load(os.path.join(GOLD, "rays.npz"))
    off = g["offsets"]
    keys = {}
    for i in range(g["N"].size):
        keys.setdefault((int(g["N"][i]), float(g["center"][i])), []).append(i)
    for (N, ctr), rows in keys.items():
        print(f"golden rays N={N} center={ctr} n={len(rows)}", flush=True)

        def expect(k, rows=rows):
            i = rows[k]
            a, b = off[i], off[i + 1]
            return g["idx"][a:b], g["len"][a:b], g["len2"][a:b]

        _check_group(Projector, N, ctr, g["theta"][rows], g["col"][rows], expect)


@pytest.mark.parametrize("N", [31, 256, 1024, 2048])
def test_random_rays_vs_oracle(gpu, oracle, N):
    from trace_amd import Projector
    rng = np.random.default_rng(N)
    th = np.concatenate([adversarial_angles(), rng.uniform(-7, 7, 60).astype(np.float32),
                         rng.uniform(0, np.pi, 60).astype(np.float32)])
    cols = rng.integers(0, N, th.size).astype(np.int32)
    cols[:4] = [0, N - 1, N // 2, 1]
    for ctr in (N / 2.0, 0.0):
        print(f"random rays N={N} center={ctr}", flush=True)
        exp = [oracle.trace_ray(N, N, ctr, t, int(c)) for t, c in zip(th, cols)]
        _check_group(Projector, N, ctr, th, cols, lambda i: exp[i])


def test_fast_sqrt_mode(gpu, oracle):
    """v_sqrt_f32 mode: same indices and len^2, lengths within 2 ulp."""
    from trace_amd import Projector
    N = 512
    rng = np.random.default_rng(5)
    th = rng.uniform(0, np.pi, 200).astype(np.float32)
    cols = rng.integers(0, N, th.size).astype(np.int32)
    exp = [oracle.trace_ray(N, N, N / 2, t, int(c)) for t, c in zip(th, cols)]
    _check_group(Projector, N, N / 2, th, cols, lambda i: exp[i], exact=False)


def test_segment_count(gpu, oracle):
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    for N, P in [(64, 45), (256, 180)]:
        th = angles_rad(P)
        with Projector(th, 1, N, N / 2) as pj:
            nnz, oor = pj.count_segments()
        e_nnz, e_oor = oracle.count_segments(P, N, N / 2, th)
        assert (nnz, oor) == (e_nnz, e_oor)
