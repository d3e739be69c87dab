"""Pins of the oracle that do not depend on the reference: closed forms,
invariants, a brute-force fp64 geometry, and the fp64 twin."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import Params

KPI = np.float32(3.14159265358979)


def test_theta_zero_rows(oracle):
    """theta = 0: horizontal rays through row centres; every segment has
    length 1 and index iy*N + ix, ix ascending."""
    for N in (4, 9, 64):
        for col in range(N):
            idx, ln, _ = oracle.trace_ray(N, N, N / 2, np.float32(0.0), col)
            assert idx.size == N
            assert np.all(ln == 1.0)
            assert np.array_equal(idx, col * N + np.arange(N))


def test_vertical_columns(oracle):
    """theta = float(pi/2): vertical rays; column index fixed, rows ascend."""
    N = 32
    th = np.float32(KPI / 2)
    for col in (0, 5, 31):
        idx, ln, _ = oracle.trace_ray(N, N, N / 2, th, col)
        ix = idx % N
        assert np.all(ix == ix[0])
        assert np.array_equal(idx // N, np.arange(N)[: idx.size]) or np.array_equal(
            idx // N, np.arange(N)[::-1][: idx.size])
        assert abs(ln.sum() - N) < 0.25


def _exact_lengths(N, theta, yi):
    """fp64 Siddon: per-pixel intersection lengths of the line
    {q : q.(-sin, cos) = yi} with the N x N grid; also entry/exit points."""
    th = float(theta)
    s, c = np.sin(th), np.cos(th)
    d = np.array([c, s])
    n = np.array([-s, c])
    q0 = yi * n
    h = N / 2
    us = []
    for ax in range(2):
        if abs(d[ax]) > 1e-12:
            us += [(g - q0[ax]) / d[ax] for g in np.arange(-h, h + 1)]
    us = np.sort(np.array(us))
    out, pts = {}, []
    for a, b in zip(us[:-1], us[1:]):
        m = q0 + 0.5 * (a + b) * d
        if -h < m[0] < h and -h < m[1] < h and b > a:
            ix, iy = int(np.floor(m[0] + h)), int(np.floor(m[1] + h))
            out[ix + iy * N] = out.get(ix + iy * N, 0.0) + (b - a)
            pts += [q0 + a * d, q0 + b * d]
    return out, pts


@pytest.mark.parametrize("N", [16, 64])
def test_bruteforce_geometry(oracle, N):
    """Per-pixel lengths agree with exact fp64 geometry to the crossing
    noise of the reference's source point at -1e6 (float ulp 0.0625 there):
    0.3 px per pixel, 3% in total, except rays entering/leaving within 0.3 px
    of a grid corner (where the reference's strict trims drop a segment)."""
    rng = np.random.default_rng(N)
    mov = oracle.mov(N, N / 2)
    checked = 0
    for _ in range(200):
        th = np.float32(rng.uniform(0, np.pi))
        col = int(rng.integers(0, N))
        yi = float(np.float32((1 - N) / 2.0 + col + mov))
        ex, pts = _exact_lengths(N, th, yi)
        corners = np.array([[-N / 2, -N / 2], [-N / 2, N / 2], [N / 2, -N / 2], [N / 2, N / 2]])
        if not pts or min(np.min(np.linalg.norm(corners - p, axis=1)) for p in (pts[0], pts[-1])) < 0.3:
            continue
        idx, ln, _ = oracle.trace_ray(N, N, N / 2, th, col)
        got = {}
        for i, l in zip(idx, ln):
            got[int(i)] = got.get(int(i), 0.0) + float(l)
        keys = set(got) | set(ex)
        assert max(abs(got.get(k, 0) - ex.get(k, 0)) for k in keys) < 0.3
        assert abs(sum(got.values()) - sum(ex.values())) < 0.03 * sum(ex.values())
        checked += 1
    assert checked > 150


def test_forward_of_ones_is_chord(oracle):
    N, P = 48, 16
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    y = oracle.forward(th, np.ones((1, N, N), np.float32), N / 2)
    for p in range(P):
        for c in range(N):
            _, ln, _ = oracle.trace_ray(N, N, N / 2, th[p], c)
            assert abs(y[p, 0, c] - ln.astype(np.float64).sum()) <= 1e-5 * max(1, ln.sum())


def test_adjoint_dot_product(oracle):
    """<A x, r> = <x, A^T r> (double adjoint vs fp32 forward)."""
    rng = np.random.default_rng(3)
    N, P, S = 40, 25, 2
    th = rng.uniform(0, np.pi, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    r = rng.normal(size=(P, S, N)).astype(np.float32)
    lhs = np.sum(oracle.forward(th, x, N / 2).astype(np.float64) * r)
    rhs = np.sum(x.astype(np.float64) * oracle.adjoint(th, r, N / 2))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)


def _problem(N=24, P=18, S=2, seed=5):
    rng = np.random.default_rng(seed)
    x_true = rng.uniform(0.2, 1.0, (S, N, N)).astype(np.float32)
    return rng, x_true


def test_fixed_points(oracle):
    """p = A x*  =>  x* is a fixed point: exactly for MLEM (num == den bit
    for bit) and SIRT (num == 0)."""
    N, P, S = 24, 18, 2
    rng, xs = _problem(N, P, S)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    p = oracle.forward(th, xs, N / 2)
    xm = oracle.reconstruct("mlem", th, p, xs, N / 2, 2)
    assert np.array_equal(xm, xs)
    xsr = oracle.reconstruct("sirt", th, p, xs, N / 2, 2)
    assert np.array_equal(xsr, xs)


def test_mlem_nonnegative_and_converging(oracle):
    N, P, S = 24, 30, 1
    rng, xs = _problem(N, P, S)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    p = oracle.forward(th, xs, N / 2)
    x0 = np.ones_like(xs)
    e_prev = None
    for it in (2, 6, 20):
        x = oracle.reconstruct("mlem", th, p, x0, N / 2, it)
        assert np.all(x >= 0)
        e = np.linalg.norm(oracle.forward(th, x, N / 2) - p)
        if e_prev is not None:
            assert e < e_prev
        e_prev = e


def test_pml_constant_image_stencil(oracle):
    """For a constant image x = c: F = 2 beta sum(w) and G = den - 4 beta c
    sum(w) per pixel class (pml.h:22-29 weights)."""
    N, P = 12, 8
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    c, beta = 0.75, 0.3
    sino = np.ones((P, 1, N), np.float32)
    x0 = np.full((1, N, N), c, np.float32)
    _, st = oracle.reconstruct("pml", th, sino, x0, N / 2, 1, Params(beta=beta), want_stages=True)
    wsum = {"in": 4 * 0.1464466094 + 4 * 0.10355339059,
            "edge": 3 * 0.226540919667 + 2 * 0.1601886205,
            "corner": 2 * 0.36939806251 + 0.26120387496}
    F, G, den = st["F"][0], st["G"][0], st["den"][0]
    for (iy, ix), kind in [((5, 5), "in"), ((0, 4), "edge"), ((4, N - 1), "edge"), ((0, 0), "corner"),
                           ((N - 1, N - 1), "corner")]:
        assert abs(F[iy, ix] - 2 * beta * wsum[kind]) < 1e-6
        assert abs(G[iy, ix] - (den[iy, ix] - 4 * beta * c * wsum[kind])) < 1e-5 * max(1, den[iy, ix])


def test_pml_small_beta_is_mlem(oracle):
    """PML's positive root tends to the MLEM update as beta -> 0 (first-order
    expansion of (-G + sqrt(G^2 - 8 num F)) / 4F with F = 2 beta).  In fp32
    the root cancels catastrophically for small F (as it does in the
    reference), so the limit is checked on the fp64 twin."""
    N, P, S = 20, 16, 1
    rng = np.random.default_rng(9)
    sino = rng.uniform(0.5, 3.0, (P, S, N)).astype(np.float32)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    x0 = np.ones((S, N, N), np.float32)
    xm = oracle.reconstruct_f64("mlem", th, sino, x0, N / 2, 3)
    xp = oracle.reconstruct_f64("pml", th, sino, x0, N / 2, 3, Params(beta=1e-7))
    assert rel_l2(xp, xm) < 1e-5
    xq = oracle.reconstruct_f64("pml", th, sino, x0, N / 2, 3, Params(beta=1e-3))
    assert rel_l2(xq, xm) > 10 * rel_l2(xp, xm)


def test_apmlr_reduces_to_pml(oracle):
    """delta, delta1 -> inf (gamma = 1), regw = 0: the APMLR in-plane weights
    equal PML's, and with apmlr_compat=0 the whole iteration matches PML."""
    N, P, S = 16, 12, 3
    rng = np.random.default_rng(4)
    sino = rng.uniform(0.5, 3.0, (P, S, N)).astype(np.float32)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    x0 = np.ones((S, N, N), np.float32)
    xp, sp = oracle.reconstruct("pml", th, sino, x0, N / 2, 2, Params(beta=0.4), want_stages=True)
    xa, sa = oracle.reconstruct("apmlr", th, sino, x0, N / 2, 2,
                                Params(beta=0.4, beta1=0.0, delta=1e30, delta1=1e30, regw=0.0,
                                       apmlr_compat=0), want_stages=True)
    assert rel_l2(sa["F"], sp["F"]) < 1e-6
    assert rel_l2(sa["G"], sp["G"]) < 1e-6
    assert rel_l2(xa, xp) < 1e-5


def test_apmlr_compat_shifts_slices(oracle):
    """Compat mode projects slice s with the image of slice s-1: with S
    identical sinogram rows, slice 0 sees the constant-1 ghost."""
    N, P, S = 12, 8, 3
    rng = np.random.default_rng(2)
    row = rng.uniform(0.5, 2.0, (P, 1, N)).astype(np.float32)
    sino = np.repeat(row, S, axis=1)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    x0 = np.ones((S, N, N), np.float32)
    prm = Params(beta=0.3, beta1=0.3, delta=0.5, delta1=0.5, regw=0.5, apmlr_compat=1)
    _, st = oracle.reconstruct("apmlr", th, sino, x0, N / 2, 1, prm, want_stages=True)
    # first iteration: every slice (and the ghost) is 1, so num is equal
    assert np.array_equal(st["num"][0], st["num"][1])


@pytest.mark.parametrize("alg", ["sirt", "mlem", "pml", "apmlr"])
def test_f64_twin_agrees(oracle, alg):
    """The fp64 twin and the fp32 restatement differ only by fp32
    rounding."""
    N, P, S, iters = 20, 16, 2, 4
    rng = np.random.default_rng(1)
    sino = rng.uniform(0.5, 3.0, (P, S, N)).astype(np.float32)
    th = oracle.deg2rad((np.arange(P) * 180.0 / P).astype(np.float32))
    x0 = np.full((S, N, N), 0.0 if alg == "sirt" else 1.0, np.float32)
    prm = Params(beta=0.3, beta1=0.2, delta=0.5, delta1=0.5, regw=0.5, apmlr_compat=1)
    x32 = oracle.reconstruct(alg, th, sino, x0, N / 2, iters, prm)
    x64 = oracle.reconstruct_f64(alg, th, sino, x0, N / 2, iters, prm)
    assert rel_l2(x32, x64) < 2e-5


@pytest.mark.parametrize("N,C", [(48, 32), (32, 48)])
def test_bruteforce_geometry_image_side_differs(oracle, N, C):
    """N != C: per-pixel lengths against exact fp64 geometry (the C columns'
    lines q.(-sin, cos) = yi across an N x N grid), as for N = C."""
    rng = np.random.default_rng(N + 7 * C)
    mov = oracle.mov(C, C / 2)
    checked = 0
    for _ in range(200):
        th = np.float32(rng.uniform(0, np.pi))
        col = int(rng.integers(0, C))
        yi = float(np.float32((1 - C) / 2.0 + col + mov))
        ex, pts = _exact_lengths(N, th, yi)
        corners = np.array([[-N / 2, -N / 2], [-N / 2, N / 2], [N / 2, -N / 2], [N / 2, N / 2]])
        if not pts or min(np.min(np.linalg.norm(corners - p, axis=1)) for p in (pts[0], pts[-1])) < 0.3:
            continue
        idx, ln, _ = oracle.trace_ray(N, C, C / 2, th, col)
        got = {}
        for i, l in zip(idx, ln):
            got[int(i)] = got.get(int(i), 0.0) + float(l)
        keys = set(got) | set(ex)
        assert max(abs(got.get(k, 0) - ex.get(k, 0)) for k in keys) < 0.3
        assert abs(sum(got.values()) - sum(ex.values())) < 0.03 * sum(ex.values())
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("N,C", [(40, 28), (28, 40)])
def test_adjoint_dot_product_image_side_differs(oracle, N, C):
    rng = np.random.default_rng(N + C)
    P, S = 25, 2
    th = rng.uniform(0, np.pi, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    r = rng.normal(size=(P, S, C)).astype(np.float32)
    lhs = np.sum(oracle.forward(th, x, C / 2, C=C).astype(np.float64) * r)
    rhs = np.sum(x.astype(np.float64) * oracle.adjoint(th, r, C / 2, N=N))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
    # A 1 is each ray's total length
    y = oracle.forward(th, np.ones((1, N, N), np.float32), C / 2, C=C)
    for p in range(0, P, 5):
        for c in range(C):
            _, ln, _ = oracle.trace_ray(N, C, C / 2, th[p], c)
            assert abs(y[p, 0, c] - ln.astype(np.float64).sum()) <= 1e-5 * max(1, ln.sum())
