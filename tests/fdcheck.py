"""Finite-difference checker for gradients through the ray transform.

A central difference of L(x) = 1/2 ||A x - y||^2 at pixel j with step h
changes only the rays through j.  For those rays the oracle's traversal
(the reference's rays and fp32 segment lengths) is re-evaluated and the
loss difference is summed in float64:

    FD_j = sum_i [1/2 (s_i + h l_ij - y_i)^2 - 1/2 (s_i - h l_ij - y_i)^2] / (2h),
    s_i = sum_k x[idx_ik] len_ik   (float64).

The other rays contribute identical terms to L(x + h e_j) and L(x - h e_j)
and cancel exactly, so this is the central difference of the full loss.
Evaluating it in float64 is what makes h = 1e-3 usable: at 2048² the fp32
ray sums (~1e3) carry rounding of the same size as h * len.
"""
from __future__ import annotations

import numpy as np


def _col0(orc, C, center):
    """Detector coordinate of column 0 (trace_data.h:97-101, sirt.cc:128)."""
    return (1.0 - C) / 2.0 + float(orc.mov(C, center))


def candidate_rays(orc, N, C, center, theta, j, margin=1.25):
    """(proj, col) of every ray whose line passes within `margin` detector
    units of pixel j's centre (a pixel's half-diagonal is 0.71, the
    reference's crossings carry < 0.15 px of noise)."""
    iy, ix = divmod(int(j), N)
    cx = -N / 2.0 + ix + 0.5
    cy = -N / 2.0 + iy + 0.5
    th = np.asarray(theta, np.float64)
    t = -cx * np.sin(th) + cy * np.cos(th) - _col0(orc, C, center)
    lo = np.floor(t - margin).astype(np.int64)
    hi = np.ceil(t + margin).astype(np.int64)
    width = int((hi - lo).max()) + 1
    proj = np.repeat(np.arange(th.size), width)
    col = (lo[:, None] + np.arange(width)[None, :]).ravel()
    keep = (col >= 0) & (col < C) & (col <= np.repeat(hi, width))
    return proj[keep], col[keep]


def rays_through_pixel(orc, N, C, center, theta, j):
    """Rays whose traversal has a segment in pixel j: (proj, col, offsets,
    idx, len) with the traced segments of exactly those rays."""
    theta = np.asarray(theta, np.float32)
    proj, col = candidate_rays(orc, N, C, center, theta, j)
    off, idx, ln = orc.trace_rays(N, C, center, theta[proj], col)
    hit = np.zeros(proj.size, bool)
    seg_ray = np.repeat(np.arange(proj.size), np.diff(off))
    hit[np.unique(seg_ray[idx == j])] = True
    keep = np.flatnonzero(hit)
    off2, idx2, ln2 = orc.trace_rays(N, C, center, theta[proj[keep]], col[keep])
    return proj[keep], col[keep], off2, idx2, ln2


def _ray_sums(off, vals):
    """Per-ray sums of `vals` over segments [off[i], off[i+1])."""
    cs = np.concatenate([[0.0], np.cumsum(vals, dtype=np.float64)])
    return cs[off[1:]] - cs[off[:-1]]


def fd_data_term(orc, N, C, center, theta, x, y, j, h=1e-3):
    """Central difference at pixel j of 1/2 ||A x - y||^2 for one slice
    (x [N][N], y [P][C]), in float64."""
    x64 = np.asarray(x, np.float64).ravel()
    y64 = np.asarray(y, np.float64)
    proj, col, off, idx, ln = rays_through_pixel(orc, N, C, center, theta, j)
    ln64 = ln.astype(np.float64)
    inside = (idx >= 0) & (idx < N * N)
    xv = np.where(inside, x64[np.clip(idx, 0, N * N - 1)], 0.0)
    s = _ray_sums(off, xv * ln64)
    lj = _ray_sums(off, np.where(idx == j, ln64, 0.0))
    r = s - y64[proj, col]
    rp = r + h * lj
    rm = r - h * lj
    return float(np.sum(0.5 * (rp * rp) - 0.5 * (rm * rm)) / (2.0 * h))


def fd_tv_term(x, j, h=1e-3, eps=1e-8):
    """Central difference at interior pixel j of the isotropic TV of x
    [N][N] (forward differences), in float64, from the 3x3 neighbourhood
    (the only terms that depend on pixel j)."""
    N = x.shape[-1]
    iy, ix = divmod(int(j), N)
    assert 1 <= iy < N - 1 and 1 <= ix < N - 1, "interior pixels only"
    patch = np.asarray(x, np.float64)[iy - 1:iy + 2, ix - 1:ix + 2].copy()

    def tv(p):
        dx = np.zeros_like(p)
        dy = np.zeros_like(p)
        dx[:, :-1] = p[:, 1:] - p[:, :-1]
        dy[:-1, :] = p[1:, :] - p[:-1, :]
        return np.sqrt(dx * dx + dy * dy + eps).sum()

    pp = patch.copy()
    pp[1, 1] += h
    pm = patch.copy()
    pm[1, 1] -= h
    return float((tv(pp) - tv(pm)) / (2.0 * h))
