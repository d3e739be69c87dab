"""Seeded synthetic inputs: ellipse phantoms and their noisy sinograms.

Geometry is the reference's (trace_data.h:97-108, sirt.cc:128-129): pixel
(ix, iy) of an N x N slice covers x in [-N/2+ix, -N/2+ix+1] (same for y) and
is stored at image[iy][ix]; the ray of detector column c at angle theta is
the line {q : q . (-sin theta, cos theta) = yi_c} with
yi_c = (1-C)/2 + c + mov, direction (cos theta, sin theta).

Phantom: a sum of K ellipses in the unit square [-1, 1]^2 (scaled by N/2 to
pixel units) with centres ~ U[-0.8, 0.8]^2, semi-axes ~ U[0.05, 0.4],
rotations ~ U[0, pi) and additive intensities ~ U[0.1, 1.0]; slice s of
seed `seed` draws from numpy PCG64(seed * 1000 + s).

Sinogram: exact line integrals of the ellipses (pixel units) along the rays,
Poisson noise counts ~ Poisson(I0 exp(-kappa p)) with kappa = 2/N, then
p~ = -ln(max(counts, 1) / I0) / kappa clamped at 0, stored fp32 [P][S][C].
The noise draws use a torch generator seeded with the same seed.
"""
from __future__ import annotations

import numpy as np


def ellipse_params(seed: int, slice_index: int, K: int) -> np.ndarray:
    """[K, 6] rows (cx, cy, a, b, phi, rho) in unit-square coordinates."""
    rng = np.random.Generator(np.random.PCG64(seed * 1000 + slice_index))
    cx = rng.uniform(-0.8, 0.8, K)
    cy = rng.uniform(-0.8, 0.8, K)
    a = rng.uniform(0.05, 0.4, K)
    b = rng.uniform(0.05, 0.4, K)
    phi = rng.uniform(0.0, np.pi, K)
    rho = rng.uniform(0.1, 1.0, K)
    return np.stack([cx, cy, a, b, phi, rho], axis=1)


def mov_of(C: int, center: float) -> float:
    """trace_data.h:97-101 (same arithmetic as the C-ABI)."""
    c = np.float32(center)
    if c <= 0:
        c = np.float32(np.float64(np.float32(C)) / 2.0 + 1.0)
    mov = np.float32(np.float64(np.float32(C)) / 2.0 - np.float64(c))
    if np.float64(mov - np.ceil(mov)) < 1e-6:
        mov = np.float32(np.float64(mov) + 1e-6)
    return float(mov)


def phantom_slices(N: int, S: int, seed: int, K: int = 20, device="cpu", first_slice: int = 0):
    """Torch tensor [S][N][N] float32 of ellipse phantoms at pixel centres
    (slices first_slice .. first_slice + S - 1 of the seed)."""
    import torch
    dev = torch.device(device)
    half = N / 2.0
    coords = (torch.arange(N, dtype=torch.float64, device=dev) - half + 0.5) / half
    Y, X = torch.meshgrid(coords, coords, indexing="ij")  # [iy][ix]
    out = torch.zeros((S, N, N), dtype=torch.float64, device=dev)
    for s in range(S):
        for cx, cy, a, b, phi, rho in ellipse_params(seed, first_slice + s, K):
            c, sn = np.cos(phi), np.sin(phi)
            dx, dy = X - cx, Y - cy
            u = dx * c + dy * sn
            v = -dx * sn + dy * c
            out[s] += rho * (((u / a) ** 2 + (v / b) ** 2) <= 1.0)
    return out.to(torch.float32)


def line_integrals(theta, C: int, S: int, seed: int, K: int = 20,
                   center: float | None = None, device="cpu", first_slice: int = 0):
    """Exact ellipse line integrals [P][S][C] (float64 tensor, pixel units)
    along the reference's rays (slices first_slice .. first_slice + S - 1)."""
    import torch
    dev = torch.device(device)
    N = C
    half = N / 2.0
    ctr = half if center is None else center
    mov = mov_of(C, ctr)
    th = torch.as_tensor(np.asarray(theta, np.float64), device=dev)
    yi = (torch.arange(C, dtype=torch.float64, device=dev) + (1 - C) / 2.0 + mov)
    nx = -torch.sin(th)[:, None]  # ray normal n = (-sin, cos)
    ny = torch.cos(th)[:, None]
    out = torch.zeros((th.numel(), S, C), dtype=torch.float64, device=dev)
    for s in range(S):
        acc = torch.zeros((th.numel(), C), dtype=torch.float64, device=dev)
        for cx, cy, a, b, phi, rho in ellipse_params(seed, first_slice + s, K):
            cxp, cyp, ap, bp = cx * half, cy * half, a * half, b * half
            c, sn = np.cos(phi), np.sin(phi)
            # normal in the ellipse frame and signed distance of the line
            n1 = nx * c + ny * sn
            n2 = -nx * sn + ny * c
            tau = yi[None, :] - (cxp * nx + cyp * ny)
            s2 = (ap * n1) ** 2 + (bp * n2) ** 2
            chord = 2.0 * ap * bp * torch.sqrt(torch.clamp(s2 - tau ** 2, min=0.0)) / s2
            acc += rho * chord
        out[:, s, :] = acc
    return out


def noisy_sinogram(theta, C: int, S: int, seed: int, I0: float, K: int = 20,
                   center: float | None = None, device="cpu", first_slice: int = 0):
    """Poisson-noisy sinogram [P][S][C] float32 (non-negative)."""
    import torch
    dev = torch.device(device)
    p = line_integrals(theta, C, S, seed, K, center, device, first_slice)
    kappa = 2.0 / C
    lam = I0 * torch.exp(-kappa * p)
    gen = torch.Generator(device=dev)
    # noise stream per (seed, first slice); the seed alone for first_slice 0
    gen.manual_seed(int(seed) if first_slice == 0 else int(seed) * 1000 + int(first_slice))
    counts = torch.poisson(lam, generator=gen)
    pt = -torch.log(torch.clamp(counts, min=1.0) / I0) / kappa
    return torch.clamp(pt, min=0.0).to(torch.float32)


def angles_deg(P: int) -> np.ndarray:
    """P angles evenly spaced in [0, 180) degrees (float32)."""
    return (np.arange(P, dtype=np.float64) * 180.0 / P).astype(np.float32)


def angles_rad(P: int) -> np.ndarray:
    """The same angles through the reference's degree conversion
    (trace_utils.cc:10-24)."""
    kpi = np.float32(3.14159265358979)
    d = angles_deg(P)
    return ((d * kpi).astype(np.float64) / 180.0).astype(np.float32)
