"""Differentiable ray transform for PyTorch.

The forward projector (`trace_forward`, y = A x) and the hand-written
adjoint kernel (`trace_adjoint`, z = A^T r) are exposed as a pair of
`torch.autograd.Function`s, each the other's backward:

    d<A x, g>/dx = A^T g          d<A^T r, g>/dr = A g

A is the reference's system matrix: rays and segment lengths of
`trace_utils.cc:48-187` with the geometry of `trace_data.h:53-109`, the
product of `CalculateSimdata` (`sirt.cc:4-19`) and the scatter of
`UpdateReconReplica` (`sirt.cc:35-64`).  The kernels run on the handle's
HIP stream (the current torch stream when the operator is made); tensors
are float32 on that GPU.

On top of the operator:
  * `tv_isotropic`        - smoothed isotropic total variation;
  * `tv_adam`             - 1/2 ||A x - y||^2 + lambda TV(x) by Adam;
  * `tv_grad`             - its gradient in closed form (differentiable);
  * `sirt_unrolled`       - the reference's SIRT iteration (sirt.cc:21-64:
                            w = (p - A x)/a2, x += A^T w / A^T 1) unrolled
                            as a differentiable function, optionally
                            followed by a per-iteration proximal/learned
                            step, with one checkpoint per iteration: only
                            x_k is kept, and each iteration's projections and
                            step activations are recomputed on the backward
                            pass.  (The projector's own backward needs no
                            saved tensors - A is linear - so the memory the
                            checkpoints save is that of the non-linear
                            per-iteration steps.)
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from .recon import Projector


class RayTransform:
    """A for one geometry: `theta_rad` [P] angles, N x N images, C = N
    detector columns, `slices` images per call ([N][N] when slices == 1,
    else [S][N][N]; sinograms [P][C] or [P][S][C])."""

    def __init__(self, theta_rad, N: int, slices: int = 1, center: float | None = None,
                 device: int = 0, exact_sqrt: bool = True):
        self.N = int(N)
        self.S = int(slices)
        self.center = float(N / 2.0 if center is None else center)
        self.device = int(device)
        self.pj = Projector(np.asarray(theta_rad, dtype=np.float32), self.S, self.N,
                            self.center, device=self.device, exact_sqrt=exact_sqrt)
        self.P = self.pj.P
        self.C = self.pj.C
        self._a2 = None
        self._den = None

    def close(self):
        self.pj.close()

    # -- shapes -------------------------------------------------------------
    def _image3(self, x):
        if x.dim() == 2 and self.S == 1:
            return x.reshape(1, self.N, self.N)
        if x.dim() == 3 and tuple(x.shape) == (self.S, self.N, self.N):
            return x
        raise ValueError(f"image of shape {tuple(x.shape)} does not match "
                         f"{self.S} slice(s) of {self.N}x{self.N}")

    def _sino3(self, r):
        if r.dim() == 2 and self.S == 1:
            return r.reshape(self.P, 1, self.C)
        if r.dim() == 3 and tuple(r.shape) == (self.P, self.S, self.C):
            return r
        raise ValueError(f"sinogram of shape {tuple(r.shape)} does not match "
                         f"[{self.P}][{self.S}][{self.C}]")

    def _check(self, t):
        if t.dtype != torch.float32:
            raise TypeError("the ray transform computes in float32")
        if not t.is_cuda or t.device.index != self.device:
            raise TypeError(f"expected a tensor on cuda:{self.device}")

    # -- raw (non-differentiable) products ----------------------------------
    def apply_A(self, x):
        self._check(x)
        y = self.pj.forward(self._image3(x).contiguous())
        return y.reshape(self.P, self.C) if (x.dim() == 2) else y

    def apply_AT(self, r):
        self._check(r)
        z = self.pj.adjoint(self._sino3(r).contiguous())
        return z.reshape(self.N, self.N) if (r.dim() == 2) else z

    # -- differentiable -------------------------------------------------------
    def __call__(self, x):
        """y = A x (differentiable in x)."""
        return _Project.apply(x, self)

    def T(self, r):
        """z = A^T r (differentiable in r)."""
        return _Backproject.apply(r, self)

    # -- geometry constants ---------------------------------------------------
    def row_norms2(self):
        """a2 = sum of squared segment lengths per ray, [P][C] (the SIRT
        normaliser, sirt.cc:47-50)."""
        if self._a2 is None:
            self._a2 = self.pj.ray_sums()[1]
        return self._a2

    def column_sums(self):
        """A^T 1 of one slice, [N][N] (sirt.cc:62)."""
        if self._den is None:
            self._den = self.pj.column_sums()
        return self._den


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op: RayTransform):
        ctx.op = op
        return op.apply_A(x.detach())

    @staticmethod
    def backward(ctx, gy):
        op = ctx.op
        return op.apply_AT(gy.contiguous()), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, op: RayTransform):
        ctx.op = op
        return op.apply_AT(r.detach())

    @staticmethod
    def backward(ctx, gz):
        op = ctx.op
        return op.apply_A(gz.contiguous()), None


def project(x, op: RayTransform):
    return _Project.apply(x, op)


def backproject(r, op: RayTransform):
    return _Backproject.apply(r, op)


# ---------------------------------------------------------------------------
# Regularised least squares
# ---------------------------------------------------------------------------

def tv_isotropic(x, eps: float = 1e-8):
    """sum_ij sqrt(dx^2 + dy^2 + eps) with forward differences along the
    last two axes (zero difference past the last row/column).  Works in any
    float dtype on any device."""
    dx = torch.zeros_like(x)
    dy = torch.zeros_like(x)
    dx[..., :, :-1] = x[..., :, 1:] - x[..., :, :-1]
    dy[..., :-1, :] = x[..., 1:, :] - x[..., :-1, :]
    return torch.sqrt(dx * dx + dy * dy + eps).sum()


def tv_grad(x, eps: float = 1e-8):
    """d tv_isotropic / dx in closed form (differentiable): with
    n = sqrt(dx^2 + dy^2 + eps), px = dx/n, py = dy/n,
    g = -(px + py) + px shifted one column right + py shifted one row down."""
    dx = torch.zeros_like(x)
    dy = torch.zeros_like(x)
    dx[..., :, :-1] = x[..., :, 1:] - x[..., :, :-1]
    dy[..., :-1, :] = x[..., 1:, :] - x[..., :-1, :]
    n = torch.sqrt(dx * dx + dy * dy + eps)
    px = dx / n
    py = dy / n
    g = -(px + py)
    g = g + torch.nn.functional.pad(px[..., :, :-1], (1, 0, 0, 0))
    g = g + torch.nn.functional.pad(py[..., :-1, :], (0, 0, 1, 0))
    return g


def ls_tv_loss(op: RayTransform, x, y, tv_weight: float = 1e-3, eps: float = 1e-8):
    """1/2 ||A x - y||^2 + tv_weight * TV(x)."""
    r = op(x) - y
    return 0.5 * (r * r).sum() + tv_weight * tv_isotropic(x, eps)


def tv_adam(op: RayTransform, y, x0=None, steps: int = 100, lr: float = 1e-2,
            tv_weight: float = 1e-3, eps: float = 1e-8, callback=None):
    """Minimises 1/2 ||A x - y||^2 + tv_weight TV(x) with Adam from x0
    (zeros by default); returns (x, losses)."""
    shape = (op.N, op.N) if op.S == 1 else (op.S, op.N, op.N)
    x = (torch.zeros(shape, dtype=torch.float32, device=y.device) if x0 is None
         else x0.detach().clone()).requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    losses = []
    for k in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = ls_tv_loss(op, x, y, tv_weight, eps)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
        if callback is not None:
            callback(k, x, loss)
    return x.detach(), torch.stack(losses).cpu().numpy()


# ---------------------------------------------------------------------------
# Unrolled SIRT
# ---------------------------------------------------------------------------

def _sirt_step(op, x, sino, a2, inv_den, lam, prox, k):
    """One reference SIRT iteration (sirt.cc:21-64), differentiable:
    w = (p - A x)/a2 per ray, x += lam * (A^T w) / (A^T 1); then the
    optional per-iteration step prox(x, k)."""
    w = (sino - op(x)) / a2
    x = x + lam * op.T(w) * inv_den
    return x if prox is None else prox(x, k)


def sirt_unrolled(op: RayTransform, sino, iters: int, x0=None, relax=None, prox=None,
                  use_checkpoint: bool = True):
    """`iters` SIRT iterations from x0 (zeros, the reference's SIRT start)
    as one differentiable function of (sino, x0, relax, prox parameters).
    `relax` [iters] are per-iteration step factors (1 = the reference's
    update); `prox(x, k)`, if given, runs after update k (a proximal or
    learned step, e.g. `lambda x, k: x - tau[k] * tv_grad(x)`).  With
    `use_checkpoint` each iteration is one checkpoint: only the iterates
    are stored and the iteration is recomputed during backward."""
    shape = (op.N, op.N) if op.S == 1 else (op.S, op.N, op.N)
    x = (torch.zeros(shape, dtype=torch.float32, device=sino.device) if x0 is None else x0)
    a2 = op.row_norms2()
    a2 = torch.where(a2 > 0, a2, torch.ones_like(a2))  # rays missing the image
    a2 = a2 if op.S == 1 and sino.dim() == 2 else a2.unsqueeze(1)
    den = op.column_sums()
    inv_den = torch.where(den > 0, 1.0 / den, torch.zeros_like(den))
    if relax is None:
        relax = torch.ones(iters, dtype=torch.float32, device=sino.device)
    for k in range(iters):
        if use_checkpoint and torch.is_grad_enabled():
            x = checkpoint(_sirt_step, op, x, sino, a2, inv_den, relax[k], prox, k,
                           use_reentrant=False)
        else:
            x = _sirt_step(op, x, sino, a2, inv_den, relax[k], prox, k)
    return x
