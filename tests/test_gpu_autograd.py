"""The differentiable ray transform on the GPU.

* adjoint dot-product test <A x, y> = <x, A^T y> to 1e-5 relative at
  2048², 1500 angles (and smaller geometries);
* central finite differences (h = 1e-3, evaluated in float64 over the
  rays each pixel perturbs, see fdcheck.py) of 1/2||Ax - y||^2 + 1e-3 TV(x)
  on 64 seeded pixels at 2048², 1500 angles;
* the autograd gradient against the oracle's restatement of the
  reference backprojector applied to the oracle's (reference-exact)
  residual;
* a dense-matrix PyTorch fp32 reference of the same loss and Adam run;
* unrolled SIRT: agrees with the plain SIRT kernel path, checkpointed
  backward == plain backward, vector-Jacobian products pass the adjoint
  identity, and checkpointing cuts the stored activations.
"""
import numpy as np
import pytest

import fdcheck
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _geom(N, P):
    from trace_amd.phantom import angles_rad
    return angles_rad(P), N / 2.0


def _op(theta, N, S=1, center=None):
    from trace_amd.autograd import RayTransform
    return RayTransform(theta, N, slices=S, center=center)


@pytest.mark.parametrize("N,P,S", [(2048, 1500, 1), (256, 180, 1), (512, 360, 3), (97, 64, 2)])
def test_dot_product(gpu, N, P, S):
    torch = gpu
    theta, center = _geom(N, P)
    print(f"dot product N={N} P={P} S={S}", flush=True)
    op = _op(theta, N, S, center)
    g = torch.Generator(device="cuda").manual_seed(N + P)
    shape_x = (N, N) if S == 1 else (S, N, N)
    shape_y = (P, N) if S == 1 else (P, S, N)
    x = torch.rand(shape_x, device="cuda", generator=g)
    y = torch.rand(shape_y, device="cuda", generator=g) - 0.25
    lhs = (op.apply_A(x).double() * y.double()).sum().item()
    rhs = (x.double() * op.apply_AT(y).double()).sum().item()
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"<Ax,y> = {lhs:.10e}  <x,A^T y> = {rhs:.10e}  rel {err:.2e}")
    assert err <= 1e-5
    op.close()


def _config4_data(N=2048, P=1500, seed=4):
    import torch
    from trace_amd.phantom import noisy_sinogram, phantom_slices
    theta, center = _geom(N, P)
    y = noisy_sinogram(theta, N, 1, seed=seed, I0=1e4, K=20, device="cuda")[:, 0, :].contiguous()
    x_true = phantom_slices(N, 1, seed=seed, K=20, device="cuda")[0].float()
    return theta, center, x_true, y


def test_finite_differences_config4(gpu, oracle):
    """Gradient of 1/2||Ax - y||^2 + 1e-3 TV(x) by autograd vs central
    differences (h = 1e-3) on 64 seeded interior pixels, 2048², 1500 angles."""
    torch = gpu
    from trace_amd.autograd import ls_tv_loss
    N, P = 2048, 1500
    theta, center, x_true, y = _config4_data(N, P)
    g = torch.Generator(device="cuda").manual_seed(44)
    x = (x_true + 0.05 * torch.randn(x_true.shape, device="cuda", generator=g)).contiguous()
    op = _op(theta, N, 1, center)
    xv = x.clone().requires_grad_(True)
    print("finite differences: autograd gradient", flush=True)
    ls_tv_loss(op, xv, y, tv_weight=1e-3).backward()
    grad = xv.grad.double().cpu().numpy().ravel()
    xh = x.cpu().numpy()
    yh = y.cpu().numpy()
    rng = np.random.default_rng(4)
    iy = rng.integers(8, N - 8, 64)
    ix = rng.integers(8, N - 8, 64)
    worst = 0.0
    scale = np.median(np.abs(grad))
    for k, j in enumerate(iy * N + ix):
        fd = (fdcheck.fd_data_term(oracle, N, N, center, theta, xh, yh, j, h=1e-3)
              + 1e-3 * fdcheck.fd_tv_term(xh, j, h=1e-3))
        e = abs(grad[j] - fd) / max(abs(fd), scale)
        worst = max(worst, e)
        if k < 4:
            print(f"pixel {j}: autograd {grad[j]:.6e}  FD {fd:.6e}  rel {e:.2e}", flush=True)
    print(f"64 pixels: worst relative difference {worst:.2e} (median |grad| {scale:.3e})")
    assert worst <= 1e-3
    op.close()


@pytest.mark.parametrize("N,P,S", [(256, 180, 1), (512, 360, 2)])
def test_gradient_vs_reference_backprojector(gpu, oracle, N, P, S):
    """autograd d/dx 1/2||Ax - y||^2 = A^T (A x - y) against the oracle:
    or_forward (bit-exact to the reference's CalculateSimdata) for the
    residual and or_adjoint (the reference's scatter, sirt.cc:61, summed in
    float64) for the backprojection."""
    torch = gpu
    theta, center = _geom(N, P)
    rng = np.random.default_rng(N)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    y = rng.uniform(0, 200, (P, S, N)).astype(np.float32)
    op = _op(theta, N, S, center)
    xt = torch.from_numpy(x).cuda()
    xt = (xt[0] if S == 1 else xt).clone().requires_grad_(True)
    yt = torch.from_numpy(y).cuda()
    yt = yt[:, 0, :].contiguous() if S == 1 else yt
    r = op(xt) - yt
    (0.5 * (r * r).sum()).backward()
    g = xt.grad.cpu().numpy().reshape(S, N, N)
    res = oracle.forward(theta, x, center) - y
    gref = oracle.adjoint(theta, res, center)
    e = rel_l2(g, gref)
    print(f"N={N} P={P} S={S}: gradient vs reference backprojector rel L2 {e:.2e}")
    assert e <= 1e-5
    op.close()


def _dense_A(oracle, N, P, theta, center):
    A = np.zeros((P * N, N * N), np.float32)
    for p in range(P):
        for c in range(N):
            idx, ln, _ = oracle.trace_ray(N, N, center, float(theta[p]), c)
            np.add.at(A[p * N + c], idx, ln)
    return A


def test_tv_adam_vs_dense_torch(gpu, oracle):
    """100 Adam steps (lr 1e-2) on 1/2||Ax - y||^2 + 1e-3 TV(x) through the
    HIP operator and through a dense fp32 matrix built from the oracle's
    rays (plain PyTorch reference of the same loss)."""
    torch = gpu
    from trace_amd.autograd import tv_adam, tv_isotropic
    N, P = 64, 48
    theta, center = _geom(N, P)
    from trace_amd.phantom import noisy_sinogram
    y = noisy_sinogram(theta, N, 1, seed=4, I0=1e4, K=20, device="cuda")[:, 0, :].contiguous()
    op = _op(theta, N, 1, center)
    x_hip, losses = tv_adam(op, y, steps=100, lr=1e-2, tv_weight=1e-3)
    A = torch.from_numpy(_dense_A(oracle, N, P, theta, center)).cuda()
    x = torch.zeros((N, N), device="cuda", requires_grad=True)
    opt = torch.optim.Adam([x], lr=1e-2)
    ref_losses = []
    for _ in range(100):
        opt.zero_grad()
        r = (A @ x.reshape(-1)).reshape(P, N) - y
        loss = 0.5 * (r * r).sum() + 1e-3 * tv_isotropic(x)
        loss.backward()
        opt.step()
        ref_losses.append(loss.item())
    e = rel_l2(x_hip.cpu().numpy(), x.detach().cpu().numpy())
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e} (dense {ref_losses[-1]:.4e}); "
          f"x rel L2 vs dense {e:.2e}")
    assert losses[-1] < 0.5 * losses[0]
    assert abs(losses[-1] - ref_losses[-1]) <= 1e-3 * ref_losses[-1]
    assert e <= 1e-3
    op.close()


def test_unrolled_sirt_matches_kernel_path(gpu):
    """10 unrolled SIRT iterations (no grad) vs trace_reconstruct SIRT."""
    torch = gpu
    from trace_amd import Projector
    from trace_amd.autograd import sirt_unrolled
    from trace_amd.phantom import noisy_sinogram
    N, P, S = 256, 180, 2
    theta, center = _geom(N, P)
    sino = noisy_sinogram(theta, N, S, seed=0, I0=1e4, K=10, device="cuda")
    op = _op(theta, N, S, center)
    with torch.no_grad():
        x_unrolled = sirt_unrolled(op, sino, 10)
    with Projector(theta, S, N, center) as pj:
        x_kernel = torch.zeros((S, N, N), device="cuda")
        pj.reconstruct("sirt", 10, sino, x_kernel)
    e = rel_l2(x_unrolled.cpu().numpy(), x_kernel.cpu().numpy())
    print(f"unrolled vs kernel SIRT: rel L2 {e:.2e}")
    assert e <= 1e-4
    op.close()


def test_unrolled_sirt_checkpoint_gradients(gpu):
    """Checkpointed and plain backward give identical gradients with
    respect to the sinogram and the per-iteration step factors, and the
    sinogram gradient is the adjoint of the (linear) unrolled map:
    <M dp, v> = <dp, M^T v>."""
    torch = gpu
    from trace_amd.autograd import sirt_unrolled
    from trace_amd.phantom import noisy_sinogram
    N, P, S, K = 256, 180, 1, 10
    theta, center = _geom(N, P)
    sino = noisy_sinogram(theta, N, S, seed=0, I0=1e4, K=10, device="cuda")[:, 0, :].contiguous()
    op = _op(theta, N, S, center)
    g = torch.Generator(device="cuda").manual_seed(7)
    v = torch.randn((N, N), device="cuda", generator=g)
    grads = []
    for ck in (True, False):
        p = sino.clone().requires_grad_(True)
        lam = torch.ones(K, device="cuda", requires_grad=True)
        x = sirt_unrolled(op, p, K, relax=lam, use_checkpoint=ck)
        (x * v).sum().backward()
        grads.append((p.grad.clone(), lam.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])
    dp = torch.randn(sino.shape, device="cuda", generator=g)
    with torch.no_grad():
        m_dp = sirt_unrolled(op, dp, K)
    lhs = (m_dp.double() * v.double()).sum().item()
    rhs = (dp.double() * grads[0][0].double()).sum().item()
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"unrolled SIRT adjoint identity: {lhs:.8e} vs {rhs:.8e}, rel {err:.2e}")
    assert err <= 1e-4
    op.close()


def test_unrolled_sirt_checkpoint_memory(gpu):
    """Memory kept for backward after the forward pass of a 10-iteration
    unrolled SIRT with a per-iteration TV step: checkpointing keeps the
    iterates only; without it every iteration's step activations stay.
    Gradients are the same either way."""
    torch = gpu
    from trace_amd.autograd import sirt_unrolled, tv_grad
    N, P, S, K = 512, 360, 2, 10
    theta, center = _geom(N, P)
    op = _op(theta, N, S, center)
    sino = torch.rand((P, S, N), device="cuda") * 100
    op.row_norms2(), op.column_sums()
    kept, grads = {}, {}
    for ck in (True, False):
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        p = sino.clone().requires_grad_(True)
        tau = torch.full((K,), 1e-2, device="cuda", requires_grad=True)
        x = sirt_unrolled(op, p, K, prox=lambda x, k: x - tau[k] * tv_grad(x),
                          use_checkpoint=ck)
        torch.cuda.synchronize()
        kept[ck] = torch.cuda.memory_allocated() - base
        x.square().sum().backward()
        grads[ck] = (p.grad.clone(), tau.grad.clone())
        del x, p, tau
    print(f"kept after forward: checkpointed {kept[True] / 2**20:.1f} MiB, "
          f"plain {kept[False] / 2**20:.1f} MiB")
    assert kept[True] * 2 < kept[False]
    assert rel_l2(grads[True][0].cpu().numpy(), grads[False][0].cpu().numpy()) <= 1e-6
    assert rel_l2(grads[True][1].cpu().numpy(), grads[False][1].cpu().numpy()) <= 1e-6
    op.close()
