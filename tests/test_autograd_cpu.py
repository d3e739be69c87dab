"""CPU checks of the differentiable layer's host logic and of the
finite-difference checker the GPU gradient tests rely on."""
import numpy as np
import pytest

import fdcheck


def _brute_rays_through(orc, N, C, center, theta, j):
    out = set()
    for p, q in enumerate(theta):
        for c in range(C):
            idx, _, _ = orc.trace_ray(N, C, center, float(q), c)
            if (idx == j).any():
                out.add((p, c))
    return out


@pytest.mark.parametrize("N,P,center", [(24, 30, 12.0), (17, 19, 8.5), (32, 40, 15.25)])
def test_rays_through_pixel_is_complete(oracle, N, P, center):
    """The candidate window around a pixel finds exactly the rays whose
    traversal has a segment there (brute force over every ray)."""
    theta = (np.arange(P) * np.pi / P).astype(np.float32)
    rng = np.random.default_rng(N)
    for j in rng.choice(N * N, 6, replace=False):
        proj, col, off, idx, ln = fdcheck.rays_through_pixel(oracle, N, N, center, theta, j)
        got = set(zip(proj.tolist(), col.tolist()))
        assert got == _brute_rays_through(oracle, N, N, center, theta, j)
        for i in range(proj.size):  # every kept ray really crosses j
            assert (idx[off[i]:off[i + 1]] == j).any()


def test_fd_data_term_matches_exact_derivative(oracle):
    """For the quadratic data term the central difference equals the exact
    derivative sum_i r_i l_ij; computed here from a dense A built from the
    oracle's rays (an independent assembly)."""
    N, P = 20, 24
    center = N / 2.0
    theta = (np.arange(P) * np.pi / P).astype(np.float32)
    A = np.zeros((P * N, N * N))
    for p in range(P):
        for c in range(N):
            idx, ln, _ = oracle.trace_ray(N, N, center, float(theta[p]), c)
            np.add.at(A[p * N + c], idx, ln.astype(np.float64))
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, (N, N))
    y = rng.uniform(0, 10, (P, N))
    g = A.T @ (A @ x.ravel() - y.ravel())
    for j in rng.choice(N * N, 10, replace=False):
        fd = fdcheck.fd_data_term(oracle, N, N, center, theta, x, y, j, h=1e-3)
        assert abs(fd - g[j]) <= 1e-9 * max(1.0, abs(g[j]))


def test_fd_data_term_catches_wrong_gradient(oracle):
    """A planted mistake (one ray dropped from the sum) is detected."""
    N, P = 16, 12
    theta = (np.arange(P) * np.pi / P).astype(np.float32)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (N, N))
    y = rng.uniform(0, 10, (P, N))
    j = 7 * N + 9
    proj, col, off, idx, ln = fdcheck.rays_through_pixel(oracle, N, N, N / 2.0, theta, j)
    full = fdcheck.fd_data_term(oracle, N, N, N / 2.0, theta, x, y, j)
    # the same sum without the first ray
    s = np.array([np.sum(x.ravel()[idx[off[i]:off[i + 1]]] * ln[off[i]:off[i + 1]])
                  for i in range(proj.size)])
    lj = np.array([np.sum(ln[off[i]:off[i + 1]][idx[off[i]:off[i + 1]] == j])
                   for i in range(proj.size)])
    r = s - y[proj, col]
    assert abs(np.sum(r * lj) - full) < 1e-6 * abs(full)
    assert abs(np.sum(r[1:] * lj[1:]) - full) > 1e-3 * abs(full)


def test_fd_tv_matches_autograd_float64():
    torch = pytest.importorskip("torch")
    from trace_amd.autograd import tv_isotropic
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, (12, 12))
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    tv_isotropic(xt).backward()
    g = xt.grad.numpy().ravel()
    for j in [13, 40, 77, 130]:
        fd = fdcheck.fd_tv_term(x, j, h=1e-4)
        assert abs(fd - g[j]) < 1e-6 * max(1.0, abs(g[j]))


def test_tv_grad_closed_form_float64():
    torch = pytest.importorskip("torch")
    from trace_amd.autograd import tv_grad, tv_isotropic
    rng = np.random.default_rng(2)
    x = torch.tensor(rng.uniform(0, 1, (2, 9, 11)), dtype=torch.float64, requires_grad=True)
    tv_isotropic(x).backward()
    g = tv_grad(x.detach())
    assert torch.allclose(g, x.grad, rtol=1e-12, atol=1e-12)


def test_autograd_module_imports_without_gpu():
    pytest.importorskip("torch")
    import trace_amd.autograd as ag
    for name in ("RayTransform", "project", "backproject", "tv_adam",
                 "sirt_unrolled", "ls_tv_loss"):
        assert hasattr(ag, name)
