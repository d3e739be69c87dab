"""Forward and adjoint operators on the GPU against the oracle: forward
projection bit-identical in exact mode (same segment order, no FMA); the
adjoint and column sums within fp32 atomic-order rounding; the dot-product
identity <Ax, r> = <x, A^T r> to 1e-5."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


@pytest.mark.parametrize("N,P,S", [(33, 21, 2), (64, 45, 3), (256, 90, 1)])
def test_forward_bitexact(gpu, oracle, N, P, S):
    torch = gpu
    from trace_amd import Projector
    rng = np.random.default_rng(N + P)
    th = rng.uniform(-0.5, 3.5, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    print(f"forward N={N} P={P} S={S}", flush=True)
    with Projector(th, S, N, N / 2) as pj:
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(_bits(y), _bits(oracle.forward(th, x, N / 2)))


def test_forward_fast_sqrt_close(gpu, oracle):
    torch = gpu
    from trace_amd import Projector
    N, P, S = 128, 60, 2
    rng = np.random.default_rng(1)
    th = rng.uniform(0, np.pi, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    with Projector(th, S, N, N / 2, exact_sqrt=False) as pj:
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel_l2(y, oracle.forward(th, x, N / 2)) < 1e-6


@pytest.mark.parametrize("N,P,S", [(40, 30, 2), (128, 64, 1)])
def test_adjoint_and_dot(gpu, oracle, N, P, S):
    torch = gpu
    from trace_amd import Projector
    rng = np.random.default_rng(N)
    th = rng.uniform(0, np.pi, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    r = rng.normal(size=(P, S, N)).astype(np.float32)
    print(f"adjoint N={N} P={P} S={S}", flush=True)
    with Projector(th, S, N, N / 2) as pj:
        z = pj.adjoint(torch.from_numpy(r).cuda()).cpu().numpy()
        y = pj.forward(torch.from_numpy(x).cuda()).cpu().numpy()
        den = pj.column_sums().cpu().numpy()
    assert rel_l2(z, oracle.adjoint(th, r, N / 2)) < 1e-6
    lhs = np.sum(y.astype(np.float64) * r)
    rhs = np.sum(x.astype(np.float64) * z.astype(np.float64))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
    ones = np.ones((P, 1, N), np.float32)
    assert rel_l2(den, oracle.adjoint(th, ones, N / 2)[0]) < 1e-6


@pytest.mark.parametrize("handle_on_side", [False, True])
def test_calls_ordered_against_another_current_stream(gpu, oracle, handle_on_side):
    """A handle made on one stream, called while torch's current stream is
    another one whose queued work produces the input: the call sees that
    input, and work queued after it on the current stream sees the output.
    Both ways round: handle on the default stream called from a side stream,
    and handle on a side stream called from the default stream."""
    torch = gpu
    from trace_amd import Projector
    N, P, S = 256, 120, 4
    rng = np.random.default_rng(3)
    th = rng.uniform(0, np.pi, P).astype(np.float32)
    x = rng.uniform(0, 1, (S, N, N)).astype(np.float32)
    want = oracle.forward(th, x, N / 2)
    side = torch.cuda.Stream()
    xs = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    if handle_on_side:
        with torch.cuda.stream(side):
            pj = Projector(th, S, N, N / 2)
        caller = torch.cuda.default_stream()
    else:
        pj = Projector(th, S, N, N / 2)
        caller = side
    with pj:
        with torch.cuda.stream(caller):
            # a long producer on the calling stream, then the projection
            xi = torch.zeros_like(xs)
            for _ in range(50):
                xi.mul_(0.5)
            xi.add_(xs)
            y = pj.forward(xi)
            y2 = y * 2.0  # consumer queued on the calling stream
        caller.synchronize()
        got = y.cpu().numpy()
        got2 = y2.cpu().numpy()
    assert np.array_equal(_bits(got), _bits(want))
    assert np.array_equal(_bits(got2), _bits(2.0 * want))
