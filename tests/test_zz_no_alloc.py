"""trace_forward_noalloc, the forward projection that allocates nothing
(its scratch is a static device array), against trace_forward: bit for bit,
on two calls in a row, the second with a smaller geometry, with no library
allocation made by either call; and a geometry larger than the scratch is
refused before anything runs."""
import ctypes as ct

import numpy as np
import pytest

from trace_amd import _capi

pytestmark = pytest.mark.gpu


def _noalloc(pj, x):
    import torch
    y = torch.full((pj.P, pj.S, pj.C), -1.0, device="cuda")
    torch.cuda.synchronize()
    L = _capi.lib()
    live = L.trace_debug_live_allocations()
    rc = L.trace_forward_noalloc(pj.handle, x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    assert rc == _capi.TRACE_OK, L.trace_last_error()
    assert L.trace_debug_live_allocations() == live
    return y


def test_forward_noalloc_matches_forward(gpu):
    torch = gpu
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    rng = np.random.default_rng(7)
    # the second call is smaller than the first, and its handle has never run
    # trace_forward when trace_forward_noalloc is called
    for N, P, S, first_noalloc in [(512, 360, 3, False), (160, 100, 2, True)]:
        print(f"noalloc forward N={N} P={P} S={S}", flush=True)
        th = angles_rad(P)
        x = torch.from_numpy(rng.uniform(0, 1, (S, N, N)).astype(np.float32)).cuda()
        with Projector(th, S, N, N / 2.0) as pj:
            if first_noalloc:
                y = _noalloc(pj, x)
                y_ref = pj.forward(x)
            else:
                y_ref = pj.forward(x)
                y = _noalloc(pj, x)
            torch.cuda.synchronize()
            a = y.cpu().numpy().view(np.uint32)
            b = y_ref.cpu().numpy().view(np.uint32)
        assert np.array_equal(a, b)


def test_forward_noalloc_refuses_oversized_geometry(gpu):
    torch = gpu
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad
    N, P = 4096, 1200  # ray table alone: 157 MB
    with Projector(angles_rad(P), 1, N, N / 2.0) as pj:
        L = _capi.lib()
        x = torch.zeros((1, 8), device="cuda")  # never read
        rc = L.trace_forward_noalloc(pj.handle, x.data_ptr(), x.data_ptr())
        assert rc == _capi.TRACE_E_INVALID
        assert "128 MiB" in L.trace_last_error().decode()
