"""Failure paths of the C-ABI on a live handle, forced from the host: every
one must come back as a status code, with no device memory left behind.

The reference has no error handling of its own beyond exceptions that end
the process (trace_data.h:88, pml.cc:18-20 divides by F = 0 when beta = 0);
the cases are the ones a host can force: beta <= 0 for PML, iters < 0, an
unknown algorithm, NaN/Inf in the sinogram, and a failed device allocation
(trace_debug_fail_alloc) at every allocation of a create + reconstruct."""
import ctypes as ct

import numpy as np
import pytest

from trace_amd import _capi
from trace_amd._capi import TraceGeom, TraceParams

pytestmark = pytest.mark.gpu


def _live():
    return _capi.lib().trace_debug_live_allocations()


@pytest.fixture
def live_handle(gpu):
    import torch
    from trace_amd import Projector
    from trace_amd.phantom import angles_rad, noisy_sinogram
    th = angles_rad(24)
    sino = noisy_sinogram(th, 32, 2, seed=1, I0=1e4, K=4)
    pj = Projector(th, 2, 32, 16.0)
    ds = sino.cuda()
    dx = torch.ones((2, 32, 32), device="cuda")
    pj.reconstruct("mlem", 1, ds, dx)  # the handle is live and has run
    torch.cuda.synchronize()
    yield pj, ds, dx
    pj.close()


def _rc(pj, alg, iters, ds, dx, prm):
    return _capi.lib().trace_reconstruct(pj.handle, alg, iters, ds.data_ptr(), dx.data_ptr(),
                                         ct.byref(prm))


@pytest.mark.parametrize("beta", [0.0, -1.0, float("nan"), float("inf")])
def test_pml_beta_must_be_positive(live_handle, beta):
    pj, ds, dx = live_handle
    x0 = dx.clone()
    prm = TraceParams(beta, 1, 1, 1, 1, 1)
    for alg in (_capi.TRACE_PML, _capi.TRACE_APMLR):
        assert _rc(pj, alg, 2, ds, dx, prm) == _capi.TRACE_E_INVALID
        assert "beta" in _capi.lib().trace_last_error().decode()
    assert bool((dx == x0).all())  # nothing ran


@pytest.mark.parametrize("field,value", [("delta", 0.0), ("delta1", -1.0), ("regw", -1.0),
                                         ("beta1", float("nan")), ("apmlr_compat", 2)])
def test_apmlr_parameters(live_handle, field, value):
    pj, ds, dx = live_handle
    prm = TraceParams(0.1, 1, 0.01, 1, 1, 1)
    setattr(prm, field, value)
    assert _rc(pj, _capi.TRACE_APMLR, 1, ds, dx, prm) == _capi.TRACE_E_INVALID


def test_negative_iterations(live_handle):
    pj, ds, dx = live_handle
    prm = TraceParams(0.1, 1, 1, 1, 1, 1)
    for alg in range(4):
        assert _rc(pj, alg, -1, ds, dx, prm) == _capi.TRACE_E_INVALID
    assert "iters" in _capi.lib().trace_last_error().decode()


@pytest.mark.parametrize("alg", [-1, 4, 1 << 20])
def test_unknown_algorithm(live_handle, alg):
    pj, ds, dx = live_handle
    prm = TraceParams(0.1, 1, 1, 1, 1, 1)
    assert _rc(pj, alg, 1, ds, dx, prm) == _capi.TRACE_E_UNSUPPORTED


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
def test_nonfinite_sinogram(live_handle, bad):
    pj, ds, dx = live_handle
    d2 = ds.clone()
    d2[3, 1, 7] = bad
    x0 = dx.clone()
    prm = TraceParams(0.1, 1, 1, 1, 1, 1)
    assert _rc(pj, _capi.TRACE_MLEM, 2, d2, dx, prm) == _capi.TRACE_E_NONFINITE
    assert bool((dx == x0).all())
    # the handle still works afterwards
    assert _rc(pj, _capi.TRACE_MLEM, 1, ds, dx, prm) == _capi.TRACE_OK


def test_null_buffers_on_live_handle(live_handle):
    pj, ds, dx = live_handle
    L = _capi.lib()
    prm = TraceParams(0.1, 1, 1, 1, 1, 1)
    assert L.trace_reconstruct(pj.handle, 1, 1, None, dx.data_ptr(), ct.byref(prm)) \
        == _capi.TRACE_E_INVALID
    assert L.trace_forward(pj.handle, None, ds.data_ptr()) == _capi.TRACE_E_INVALID
    assert L.trace_adjoint(pj.handle, ds.data_ptr(), None) == _capi.TRACE_E_INVALID
    assert L.trace_set_option(pj.handle, 999, 1) == _capi.TRACE_E_UNSUPPORTED
    assert L.trace_set_option(pj.handle, _capi.TRACE_OPT_ORDERED_SUMS, 3) == _capi.TRACE_E_INVALID


def _create_and_run(alg, prm):
    """trace_create + trace_reconstruct + trace_forward on a small geometry;
    returns the first failing status (or TRACE_OK) and destroys the handle."""
    import torch
    from trace_amd.phantom import angles_rad
    L = _capi.lib()
    th = angles_rad(20)
    g = TraceGeom(20, 3, 24, 24, 12.0, th.ctypes.data_as(ct.POINTER(ct.c_float)))
    h = ct.c_void_p()
    rc = L.trace_create(ct.byref(h), ct.byref(g), 0, None)
    if rc != _capi.TRACE_OK:
        assert not h.value  # nothing handed out on failure
        return rc
    sino = torch.rand((20, 3, 24), device="cuda") + 0.5
    x = torch.ones((3, 24, 24), device="cuda")
    y = torch.empty((20, 3, 24), device="cuda")
    try:
        rc = L.trace_reconstruct(h, alg, 2, sino.data_ptr(), x.data_ptr(), ct.byref(prm))
        if rc == _capi.TRACE_OK:
            rc = L.trace_forward(h, x.data_ptr(), y.data_ptr())
        torch.cuda.synchronize()
    finally:
        L.trace_destroy(h)
    return rc


@pytest.mark.parametrize("alg", [_capi.TRACE_MLEM, _capi.TRACE_PML, _capi.TRACE_APMLR])
def test_allocation_failure_at_every_step(gpu, alg):
    """Each device allocation of create + reconstruct + forward, made to fail
    in turn: TRACE_E_NOMEM every time, and no allocation left live."""
    L = _capi.lib()
    prm = TraceParams(0.1, 1, 0.01, 1, 1, 1)
    base = _live()
    assert _create_and_run(alg, prm) == _capi.TRACE_OK
    assert _live() == base
    failures = 0
    for k in range(64):
        L.trace_debug_fail_alloc(k)
        rc = _create_and_run(alg, prm)
        L.trace_debug_fail_alloc(-1)
        assert _live() == base, f"allocation {k}: {_live() - base} left live"
        if rc == _capi.TRACE_OK:
            break  # k exceeds the number of allocations
        assert rc == _capi.TRACE_E_NOMEM, (k, rc, L.trace_last_error())
        failures += 1
    assert failures >= 8  # create: 5 tables; reconstruct: rays, count, work buffers


def test_batch_allocation_failure(gpu):
    import torch
    from trace_amd._capi import TraceBatchSlice
    from trace_amd.phantom import angles_rad
    L = _capi.lib()
    base = _live()
    ths = [angles_rad(12), angles_rad(16)]
    arr = (TraceBatchSlice * 2)()
    keep = []
    for a, th, C in zip(arr, ths, (16, 24)):
        s = torch.rand((th.size, C), device="cuda") + 0.5
        x = torch.ones((C, C), device="cuda")
        keep += [s, x]
        a.num_projs, a.num_cols, a.center, a.iters, a.num_grids = th.size, C, C / 2, 2, 0
        a.theta = th.ctypes.data_as(ct.POINTER(ct.c_float))
        a.alg = _capi.TRACE_MLEM
        a.sino, a.recon = s.data_ptr(), x.data_ptr()
    failures = 0
    for k in range(96):
        L.trace_debug_fail_alloc(k)
        h = ct.c_void_p()
        rc = L.trace_batch_create(ct.byref(h), arr, 2, 0, None)
        L.trace_debug_fail_alloc(-1)
        if rc == _capi.TRACE_OK:
            assert L.trace_batch_run(h, 0) == _capi.TRACE_OK
            torch.cuda.synchronize()
            L.trace_batch_destroy(h)
            assert _live() == base
            break
        assert rc == _capi.TRACE_E_NOMEM, (k, rc, L.trace_last_error())
        assert not h.value
        assert _live() == base
        failures += 1
    assert failures >= 10
