"""The C-ABI library on the CPU: it loads, exports every declared symbol,
its host helpers match the oracle bit for bit, and argument validation
returns status codes (these paths return before any HIP call)."""
import ctypes as ct
import os

import numpy as np
import pytest

from trace_amd import _capi
from trace_amd._capi import TraceGeom, TraceParams


def test_library_exports_header_symbols():
    L = _capi.lib()
    names = _capi.header_symbols()
    assert "trace_create" in names and "trace_reconstruct" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_helpers_match_oracle(oracle):
    L = _capi.lib()
    rng = np.random.default_rng(0)
    deg = rng.uniform(-720, 720, 4000).astype(np.float32)
    out = np.empty_like(deg)
    L.trace_deg2rad(deg.ctypes.data, out.ctypes.data, deg.size)
    assert np.array_equal(out.view(np.int32), oracle.deg2rad(deg).view(np.int32))
    v = rng.normal(size=1000).astype(np.float32)
    w = v.copy()
    L.trace_absolute(w.ctypes.data, w.size)
    assert np.array_equal(w, np.abs(v))


def _geom(P=4, S=1, C=8, N=8, center=4.0, theta=None):
    th = np.linspace(0, 1, P).astype(np.float32) if theta is None else theta
    g = TraceGeom(P, S, C, N, center, th.ctypes.data_as(ct.POINTER(ct.c_float)))
    return g, th


def _create(g, dev=0):
    h = ct.c_void_p()
    rc = _capi.lib().trace_create(ct.byref(h), ct.byref(g) if g is not None else None, dev, None)
    return rc, h


@pytest.mark.parametrize("kw", [
    dict(P=0), dict(S=0), dict(C=1, N=1), dict(N=1), dict(C=40000, N=40000), dict(N=40000),
    dict(center=float("nan")), dict(theta=np.array([0, np.nan, 1, 2], np.float32)),
])
def test_create_rejects_bad_geometry(kw):
    g, th = _geom(**kw)
    rc, h = _create(g)
    assert rc == _capi.TRACE_E_INVALID
    assert not h.value
    assert _capi.lib().trace_last_error()


def test_create_rejects_null_and_negative_device():
    L = _capi.lib()
    assert L.trace_create(None, None, 0, None) == _capi.TRACE_E_INVALID
    rc, _ = _create(None)
    assert rc == _capi.TRACE_E_INVALID
    g, th = _geom()
    rc, _ = _create(g, dev=-1)
    assert rc == _capi.TRACE_E_INVALID


def test_null_handle_calls():
    L = _capi.lib()
    prm = TraceParams(1, 1, 1, 1, 1, 1)
    assert L.trace_reconstruct(None, 0, 1, None, None, ct.byref(prm)) == _capi.TRACE_E_INVALID
    assert L.trace_forward(None, None, None) == _capi.TRACE_E_INVALID
    assert L.trace_adjoint(None, None, None) == _capi.TRACE_E_INVALID
    assert L.trace_set_option(None, 1, 1) == _capi.TRACE_E_INVALID
    assert L.trace_count_segments(None, None, None) == _capi.TRACE_E_INVALID
    L.trace_destroy(None)  # accepted


def test_batch_argument_validation():
    """Batch creation rejects bad slices before any HIP call."""
    L = _capi.lib()
    from trace_amd._capi import TraceBatchSlice
    h = ct.c_void_p()
    th = np.zeros(4, np.float32)
    arr = (TraceBatchSlice * 2)()
    for a in arr:
        a.num_projs, a.num_cols, a.center, a.iters = 4, 8, 4.0, 3
        a.theta = th.ctypes.data_as(ct.POINTER(ct.c_float))
        a.alg = _capi.TRACE_SIRT
        a.sino, a.recon = 16, 32  # never dereferenced on these paths
    assert L.trace_batch_create(None, arr, 2, 0, None) == _capi.TRACE_E_INVALID
    assert L.trace_batch_create(ct.byref(h), arr, 0, 0, None) == _capi.TRACE_E_INVALID
    arr[1].alg = _capi.TRACE_PML
    assert L.trace_batch_create(ct.byref(h), arr, 2, 0, None) == _capi.TRACE_E_UNSUPPORTED
    assert "SIRT or MLEM" in L.trace_last_error().decode()
    arr[1].alg = _capi.TRACE_MLEM
    arr[0].iters = -1
    assert L.trace_batch_create(ct.byref(h), arr, 2, 0, None) == _capi.TRACE_E_INVALID
    arr[0].iters = 3
    arr[0].sino = None
    assert L.trace_batch_create(ct.byref(h), arr, 2, 0, None) == _capi.TRACE_E_INVALID
    assert L.trace_batch_run(None, 0) == _capi.TRACE_E_INVALID
    assert L.trace_batch_set_launch_shape(None, 8, 32, 32) == _capi.TRACE_E_INVALID
    L.trace_batch_destroy(None)  # accepted


def test_python_wrapper_raises():
    from trace_amd import Projector, TraceInvalid
    with pytest.raises(TraceInvalid):
        Projector(np.zeros(4, np.float32), 1, 8, 4.0, num_grids=1)
    from trace_amd.recon import _run
    with pytest.raises(ValueError):
        _run("sirt", np.zeros(3, np.float32), np.zeros((4, 1, 8), np.float32), 0.0, 1, 4.0, None, 0)


def test_no_packed_fp32_in_device_code():
    """The build writes the gfx950 assembly of every kernel source next to
    the objects and refuses packed-fp32 VALU instructions; check the record."""
    from trace_amd import build
    for src in build.SOURCES:
        s = os.path.join(build.HERE, "_build", src + ".gfx950.s")
        assert os.path.exists(s), s
        text = open(s).read()
        for op in ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32"):
            assert op not in text, (src, op)
