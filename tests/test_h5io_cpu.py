"""HDF5 input/output library (include/trace_h5io.h) on the CPU: exported
symbols, round trips, the sinogram slice read of ReadSlices
(trace_h5io.cc:135-266), type conversion, and error statuses."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from trace_amd import h5io

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "trace_h5io.h")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


def test_library_exports_header_symbols():
    names = sorted(set(re.findall(r"\b(trace_h5_[a-z0-9_]+)\s*\(", open(HEADER).read())))
    assert {"trace_h5_dims", "trace_h5_read", "trace_h5_read_slices", "trace_h5_write"} <= set(names)
    L = h5io.lib()
    assert [n for n in names if not hasattr(L, n)] == []


def test_round_trip_and_slices(tmp_path):
    rng = np.random.default_rng(0)
    P, S, C = 7, 5, 11
    sino = rng.standard_normal((P, S, C)).astype(np.float32)
    theta = np.linspace(0, 180, P, endpoint=False).astype(np.float32)
    f = str(tmp_path / "in.h5")
    h5io.write_input(f, sino, theta)
    assert h5io.dims(f, "/exchange/data") == (P, S, C)
    assert h5io.dims(f, "/exchange/theta") == (P,)
    assert np.array_equal(h5io.read(f, "/exchange/theta"), theta)
    assert np.array_equal(h5io.read(f, "/exchange/data"), sino)
    assert np.array_equal(h5io.read_slices(f, "/exchange/data"), sino)
    for b, n in [(0, 1), (1, 3), (4, 1), (0, 5), (2, 3)]:
        got = h5io.read_slices(f, "/exchange/data", b, n)
        assert got.shape == (P, n, C)
        assert np.array_equal(got, sino[:, b:b + n, :]), (b, n)


@pytest.mark.parametrize("dtype", [np.float64, np.uint16, np.int32])
def test_values_are_converted_to_float(tmp_path, dtype):
    """A float64 theta or an integer sinogram is read as float values (the
    reference would reinterpret the stored bytes as float)."""
    f = str(tmp_path / "t.h5")
    a = (np.arange(24).reshape(2, 3, 4) * 3 + 1).astype(dtype)
    if dtype == np.float64:
        a = a / 7.0
    h5io.write(f, "/exchange/data", a, truncate=True)
    assert np.array_equal(h5io.read(f, "/exchange/data"), a.astype(np.float32))
    assert np.array_equal(h5io.read_slices(f, "/exchange/data", 1, 2),
                          a.astype(np.float32)[:, 1:3])


def test_truncate_replaces_and_append_adds(tmp_path):
    f = str(tmp_path / "o.h5")
    h5io.write(f, "/a", np.ones(3, np.float32), truncate=True)
    h5io.write(f, "/g/b", np.zeros((2, 2), np.float32))
    assert h5io.dims(f, "/a") == (3,) and h5io.dims(f, "/g/b") == (2, 2)
    with pytest.raises(h5io.H5Error) as e:  # dataset exists already
        h5io.write(f, "/a", np.ones(3, np.float32))
    assert e.value.code == h5io.TRACE_H5_E_IO
    h5io.write(f, "/c", np.full(4, 2.0, np.float32), truncate=True)
    assert h5io.dims(f, "/c") == (4,)
    with pytest.raises(h5io.H5Error):
        h5io.dims(f, "/a")  # gone with the truncation


def test_error_statuses(tmp_path):
    f = str(tmp_path / "e.h5")
    h5io.write_input(f, np.zeros((3, 2, 4), np.float32), np.zeros(3, np.float32))
    cases = [
        (lambda: h5io.read(str(tmp_path / "missing.h5"), "/exchange/data"), h5io.TRACE_H5_E_IO,
         "unable to open HDF5 file"),
        (lambda: h5io.read(f, "/exchange/nothing"), h5io.TRACE_H5_E_IO, "unable to open dataset"),
        (lambda: h5io.read_slices(f, "/exchange/data", 1, 2), h5io.TRACE_H5_E_INVALID, "outside"),
        (lambda: h5io.read_slices(f, "/exchange/data", -1, 1), h5io.TRACE_H5_E_INVALID, "outside"),
        (lambda: h5io.read_slices(f, "/exchange/data", 0, 0), h5io.TRACE_H5_E_INVALID, "outside"),
        (lambda: h5io.read_slices(f, "/exchange/theta", 0, 1), h5io.TRACE_H5_E_INVALID, "3-D"),
    ]
    for fn, code, text in cases:
        with pytest.raises(h5io.H5Error) as e:
            fn()
        assert e.value.code == code and text in str(e.value), str(e.value)
    # element count must match the buffer
    import ctypes as ct
    buf = np.zeros(5, np.float32)
    rc = h5io.lib().trace_h5_read(f.encode(), b"/exchange/theta", buf.ctypes.data, 5)
    assert rc == h5io.TRACE_H5_E_INVALID
    assert b"holds 3 values" in h5io.lib().trace_h5_last_error()
    rc = h5io.lib().trace_h5_dims(None, b"/a", ct.byref(ct.c_int()), (ct.c_int64 * 8)())
    assert rc == h5io.TRACE_H5_E_INVALID


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_files_read_by_standard_tools(tmp_path):
    """The files are plain HDF5: h5dump reports the layout the reference's
    readers expect (3-D float32 /exchange/data, 1-D /exchange/theta)."""
    f = str(tmp_path / "in.h5")
    h5io.write_input(f, np.ones((6, 2, 8), np.float32), np.arange(6, dtype=np.float32))
    out = subprocess.run([H5DUMP, "-H", f], capture_output=True, text=True, check=True).stdout
    assert re.search(r'DATASET "data" \{\s*DATATYPE\s+H5T_IEEE_F32LE\s*DATASPACE\s+SIMPLE '
                     r'\{ \( 6, 2, 8 \)', out), out
    assert re.search(r'DATASET "theta" \{\s*DATATYPE\s+H5T_IEEE_F32LE\s*DATASPACE\s+SIMPLE '
                     r'\{ \( 6 \)', out), out
