"""HDF5 datasets as numpy arrays through libtrace_h5io.so (include/trace_h5io.h).

The reconstruction programs read the sinogram ``/exchange/data`` [P][S][C]
and the angles ``/exchange/theta`` [P] and write the image ``/data``
[S][N][N] (trace_h5io.cc:135-266, 459-503); these functions read and write
the same files from Python without h5py.  Values are read as float32
whatever the stored numeric type.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrace_h5io.so")
BIN = os.path.join(HERE, "bin")
PROGRAMS = ("sirt", "mlem", "pml", "apmlr", "trace")

TRACE_H5_OK = 0
TRACE_H5_E_INVALID = -1
TRACE_H5_E_IO = -2
TRACE_H5_E_TYPE = -3
MAX_DIMS = 8
_TYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.uint16): 2,
          np.dtype(np.int32): 3}

_lib = None


class H5Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run trace_amd/build.py)")
        L = ct.CDLL(LIB_PATH)
        i64p = ct.POINTER(ct.c_int64)
        L.trace_h5_dims.argtypes = [ct.c_char_p, ct.c_char_p, ct.POINTER(ct.c_int), i64p]
        L.trace_h5_read.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_void_p, ct.c_int64]
        L.trace_h5_read_slices.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int,
                                           ct.c_void_p]
        L.trace_h5_write.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_int, i64p, ct.c_int,
                                     ct.c_void_p, ct.c_int]
        L.trace_h5_last_error.restype = ct.c_char_p
        for f in ("trace_h5_dims", "trace_h5_read", "trace_h5_read_slices", "trace_h5_write"):
            getattr(L, f).restype = ct.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != TRACE_H5_OK:
        raise H5Error(rc, lib().trace_h5_last_error().decode())


def _b(s: str) -> bytes:
    return os.fsencode(s)


def dims(path: str, dataset: str) -> tuple:
    """Extent of a dataset (ReadMetadata, trace_h5io.cc:24-83)."""
    nd = ct.c_int(0)
    d = (ct.c_int64 * MAX_DIMS)()
    _check(lib().trace_h5_dims(_b(path), _b(dataset), ct.byref(nd), d))
    return tuple(int(d[i]) for i in range(nd.value))


def read(path: str, dataset: str) -> np.ndarray:
    """The whole dataset as float32 (ReadTheta, trace_h5io.cc:85-133)."""
    shape = dims(path, dataset)
    out = np.empty(shape, np.float32)
    _check(lib().trace_h5_read(_b(path), _b(dataset), out.ctypes.data, out.size))
    return out


def read_slices(path: str, dataset: str, begin: int = 0, count: int | None = None) -> np.ndarray:
    """Slices [begin, begin+count) of a [P][S][C] sinogram as float32 [P][count][C]
    (ReadSlices, trace_h5io.cc:135-266)."""
    shape = dims(path, dataset)
    if len(shape) != 3:
        raise H5Error(TRACE_H5_E_INVALID, f"sinogram dataset '{dataset}' must be 3-D "
                                          f"[P][S][C], has rank {len(shape)}")
    P, S, C = shape
    if count is None:
        count = S - begin
    out = np.empty((P, max(count, 0), C), np.float32)
    _check(lib().trace_h5_read_slices(_b(path), _b(dataset), int(begin), int(count),
                                      out.ctypes.data))
    return out


def write(path: str, dataset: str, data, truncate: bool = False) -> None:
    """Writes `data` (float32, float64, uint16 or int32) as a new dataset;
    truncate=True replaces the file, as the reconstruction programs do with
    their output (trace_h5io.cc:379-383)."""
    a = np.ascontiguousarray(data)
    if a.dtype not in _TYPES:
        raise TypeError(f"unsupported element type {a.dtype}")
    d = (ct.c_int64 * max(a.ndim, 1))(*a.shape) if a.ndim else (ct.c_int64 * 1)(1)
    _check(lib().trace_h5_write(_b(path), _b(dataset), max(a.ndim, 1), d, _TYPES[a.dtype],
                                a.ctypes.data, 1 if truncate else 0))


def write_input(path: str, sino, theta, data_path: str = "/exchange/data",
                theta_path: str = "/exchange/theta") -> None:
    """One input file holding the sinogram [P][S][C] and the angles [P], the
    layout every reconstruction program reads."""
    write(path, data_path, np.asarray(sino, np.float32), truncate=True)
    write(path, theta_path, np.asarray(theta, np.float32))


def program(name: str) -> str:
    """Path of a built reconstruction program (sirt, mlem, pml, apmlr, trace)."""
    if name not in PROGRAMS:
        raise ValueError(f"unknown program {name!r}; one of {PROGRAMS}")
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built (run trace_amd/build.py)")
    return p
