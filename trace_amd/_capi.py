"""ctypes binding of the C-ABI in include/trace_hip.h (libtrace_hip.so).

The library is loaded from this package directory.  There is no fallback:
if the library is missing, every entry raises.
"""
from __future__ import annotations

import ctypes as ct
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrace_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "trace_hip.h")

TRACE_OK = 0
TRACE_E_INVALID = -1
TRACE_E_UNSUPPORTED = -2
TRACE_E_HIP = -3
TRACE_E_NOMEM = -4
TRACE_E_NONFINITE = -5

TRACE_SIRT, TRACE_MLEM, TRACE_PML, TRACE_APMLR = 0, 1, 2, 3
ALGS = {"sirt": TRACE_SIRT, "mlem": TRACE_MLEM, "pml": TRACE_PML, "apmlr": TRACE_APMLR}

TRACE_OPT_EXACT_SQRT = 1
TRACE_OPT_PHASE_TIMING = 2
TRACE_OPT_BACKPROJECTOR = 3
TRACE_OPT_FWD_BUNDLE = 4
TRACE_OPT_FWD_LANES = 5
TRACE_OPT_ORDERED_SUMS = 6
TRACE_OPT_ORDERED_TILE = 7
TRACE_OPT_FWD_STAGED = 8
PHASES = ("forward", "back", "update", "setup")


class TraceGeom(ct.Structure):
    _fields_ = [
        ("num_projs", ct.c_int),
        ("num_slices", ct.c_int),
        ("num_cols", ct.c_int),
        ("num_grids", ct.c_int),
        ("center", ct.c_float),
        ("theta", ct.POINTER(ct.c_float)),
    ]


class TraceParams(ct.Structure):
    _fields_ = [
        ("beta", ct.c_float),
        ("beta1", ct.c_float),
        ("delta", ct.c_float),
        ("delta1", ct.c_float),
        ("regw", ct.c_float),
        ("apmlr_compat", ct.c_int),
    ]


class TraceBatchSlice(ct.Structure):
    _fields_ = [
        ("num_projs", ct.c_int),
        ("num_cols", ct.c_int),
        ("center", ct.c_float),
        ("theta", ct.POINTER(ct.c_float)),
        ("alg", ct.c_int),
        ("iters", ct.c_int),
        ("sino", ct.c_void_p),
        ("recon", ct.c_void_p),
        ("num_grids", ct.c_int),
    ]


class TraceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"trace error {code}: {msg}")
        self.code = code


class TraceInvalid(TraceError, ValueError):
    pass


class TraceUnsupported(TraceError, NotImplementedError):
    pass


class TraceNoMem(TraceError, MemoryError):
    pass


class TraceNonFinite(TraceError, ValueError):
    pass


_EXC = {
    TRACE_E_INVALID: TraceInvalid,
    TRACE_E_UNSUPPORTED: TraceUnsupported,
    TRACE_E_NOMEM: TraceNoMem,
    TRACE_E_NONFINITE: TraceNonFinite,
}

_lib = None

_vp = ct.c_void_p
_SIGS = {
    "trace_create": (ct.c_int, [ct.POINTER(_vp), ct.POINTER(TraceGeom), ct.c_int, _vp]),
    "trace_destroy": (None, [_vp]),
    "trace_set_option": (ct.c_int, [_vp, ct.c_int, ct.c_int]),
    "trace_get_option": (ct.c_int, [_vp, ct.c_int, ct.POINTER(ct.c_int)]),
    "trace_set_launch_shape": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int]),
    "trace_get_launch_shape": (ct.c_int, [_vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                          ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
    "trace_segment_counts": (ct.c_int, [_vp, _vp]),
    "trace_reconstruct": (ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, _vp,
                                     ct.POINTER(TraceParams)]),
    "trace_reconstruct_host": (ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, _vp,
                                          ct.POINTER(TraceParams)]),
    "trace_forward": (ct.c_int, [_vp, _vp, _vp]),
    "trace_forward_noalloc": (ct.c_int, [_vp, _vp, _vp]),
    "trace_adjoint": (ct.c_int, [_vp, _vp, _vp]),
    "trace_column_sums": (ct.c_int, [_vp, _vp]),
    "trace_ray_sums": (ct.c_int, [_vp, _vp, _vp]),
    "trace_dump_rays": (ct.c_int, [_vp, ct.c_int, _vp, _vp, ct.c_int, _vp, _vp,
                                   _vp, _vp]),
    "trace_count_segments": (ct.c_int, [_vp, ct.POINTER(ct.c_int64),
                                        ct.POINTER(ct.c_int64)]),
    "trace_phase_times": (ct.c_int, [_vp, ct.POINTER(ct.c_double),
                                     ct.POINTER(ct.c_int64), ct.c_int]),
    "trace_deg2rad": (None, [_vp, _vp, ct.c_int]),
    "trace_absolute": (None, [_vp, ct.c_int64]),
    "trace_batch_create": (ct.c_int, [ct.POINTER(_vp), ct.POINTER(TraceBatchSlice), ct.c_int,
                                       ct.c_int, _vp]),
    "trace_batch_set_launch_shape": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int]),
    "trace_batch_get_launch_shape": (ct.c_int, [_vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                                ct.POINTER(ct.c_int)]),
    "trace_batch_set_option": (ct.c_int, [_vp, ct.c_int, ct.c_int]),
    "trace_batch_run": (ct.c_int, [_vp, ct.c_int]),
    "trace_batch_prepare": (ct.c_int, [_vp, ct.c_int]),
    "trace_batch_reconstruct": (ct.c_int, [ct.POINTER(TraceBatchSlice), ct.c_int, ct.c_int, _vp]),
    "trace_batch_destroy": (None, [_vp]),
    "trace_debug_fail_alloc": (None, [ct.c_int]),
    "trace_debug_live_allocations": (ct.c_int64, []),
    "trace_last_error": (ct.c_char_p, []),
    "trace_version": (ct.c_char_p, []),
}


def header_symbols() -> list[str]:
    """Function names declared in include/trace_hip.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(trace_[a-z0-9_]+)\s*\(", text)))


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m trace_amd.build`")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != TRACE_OK:
        msg = lib().trace_last_error().decode(errors="replace")
        raise _EXC.get(rc, TraceError)(rc, msg)
