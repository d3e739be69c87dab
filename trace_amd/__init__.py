"""MI355X (gfx950) iterative tomographic reconstruction hot path.

Ray-traced forward projection, backprojection and the SIRT / MLEM / PML /
APMLR image updates of Trace, as hand-written HIP kernels behind the C-ABI
in ``include/trace_hip.h``.  See :mod:`trace_amd.recon`.
"""
from .recon import (  # noqa: F401
    Projector, RegParams, absolute, apmlr, degree_to_radian, mlem, pml, sirt)
from ._capi import TraceError, TraceInvalid, TraceNonFinite, TraceUnsupported  # noqa: F401

__all__ = [
    "Projector", "RegParams", "absolute", "apmlr", "degree_to_radian", "mlem",
    "pml", "sirt", "TraceError", "TraceInvalid", "TraceNonFinite",
    "TraceUnsupported",
]
