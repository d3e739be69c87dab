"""Host-side interface of the reconstruction hot path.

Two levels, both over the C-ABI (``include/trace_hip.h``):

* :class:`Projector` -- one slice stack's geometry on one GPU (the role of
  the reference's ``TraceMetadata``, trace_data.h:53-109): forward and
  adjoint projection, column sums, and the iteration loop on device
  tensors.  PyTorch provides device memory and the stream only.
* :func:`sirt`, :func:`mlem`, :func:`pml`, :func:`apmlr` -- the reference
  applications' reconstruction step (sirt_main.cc:120-202,
  mlem_main.cc:113-182, pml_main.cc:120-193, apmlr_main.cc:150-244) with the
  HDF5 reads replaced by arrays: the same per-application preprocessing and
  initial values, the same argument meaning, and errors raised as
  exceptions like the reference's ``std::invalid_argument``.
"""
from __future__ import annotations

import contextlib
import ctypes as ct
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import ALGS, TraceGeom, TraceParams, check, lib

KPI = np.float32(3.14159265358979)  # trace_utils.h:8


@dataclass
class RegParams:
    """Regularisation parameters; defaults are the reference CLI defaults
    (pml_main.cc:48-49, apmlr_main.cc:22-26)."""

    beta: float = 10.0
    beta1: float = 1.0
    delta: float = 1.0
    delta1: float = 1.0
    regw: float = 1.0
    apmlr_compat: bool = True

    def c(self) -> TraceParams:
        return TraceParams(self.beta, self.beta1, self.delta, self.delta1,
                           self.regw, 1 if self.apmlr_compat else 0)


def degree_to_radian(theta_deg) -> np.ndarray:
    """trace_utils::DegreeToRadian (trace_utils.cc:10-24), bit for bit."""
    d = np.ascontiguousarray(theta_deg, dtype=np.float32)
    out = np.empty_like(d)
    lib().trace_deg2rad(d.ctypes.data, out.ctypes.data, d.size)
    return out


def absolute(a) -> np.ndarray:
    """trace_utils::Absolute (trace_utils.cc:4-8) on a copy."""
    v = np.array(a, dtype=np.float32, copy=True, order="C")
    lib().trace_absolute(v.ctypes.data, v.size)
    return v


def _dptr(t) -> int:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (HIP) torch tensor")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError("expected a contiguous float32 tensor")
    return t.data_ptr()


class Projector:
    """Geometry of one slice stack on one GPU (a C-ABI handle).

    The handle's kernels run on the stream given at construction (by
    default torch's current stream at that time).  A call made while torch's
    current stream is another one is ordered against it: the handle's
    stream first waits for the work queued so far on the current stream
    (which produced the inputs), and the current stream then waits for the
    call's kernels (which produce the outputs).  A handle on the null
    stream called from another stream synchronises instead."""

    def __init__(self, theta_rad, num_slices: int, num_cols: int,
                 center: float = 0.0, device: int = 0, stream=None,
                 exact_sqrt: bool = True, num_grids: int | None = None):
        theta = np.ascontiguousarray(theta_rad, dtype=np.float32)
        self.theta = theta
        self.P = int(theta.size)
        self.S = int(num_slices)
        self.C = int(num_cols)
        self.N = int(num_cols if num_grids is None else num_grids)
        self.center = float(center)
        self.device = int(device)
        if stream is None:
            stream = 0
            try:
                import torch
                if torch.cuda.is_available():
                    stream = torch.cuda.current_stream(self.device).cuda_stream
            except Exception:  # pragma: no cover - torch is plumbing only
                stream = 0
        self.stream = stream
        g = TraceGeom(self.P, self.S, self.C, self.N, self.center,
                      theta.ctypes.data_as(ct.POINTER(ct.c_float)))
        h = ct.c_void_p()
        check(lib().trace_create(ct.byref(h), ct.byref(g), self.device,
                                 ct.c_void_p(stream)))
        self._h = h
        if not exact_sqrt:
            self.set_exact_sqrt(False)

    @contextlib.contextmanager
    def _ordered(self):
        """Orders a call against torch's current stream (see the class).
        A handle on the null stream is ordered by synchronising: the null
        stream the library launches on need not be the stream torch means by
        0, so events recorded through torch do not order it."""
        import torch
        cur = torch.cuda.current_stream(self.device)
        if cur.cuda_stream == self.stream:
            yield
            return
        if not self.stream:
            cur.synchronize()
            try:
                yield
            finally:
                torch.cuda.synchronize(self.device)
            return
        hs = getattr(self, "_hstream", None)
        if hs is None:
            hs = torch.cuda.ExternalStream(self.stream, device=self.device)
            self._hstream = hs
        hs.wait_stream(cur)
        try:
            yield
        finally:
            cur.wait_stream(hs)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().trace_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_exact_sqrt(self, on: bool):
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_EXACT_SQRT, 1 if on else 0))

    def set_backprojector(self, kind: str):
        """'tile' (LDS-tiled, default) or 'atomic' (whole-ray global atomics)."""
        v = {"tile": 0, "atomic": 1}[kind]
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_BACKPROJECTOR, v))

    def set_launch_shape(self, fwd_batch: int, tile_w: int, tile_h: int, tile_batch: int):
        check(lib().trace_set_launch_shape(self._h, fwd_batch, tile_w, tile_h, tile_batch))

    def set_fwd_bundle(self, angles: int):
        """Adjacent angles per forward-projector wavefront (1, 2, 4, 8, 16)."""
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_FWD_BUNDLE, int(angles)))
        self._fwd_bundle = int(angles)

    def set_fwd_lanes(self, lanes: int):
        """Lanes per ray of the forward projector (1, 2, 4); the slice batch
        must be compiled for it (2: 8/16/32, 4: 16/32)."""
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_FWD_LANES, int(lanes)))
        self._fwd_lanes = int(lanes)

    def set_ordered_sums(self, level: int):
        """Order of the backprojected float sums (TRACE_OPT_ORDERED_SUMS):
        0 all in fixed point; 1 (default) den and the PML/APMLR num in the
        reference's order; 2 also the SIRT/MLEM num."""
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_ORDERED_SUMS, int(level)))

    def set_ordered_tile(self, on: bool):
        """Tile (default) or per-pixel form of the ordered sums (same bits)."""
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_ORDERED_TILE, 1 if on else 0))

    def set_fwd_staged(self, on: bool):
        """Slab-staged (LDS, the default) or whole-ray form of the forward
        projector at 32-slice batches (TRACE_OPT_FWD_STAGED; same bits either
        way)."""
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_FWD_STAGED, 1 if on else 0))

    def get_option(self, option: int) -> int:
        v = ct.c_int()
        check(lib().trace_get_option(self._h, int(option), ct.byref(v)))
        return v.value

    def fwd_lanes(self) -> int:
        return self.get_option(_capi.TRACE_OPT_FWD_LANES)

    def fwd_bundle(self) -> int:
        return self.get_option(_capi.TRACE_OPT_FWD_BUNDLE)

    def launch_shape(self) -> tuple:
        """(forward slice batch, tile_w, tile_h, tile slice batch)."""
        v = [ct.c_int() for _ in range(4)]
        check(lib().trace_get_launch_shape(self._h, *[ct.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def segment_counts(self, out=None):
        """Segments per pixel of one slice, [N][N] (float, integer-valued)."""
        d = self._empty(self.N, self.N) if out is None else out
        with self._ordered():
            check(lib().trace_segment_counts(self._h, _dptr(d)))
        return d

    def set_phase_timing(self, on: bool):
        check(lib().trace_set_option(self._h, _capi.TRACE_OPT_PHASE_TIMING, 1 if on else 0))

    def phase_times(self, reset: bool = True) -> dict:
        """{phase: (milliseconds, launches)} of the kernels recorded by the
        phase timer (HIP events on this handle's stream)."""
        ms = (ct.c_double * 4)()
        n = (ct.c_int64 * 4)()
        check(lib().trace_phase_times(self._h, ms, n, 1 if reset else 0))
        return {name: (ms[i], n[i]) for i, name in enumerate(_capi.PHASES)}

    # -- tensors -------------------------------------------------------
    def _empty(self, *shape):
        import torch
        return torch.empty(shape, dtype=torch.float32, device=f"cuda:{self.device}")

    def forward(self, x, out=None):
        """y = A x; x [S][N][N] -> y [P][S][C]."""
        y = self._empty(self.P, self.S, self.C) if out is None else out
        with self._ordered():
            check(lib().trace_forward(self._h, _dptr(x), _dptr(y)))
        return y

    def adjoint(self, r, out=None):
        """z = A^T r; r [P][S][C] -> z [S][N][N]."""
        z = self._empty(self.S, self.N, self.N) if out is None else out
        with self._ordered():
            check(lib().trace_adjoint(self._h, _dptr(r), _dptr(z)))
        return z

    def column_sums(self, out=None):
        """A^T 1 for one slice, [N][N]."""
        d = self._empty(self.N, self.N) if out is None else out
        with self._ordered():
            check(lib().trace_column_sums(self._h, _dptr(d)))
        return d

    def ray_sums(self):
        """(sum len, sum len^2) per ray of one slice, each [P][C]."""
        a1 = self._empty(self.P, self.C)
        a2 = self._empty(self.P, self.C)
        with self._ordered():
            check(lib().trace_ray_sums(self._h, _dptr(a1), _dptr(a2)))
        return a1, a2

    def reconstruct(self, alg: str, iters: int, sino, recon,
                    params: RegParams | None = None):
        """Runs `iters` iterations in place on device tensors
        sino [P][S][C] and recon [S][N][N]."""
        a = ALGS[alg] if isinstance(alg, str) else int(alg)
        prm = (params or RegParams()).c()
        with self._ordered():
            check(lib().trace_reconstruct(self._h, a, int(iters), _dptr(sino),
                                          _dptr(recon), ct.byref(prm)))
        return recon

    def reconstruct_host(self, alg: str, iters: int, sino: np.ndarray,
                         recon: np.ndarray, params: RegParams | None = None):
        """Same on host arrays (copies over PCIe); recon is updated."""
        a = ALGS[alg] if isinstance(alg, str) else int(alg)
        s = np.ascontiguousarray(sino, dtype=np.float32)
        if recon.dtype != np.float32 or not recon.flags.c_contiguous:
            raise TypeError("recon must be a C-contiguous float32 array")
        prm = (params or RegParams()).c()
        check(lib().trace_reconstruct_host(self._h, a, int(iters), s.ctypes.data,
                                           recon.ctypes.data, ct.byref(prm)))
        return recon

    def dump_rays(self, proj, col, max_seg: int | None = None):
        """Segments of rays (proj[i], col[i]) as computed by the kernels."""
        import torch
        proj = np.ascontiguousarray(proj, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        n = proj.size
        ms = int(max_seg or (2 * self.N + 2))
        dev = f"cuda:{self.device}"
        dp = torch.from_numpy(proj).to(dev)
        dc = torch.from_numpy(col).to(dev)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        idx = torch.zeros((n, ms), dtype=torch.int32, device=dev)
        ln = torch.zeros((n, ms), dtype=torch.float32, device=dev)
        l2 = torch.zeros((n, ms), dtype=torch.float32, device=dev)
        check(lib().trace_dump_rays(self._h, n, dp.data_ptr(), dc.data_ptr(), ms,
                                    counts.data_ptr(), idx.data_ptr(),
                                    ln.data_ptr(), l2.data_ptr()))
        torch.cuda.synchronize(self.device)
        return counts.cpu().numpy(), idx.cpu().numpy(), ln.cpu().numpy(), l2.cpu().numpy()

    def count_segments(self):
        nnz = ct.c_int64(0)
        oor = ct.c_int64(0)
        check(lib().trace_count_segments(self._h, ct.byref(nnz), ct.byref(oor)))
        return int(nnz.value), int(oor.value)


# ---------------------------------------------------------------------------
# Application-level calls (the reference mains' reconstruction step)
# ---------------------------------------------------------------------------

def _run(alg, theta_rad, sino, init, iters, center, params, device):
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    if sino.ndim != 3:
        raise ValueError("sinogram must be [projections][slices][columns]")
    P, S, C = sino.shape
    if np.asarray(theta_rad).size != P:
        raise ValueError("theta length differs from the number of projections")
    recon = np.full((S, C, C), init, dtype=np.float32)
    with Projector(theta_rad, S, C, center, device=device) as pj:
        pj.reconstruct_host(alg, iters, sino, recon, params)
    return recon


def sirt(theta_deg, sino, iters: int, center: float = 0.0, device: int = 0):
    """sirt_main.cc:120-202: degrees -> radians, initial image 0."""
    return _run("sirt", degree_to_radian(theta_deg), sino, 0.0, iters, center,
                None, device)


def mlem(theta_deg, sino, iters: int, center: float = 0.0, device: int = 0):
    """mlem_main.cc:113-182: degrees -> radians, |sinogram|, initial 1."""
    return _run("mlem", degree_to_radian(theta_deg), absolute(sino), 1.0, iters,
                center, None, device)


def pml(theta, sino, iters: int, center: float = 0.0, beta: float = 10.0,
        device: int = 0):
    """pml_main.cc:120-193: theta used as given (the degree conversion is
    commented out at pml_main.cc:121), |sinogram|, initial 1."""
    return _run("pml", np.asarray(theta, np.float32), absolute(sino), 1.0, iters,
                center, RegParams(beta=beta), device)


def apmlr(theta_deg, sino, iters: int, center: float = 0.0, beta0: float = 10.0,
          beta1: float = 1.0, delta0: float = 1.0, delta1: float = 1.0,
          regw: float = 1.0, compat: bool = True, device: int = 0):
    """apmlr_main.cc:150-244 at one rank: degrees -> radians, |sinogram|,
    initial 1, constant-1 ghost slices."""
    return _run("apmlr", degree_to_radian(theta_deg), absolute(sino), 1.0, iters,
                center, RegParams(beta0, beta1, delta0, delta1, regw, compat),
                device)
