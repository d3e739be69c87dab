"""ctypes bindings for the oracle libraries (test infrastructure only)."""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtrace_ref.so")

SIRT, MLEM, PML, APMLR = 0, 1, 2, 3
ALG_NAMES = {"sirt": SIRT, "mlem": MLEM, "pml": PML, "apmlr": APMLR}

_fp = ct.POINTER(ct.c_float)
_ip = ct.POINTER(ct.c_int)
_dp = ct.POINTER(ct.c_double)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, kind=_fp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class OrParams(ct.Structure):
    _fields_ = [
        ("beta", ct.c_float),
        ("beta1", ct.c_float),
        ("delta", ct.c_float),
        ("delta1", ct.c_float),
        ("regw", ct.c_float),
        ("apmlr_compat", ct.c_int),
    ]


@dataclass
class Params:
    """Regularisation parameters (defaults: pml_main.cc:48-49,
    apmlr_main.cc:22-26)."""

    beta: float = 10.0
    beta1: float = 1.0
    delta: float = 1.0
    delta1: float = 1.0
    regw: float = 1.0
    apmlr_compat: int = 1

    def c(self) -> OrParams:
        return OrParams(self.beta, self.beta1, self.delta, self.delta1,
                        self.regw, int(self.apmlr_compat))


class Oracle:
    """The C restatement (oracle/trace_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        L = ct.CDLL(path)
        L.or_deg2rad.argtypes = [_fp, _fp, ct.c_int]
        L.or_absolute.argtypes = [_fp, ct.c_int64]
        L.or_quadrant.argtypes = [ct.c_float]
        L.or_quadrant.restype = ct.c_int
        L.or_mov.argtypes = [ct.c_int, ct.c_float]
        L.or_mov.restype = ct.c_float
        L.or_yi.argtypes = [ct.c_int, ct.c_int, ct.c_float]
        L.or_yi.restype = ct.c_float
        L.or_trace_ray.argtypes = [ct.c_int, ct.c_int, ct.c_float, ct.c_float,
                                   ct.c_int, _ip, _fp, _fp]
        L.or_trace_ray.restype = ct.c_int
        L.or_trace_rays.argtypes = [ct.c_int, ct.c_int, ct.c_float, ct.c_int, _fp,
                                    _ip, ct.POINTER(ct.c_int64), _ip, _fp, ct.c_int64]
        L.or_trace_rays.restype = ct.c_int64
        L.or_ray_points.argtypes = [ct.c_int, ct.c_int, ct.c_float, ct.c_float,
                                    ct.c_int, _fp, _fp]
        L.or_ray_points.restype = ct.c_int
        L.or_forward.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, _fp,
                                 _fp, _fp]
        L.or_forward.restype = ct.c_int
        L.or_adjoint.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, _fp,
                                 _fp, _dp]
        L.or_adjoint.restype = ct.c_int
        L.or_count_segments.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_float, _fp,
                                        ct.POINTER(ct.c_int64)]
        L.or_count_segments.restype = ct.c_int64
        L.or_reconstruct.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                     _fp, ct.c_float, ct.c_int, _fp, _fp,
                                     ct.POINTER(OrParams), _fp, _fp, _fp, _fp]
        L.or_reconstruct.restype = ct.c_int
        L.or_reconstruct_f64.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                         _fp, ct.c_float, ct.c_int, _fp, _dp,
                                         ct.POINTER(OrParams)]
        L.or_reconstruct_f64.restype = ct.c_int
        self.L = L

    # -- preprocessing -------------------------------------------------
    def deg2rad(self, deg):
        deg = _f32(deg)
        out = np.empty_like(deg)
        self.L.or_deg2rad(_ptr(deg), _ptr(out), deg.size)
        return out

    def absolute(self, a):
        a = _f32(a).copy()
        self.L.or_absolute(_ptr(a), a.size)
        return a

    def quadrant(self, q: float) -> int:
        return self.L.or_quadrant(float(q))

    def mov(self, C: int, center: float) -> float:
        return self.L.or_mov(C, center)

    # -- traversal -----------------------------------------------------
    def trace_ray(self, N, C, center, theta, col):
        idx = np.zeros(2 * N + 2, np.int32)
        ln = np.zeros(2 * N + 2, np.float32)
        l2 = np.zeros(2 * N + 2, np.float32)
        n = self.L.or_trace_ray(N, C, center, float(theta), col, _ptr(idx, _ip),
                                _ptr(ln), _ptr(l2))
        return idx[:n].copy(), ln[:n].copy(), l2[:n].copy()

    def trace_rays(self, N, C, center, theta, col):
        """Rays (theta[i], col[i]): (offsets [n+1], idx, len) with ray i's
        segments at [offsets[i], offsets[i+1])."""
        th = _f32(np.asarray(theta).ravel())
        cl = np.ascontiguousarray(np.asarray(col).ravel(), dtype=np.int32)
        n = th.size
        cap = int(n) * (2 * N + 2)
        off = np.zeros(n + 1, np.int64)
        idx = np.zeros(cap, np.int32)
        ln = np.zeros(cap, np.float32)
        tot = self.L.or_trace_rays(N, C, center, n, _ptr(th), _ptr(cl, _ip),
                                   off.ctypes.data_as(ct.POINTER(ct.c_int64)),
                                   _ptr(idx, _ip), _ptr(ln), cap)
        assert tot >= 0
        return off, idx[:tot].copy(), ln[:tot].copy()

    def ray_points(self, N, C, center, theta, col):
        px = np.zeros(2 * N + 2, np.float32)
        py = np.zeros(2 * N + 2, np.float32)
        n = self.L.or_ray_points(N, C, center, float(theta), col, _ptr(px), _ptr(py))
        return px[:n].copy(), py[:n].copy()

    def count_segments(self, P, C, center, theta, N=None):
        th = _f32(theta)
        oor = ct.c_int64(0)
        n = self.L.or_count_segments(P, C, C if N is None else N, center, _ptr(th),
                                     ct.byref(oor))
        return int(n), int(oor.value)

    # -- operators -----------------------------------------------------
    def forward(self, theta, x, center, C=None):
        """y [P][S][C] = A x for x [S][N][N]; C detector columns (default N)."""
        x = _f32(x)
        S, N, _ = x.shape
        C = N if C is None else int(C)
        th = _f32(theta)
        P = th.size
        y = np.zeros((P, S, C), np.float32)
        rc = self.L.or_forward(P, S, C, N, center, _ptr(th), _ptr(x), _ptr(y))
        assert rc == 0
        return y

    def adjoint(self, theta, r, center, N=None):
        """z [S][N][N] = A^T r for r [P][S][C]; N image side (default C)."""
        r = _f32(r)
        P, S, C = r.shape
        N = C if N is None else int(N)
        th = _f32(theta)
        z = np.zeros((S, N, N), np.float64)
        rc = self.L.or_adjoint(P, S, C, N, center, _ptr(th), _ptr(r), _ptr(z, _dp))
        assert rc == 0
        return z

    # -- reconstruction --------------------------------------------------
    def reconstruct(self, alg, theta_rad, sino, recon0, center, iters,
                    params: Params | None = None, want_stages=False):
        """One-thread restatement of the reference loop.  theta in radians,
        sino [P][S][C] already preprocessed, recon0 [S][N][N] initial image."""
        alg = ALG_NAMES.get(alg, alg)
        sino = _f32(sino)
        P, S, C = sino.shape
        th = _f32(theta_rad)
        x = _f32(recon0).copy()
        N = x.shape[-1]
        prm = (params or Params()).c()
        num = den = F = G = None
        if want_stages:
            num = np.zeros((S, N, N), np.float32)
            den = np.zeros_like(num)
            if alg in (PML, APMLR):
                F = np.zeros_like(num)
                G = np.zeros_like(num)
        rc = self.L.or_reconstruct(alg, P, S, C, N, _ptr(th), center, iters,
                                   _ptr(sino), _ptr(x), ct.byref(prm), _ptr(num),
                                   _ptr(den), _ptr(F), _ptr(G))
        assert rc == 0
        if want_stages:
            return x, dict(num=num, den=den, F=F, G=G)
        return x

    def reconstruct_f64(self, alg, theta_rad, sino, recon0, center, iters,
                        params: Params | None = None):
        alg = ALG_NAMES.get(alg, alg)
        sino = _f32(sino)
        P, S, C = sino.shape
        th = _f32(theta_rad)
        x = np.ascontiguousarray(recon0, dtype=np.float64).copy()
        N = x.shape[-1]
        prm = (params or Params()).c()
        rc = self.L.or_reconstruct_f64(alg, P, S, C, N, _ptr(th), center, iters,
                                       _ptr(sino), _ptr(x, _dp), ct.byref(prm))
        assert rc == 0
        return x


class RefLib:
    """The compiled reference (oracle/_ref/libtrace_ref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = ct.CDLL(path)
        L.ref_trace_ray.argtypes = [ct.c_int, ct.c_float, ct.c_float, ct.c_int,
                                    _ip, _fp, _fp]
        L.ref_trace_ray.restype = ct.c_int
        L.ref_trace_ray_n.argtypes = [ct.c_int, ct.c_int, ct.c_float, ct.c_float, ct.c_int,
                                      _ip, _fp, _fp]
        L.ref_trace_ray_n.restype = ct.c_int
        L.ref_forward.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_float, _fp,
                                  _fp, _fp]
        L.ref_forward.restype = ct.c_int
        L.ref_deg2rad.argtypes = [_fp, _fp, ct.c_int]
        L.ref_quadrant.argtypes = [ct.c_float]
        L.ref_quadrant.restype = ct.c_int
        L.ref_mov.argtypes = [ct.c_int, ct.c_float]
        L.ref_mov.restype = ct.c_float
        L.ref_compiler.argtypes = []
        L.ref_compiler.restype = ct.c_char_p
        common = [ct.c_int, ct.c_int, ct.c_int, _fp, ct.c_float, ct.c_int,
                  ct.c_int, _fp, _fp]
        L.ref_sirt.argtypes = common + [_fp, _fp]
        L.ref_mlem.argtypes = common + [_fp, _fp]
        L.ref_pml.argtypes = common + [ct.c_float, _fp, _fp, _fp, _fp]
        L.ref_apmlr.argtypes = common + [ct.c_float] * 5 + [_fp, _fp, _fp, _fp]
        for f in (L.ref_sirt, L.ref_mlem, L.ref_pml, L.ref_apmlr):
            f.restype = ct.c_int
        self.L = L

    def trace_ray(self, C, center, theta, col):
        idx = np.zeros(2 * C + 2, np.int32)
        ln = np.zeros(2 * C + 2, np.float32)
        l2 = np.zeros(2 * C + 2, np.float32)
        n = self.L.ref_trace_ray(C, center, float(theta), col, _ptr(idx, _ip),
                                 _ptr(ln), _ptr(l2))
        return idx[:n].copy(), ln[:n].copy(), l2[:n].copy()

    def trace_ray_n(self, N, C, center, theta, col):
        """A ray of C detector columns on an N x N grid."""
        idx = np.zeros(2 * N + 2, np.int32)
        ln = np.zeros(2 * N + 2, np.float32)
        l2 = np.zeros(2 * N + 2, np.float32)
        n = self.L.ref_trace_ray_n(N, C, center, float(theta), col, _ptr(idx, _ip),
                                   _ptr(ln), _ptr(l2))
        return idx[:n].copy(), ln[:n].copy(), l2[:n].copy()

    def forward(self, theta, x, center):
        x = _f32(x)
        S, N, _ = x.shape
        th = _f32(theta)
        P = th.size
        y = np.zeros((P, S, N), np.float32)
        self.L.ref_forward(P, S, N, center, _ptr(th), _ptr(x), _ptr(y))
        return y

    def deg2rad(self, deg):
        deg = _f32(deg)
        out = np.empty_like(deg)
        self.L.ref_deg2rad(_ptr(deg), _ptr(out), deg.size)
        return out

    def quadrant(self, q):
        return self.L.ref_quadrant(float(q))

    def compiler(self) -> str:
        return self.L.ref_compiler().decode()

    def mov(self, C, center):
        return self.L.ref_mov(C, center)

    def run(self, alg, theta_file, sino, center, iters, threads=1,
            params: Params | None = None, want_stages=False):
        """Run a reference app loop.  theta_file/sino are the values as the
        app reads them from HDF5 (degrees except for PML); the app applies its
        own preprocessing.  Returns the recon [S][N][N] (and stages)."""
        alg = ALG_NAMES.get(alg, alg)
        sino = _f32(sino)
        P, S, C = sino.shape
        th = _f32(theta_file)
        out = np.zeros((S, C, C), np.float32)
        prm = params or Params()
        st = {}
        if want_stages:
            st = dict(num=np.zeros_like(out), den=np.zeros_like(out))
            if alg in (PML, APMLR):
                st.update(F=np.zeros_like(out), G=np.zeros_like(out))
        g = lambda k: _ptr(st.get(k))  # noqa: E731
        args = (P, S, C, _ptr(th), center, iters, threads, _ptr(sino), _ptr(out))
        if alg == SIRT:
            rc = self.L.ref_sirt(*args, g("num"), g("den"))
        elif alg == MLEM:
            rc = self.L.ref_mlem(*args, g("num"), g("den"))
        elif alg == PML:
            rc = self.L.ref_pml(*args, prm.beta, g("num"), g("den"), g("F"), g("G"))
        else:
            rc = self.L.ref_apmlr(*args, prm.beta, prm.beta1, prm.delta,
                                  prm.delta1, prm.regw, g("num"), g("den"),
                                  g("F"), g("G"))
        assert rc == 0
        if want_stages:
            return out, st
        return out


_oracle = None
_ref = None


def load_oracle() -> Oracle:
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref() -> RefLib:
    global _ref
    if _ref is None:
        _ref = RefLib()
    return _ref
