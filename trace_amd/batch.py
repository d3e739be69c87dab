"""Batched reconstruction of many slices with different geometries.

`Batch` wraps the C-ABI's trace_batch_*: every slice has its own detector
width C (image C x C), angles, centre, algorithm (SIRT or MLEM) and
iteration count; one `run()` is one HIP-graph launch whose persistent
kernels work through all slices together.  Each slice's result equals
`Projector(theta, 1, C, center).reconstruct(alg, iters, ...)` for that
slice alone.

`uneven_slices` makes the uneven workload: widths drawn from
{256, 512, 1024, 1536, 2048}, angle counts from [180, 1800], SIRT or MLEM
with equal probability, seeded ellipse phantoms.  `autotune_batch` picks the
ray-bundle width and backprojector tile on a sub-batch.
"""
from __future__ import annotations

import ctypes as ct
import time
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import TraceBatchSlice, check, lib
from .recon import _dptr

BATCH_TILES = ((16, 16), (32, 32), (64, 64))
BATCH_BUNDLES = (1, 2, 4, 8, 16)


@dataclass
class BatchSlice:
    theta: np.ndarray   # [P] radians
    sino: object        # torch float32 [P][C] on the GPU
    recon: object       # torch float32 [N][N] on the GPU (initial image in, result out);
                        # N = C in the reference's programs
    alg: str = "mlem"   # "sirt" or "mlem"
    iters: int = 25
    center: float = 0.0  # <= 0: C/2 + 1 (the reference's default)


class Batch:
    def __init__(self, slices: list[BatchSlice], device: int = 0, stream=None,
                 exact_sqrt: bool = True):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        self.device = int(device)
        self.slices = list(slices)
        self._theta = [np.ascontiguousarray(s.theta, dtype=np.float32) for s in self.slices]
        arr = (TraceBatchSlice * len(self.slices))()
        for i, s in enumerate(self.slices):
            P, C = s.sino.shape
            N = int(s.recon.shape[0])
            if s.recon.dim() != 2 or tuple(s.recon.shape) != (N, N):
                raise ValueError(f"slice {i}: recon must be a square [N][N] image")
            if self._theta[i].size != P:
                raise ValueError(f"slice {i}: theta length differs from the projections")
            arr[i].num_projs = P
            arr[i].num_cols = C
            arr[i].num_grids = N
            arr[i].center = float(s.center)
            arr[i].theta = self._theta[i].ctypes.data_as(ct.POINTER(ct.c_float))
            arr[i].alg = _capi.ALGS[s.alg] if isinstance(s.alg, str) else int(s.alg)
            arr[i].iters = int(s.iters)
            arr[i].sino = _dptr(s.sino)
            arr[i].recon = _dptr(s.recon)
        h = ct.c_void_p()
        check(lib().trace_batch_create(ct.byref(h), arr, len(self.slices), self.device,
                                       ct.c_void_p(stream)))
        self._h = h
        if not exact_sqrt:
            check(lib().trace_batch_set_option(self._h, _capi.TRACE_OPT_EXACT_SQRT, 0))

    def set_launch_shape(self, bundle_angles: int, tile_w: int, tile_h: int):
        check(lib().trace_batch_set_launch_shape(self._h, bundle_angles, tile_w, tile_h))

    def launch_shape(self) -> tuple:
        v = [ct.c_int() for _ in range(3)]
        check(lib().trace_batch_get_launch_shape(self._h, *[ct.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def prepare(self, max_iters: int = 0):
        """Captures and instantiates the graph of run(max_iters) now."""
        check(lib().trace_batch_prepare(self._h, int(max_iters)))

    def run(self, max_iters: int = 0):
        """All slices' iterations (at most max_iters each if > 0): one graph
        launch, asynchronous on the stream."""
        check(lib().trace_batch_run(self._h, int(max_iters)))

    def close(self):
        if getattr(self, "_h", None):
            lib().trace_batch_destroy(self._h)
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


def uneven_specs(n: int, seed: int = 3, widths=(256, 512, 1024, 1536, 2048),
                 angles=(180, 1800)):
    """[(C, P, alg)] for n slices: C uniform over `widths`, P uniform over
    [angles[0], angles[1]], SIRT or MLEM with equal probability."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = rng.choice(np.asarray(widths), n)
    P = rng.integers(angles[0], angles[1] + 1, n)
    alg = np.where(rng.random(n) < 0.5, "sirt", "mlem")
    return [(int(c), int(p), str(a)) for c, p, a in zip(C, P, alg)]


def uneven_slices(specs, seed: int = 3, iters: int = 25, I0: float = 1e4, K: int = 20,
                  device=0) -> list[BatchSlice]:
    """Slice i: seeded ellipse phantom (sub-seed seed*1000 + i) of width C,
    P angles evenly spaced in [0, pi), Poisson noise at I0, centre C/2; the
    initial image is the reference main's (0 for SIRT, 1 for MLEM)."""
    import torch
    from .phantom import angles_rad, noisy_sinogram
    dev = torch.device(f"cuda:{device}")
    out = []
    for i, (C, P, alg) in enumerate(specs):
        th = angles_rad(P)
        sino = noisy_sinogram(th, C, 1, seed=seed, I0=I0, K=K, device=dev,
                              first_slice=i)[:, 0, :].contiguous()
        x0 = torch.full((C, C), 0.0 if alg == "sirt" else 1.0, dtype=torch.float32, device=dev)
        out.append(BatchSlice(th, sino, x0, alg, iters, C / 2.0))
    return out


def autotune_batch(slices: list[BatchSlice], bundles=BATCH_BUNDLES, tiles=BATCH_TILES,
                   sample: int = 40, iters: int = 2, seed: int = 0, verbose: bool = False):
    """Times every (bundle, tile) on a sub-batch of up to `sample` slices
    (a seeded random subset; copies of their images) for `iters`
    iterations and returns {"bundle", "tile", "ms": {...}}."""
    import torch
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(slices), min(sample, len(slices)), replace=False)
    sub = [BatchSlice(slices[i].theta, slices[i].sino, slices[i].recon.clone(),
                      slices[i].alg, slices[i].iters, slices[i].center) for i in idx]
    init = [s.recon.clone() for s in sub]
    ms = {}
    with Batch(sub, device=sub[0].recon.device.index) as b:
        for ba in bundles:
            for tw, th in tiles:
                b.set_launch_shape(ba, tw, th)
                b.run(iters)  # graph capture + warm caches
                torch.cuda.synchronize()
                for s, x0 in zip(sub, init):
                    s.recon.copy_(x0)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                b.run(iters)
                torch.cuda.synchronize()
                ms[(ba, tw, th)] = (time.perf_counter() - t0) * 1e3
                if verbose:
                    print(f"[autotune_batch] bundle {ba} tile {tw}x{th}: "
                          f"{ms[(ba, tw, th)]:.1f} ms", flush=True)
    best = min(ms, key=ms.get)
    return {"bundle": best[0], "tile": (best[1], best[2]),
            "ms": {f"{k[0]}/{k[1]}x{k[2]}": v for k, v in ms.items()},
            "sample": [int(i) for i in idx]}
