"""Launch-shape autotuner for the forward projector and the tile
backprojector.

The forward projector's shape is its slice batch (slices gathered per
segment, 1 to 32) and ray-bundle width (at 32-slice batches its slab-staged
form, the default there, has a single shape); the backprojector's is its image tile and slice batch
(tile_w x tile_h pixels x tile_batch slices of LDS accumulators).  The two
kernels are independent, so each is tuned on its own: every candidate runs
one iteration of the real workload and the phase timer (HIP events on the
handle's stream) gives the kernel's time.  Results are cached per
(device name, N, P, S, algorithm).
"""
from __future__ import annotations

import json
import os

from . import _capi

FWD_BATCHES = (1, 2, 4, 8, 16, 32)
BUNDLES = (1, 2, 4, 8, 16)
TILES = ((16, 16, 8), (16, 16, 16), (16, 16, 4), (24, 24, 8), (32, 16, 8),
         (16, 32, 8), (32, 32, 4), (32, 32, 8), (8, 8, 16), (64, 64, 2),
         (32, 32, 1), (64, 64, 1), (32, 32, 2), (16, 16, 2), (16, 16, 1),
         (16, 16, 32), (32, 32, 16), (32, 16, 16), (24, 24, 16))


def _pow2_ceil(n: int) -> int:
    p = 1
    while p < n:
        p *= 2
    return p


def batch_candidates(batches, S: int):
    """Slice batches worth timing for a stack of S slices: no more than one
    batch of padding (<= max(4, S rounded up to a power of two)) and no more
    than 8 launches' worth of repeated traversal (>= S / 8)."""
    hi = max(4, _pow2_ceil(S))
    c = [b for b in batches if b <= hi and 8 * b >= S]
    return c or [max(batches)]

_CACHE: dict = {}


# (slice batch, lanes per ray) pairs compiled into the forward projector
FWD_SHAPES = {(1, 1), (2, 1), (4, 1), (8, 1), (16, 1), (32, 1), (8, 2), (16, 2), (32, 2),
              (16, 4), (32, 4)}
LANES = (1, 2, 4)


def apply_shape(pj, res: dict):
    """Sets {"shape": (fwd_batch, tile_w, tile_h, tile_batch), "bundle",
    "lanes"} on projector `pj`."""
    pj.set_fwd_bundle(1)
    pj.set_launch_shape(*res["shape"])
    pj.set_fwd_lanes(res.get("lanes", 1))
    pj.set_fwd_bundle(res["bundle"])


def _key(pj, alg):
    import torch
    name = torch.cuda.get_device_name(pj.device)
    return f"{name}|N{pj.N}|P{pj.P}|S{pj.S}|{alg}"


def autotune(pj, sino, recon, alg: str = "mlem", params=None,
             fwd_batches=FWD_BATCHES, bundles=BUNDLES, tiles=TILES,
             cache_file: str | None = None, verbose: bool = False,
             lanes_per_ray=LANES) -> dict:
    """Times the candidates on (sino, a copy of recon) and sets the best
    shape on projector `pj`.  Returns {"shape": (fb, tw, th, tb),
    "bundle": angles, "lanes": lanes per ray, "forward_ms": {...},
    "back_ms": {...}}."""
    import torch
    key = _key(pj, alg)
    if cache_file and not _CACHE and os.path.exists(cache_file):
        try:
            with open(cache_file) as f:
                _CACHE.update(json.load(f))
        except (OSError, ValueError):
            pass
    if key in _CACHE:
        res = _CACHE[key]
        apply_shape(pj, res)
        return res
    work = recon.clone()
    fwd_ms, back_ms = {}, {}
    tile_batches = batch_candidates(sorted({t[2] for t in tiles}), pj.S)
    base_tile = next(t for t in tiles if t[2] in tile_batches)
    staged = pj.get_option(_capi.TRACE_OPT_FWD_STAGED) == 1
    staged_pair = next(((ln, ba) for ln in lanes_per_ray for ba in bundles
                        if (32, ln) in FWD_SHAPES and ba * ln <= 64), None)
    pj.set_phase_timing(True)
    try:
        for fb in batch_candidates(fwd_batches, pj.S):
            for lanes in lanes_per_ray:
                if (fb, lanes) not in FWD_SHAPES:
                    continue
                for ba in bundles:
                    if ba * lanes > 64:
                        continue
                    # the staged form (32-slice batches) has one launch shape
                    if fb == 32 and staged and (lanes, ba) != staged_pair:
                        continue
                    apply_shape(pj, {"shape": (fb, *base_tile), "bundle": ba, "lanes": lanes})
                    work.copy_(recon)
                    pj.reconstruct(alg, 1, sino, work, params)  # allocation, warm caches
                    pj.phase_times(reset=True)
                    work.copy_(recon)
                    pj.reconstruct(alg, 1, sino, work, params)
                    fwd_ms[(fb, lanes, ba)] = pj.phase_times(reset=True)["forward"][0]
                    if verbose:
                        print(f"[autotune] forward batch {fb} lanes {lanes} bundle {ba}: "
                              f"{fwd_ms[(fb, lanes, ba)]:.2f} ms", flush=True)
        best_fb, best_lanes, best_ba = min(fwd_ms, key=fwd_ms.get)
        apply_shape(pj, {"shape": (best_fb, *base_tile), "bundle": best_ba, "lanes": best_lanes})
        for t in tiles:
            if t[2] not in tile_batches:
                continue
            apply_shape(pj, {"shape": (best_fb, *t), "bundle": best_ba, "lanes": best_lanes})
            work.copy_(recon)
            pj.reconstruct(alg, 1, sino, work, params)
            back_ms[t] = pj.phase_times(reset=True)["back"][0]
            if verbose:
                print(f"[autotune] tile {t}: {back_ms[t]:.2f} ms", flush=True)
        best_tile = min(back_ms, key=back_ms.get)
    finally:
        pj.set_phase_timing(False)
        torch.cuda.synchronize(pj.device)
    shape = (best_fb, *best_tile)
    res = {"shape": list(shape), "bundle": best_ba, "lanes": best_lanes,
           "forward_ms": {f"{k[0]}x{k[1]}x{k[2]}": v for k, v in fwd_ms.items()},
           "back_ms": {"x".join(map(str, k)): v for k, v in back_ms.items()}}
    apply_shape(pj, res)
    _CACHE[key] = res
    if cache_file:
        try:
            with open(cache_file, "w") as f:
                json.dump(_CACHE, f, indent=1)
        except OSError:
            pass
    return res
