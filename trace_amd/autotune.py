"""Launch-shape autotuner for the forward projector and the tile
backprojector.

The forward projector's shape is its slice batch (slices gathered per
segment, 4/8/16); the backprojector's is its image tile and slice batch
(tile_w x tile_h pixels x tile_batch slices of LDS accumulators).  The two
kernels are independent, so each is tuned on its own: every candidate runs
one iteration of the real workload and the phase timer (HIP events on the
handle's stream) gives the kernel's time.  Results are cached per
(device name, N, P, S, algorithm).
"""
from __future__ import annotations

import json
import os

FWD_BATCHES = (4, 8, 16, 32)
TILES = ((16, 16, 8), (16, 16, 16), (16, 16, 4), (24, 24, 8), (32, 16, 8),
         (16, 32, 8), (32, 32, 4), (32, 32, 8), (8, 8, 16), (64, 64, 2))

_CACHE: dict = {}


def _key(pj, alg):
    import torch
    name = torch.cuda.get_device_name(pj.device)
    return f"{name}|N{pj.N}|P{pj.P}|S{pj.S}|{alg}"


def autotune(pj, sino, recon, alg: str = "mlem", params=None,
             fwd_batches=FWD_BATCHES, tiles=TILES, cache_file: str | None = None,
             verbose: bool = False) -> dict:
    """Times the candidates on (sino, a copy of recon) and sets the best
    shape on projector `pj`.  Returns {"shape": (fb, tw, th, tb),
    "forward_ms": {...}, "back_ms": {...}}."""
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
        pj.set_launch_shape(*res["shape"])
        return res
    work = recon.clone()
    fwd_ms, back_ms = {}, {}
    base_tile = tiles[0]
    pj.set_phase_timing(True)
    try:
        for fb in fwd_batches:
            pj.set_launch_shape(fb, *base_tile)
            work.copy_(recon)
            pj.reconstruct(alg, 1, sino, work, params)  # allocation, warm caches
            pj.phase_times(reset=True)
            work.copy_(recon)
            pj.reconstruct(alg, 1, sino, work, params)
            fwd_ms[fb] = pj.phase_times(reset=True)["forward"][0]
            if verbose:
                print(f"[autotune] forward batch {fb}: {fwd_ms[fb]:.2f} ms", flush=True)
        best_fb = min(fwd_ms, key=fwd_ms.get)
        for t in tiles:
            pj.set_launch_shape(best_fb, *t)
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
    pj.set_launch_shape(*shape)
    res = {"shape": list(shape),
           "forward_ms": {str(k): v for k, v in fwd_ms.items()},
           "back_ms": {"x".join(map(str, k)): v for k, v in back_ms.items()}}
    _CACHE[key] = res
    if cache_file:
        try:
            with open(cache_file, "w") as f:
                json.dump(_CACHE, f, indent=1)
        except OSError:
            pass
    return res
