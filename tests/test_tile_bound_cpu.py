"""The tile backprojector's host bound on the columns per tile and angle
(tile_count_bound, kernels.hip) against the device's column range
(tile_columns, traverse.h), restated here in float32 with the device's
operation order: the range must never hold more columns than the bound,
or the outermost column's segments would be dropped without an error.

Checked at the size limits (N = C = 32768) and smaller sides, for every
compiled tile shape, on the edge tiles and a seeded sample of the others
(the float rounding of the projections grows with |t|, so the image's far
corners matter most), at adversarial angles (axis-aligned and diagonal float
neighbours, angles outside [0, pi)) and a sweep, with the extra strip column
of the wrap pass (one more slot)."""
import os
import re

import numpy as np
import pytest

from conftest import adversarial_angles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "trace_amd", "csrc")


def _margins():
    t = open(os.path.join(CSRC, "traverse.h")).read()
    col = float(re.search(r"kColumnMargin = ([0-9.]+)f;", t).group(1))
    cnt = float(re.search(r"kCountMargin = ([0-9.]+);", t).group(1))
    return np.float32(col), cnt


def _tile_shapes():
    t = open(os.path.join(CSRC, "internal.h")).read()
    body = t[t.index("#define TRACE_BP_SHAPES"):]
    body = body[:body.index("\n\n")]
    shapes = {(int(a), int(b)) for a, b, _ in re.findall(r"X\((\d+), (\d+), (\d+)\)", body)}
    shapes |= {(16, 16), (32, 32), (64, 64)}  # batch tiles
    return sorted(shapes)


def count_bound(sinq, cosq, TW, TH, cnt_margin):
    """tile_count_bound: floor(TW|sin| + TH|cos| + 2 margin) + 1, in double."""
    ln = TW * abs(float(sinq)) + TH * abs(float(cosq)) + 2.0 * cnt_margin
    return int(np.floor(ln)) + 1


def device_counts(sinq, cosq, x0, x1, y0, y1, yi0, margin):
    """cmax - cmin + 1 of tile_columns for arrays of tile corners (float32,
    no contraction: t = (-x * s) + (y * c))."""
    f = np.float32
    s, c = f(sinq), f(cosq)
    t00 = (-x0) * s + y0 * c
    t01 = (-x0) * s + y1 * c
    t10 = (-x1) * s + y0 * c
    t11 = (-x1) * s + y1 * c
    tmin = np.minimum(np.minimum(t00, t01), np.minimum(t10, t11))
    tmax = np.maximum(np.maximum(t00, t01), np.maximum(t10, t11))
    cmin = np.ceil((tmin - margin) - yi0).astype(np.int64)
    cmax = np.floor((tmax + margin) - yi0).astype(np.int64)
    return cmax - cmin + 1


def _angles():
    kpi = np.float32(3.14159265358979)
    sweep = (np.arange(360, dtype=np.float64) * np.pi / 360).astype(np.float32)
    near45 = kpi / np.float32(4) + np.arange(-50, 51, dtype=np.float32) * np.float32(1e-7)
    return np.concatenate([adversarial_angles(), sweep, near45])


@pytest.mark.parametrize("N", [32768, 20001, 4096, 257])
def test_device_column_range_within_host_bound(N):
    margin, cnt_margin = _margins()
    from trace_amd.phantom import mov_of
    C = N
    mov = np.float32(mov_of(C, N / 2.0))
    yi0 = np.float32(np.float64(1 - C) / 2.0 + np.float64(mov))
    ang = _angles()
    sinq = np.sin(ang.astype(np.float64)).astype(np.float32)
    cosq = np.cos(ang.astype(np.float64)).astype(np.float32)
    gl = ((2 * np.arange(-1, N + 1, dtype=np.int64) - N).astype(np.float32) * np.float32(0.5))
    worst = -10
    rng = np.random.default_rng(N)
    for TW, TH in _tile_shapes():
        nx, ny = -(-N // TW), -(-N // TH)
        # edge tiles and a sample of the rest (pairs (tx, ty) as rows)
        ix = np.concatenate([np.arange(nx), np.arange(nx), np.zeros(ny, np.int64),
                             np.full(ny, nx - 1), rng.integers(0, nx, 4096)])
        iy = np.concatenate([np.zeros(nx, np.int64), np.full(nx, ny - 1), np.arange(ny),
                             np.arange(ny), rng.integers(0, ny, 4096)])
        tx0 = ix * TW
        ty0 = iy * TH
        # the wrap pass walks the strip [N-1, N] x [ty0-1, ty1]: one slot more
        regions = [(tx0, np.minimum(tx0 + TW, N), ty0, np.minimum(ty0 + TH, N), 0),
                   (np.full_like(ty0, N - 1), np.full_like(ty0, N), ty0 - 1,
                    np.minimum(ty0 + TH, N), 1)]
        for ax0, ax1, ay0, ay1, extra in regions:
            X0, X1, Y0, Y1 = gl[ax0 + 1], gl[ax1 + 1], gl[ay0 + 1], gl[ay1 + 1]
            for s, c in zip(sinq, cosq):
                got = int(device_counts(s, c, X0, X1, Y0, Y1, yi0, margin).max())
                bound = count_bound(s, c, TW, TH, cnt_margin) + extra
                worst = max(worst, got - bound)
                assert got <= bound, (N, TW, TH, float(s), float(c), got, bound)
    print(f"N={N}: largest range minus bound {worst}")


def test_a_wider_device_margin_would_break_the_bound():
    """The check has teeth: a device margin of 0.55 pixel (more than the
    host's 0.5) exceeds the bound on the same tiles and angles."""
    _, cnt_margin = _margins()
    N = 4096
    from trace_amd.phantom import mov_of
    mov = np.float32(mov_of(N, N / 2.0))
    yi0 = np.float32(np.float64(1 - N) / 2.0 + np.float64(mov))
    gl = ((2 * np.arange(0, N + 1, dtype=np.int64) - N).astype(np.float32) * np.float32(0.5))
    t = np.arange(0, N, 16)
    X0, X1 = gl[t][:, None], gl[np.minimum(t + 16, N)][:, None]
    Y0, Y1 = gl[t][None, :], gl[np.minimum(t + 16, N)][None, :]
    over = 0
    for a in _angles()[::7]:
        s, c = np.float32(np.sin(np.float64(a))), np.float32(np.cos(np.float64(a)))
        got = device_counts(s, c, X0, X1, Y0, Y1, yi0, np.float32(0.55))
        over += int((got > count_bound(s, c, 16, 16, cnt_margin)).sum())
    assert over > 0
