// traverse.h -- on-the-fly ray traversal for gfx950, bit-compatible with the
// reference's per-ray pipeline
//   CalculateCoordinates  trace_utils.cc:48-80
//   MergeTrimCoordinates  trace_utils.cc:82-111
//   SortIntersectionPoints trace_utils.cc:113-158
//   CalculateDistanceLengths trace_utils.cc:160-187
// without materialising any per-ray array.
//
// The reference evaluates all 2(N+1) grid-line crossings, keeps those strictly
// inside the grid square, and merges the two lists by x.  Here:
//
// * the crossing with grid line n is recomputed from the same global formula
//   (islope*(gy[n]-srcy)+srcx, slope*(gx[n]-srcx)+srcy) whenever it is
//   needed, with FMA contraction off, so every coordinate is bit-identical;
// * each crossing sequence is monotone in n (every step is a correctly
//   rounded, monotone function of the previous value), so the trimmed lists
//   are contiguous index ranges, found by binary search in O(log N) instead
//   of the reference's O(N) scan;
// * the merge runs as a two-pointer walk that emits one segment per step.
//
// The pixel index uses the reference's truncating conversion, so an index
// can fall on N (a midpoint rounding onto the far grid line); callers guard
// the image bounds.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace trace_hip {

// Per-angle constants computed on the host with the host libm, exactly as
// Reduce does per projection row (sirt.cc:117-120).
struct AngleInfo {
  float sinq;
  float cosq;
  int quadrant;  // trace_utils.cc:41-46
  int pad;
};

struct GridInfo {
  int N;         // grid side (= C)
  int C;         // detector columns
  float gx0;     // gridx[0]
  float gxN;     // gridx[N]
  float mg;      // (float)(N / 2.0), trace_utils.cc:220
};

// gridx[n] = gridy[n] = (float)(-N/2.0 + n) (trace_data.h:103-108), written
// so that it is exact for both parities of N.
__device__ __forceinline__ float grid_line(int N, int n) {
  return (float)(2 * n - N) * 0.5f;
}

template <bool kExactSqrt>
__device__ __forceinline__ float seg_sqrt(float v) {
  if constexpr (kExactSqrt) {
    return __builtin_sqrtf(v);  // correctly rounded (IEEE) expansion
  } else {
    return __builtin_amdgcn_sqrtf(v);  // v_sqrt_f32, <= 1 ulp
  }
}

// Ray set-up (trace_utils.cc:63-75 with xi = -1e6, yi from sirt.cc:129).
struct RaySetup {
  float srcx, srcy, slope, islope;
};

__device__ __forceinline__ RaySetup ray_setup(float yi, float sinq, float cosq) {
  const float xi = -1e6f;
  RaySetup r;
  float xc = xi * cosq;
  float ys = yi * sinq;
  float xs = xi * sinq;
  float yc = yi * cosq;
  r.srcx = xc - ys;
  r.srcy = xs + yc;
  float detx = -(xc + ys);
  float dety = -xs + yc;
  r.slope = (r.srcy - dety) / (r.srcx - detx);
  r.islope = 1.0f / r.slope;
  return r;
}

// Crossing of the ray with horizontal line y = gy[n] (the A list) and with
// vertical line x = gx[n] (the B list).
__device__ __forceinline__ float cross_a(const RaySetup &r, int N, int n) {
  return r.islope * (grid_line(N, n) - r.srcy) + r.srcx;
}
__device__ __forceinline__ float cross_b(const RaySetup &r, int N, int n) {
  return r.slope * (grid_line(N, n) - r.srcx) + r.srcy;
}

// Contiguous range [lo, hi] of n in [0, N] whose crossing lies strictly
// inside (gx0, gxN).  `k` is the multiplier of the sequence (islope or
// slope), `b0`/`b1` the subtrahend and addend (srcy, srcx for A).  Returns
// the range length (0 if empty).
__device__ __forceinline__ int trim_range(float k, float b0, float b1, int N,
                                          float gx0, float gxN, int &lo) {
  // A non-finite multiplier or offset makes every crossing inf or NaN, which
  // the strict trims of trace_utils.cc:98-109 all reject.
  if (!__builtin_isfinite(k) || !__builtin_isfinite(b0) || !__builtin_isfinite(b1)) {
    lo = 0;
    return 0;
  }
  auto f = [&](int n) { return k * (grid_line(N, n) - b0) + b1; };
  const bool inc = (k >= 0.0f);  // +-0 gives a constant sequence
  // first/last index satisfying f > gx0 and f < gxN (monotone predicates)
  int a, b;
  if (inc) {
    // lower end: first n with f(n) > gx0
    int L = 0, H = N + 1;
    while (L < H) { int m = (L + H) >> 1; if (f(m) > gx0) H = m; else L = m + 1; }
    a = L;
    // upper end: last n with f(n) < gxN
    L = -1; H = N;
    while (L < H) { int m = (L + H + 1) >> 1; if (f(m) < gxN) L = m; else H = m - 1; }
    b = L;
  } else {
    // decreasing: f > gx0 holds on a prefix, f < gxN on a suffix
    int L = -1, H = N;
    while (L < H) { int m = (L + H + 1) >> 1; if (f(m) > gx0) L = m; else H = m - 1; }
    b = L;
    L = 0; H = N + 1;
    while (L < H) { int m = (L + H) >> 1; if (f(m) < gxN) H = m; else L = m + 1; }
    a = L;
  }
  lo = a;
  return (b >= a) ? (b - a + 1) : 0;
}

// Walks one ray and calls visit(idx, len, len2) for each of its len-1
// segments in the reference's order.  Returns the segment count.
template <bool kExactSqrt, typename Visit>
__device__ __forceinline__ int traverse_ray(const GridInfo &g, float yi,
                                            const AngleInfo &ang, Visit &&visit) {
  const int N = g.N;
  RaySetup r = ray_setup(yi, ang.sinq, ang.cosq);
  int alo, blo;
  const int alen = trim_range(r.islope, r.srcy, r.srcx, N, g.gx0, g.gxN, alo);
  const int blen = trim_range(r.slope, r.srcx, r.srcy, N, g.gx0, g.gxN, blo);
  // A is walked forward when quadrant==1, backward otherwise
  // (trace_utils.cc:175,183,197).
  const int dA = ang.quadrant ? 1 : -1;
  int nA = ang.quadrant ? alo : (alo + alen - 1);
  int nB = blo;
  int i = 0, j = 0;
  // Head of each list.  Heads past the end of a list are computed but never
  // selected.
  float ax = cross_a(r, N, nA), ay = grid_line(N, nA);
  float bx = grid_line(N, nB), by = cross_b(r, N, nB);
  float px = 0.f, py = 0.f;
  const int total = alen + blen;
  int nseg = 0;
  for (int k = 0; k < total; ++k) {
    // merge step: A wins only on a strict ax < bx (trace_utils.cc:184)
    const bool takeA = (i < alen) && (j >= blen || ax < bx);
    const float qx = takeA ? ax : bx;
    const float qy = takeA ? ay : by;
    // advance the list that was consumed and recompute its head with the
    // same formula for either list (branch-free across the wavefront)
    i += takeA ? 1 : 0;
    j += takeA ? 0 : 1;
    nA += takeA ? dA : 0;
    nB += takeA ? 0 : 1;
    const int nn = takeA ? nA : nB;
    const float kk = takeA ? r.islope : r.slope;
    const float b0 = takeA ? r.srcy : r.srcx;
    const float b1 = takeA ? r.srcx : r.srcy;
    const float gl = grid_line(N, nn);
    const float cr = kk * (gl - b0) + b1;
    ax = takeA ? cr : ax;
    ay = takeA ? gl : ay;
    bx = takeA ? bx : gl;
    by = takeA ? by : cr;
    if (k > 0) {
      const float dx = qx - px;
      const float dy = qy - py;
      const float mx = (qx + px) * 0.5f;
      const float my = (qy + py) * 0.5f;
      const float l2 = dx * dx + dy * dy;
      const float l = seg_sqrt<kExactSqrt>(l2);
      const int ix = (int)(mx + g.mg);
      const int iy = (int)(my + g.mg);
      visit(ix + iy * N, l, l2);
      ++nseg;
    }
    px = qx; py = qy;
  }
  return nseg;
}

}  // namespace trace_hip
