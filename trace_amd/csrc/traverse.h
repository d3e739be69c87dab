// traverse.h -- on-the-fly ray traversal for gfx950, bit-compatible with the
// reference's per-ray pipeline
//   CalculateCoordinates     trace_utils.cc:48-80
//   MergeTrimCoordinates     trace_utils.cc:82-111
//   SortIntersectionPoints   trace_utils.cc:113-158
//   CalculateDistanceLengths trace_utils.cc:160-187
// without materialising any per-ray array.
//
// The reference evaluates all 2(N+1) grid-line crossings of a ray, keeps
// those strictly inside the grid square, and merges the two lists by x.
// Here:
//
// * the crossing with grid line n is recomputed from the same global formula
//   (islope*(gy[n]-srcy)+srcx for the A list, slope*(gx[n]-srcx)+srcy for
//   the B list) whenever it is needed, with FMA contraction off, so every
//   coordinate is bit-identical to the reference's;
// * each crossing sequence is monotone in n (every step is a correctly
//   rounded, monotone function of the previous value), so the trimmed lists
//   are contiguous index ranges, found once per ray by binary search
//   (RayInfo) instead of the reference's O(N) scan;
// * the merge is a two-pointer walk that emits one segment per step and can
//   be resumed at any element: when the A list is ascending in walk order
//   (always, for angles in [0, pi)), the merged list is sorted by x with ties
//   going to B, so the state just after any element is known in closed form.
//   This lets an image-tile kernel walk only the part of a ray near its tile.
//
// The pixel index uses the reference's truncating conversion, so a midpoint
// that rounds onto the far grid line gives ix == N, i.e. pixel 0 of the next
// row, as in the reference; callers guard the image bounds.
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
  int N;      // grid side (= C)
  int C;      // detector columns
  float gx0;  // gridx[0]
  float gxN;  // gridx[N]
  float mg;   // (float)(N / 2.0), trace_utils.cc:220
};

// Per-ray constants: source point, slopes and the trimmed list ranges
// (A: horizontal-line crossings n in [alo, alo+alen); B: vertical-line
// crossings n in [blo, blo+blen)).  Computed once per geometry.  Three
// flag bits ride on alo (see ray_flags): kRayWhole, the tile walks must take
// the whole ray; kRayWraps, the ray has a segment with ix == N; kRay1D, only
// the primary axis may bound the ray's tile windows.
struct RayInfo {
  float srcx, srcy, slope, islope;
  int alo, alen, blo, blen;
};

constexpr int kRayWhole = 1 << 30;
constexpr int kRayWraps = 1 << 29;
constexpr int kRay1D = 1 << 28;

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

// Contiguous range of n in [0, N] whose crossing k*(gl(n)-b0)+b1 lies
// strictly inside (gx0, gxN) (trace_utils.cc:98-109).  Returns the length.
__device__ __forceinline__ int trim_range(float k, float b0, float b1, int N,
                                          float gx0, float gxN, int &lo) {
  // A non-finite multiplier or offset makes every crossing inf or NaN, which
  // the strict trims all reject.
  if (!__builtin_isfinite(k) || !__builtin_isfinite(b0) || !__builtin_isfinite(b1)) {
    lo = 0;
    return 0;
  }
  auto f = [&](int n) { return k * (grid_line(N, n) - b0) + b1; };
  const bool inc = (k >= 0.0f);  // +-0 gives a constant sequence
  int a, b;
  if (inc) {
    int L = 0, H = N + 1;  // first n with f(n) > gx0
    while (L < H) { int m = (L + H) >> 1; if (f(m) > gx0) H = m; else L = m + 1; }
    a = L;
    L = -1; H = N;  // last n with f(n) < gxN
    while (L < H) { int m = (L + H + 1) >> 1; if (f(m) < gxN) L = m; else H = m - 1; }
    b = L;
  } else {
    int L = -1, H = N;  // last n with f(n) > gx0
    while (L < H) { int m = (L + H + 1) >> 1; if (f(m) > gx0) L = m; else H = m - 1; }
    b = L;
    L = 0; H = N + 1;  // first n with f(n) < gxN
    while (L < H) { int m = (L + H) >> 1; if (f(m) < gxN) H = m; else L = m + 1; }
    a = L;
  }
  lo = a;
  return (b >= a) ? (b - a + 1) : 0;
}

// Ray set-up (trace_utils.cc:63-75 with xi = -1e6, yi from sirt.cc:129) and
// trims.
__device__ __forceinline__ RayInfo make_ray(const GridInfo &g, float yi, float sinq,
                                            float cosq) {
  const float xi = -1e6f;
  RayInfo r;
  const float xc = xi * cosq;
  const float ys = yi * sinq;
  const float xs = xi * sinq;
  const float yc = yi * cosq;
  r.srcx = xc - ys;
  r.srcy = xs + yc;
  const float detx = -(xc + ys);
  const float dety = -xs + yc;
  r.slope = (r.srcy - dety) / (r.srcx - detx);
  r.islope = 1.0f / r.slope;
  r.alen = trim_range(r.islope, r.srcy, r.srcx, g.N, g.gx0, g.gxN, r.alo);
  r.blen = trim_range(r.slope, r.srcx, r.srcy, g.N, g.gx0, g.gxN, r.blo);
  return r;
}

// Two-pointer merge of the A and B lists of one ray (trace_utils.cc:113-158),
// resumable at any element.
struct Walker {
  float srcx, srcy, slope, islope;
  int N, alo, alen, blo, blen, dA;
  bool fwd;
  bool whole;           // windows are unsafe for this ray: walk it whole
  bool one_axis;        // only the primary axis may bound its windows
  int i, j;             // A/B elements before the position seek() set
  int nA, nB;           // grid lines of the heads (next() advances these only)
  int nAend, nBend;     // the head lines once a list is exhausted
  float ax, ay, bx, by; // heads (computed even past the end, never selected)

  __device__ __forceinline__ void init(const RayInfo &r, int N_, int quadrant) {
    srcx = r.srcx; srcy = r.srcy; slope = r.slope; islope = r.islope;
    N = N_; alo = r.alo & ~(kRayWhole | kRayWraps | kRay1D); alen = r.alen; blo = r.blo; blen = r.blen;
    whole = (r.alo & kRayWhole) != 0;
    one_axis = (r.alo & kRay1D) != 0;
    fwd = quadrant != 0;  // A walked forward iff quadrant == 1 (:175,183,197)
    dA = fwd ? 1 : -1;
    nAend = fwd ? alo + alen : alo - 1;  // a_line(alen)
    nBend = blo + blen;
  }
  __device__ __forceinline__ int a_line(int wi) const {
    return fwd ? alo + wi : alo + alen - 1 - wi;
  }
  __device__ __forceinline__ float cross_a(int n) const {
    return islope * (grid_line(N, n) - srcy) + srcx;
  }
  __device__ __forceinline__ float cross_b(int n) const {
    return slope * (grid_line(N, n) - srcx) + srcy;
  }
  // A list ascending in walk order (then the merged list is sorted by x).
  __device__ __forceinline__ bool a_sorted() const {
    return alen < 2 || cross_a(a_line(0)) <= cross_a(a_line(alen - 1));
  }
  // Positions the walker after i_ A and j_ B elements.
  __device__ __forceinline__ void seek(int i_, int j_) {
    i = i_; j = j_;
    nA = a_line(i); nB = blo + j;
    ax = cross_a(nA); ay = grid_line(N, nA);
    bx = grid_line(N, nB); by = cross_b(nB);
  }
  // Number of A elements (walk order, assumed ascending) with x < v: the
  // first walk index whose crossing is >= v.  It is estimated from the line
  // whose crossing lies at x = v and confirmed by comparing the crossings
  // on either side (exact: the sequence is monotone); an estimate more than
  // a few elements off (ill-conditioned near-vertical rays) falls back to a
  // binary search.
  __device__ __forceinline__ int count_a_lt(float v) const {
    const float t = (v - srcx) * slope + srcy + 0.5f * (float)N;
    if (alen > 0 && __builtin_isfinite(t) && __builtin_fabsf(t) < 1.0e7f) {
      const int n = (int)floorf(t);
      int i = fwd ? n - alo : alo + alen - 1 - n;
      i = max(0, min(i, alen));
      for (int step = 0; step < 4; ++step) {
        const bool below = (i == 0) || cross_a(a_line(i - 1)) < v;
        const bool above = (i == alen) || !(cross_a(a_line(i)) < v);
        if (below && above) return i;
        i += below ? 1 : -1;
      }
    }
    int L = 0, H = alen;
    while (L < H) { int m = (L + H) >> 1; if (cross_a(a_line(m)) < v) L = m + 1; else H = m; }
    return L;
  }
  // Number of B elements with x <= v.  The B elements are the exact grid
  // lines blo .. blo + blen - 1, so the count is the last line at or left of
  // v, found from floor(v + N/2) and corrected by exact comparisons (the
  // float sum is off by less than one line).  NaN counts nothing.
  __device__ __forceinline__ int count_b_le(float v) const {
    if (blen == 0 || !(v >= grid_line(N, blo))) return 0;
    const int last = blo + blen - 1;
    if (v >= grid_line(N, last)) return blen;
    int n = (int)floorf(v + 0.5f * (float)N);
    n = max(blo, min(n, last));
    while (grid_line(N, n + 1) <= v) ++n;  // stops at last: v < gl(last)
    while (grid_line(N, n) > v) --n;       // stops at blo: v >= gl(blo)
    return n - blo + 1;
  }
  // Whether the next merged element is A's head: A wins only on a strict
  // ax < bx (trace_utils.cc:184).  The heads' lines tell whether a list is
  // exhausted (i == alen iff nA == a_line(alen); the counted walks never
  // take more than alen + blen elements, so j <= blen and j == blen iff
  // nB == blo + blen): no element counters are kept per step.
  __device__ __forceinline__ bool take_a() const {
    return (nA != nAend) && (nB == nBend || ax < bx);
  }
  // Emits the next merged element.  Branch-free across the wavefront.
  __device__ __forceinline__ void next(float &qx, float &qy) {
    const bool takeA = take_a();
    qx = takeA ? ax : bx;
    qy = takeA ? ay : by;
    nA += takeA ? dA : 0;
    nB += takeA ? 0 : 1;
    const int nn = takeA ? nA : nB;
    const float kk = takeA ? islope : slope;
    const float b0 = takeA ? srcy : srcx;
    const float b1 = takeA ? srcx : srcy;
    const float gl = grid_line(N, nn);
    const float cr = kk * (gl - b0) + b1;
    ax = takeA ? cr : ax;
    ay = takeA ? gl : ay;
    bx = takeA ? bx : gl;
    by = takeA ? by : cr;
  }
};

// Segment between consecutive merged points (trace_utils.cc:160-187).
template <bool kExactSqrt>
__device__ __forceinline__ void segment(float px, float py, float qx, float qy, float mg,
                                        int &ix, int &iy, float &len, float &len2) {
  const float dx = qx - px;
  const float dy = qy - py;
  const float mx = (qx + px) * 0.5f;
  const float my = (qy + py) * 0.5f;
  len2 = dx * dx + dy * dy;
  len = seg_sqrt<kExactSqrt>(len2);
  ix = (int)(mx + mg);
  iy = (int)(my + mg);
}

// Takes `pts` merged points from the walker's state, with an optional
// previous point; visit(ix, iy, len, len2) per segment.  Counted loop.
template <bool kExactSqrt, typename Visit>
__device__ __forceinline__ int walk_points(Walker &w, float mg, bool has_prev, float px,
                                           float py, int pts, Visit &&visit) {
  if (pts <= 0) return 0;
  if (!has_prev) {
    w.next(px, py);
    --pts;
  }
  for (int k = 0; k < pts; ++k) {
    float qx, qy;
    w.next(qx, qy);
    int ix, iy;
    float l, l2;
    segment<kExactSqrt>(px, py, qx, qy, mg, ix, iy, l, l2);
    visit(ix, iy, l, l2);
    px = qx;
    py = qy;
  }
  return pts;
}

// Whole ray in the reference's order.  A whole walk takes exactly
// alen + blen merged points (each step takes one of either list), so the
// loop is counted, with the first point peeled: no per-step branch.
template <bool kExactSqrt, typename Visit>
__device__ __forceinline__ int traverse_full(const GridInfo &g, const RayInfo &r, int quadrant,
                                             Visit &&visit) {
  Walker w;
  w.init(r, g.N, quadrant);
  w.seek(0, 0);
  const int n = w.alen + w.blen;
  if (n < 2) return 0;
  float px, py;
  w.next(px, py);
  for (int k = 1; k < n; ++k) {
    float qx, qy;
    w.next(qx, qy);
    int ix, iy;
    float l, l2;
    segment<kExactSqrt>(px, py, qx, qy, g.mg, ix, iy, l, l2);
    visit(ix, iy, l, l2);
    px = qx;
    py = qy;
  }
  return n - 1;
}

// The window consistency check of ray_flags on one axis (on_a: rows, where
// the A points are exact), fed the merged points in order.
struct AxisCheck {
  bool on_a, have_exact = false, pending = false, safe = true;
  float last = 0.f, cmin = 0.f, cmax = 0.f;
  __device__ __forceinline__ explicit AxisCheck(bool a) : on_a(a) {}
  __device__ __forceinline__ void point(bool takeA, float qx, float qy) {
    const float v = on_a ? qy : qx;
    if (takeA == on_a) {  // exact point: a grid line of this axis
      if (pending) {
        const float lo = have_exact ? fminf(last, v) - 0.5f : v - 1.5f;
        const float hi = have_exact ? fmaxf(last, v) + 0.5f : v + 1.5f;
        safe &= (cmin >= lo && cmax <= hi);
      }
      have_exact = true;
      pending = false;
      last = v;
    } else {
      cmin = pending ? fminf(cmin, v) : v;
      cmax = pending ? fmaxf(cmax, v) : v;
      pending = true;
    }
  }
  __device__ __forceinline__ bool done(int n) const {
    if (!pending) return safe;
    return safe && (have_exact ? (cmin >= last - 1.5f && cmax <= last + 1.5f) : (n < 2));
  }
};

// Flags of a ray, from one walk of its merged points.
//
// kRayWraps: some segment's midpoint rounds onto the right image edge, so
// its pixel index ix == N is pixel 0 of the next row (the reference's
// truncation, trace_utils.cc:177-185).  Generic rays never do: their
// crossings carry the 1/16-pixel granularity of the far source point.
//
// kRayWhole: a window walk of the ray would not be exact.  The window of a tile is a
// range of grid lines along one axis (rows for |islope| <= 1, where the
// A points carry exact row coordinates, columns otherwise).  The points of
// the other list carry a computed coordinate along that axis; it must lie
// within half a line of the span of the two exact points it is merged
// between (1.5 lines before the first or after the last exact point), so
// that every segment with its midpoint in the tile lies between exact
// points inside the window and its margin.  Generic rays are off by less
// than 0.2 line.  A ray lying within float rounding of a grid line (an odd
// C - N, or a centre on a half pixel, at an axis-aligned angle) can have a
// crossing of the parallel lines computed tens of pixels away; such rays are
// walked whole by the tile kernels.
//
// kRay1D: the same check fails on the second axis (columns for |islope| <=
// 1, rows otherwise), so the ray's windows must not be intersected with a
// window on that axis.
__device__ __forceinline__ int ray_flags(const GridInfo &g, const RayInfo &r, int quadrant) {
  Walker w;
  w.init(r, g.N, quadrant);
  const bool sorted = w.a_sorted();  // unsorted rays are walked whole in any case
  w.seek(0, 0);
  const bool on_a = __builtin_fabsf(r.islope) <= 1.0f;
  AxisCheck prim(on_a), sec(!on_a);
  bool wraps = false;
  float px = 0.f;
  const int n = w.alen + w.blen;
  for (int k = 0; k < n; ++k) {
    const bool takeA = w.take_a();
    float qx, qy;
    w.next(qx, qy);
    if (k > 0) {
      const float mx = (qx + px) * 0.5f;  // as segment()
      wraps |= ((int)(mx + g.mg) == g.N);
    }
    px = qx;
    prim.point(takeA, qx, qy);
    sec.point(takeA, qx, qy);
  }
  const bool whole = sorted && !prim.done(n);
  const bool one_axis = !sec.done(n);
  return (whole ? kRayWhole : 0) | (wraps ? kRayWraps : 0) | (one_axis ? kRay1D : 0);
}

// Rays whose merged points are used for a window of grid lines: every
// segment whose midpoint falls in the lines' span [lo, hi] is produced,
// together with a few neighbours (the callers filter by pixel).  One line
// of margin suffices for the rays ray_flags passes: the points merged
// between two consecutive exact points lie within half a line of their
// span, so a segment with its midpoint in row r lies between exact points
// on lines r-1 .. r+2 (1.5 lines before the first or after the last exact
// point: within one line of margin as well).
constexpr int kWindowMargin = 1;

// Segments of the ray that can have their midpoint in pixel columns
// [tx0, tx1) x rows [ty0, ty1).  The window of the A list (rows) is the A
// elements on lines [ty0 - M, ty1 + M], that of the B list (columns) the B
// elements on lines [tx0 - M, tx1 + M].  The walk covers the merged points
// from just after the later of the elements preceding the two windows to the
// earlier of their last elements (kTwoAxes), or of the primary window alone
// (rows for |islope| <= 1, where the A points carry exact row coordinates;
// columns otherwise) without kTwoAxes and for rays flagged kRay1D.  Two axes
// walk about 30% fewer segments, but the rays of a tile then differ in walk
// length, which costs a wavefront unless its lanes hold rays of similar
// length.  Consecutive merged points lie
// within about one line of each other along an axis that ray_flags passed,
// so the margin covers every segment of the tile.  Rays whose A list is not
// ascending (angles outside [0, pi) where the quadrant rule and the slope
// sign disagree), or whose primary axis fails, are walked whole.
template <bool kExactSqrt, bool kTwoAxes, typename Visit>
__device__ __forceinline__ int traverse_window(const GridInfo &g, const RayInfo &r, int quadrant,
                                               int tx0, int tx1, int ty0, int ty1,
                                               Visit &&visit) {
  Walker w;
  w.init(r, g.N, quadrant);
  if (!w.a_sorted() || w.whole) {
    w.seek(0, 0);
    return walk_points<kExactSqrt>(w, g.mg, false, 0.f, 0.f, w.alen + w.blen, visit);
  }
  const int M = kWindowMargin;
  const bool on_a = __builtin_fabsf(r.islope) <= 1.0f;
  const bool useA = on_a || (kTwoAxes && !w.one_axis);
  const bool useB = !on_a || (kTwoAxes && !w.one_axis);
  // A window (rows [ty0 - M, ty1 + M]) as walk indices [iS, iE)
  int iS, iE;
  if (w.fwd) {
    iS = (ty0 - M) - w.alo;
    iE = (ty1 + M) - w.alo + 1;
  } else {
    iS = (w.alo + w.alen - 1) - (ty1 + M);
    iE = (w.alo + w.alen - 1) - (ty0 - M) + 1;
  }
  iS = max(0, min(iS, w.alen));
  iE = max(0, min(iE, w.alen));
  // B window (columns [tx0 - M, tx1 + M]) as indices [jS, jE)
  int jS = (tx0 - M) - w.blo;
  int jE = (tx1 + M) - w.blo + 1;
  jS = max(0, min(jS, w.blen));
  jE = max(0, min(jE, w.blen));
  if ((useA && iS >= iE) || (useB && jS >= jE)) return 0;
  // Start just after the later of A[iS-1] and B[jS-1] in merge order (x
  // ascending; on a tie B is merged first, so A's position is the later).
  const bool hasA = useA && iS > 0, hasB = useB && jS > 0;
  const int nA0 = w.a_line(hasA ? iS - 1 : 0), nB0 = w.blo + (hasB ? jS - 1 : 0);
  const float sa = w.cross_a(nA0), sb = grid_line(g.N, nB0);
  bool has_prev = false;
  float px = 0.f, py = 0.f;
  int i0 = 0, j0 = 0;
  if (hasA && (!hasB || sa >= sb)) {
    px = sa;
    py = grid_line(g.N, nA0);
    i0 = iS;
    j0 = w.count_b_le(px);  // B elements merged before A[iS-1]
    has_prev = true;
  } else if (hasB) {
    px = sb;
    py = w.cross_b(nB0);
    i0 = w.count_a_lt(px);  // A elements merged before B[jS-1] (strict <)
    j0 = jS;
    has_prev = true;
  }
  // Stop after the earlier of A[iE-1] and B[jE-1] (a window that reaches the
  // end of its list does not bound the walk).  The merged position of the
  // stop element follows from the same exact counts as the start: A[iE-1]
  // comes after the count_b_le(its x) B elements, B[jE-1] after the
  // count_a_lt(its x) A elements.  So the walk takes a known number of
  // points, and its loop is counted (no per-step exit test).
  const bool endA = useA && iE < w.alen, endB = useB && jE < w.blen;
  const float ea = w.cross_a(w.a_line(endA ? iE - 1 : 0));
  const float eb = grid_line(g.N, w.blo + (endB ? jE - 1 : 0));
  int i_end = w.alen, j_end = w.blen;
  if (endA && (!endB || ea < eb)) {
    i_end = iE;
    j_end = w.count_b_le(ea);
  } else if (endB) {
    i_end = w.count_a_lt(eb);
    j_end = jE;
  }
  const int pts = min((i_end - i0) + (j_end - j0), w.alen + w.blen - i0 - j0);
  w.seek(i0, j0);
  return walk_points<kExactSqrt>(w, g.mg, has_prev, px, py, pts, visit);
}

// Columns whose ray can touch tile [x0,x1]x[y0,y1] at angle p: the ray of
// column c is the line q.(-sin, cos) = yi0 + c.  Its merged points lie
// within 0.1 pixel of that line across it (a crossing's error lies along
// the computed coordinate, which for an ill-conditioned crossing is nearly
// the ray's own direction), so a margin of kColumnMargin pixels on the
// tile's projection is ample.  The host's count bound allows a margin of
// half a pixel; the float projections below carry errors below 0.01 pixel
// for N <= 32768, far below the 0.05 pixel kept in hand on each side, so
// their range never exceeds that bound.
constexpr float kColumnMargin = 0.45f;
constexpr double kCountMargin = 0.5;
__device__ __forceinline__ void tile_columns(const AngleInfo &a, float x0, float x1, float y0,
                                             float y1, float yi0, int &cmin, int &cmax) {
  const float s = a.sinq, c = a.cosq;
  const float t00 = -x0 * s + y0 * c, t01 = -x0 * s + y1 * c;
  const float t10 = -x1 * s + y0 * c, t11 = -x1 * s + y1 * c;
  const float tmin = fminf(fminf(t00, t01), fminf(t10, t11));
  const float tmax = fmaxf(fmaxf(t00, t01), fmaxf(t10, t11));
  cmin = (int)ceilf(tmin - kColumnMargin - yi0);
  cmax = (int)floorf(tmax + kColumnMargin - yi0);
}

}  // namespace trace_hip
