// kernels.hip -- gfx950 kernels of the reconstruction hot path.
//
// Data layout in HBM (per handle):
//   x      [S][N][N]        image, the reference's layout (idx = ix + iy*N)
//   xb     [S/SB][N*N][SB]  slice-batched copy of x read by the forward
//                           projector: one segment gathers SB slices as
//                           contiguous float4s
//   w      [P][S][C]        per-ray weights written by the forward projector
//   z      [S][N][N]        A^T w written by the tile backprojector
//   den    [N][N]           A^T 1 (iteration invariant, one slice's geometry)
//   rays   [P][C]           RayInfo: source point, slopes, trimmed ranges
//
// Forward projector: 1, 2 or 4 lanes per ray (p, c) and slice batch, each
// summing its share of the batch's slices; a wavefront holds a bundle of
// adjacent angles x adjacent detector columns.  Each lane walks the whole
// ray in the reference's order and accumulates y = sum x[idx]*len per
// slice, so y is bit-identical to CalculateSimdata (sirt.cc:4-19).
//
// Backprojector: one workgroup per (image tile, slice batch) owns the tile's
// accumulators in LDS for all angles.  Its lanes take the (angle, column)
// pairs whose ray can touch the tile, walk only the part of the ray near the
// tile (traverse_window) and add the fixed-point terms of len*w into LDS
// with integer atomics; the tile is then written to HBM with coalesced
// stores.  No global atomics, no per-thread image replicas (the reference's
// DISP replicas + tree merge, disp_engine_reduction.h:104-185).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"

#pragma clang fp contract(off)

namespace trace_hip {

namespace {

constexpr int kRayBlock = 256;

constexpr int kPixBlock = 256;
constexpr int kTileBlock = 256;

__device__ __forceinline__ void atomic_add_f32(float *p, float v) {
  unsafeAtomicAdd(p, v);  // hardware global float atomic, relaxed
}

// ---------------------------------------------------------------- rays --

__global__ void __launch_bounds__(kRayBlock)
k_ray_info(GridInfo g, const AngleInfo *__restrict__ ang, const float *__restrict__ yi,
           int P, int flags, RayInfo *__restrict__ rays) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  if (t >= (int64_t)P * g.C) return;
  const int c = (int)(t % g.C), p = (int)(t / g.C);
  RayInfo r = make_ray(g, yi[c], ang[p].sinq, ang[p].cosq);
  // the flags walk each ray once; whole-ray walks (the forward projector)
  // do not read them
  if (flags) r.alo |= ray_flags(g, r, ang[p].quadrant);
  rays[t] = r;
}

// ------------------------------------------------------------- forward --

enum FwdMode { kFwdY = 0, kFwdSIRT = 1, kFwdEM = 2 };

template <int SB>
struct Vec {
  float v[SB];
};

template <int SB>
__device__ __forceinline__ void load_slices(const float *__restrict__ p, float (&o)[SB]) {
  static_assert(SB == 1 || SB == 2 || SB % 4 == 0, "slice batch: 1, 2 or a multiple of 4");
  if constexpr (SB == 1) {
    o[0] = p[0];
  } else if constexpr (SB == 2) {
    const float2 q = *reinterpret_cast<const float2 *>(p);
    o[0] = q.x; o[1] = q.y;
  } else {
#pragma unroll
    for (int k = 0; k < SB; k += 4) {
      const float4 q = *reinterpret_cast<const float4 *>(p + k);
      o[k] = q.x; o[k + 1] = q.y; o[k + 2] = q.z; o[k + 3] = q.w;
    }
  }
}

template <int SB>
__device__ __forceinline__ void store_slices(float *__restrict__ p, const float (&v)[SB]) {
  if constexpr (SB == 1) {
    p[0] = v[0];
  } else if constexpr (SB == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int q = 0; q < SB; q += 4)
      *reinterpret_cast<float4 *>(p + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
  }
}

// Lanes are grouped in ray bundles: a wavefront holds BA adjacent angles x
// (64 / BA) adjacent detector columns (runtime `ba`), so at every step its
// lanes sit on nearby pixels of nearby rays and share cache lines.  One lane
// per ray (p, c) and slice batch b (blockIdx.y).  kFwdY writes y[p][s][c];
// kFwdSIRT / kFwdEM write the per-ray weight of UpdateReconReplica
//   SIRT  w = (p - y)/a2, a2 = sum len^2    (sirt.cc:47-52)
//   EM    w = p/y                           (mlem.cc:50, pml.cc:236)
// to w[p][c][s] (slices innermost, read as a block by the backprojector) and
// fold max |w| over the rays with a segment in the image per slice into
// maxw (bits; zeroed beforehand).
// LPR lanes per ray (1, 2 or 4): the lanes of one ray repeat its walk and
// each gathers and sums SB/LPR of the batch's slices, so a gather touches
// LPR-fold fewer distinct cache lines per wavefront instruction (the
// address unit, not the VALU, limits the one-lane form).
template <int SB, int MODE, bool kExact, int LPR = 1>
__device__ __forceinline__ void forward_bundle(const GridInfo &g, const AngleInfo *__restrict__ ang,
                                               const RayInfo *__restrict__ rays, int P, int S,
                                               int ba_log2, const float *__restrict__ xb,
                                               const float *__restrict__ sino,
                                               float *__restrict__ out,
                                               unsigned *__restrict__ maxw, int64_t bundle,
                                               int b) {
  static_assert(LPR == 1 || LPR == 2 || LPR == 4, "lanes per ray: 1, 2 or 4");
  static_assert(SB % LPR == 0, "slice batch must split evenly over the lanes of a ray");
  constexpr int SBL = SB / LPR;  // slices per lane
  constexpr int LPR_LOG2 = LPR == 1 ? 0 : (LPR == 2 ? 1 : 2);
  const int C = g.C;
  const int lane = threadIdx.x & 63;
  const int sub = lane & (LPR - 1);    // slice group of this lane
  const int rl = lane >> LPR_LOG2;     // ray of this lane within the wavefront
  const int bc_log2 = 6 - LPR_LOG2 - ba_log2;  // columns per bundle = 64 / (LPR * BA)
  const int nbc = (C + (1 << bc_log2) - 1) >> bc_log2;
  const int p = (int)(bundle / nbc) * (1 << ba_log2) + (rl >> bc_log2);
  const int c = (int)(bundle % nbc) * (1 << bc_log2) + (rl & ((1 << bc_log2) - 1));
  const bool valid = (p < P) && (c < C);
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  const float *xbb = xb + (int64_t)b * NN * SB + sub * SBL;
  float acc[SBL];
#pragma unroll
  for (int k = 0; k < SBL; ++k) acc[k] = 0.f;
  float a2 = 0.f;
  bool hit = false;  // the ray has a segment in the image
  const int64_t t = (int64_t)p * C + c;
  if (valid) {
    const RayInfo r = rays[t];
    traverse_full<kExact>(g, r, ang[p].quadrant, [&](int ix, int iy, float l, float l2) {
      const unsigned idx = (unsigned)(ix + iy * g.N);
      const bool in = idx < NN;
      if constexpr (SBL == 1) {
        // branch-free: a segment outside the image (skipped) loads pixel 0
        // and adds +0, which leaves the sum's bits unchanged (a sum that
        // starts at +0 is never -0 in round-to-nearest)
        float v[1];
        load_slices<1>(xbb + (int64_t)(in ? idx : 0u) * SB, v);
        const float t = v[0] * l;
        acc[0] += in ? t : 0.f;
      } else if (in) {
        float v[SBL];
        load_slices<SBL>(xbb + (int64_t)idx * SB, v);
#pragma unroll
        for (int k = 0; k < SBL; ++k) acc[k] += v[k] * l;
      }
      hit |= in;
      if (MODE == kFwdSIRT) a2 += l2;
    });
  }
#pragma unroll
  for (int k = 0; k < SBL; ++k) {
    const int s = b * SB + sub * SBL + k;
    const bool live = valid && (s < S);
    unsigned m = 0;
    if (live) {
      const int64_t o = ((int64_t)p * S + s) * C + c;
      if (MODE == kFwdY) {
        out[o] = acc[k];
      } else {
        const float wv = (MODE == kFwdSIRT) ? (sino[o] - acc[k]) / a2 : sino[o] / acc[k];
        out[t * S + s] = wv;
        // A ray with no segment in the image (a shifted rotation centre
        // leaves edge columns outside) gets y = 0 and a2 = 0, so an
        // inf/NaN weight, as in the reference; it adds nothing to A^T w
        // there or here, so it must not set the slice's scale either.
        if (hit) m = __float_as_uint(__builtin_fabsf(wv));
      }
    }
    if (MODE != kFwdY) {
      // max over the lanes holding the same slice (same sub)
      for (int off = 32; off >= LPR; off >>= 1)
        m = max(m, (unsigned)__shfl_down((int)m, off, 64));
      if (lane < LPR && s < S) atomicMax(maxw + s, m);
    }
  }
}

template <int SB, int MODE, bool kExact, int LPR>
__global__ void __launch_bounds__(kRayBlock)
k_forward(GridInfo g, const AngleInfo *__restrict__ ang, const RayInfo *__restrict__ rays,
          int P, int S, int ba_log2, const float *__restrict__ xb,
          const float *__restrict__ sino, float *__restrict__ out,
          unsigned *__restrict__ maxw) {
  forward_bundle<SB, MODE, kExact, LPR>(
      g, ang, rays, P, S, ba_log2, xb, sino, out, maxw,
      xcd_chunked(blockIdx.x, gridDim.x) * (kRayBlock / 64) + (threadIdx.x >> 6), blockIdx.y);
}

// w[p][c][s] = r[p][s][c], folding max |r| per slice into maxw (zeroed
// beforehand).  One block row per (p, s).
__global__ void __launch_bounds__(256)
k_to_pcs(const float *__restrict__ r, int P, int S, int C, float *__restrict__ w,
         unsigned *__restrict__ maxw) {
  const int64_t ps = blockIdx.x;  // p * S + s
  const int s = (int)(ps % S);
  const int64_t p = ps / S;
  unsigned m = 0;
  for (int c = threadIdx.x; c < C; c += 256) {
    const float v = r[ps * C + c];
    w[(p * C + c) * S + s] = v;
    m = max(m, __float_as_uint(__builtin_fabsf(v)));
  }
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_down((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(maxw + s, m);
}

// ------------------------------------------------------- backprojector --

enum BpMode { kBpSino = 0, kBpOnes = 1, kBpCount = 2 };

struct TileArgs {
  GridInfo g;
  const AngleInfo *ang;
  const RayInfo *rays;
  const int *count;   // [P] upper bound of the columns touching any tile
  int P, S;
  int count_lds;      // 1: count[] is staged in LDS after the accumulators
  float yi0;          // detector coordinate of column 0
  const float *w;     // [P][C][S] weights (kBpSino)
  const unsigned *maxw;  // [S] bits of max |w| per slice (kBpSino)
  float maxlen;       // longest segment of the geometry
  float denmax;       // bound on sum of lengths per pixel (max of A^T 1), kBpSino
  float *z;           // [S][N][N] output (kBpOnes/kBpCount: [N][N])
  int wraps;          // some ray has a segment with ix == N (kRayWraps)
};

// LDS accumulator layout of the tile backprojector: one plane per slice,
// rows padded to TW + 1 64-bit words.  Neighbouring rays of one angle sit
// one pixel apart across their direction, so at a given step the lanes of a
// wavefront hit pixels offset by (0, +-1), (+-1, 0) or (+-1, +-1); the odd
// pitch spreads those over the banks.
// Smallest integer whose square is >= v (v >= 0).
constexpr int isqrt_ceil(int v) {
  int q = 0;
  while (q * q < v) ++q;
  return q;
}

template <int TW, int TH>
struct TilePitch {
  static constexpr int kPitch = TW + 1;
  static constexpr int kPlane = TH * kPitch;
};

// Fixed-point accumulation.  LDS float atomics (ds_add_f32) run at a small
// fraction of the rate of LDS integer atomics on gfx950, so each term len*w
// is scaled by 2^e, converted to an integer and added with an integer LDS
// atomic.  Integer addition is associative, so the result does not depend
// on the order in which lanes and waves arrive: the backprojection is
// bit-reproducible.
//   weights (kBpSino): 32-bit words, e per slice such that
//     2^e * max|w| * denmax < 2^30, denmax >= the largest sum of segment
//     lengths in one pixel (max of A^T 1); every pixel's sum then stays
//     below 2^30 + (number of terms)/2, so it cannot overflow; and
//     2^e * max|w| * maxlen < 2^22, so that every term fits the rounding
//     below.  A term is rounded to nearest even with one fma:
//     fl(len * ws + 1.5 * 2^23) is 1.5 * 2^23 + rne(len * ws) exactly (floats
//     in [2^23, 2^24) are the integers), so its bits minus those of
//     1.5 * 2^23 are the integer term (unbiased; error <= 2^-e/2 per term).
//   column sums (kBpOnes): 64-bit words, e from the longest segment
//     (|term| < 2^30), truncated.
// A slice whose weights are not all finite (y = 0 rays give inf/NaN in the
// reference too) is accumulated in float instead, in the same slot, so that
// inf/NaN propagate as they do in the reference.
__device__ __forceinline__ int fixed_exponent(float maxw, float maxlen, bool &fixed) {
  const float m = maxw * maxlen;
  fixed = __builtin_isfinite(m);
  if (!fixed) return 0;
  if (m == 0.f) return 0;
  return 29 - ilogbf(m);
}

// Scale exponent of a weights slice (kBpSino): both bounds above.
__device__ __forceinline__ int weight_exponent(float maxw, float denmax, float maxlen,
                                               bool &fixed) {
  int e = fixed_exponent(maxw, denmax, fixed);
  const float t = maxw * maxlen;
  fixed = fixed && __builtin_isfinite(t);
  if (fixed && t != 0.f) e = min(e, 21 - ilogbf(t));
  return fixed ? e : 0;
}

constexpr float kRoundMagic = 12582912.0f;        // 1.5 * 2^23
constexpr unsigned kRoundMagicBits = 0x4B400000u;  // its bits

// Bits of 1.5 * 2^23 + rne(len * ws), for |len * ws| < 2^22.
__device__ __forceinline__ unsigned fixed_bits(float len, float ws) {
  return __float_as_uint(__builtin_fmaf(len, ws, kRoundMagic));
}

// rne(len * ws) as an integer, for |len * ws| < 2^22.
__device__ __forceinline__ int fixed_term(float len, float ws) {
  return (int)(fixed_bits(len, ws) - kRoundMagicBits);
}

// Accumulator word of a mode: the weights backprojector sums 32-bit fixed
// point (or float bits for a non-finite slice); column sums and counts use
// 64-bit words.
template <int MODE>
using BpWord = typename std::conditional<MODE == kBpSino, unsigned, unsigned long long>::type;

// Packed weights accumulators (SB >= 2): slices 2k and 2k+1 share a 64-bit
// word, planes [SB/2][TH][TW + 1] of 64-bit words, the even slice in the
// low half, then one plane [TH][TW + 1] of 32-bit per-pixel term counts.
// When every slice of the batch is fixed point, one ds_add_u64 adds both
// terms as the raw fma results: with a, b the rounded terms and M the bits
// of 1.5 * 2^23, the 64-bit value 2^32 (M + b) + (M + a) is exact (both
// halves are positive and below 2^31), so after n terms the word holds
// 2^32 (n M + sum b) + n M + sum a (mod 2^64).  Subtracting n M (2^32 + 1)
// leaves 2^32 sum(b) + sum(a): its low half is sum(a) (|sum(a)| < 2^31) and
// the rest / 2^32 is sum(b).  A term then costs one fma per slice and the
// two slices one ds_add_u64; the count costs one ds_add_u32 per segment.
// Otherwise each half is its own 32-bit slot (integer or float atomics).
template <int MODE, int SB>
struct BpPacked {
  static constexpr bool value = MODE == kBpSino && SB >= 2 && SB % 2 == 0;
};

// Main loop of the tile backprojector: lanes take consecutive (angle,
// column) pairs of the flattened per-angle column ranges and add their
// segments near the tile into the LDS accumulators.  ALLFX: every slice of
// the batch uses the fixed-point path (the common case; no per-slice
// branch in the inner loop).
//
// The reference's pixel index truncates midpoint + N/2 (trace_utils.cc:
// 177-185), so a segment whose midpoint rounds onto the right image edge
// gets ix == N, i.e. pixel 0 of the next row.  That happens to every
// segment of a near-vertical ray lying within float rounding of the edge
// line x = N/2 (rays on grid lines: an odd C - N, or a half-pixel shift of
// the centre), and such a ray need not pass through the tile that holds
// its target pixels.  pass 0 adds the segments whose midpoint is in the
// tile; pass 1 (EDGE, tiles at column 0 of a geometry with such rays only)
// walks the flagged rays through the strip just left of the right edge,
// rows ty0-1 .. ty1-1, and adds only their wrapped segments.  Each segment
// is added by exactly one pass.
template <int TW, int TH, int SB, int MODE, bool kExact, bool ALLFX, bool EDGE>
__device__ __forceinline__ void bp_tile_loop(const TileArgs &a, BpWord<MODE> *acc,
                                             const int *count, int tx0, int tx1, int ty0,
                                             int ty1, int b, const int (&ex)[SB],
                                             const bool (&fx)[SB]) {
  static_assert(SB <= kWeightPad, "a slice batch of weights past the last row must be allocated");
  constexpr int PITCH = TilePitch<TW, TH>::kPitch;
  constexpr int PLANE = TilePitch<TW, TH>::kPlane;
  const int N = a.g.N;
  const int C = a.g.C;
  // region whose rays are walked, and its window in grid lines
  const int wx0 = EDGE ? N - 1 : tx0, wx1 = EDGE ? N : tx1;
  const int wy0 = EDGE ? ty0 - 1 : ty0, wy1 = ty1;
  const float x0 = grid_line(N, wx0), x1 = grid_line(N, wx1);
  const float y0 = grid_line(N, wy0), y1 = grid_line(N, wy1);
  // the strip's column range is at most one wider than a tile's
  constexpr int extra = EDGE ? 1 : 0;
  // Lane t takes flat slots t, t + kTileBlock, ... of the tile's (angle,
  // column) pairs.  One slice (kRank): slot f is rank r = f / P of angle
  // p = f % P, rank r being the r-th of the angle's column slots counted
  // from the middle of its range outward, so the lanes of a wavefront take
  // one rank at consecutive angles: rays at the same position relative to
  // the tile, whose two-axis windows have similar lengths.  Several slices:
  // slot f is column slot k of angle p in angle-major order, so the lanes
  // hold neighbouring parallel rays.  Rays of neighbouring angles at the
  // same rank walk nearly the same pixels, and their LDS atomics contend
  // for one address; that costs little with one atomic per segment, and too
  // much with one per slice pair (measured: 15.4 vs 19.8 ms for one
  // 2048^2 slice at 32x32 tiles; 532 vs 109 ms for 64 slices).
  constexpr bool kRank = (SB == 1);
  constexpr int kRanks = isqrt_ceil(TW * TW + TH * TH) + 4 + extra;  // >= count[p] + extra
  // Slot iteration: (r, p, k) -> the lane's next (angle, column slot).
  auto advance = [&](int &r, int &p, int &k) {
    if (kRank) {
      p += kTileBlock;
      while (p >= a.P) { p -= a.P; ++r; }
    } else {
      k += kTileBlock;
      while (p < a.P && k >= count[p] + extra) { k -= count[p] + extra; ++p; }
    }
  };
  auto live = [&](int r, int p) { return kRank ? r < kRanks : p < a.P; };
  // One (angle, column) pair, fetched a pair ahead of its walk: the ray
  // table entry and the weights mostly miss L2, so their loads for the next
  // pair are in flight while the current pair is walked.
  struct Pair {
    bool ok;
    int quadrant;
    RayInfo ray;
    float w[SB];
  };
  // A pair's slot resolved to (angle, column slot), and whether it exists
  // so far (the column range is checked when the pair is fetched).
  struct Slot {
    int p, k;
    bool ok;
  };
  auto resolve = [&](int r, int p, int k) {
    const bool lv = live(r, p);
    Slot o{lv ? p : 0, k, lv};
    if (kRank) {
      const int cnt = count[o.p] + extra;
      o.ok = o.ok && r < cnt;
      o.k = ((cnt - 1) >> 1) + ((r & 1) ? (r + 1) >> 1 : -(r >> 1));
    }
    return o;
  };
  // Every lane issues the same loads on every path (row 0 when it has no
  // pair; a whole slice batch of weights, which the kWeightPad floats after
  // the weights keep in bounds for the last row): with one path, the
  // compiler's wait for the current pair's data covers that pair only, and
  // not the next pair's loads issued after it.  The angle record comes in
  // loaded a pair earlier still, so no wait at all sits between a pair's
  // resolution and its loads.
  auto fetch = [&](const Slot &sl, const AngleInfo &ang, Pair &q) {
    const int pp = sl.p, k = sl.k;
    bool ok = sl.ok;
    int cmin, cmax;
    tile_columns(ang, x0, x1, y0, y1, a.yi0, cmin, cmax);
    const int c = cmin + k;
    ok = ok && c >= 0 && c < C && c <= cmax;
    q.ok = ok;
    q.quadrant = ang.quadrant;
    const int64_t t = ok ? (int64_t)pp * C + c : 0;
    q.ray = a.rays[t];
    if (MODE == kBpSino) {
      const float *wr = a.w + t * a.S + b * SB;
      if constexpr (SB % 4 == 0) {
#pragma unroll
        for (int i = 0; i < SB; i += 4) {
          float4 v;
          __builtin_memcpy(&v, wr + i, sizeof v);  // rows are 4-byte aligned
          q.w[i] = v.x, q.w[i + 1] = v.y, q.w[i + 2] = v.z, q.w[i + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < SB; ++i) q.w[i] = wr[i];
      }
    }
  };
  int r = 0, p = 0, k = threadIdx.x;
  if (kRank) {
    r = threadIdx.x / a.P;
    p = threadIdx.x % a.P;
  } else {
    while (p < a.P && k >= count[p] + extra) { k -= count[p] + extra; ++p; }
  }
  Pair cur;
  {
    const Slot s0 = resolve(r, p, k);
    fetch(s0, a.ang[s0.p], cur);
  }
  int rn = r, pn = p, kn = k;
  advance(rn, pn, kn);
  Slot sn = resolve(rn, pn, kn);
  AngleInfo angn = a.ang[sn.p];
  while (live(r, p)) {
    // the pair after next: slot and angle record
    int rnn = rn, pnn = pn, knn = kn;
    advance(rnn, pnn, knn);
    const Slot snn = resolve(rnn, pnn, knn);
    const AngleInfo angnn = a.ang[snn.p];
    Pair nxt;
    fetch(sn, angn, nxt);
    if (cur.ok) {
      // weights pre-scaled by 2^e (exact: a power-of-two scaling commutes
      // with rounding), so fl(len * ws) = fl(len * w) * 2^e per segment
      float wv[SB];
      bool any = true;
      if (MODE == kBpSino) {
        any = false;
#pragma unroll
        for (int q = 0; q < SB; ++q) {
          // slots past the stack's last slice (read from the next row) add 0
          const float wq = b * SB + q < a.S ? cur.w[q] : 0.f;
          any |= (wq != 0.f);
          wv[q] = (ALLFX || fx[q]) ? ldexpf(wq, ex[q]) : wq;
        }
      } else if (MODE == kBpOnes) {
        wv[0] = ldexpf(1.0f, ex[0]);
      }
      if (any && (!EDGE || (cur.ray.alo & kRayWraps))) {
        traverse_window<kExact, kRank>(a.g, cur.ray, cur.quadrant, wx0, wx1, wy0, wy1,
                                [&](int ix, int iy, float l, float) {
          // ix == N is pixel 0 of the next row (truncation); this pass takes
          // the wrapped segments (EDGE) or the others (the bounds test below
          // rejects the rest without a branch of its own)
          const bool wrap = (ix == N);
          const bool take = EDGE ? wrap : !wrap;
          ix = wrap ? 0 : ix;
          iy = wrap ? iy + 1 : iy;
          const unsigned lx = take ? (unsigned)(ix - tx0) : 0xffffffffu;
          const unsigned ly = (unsigned)(iy - ty0);
          if (lx < (unsigned)(tx1 - tx0) && ly < (unsigned)(ty1 - ty0)) {
            BpWord<MODE> *dst = acc + ly * PITCH + lx;
            if (MODE == kBpCount) {
              atomicAdd(dst, (BpWord<MODE>)1);
            } else if (MODE == kBpOnes) {
              // |t| < 2^30: truncating conversion, error < 2^-e
              atomicAdd(dst, (BpWord<MODE>)(long long)(int)(l * wv[0]));
            } else if (BpPacked<MODE, SB>::value) {
              // 32-bit index of the pixel's 64-bit word in pair plane 0
              unsigned *d2 = reinterpret_cast<unsigned *>(acc) + 2 * (ly * PITCH + lx);
#pragma unroll
              for (int k2 = 0; k2 < SB / 2; ++k2) {
                unsigned *dk = d2 + 2 * k2 * PLANE;
                if (ALLFX) {
                  const unsigned long long v =
                      ((unsigned long long)fixed_bits(l, wv[2 * k2 + 1]) << 32) |
                      fixed_bits(l, wv[2 * k2]);
                  atomicAdd(reinterpret_cast<unsigned long long *>(dk), v);
                } else {
#pragma unroll
                  for (int h = 0; h < 2; ++h) {
                    const int q = 2 * k2 + h;
                    if (fx[q])
                      atomicAdd(dk + h, (unsigned)fixed_term(l, wv[q]));
                    else
                      atomicAdd(reinterpret_cast<float *>(dk + h), l * wv[q]);
                  }
                }
              }
              // terms per pixel, for the magic offsets of the ALLFX words
              if (ALLFX)
                atomicAdd(reinterpret_cast<unsigned *>(acc) + SB * PLANE + ly * PITCH + lx, 1u);
            } else {
#pragma unroll
              for (int q = 0; q < SB; ++q) {
                if (ALLFX || fx[q]) {
                  atomicAdd(dst + q * PLANE, (BpWord<MODE>)fixed_term(l, wv[q]));
                } else {
                  atomicAdd(reinterpret_cast<float *>(dst + q * PLANE), l * wv[q]);
                }
              }
            }
          }
        });
      }
    }
    r = rn;
    p = pn;
    k = kn;
    cur = nxt;
    rn = rnn;
    pn = pnn;
    kn = knn;
    sn = snn;
    angn = angnn;
  }
}

// LDS words (BpWord<MODE>) of a tile's accumulators: the slice planes, and
// for packed weights the per-pixel term counts.
template <int TW, int TH, int SB, int MODE>
constexpr int bp_acc_words() {
  return (SB + (BpPacked<MODE, SB>::value ? 1 : 0)) * TilePitch<TW, TH>::kPlane;
}

// Workgroup (tile, slice batch): LDS accumulators acc[SB][TH][TW + 1] for
// all angles, then one coalesced write-out.  `acc` is the workgroup's
// dynamic LDS (bp_acc_words() words of BpWord<MODE>, then int count[P]
// when a.count_lds); the body starts and ends with a workgroup barrier, so a
// persistent workgroup may run it for one tile after another.
template <int TW, int TH, int SB, int MODE, bool kExact>
__device__ __forceinline__ void bp_tile_body(const TileArgs &a, int tile, int b,
                                             unsigned long long *lds) {
  BpWord<MODE> *acc = reinterpret_cast<BpWord<MODE> *>(lds);
  constexpr int PITCH = TilePitch<TW, TH>::kPitch;
  constexpr int PLANE = TilePitch<TW, TH>::kPlane;
  constexpr int WORDS = bp_acc_words<TW, TH, SB, MODE>();
  int *lcount = reinterpret_cast<int *>(acc + WORDS);
  const int *count = a.count_lds ? lcount : a.count;
  const int N = a.g.N;
  const int tiles_x = (N + TW - 1) / TW;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int tx0 = tx * TW, ty0 = ty * TH;
  const int tx1 = min(tx0 + TW, N), ty1 = min(ty0 + TH, N);
  __syncthreads();  // the previous tile's write-out has read acc
  for (int i = threadIdx.x; i < WORDS; i += kTileBlock) acc[i] = 0;
  if (a.count_lds)
    for (int i = threadIdx.x; i < a.P; i += kTileBlock) lcount[i] = a.count[i];
  // per-slice scale exponents (uniform across the workgroup)
  int ex[SB];
  bool fx[SB];
  bool all_fixed = true;
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int s = b * SB + q;
    if (MODE == kBpSino)
      ex[q] = weight_exponent((s < a.S) ? __uint_as_float(a.maxw[s]) : 0.f, a.denmax, a.maxlen,
                              fx[q]);
    else
      ex[q] = fixed_exponent(1.0f, a.maxlen, fx[q]);
    all_fixed &= fx[q];
  }
  __syncthreads();
  if (MODE == kBpCount || all_fixed) {
    bp_tile_loop<TW, TH, SB, MODE, kExact, true, false>(a, acc, count, tx0, tx1, ty0, ty1, b, ex,
                                                        fx);
    if (tx0 == 0 && a.wraps)
      bp_tile_loop<TW, TH, SB, MODE, kExact, true, true>(a, acc, count, tx0, tx1, ty0, ty1, b,
                                                         ex, fx);
  } else {
    bp_tile_loop<TW, TH, SB, MODE, kExact, false, false>(a, acc, count, tx0, tx1, ty0, ty1, b,
                                                         ex, fx);
    if (tx0 == 0 && a.wraps)
      bp_tile_loop<TW, TH, SB, MODE, kExact, false, true>(a, acc, count, tx0, tx1, ty0, ty1, b,
                                                          ex, fx);
  }
  __syncthreads();
  // coalesced write-out: plane by plane, row-contiguous
  const int w_ = tx1 - tx0, h_ = ty1 - ty0;
  if (BpPacked<MODE, SB>::value) {
    const unsigned long long *acc64 = reinterpret_cast<const unsigned long long *>(acc);
    const unsigned *nterms = reinterpret_cast<const unsigned *>(acc) + SB * PLANE;
#pragma unroll
    for (int k2 = 0; k2 < SB / 2; ++k2) {
      for (int i = threadIdx.x; i < w_ * h_; i += kTileBlock) {
        const int ly = i / w_, lx = i - ly * w_;
        unsigned long long v = acc64[k2 * PLANE + ly * PITCH + lx];
        // remove the magic offsets: n * (2^32 + 1) * bits(1.5 * 2^23), mod 2^64
        if (all_fixed)
          v -= (unsigned long long)nterms[ly * PITCH + lx] *
               (((unsigned long long)kRoundMagicBits << 32) | kRoundMagicBits);
        const int lo = (int)(unsigned)v;
        const int hi = (int)(((long long)v - (long long)lo) >> 32);
        const int64_t o = (int64_t)(ty0 + ly) * N + tx0 + lx;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = 2 * k2 + h;
          const int s = b * SB + q;
          if (s >= a.S) break;
          const unsigned half = h ? (unsigned)(v >> 32) : (unsigned)v;
          float out;
          if (!fx[q])
            out = __uint_as_float(half);
          else
            out = (float)ldexp((double)(all_fixed ? (h ? hi : lo) : (int)half), -ex[q]);
          a.z[(int64_t)s * N * N + o] = out;
        }
      }
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int s = b * SB + q;
    if (MODE == kBpSino && s >= a.S) break;
    float *zs = a.z + (int64_t)s * N * N;
    for (int i = threadIdx.x; i < w_ * h_; i += kTileBlock) {
      const int ly = i / w_, lx = i - ly * w_;
      const BpWord<MODE> v = acc[q * PLANE + ly * PITCH + lx];
      float out;
      if (MODE == kBpCount)
        out = (float)v;
      else if (!fx[q])
        out = __uint_as_float((unsigned)v);
      else if (MODE == kBpSino)
        out = (float)ldexp((double)(int)v, -ex[q]);
      else
        out = (float)ldexp((double)(long long)v, -ex[q]);
      zs[(int64_t)(ty0 + ly) * N + tx0 + lx] = out;
    }
  }
}

template <int TW, int TH, int SB, int MODE, bool kExact>
__global__ void __launch_bounds__(kTileBlock)
k_bp_tile(TileArgs a) {
  extern __shared__ unsigned long long acc[];
  bp_tile_body<TW, TH, SB, MODE, kExact>(a, blockIdx.x, blockIdx.y, acc);
}

// Per-slice maximum of |w| over w [P][C][S] (float bits, atomicMax on the
// non-negative bit patterns; NaN sorts above inf).  maxw must be zeroed.
__global__ void __launch_bounds__(kPixBlock)
k_absmax_slices(const float *__restrict__ w, int P, int S, int C, unsigned *__restrict__ maxw) {
  const int s = blockIdx.y;
  const int64_t n = (int64_t)P * C;
  unsigned m = 0;
  for (int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kPixBlock)
    m = max(m, __float_as_uint(__builtin_fabsf(w[i * S + s])));
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_down((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(maxw + s, m);
}

// ------------------------------------------------------- pack / update --

// xb[t/SB][j][t%SB] = src[(t + off)][j] for t < S_src_used, ghost value
// outside: slots whose source slice index is < 0 or >= nsrc get `fill`.
template <int SB>
__global__ void __launch_bounds__(kPixBlock)
k_pack(const float *__restrict__ src, int nsrc, int off, float fill, int64_t NN, int nslots,
       float *__restrict__ xb) {
  const int64_t j = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  const int b = blockIdx.y;
  if (j >= NN) return;
  float v[SB];
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int t = b * SB + q;
    const int s = t + off;
    v[q] = (t < nslots && s >= 0 && s < nsrc) ? src[(int64_t)s * NN + j] : fill;
  }
  float *dst = xb + ((int64_t)b * NN + j) * SB;
  store_slices<SB>(dst, v);
}

// SIRT (sirt.cc:21-33) / MLEM (mlem.cc:21-34) update of a batch of slices
// at pixel j, also refreshing the batched copy.
template <int SB, bool kSIRT>
__global__ void __launch_bounds__(kPixBlock)
k_update_em(float *__restrict__ x, const float *__restrict__ z, const float *__restrict__ den,
            int S, int64_t NN, float *__restrict__ xb) {
  const int64_t j = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  const int b = blockIdx.y;
  if (j >= NN) return;
  const float d = den[j];
  float v[SB];
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int s = b * SB + q;
    v[q] = 0.f;
    if (s < S) {
      const int64_t o = (int64_t)s * NN + j;
      float xv = x[o];
      if (kSIRT) xv += z[o] / d;
      else xv *= z[o] / d;
      x[o] = xv;
      v[q] = xv;
    }
  }
  float *dst = xb + ((int64_t)b * NN + j) * SB;
  store_slices<SB>(dst, v);
}

__global__ void k_fill(float *__restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// Pixel classes and neighbour orders of the regularisation stencils:
// 0 inner, 1 top row, 2 bottom row, 3 left column, 4 right column,
// 5..8 corners top-left, top-right, bottom-left, bottom-right.  The in-plane
// neighbour order per class is the reference's q order (pml.cc:45-53,
// 69-74, 87-92, 105-110, 123-128, 141-144, 156-159, 171-174, 187-190; the
// same orders appear in apmlr.cc).
__constant__ int8_t c_nb[9][8][2] = {
    {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}},
    {{1, 0}, {-1, 0}, {0, 1}, {1, 1}, {-1, 1}},
    {{1, 0}, {-1, 0}, {0, -1}, {1, -1}, {-1, -1}},
    {{1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}},
    {{-1, 0}, {0, 1}, {0, -1}, {-1, 1}, {-1, -1}},
    {{1, 0}, {0, 1}, {1, 1}},
    {{-1, 0}, {0, 1}, {-1, 1}},
    {{1, 0}, {0, -1}, {1, -1}},
    {{-1, 0}, {0, -1}, {-1, -1}},
};
__constant__ int c_nbn[9] = {8, 5, 5, 5, 5, 3, 3, 3, 3};

// PML weights, pml.h:22-29
__constant__ float c_pml_w[3][8] = {
    {0.1464466094f, 0.1464466094f, 0.1464466094f, 0.1464466094f,
     0.10355339059f, 0.10355339059f, 0.10355339059f, 0.10355339059f},
    {0.226540919667f, 0.226540919667f, 0.226540919667f, 0.1601886205f,
     0.1601886205f},
    {0.36939806251f, 0.36939806251f, 0.26120387496f}};

__device__ __forceinline__ int pixel_class(int m, int n, int N) {
  const bool top = (n == 0), bot = (n == N - 1), left = (m == 0),
             right = (m == N - 1);
  if (!(top | bot | left | right)) return 0;
  if (top && left) return 5;
  if (top && right) return 6;
  if (bot && left) return 7;
  if (bot && right) return 8;
  if (top) return 1;
  if (bot) return 2;
  if (left) return 3;
  return 4;
}

// In-plane neighbour offsets and PML weights of an inner pixel, in the
// reference's q order (pml.cc:45-53, the first row of c_nb and c_pml_w),
// as compile-time constants: inner pixels (all but the border) skip the
// per-class tables.
constexpr int kInnerDx[8] = {1, -1, 0, 0, 1, -1, 1, -1};
constexpr int kInnerDy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
constexpr float kPmlAxial = 0.1464466094f;
constexpr float kPmlDiag = 0.10355339059f;

// PML stencil and update of the SB slices of batch b at one pixel.
template <int SB, bool kInner>
__device__ __forceinline__ void pml_pixel(const float *__restrict__ x, float *__restrict__ xn,
                                          const float *__restrict__ z, int S, int N,
                                          int64_t NN, int64_t pix, int cls, float d,
                                          float beta, int num_direct, int b, float (&v)[SB]) {
  const float *w = c_pml_w[cls == 0 ? 0 : (cls <= 4 ? 1 : 2)];
  const int cnt = c_nbn[cls];
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int s = b * SB + q;
    v[q] = 0.f;
    if (s < S) {
      const int64_t j = (int64_t)s * NN + pix;
      const float x0 = x[j];
      float f = 0.f, gg = 0.f;
      if (kInner) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float xq = x[j + kInnerDx[k] + (int64_t)kInnerDy[k] * N];
          const float wk = k < 4 ? kPmlAxial : kPmlDiag;
          f += 2.0f * beta * wk;
          gg -= 2.0f * beta * wk * (x0 + xq);
        }
      } else {
        for (int k = 0; k < cnt; ++k) {
          const float xq = x[j + c_nb[cls][k][0] + (int64_t)c_nb[cls][k][1] * N];
          f += 2.0f * beta * w[k];
          gg -= 2.0f * beta * w[k] * (x0 + xq);
        }
      }
      const float G = gg + d;
      const float F = f;
      const float nm = num_direct ? z[j] : -(x0 * z[j]);
      const float xv = (-G + __builtin_sqrtf(G * G - 8.0f * nm * F)) / (4.0f * F);
      xn[j] = xv;
      v[q] = xv;
    }
  }
}

// Fused PML stencil (pml.cc:25-201, F and G from zero) and closed-form
// update (pml.cc:3-23):  x' = (-G + sqrt(G^2 - 8 num F)) / (4F).  With
// num_direct, z holds num as the reference sums it (ordered.hip); otherwise
// z = A^T w and num = -x * z (the x factor of pml.cc:243 taken out of the
// sum).
// A thread owns pixel (m, n) of the SB slices of slot batch b and also
// writes the slice-batched copy xb of the new image for the next ray pass
// (slot t = slice t; padding slots hold 0).
template <int SB>
__global__ void __launch_bounds__(kPixBlock)
k_update_pml(const float *__restrict__ x, float *__restrict__ xn,
             const float *__restrict__ z, const float *__restrict__ den,
             int S, int N, float beta, int num_direct, float *__restrict__ xb) {
  const int m = blockIdx.x * kPixBlock + threadIdx.x;
  const int n = blockIdx.y;
  const int b = blockIdx.z;
  if (m >= N) return;
  const int64_t NN = (int64_t)N * N;
  const int64_t pix = (int64_t)n * N + m;
  const int cls = pixel_class(m, n, N);
  const float d = den[pix];
  float v[SB];
  if (cls == 0)
    pml_pixel<SB, true>(x, xn, z, S, N, NN, pix, cls, d, beta, num_direct, b, v);
  else
    pml_pixel<SB, false>(x, xn, z, S, N, NN, pix, cls, d, beta, num_direct, b, v);
  store_slices<SB>(xb + ((int64_t)b * NN + pix) * SB, v);
}

struct ApmlrArgs {
  StencilWeights w;
  RegParams p;
  int compat;
  int nslots;      // slots of the batched copy
  int num_direct;  // z holds num (ordered.hip), not A^T w
};

// APMLR stencil and update of the slots of batch b at one pixel (see
// k_update_apmlr).
template <int SB, bool kInner>
__device__ __forceinline__ void apmlr_pixel(const float *__restrict__ xp,
                                            float *__restrict__ xnp,
                                            const float *__restrict__ z, int S, int N,
                                            int64_t NN, int64_t pix, int cls, float d,
                                            const ApmlrArgs &a, int b, float (&v)[SB]) {
  const int cnt = kInner ? 8 : c_nbn[cls];
  const int off = a.compat ? 0 : 1;
  const float tb = 2.0f * a.p.beta;
  const float tb1 = 2.0f * a.p.beta1;
#pragma unroll
  for (int q = 0; q < SB; ++q) {
    const int pi = b * SB + q + off;  // padded slice of slot b*SB+q
    float val = 0.f;
    if (pi >= 1 && pi <= S) {
      const int s = pi - 1;
      const int64_t jp = (int64_t)pi * NN + pix;
      const int64_t j = (int64_t)s * NN + pix;
      const bool top = (s == 0), bottom = (s == S - 1 && S - 1 != 0);
      const int zmode = top ? 1 : (bottom ? 2 : 0);
      const int nz = zmode == 0 ? 2 : 1;
      const float *w = a.w.w[zmode][kInner ? 0 : cls];
      const float x0 = xp[jp];
      float f = 0.f, gg = 0.f;
      if (kInner) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float xq = xp[jp + kInnerDx[k] + (int64_t)kInnerDy[k] * N];
          const float mg = x0 + xq;
          const float rg = x0 - xq;
          const float gm = 1.0f / (1.0f + __builtin_fabsf(rg / a.p.delta));
          f += tb * w[k] * gm;
          gg -= tb * w[k] * gm * mg;
        }
      } else {
        for (int k = 0; k < cnt; ++k) {
          const float xq = xp[jp + c_nb[cls][k][0] + (int64_t)c_nb[cls][k][1] * N];
          const float mg = x0 + xq;
          const float rg = x0 - xq;
          const float gm = 1.0f / (1.0f + __builtin_fabsf(rg / a.p.delta));
          f += tb * w[k] * gm;
          gg -= tb * w[k] * gm * mg;
        }
      }
      for (int k = 0; k < nz; ++k) {
        const int64_t dz = (zmode == 2 || k == 1) ? -NN : NN;
        const float xq = xp[jp + dz];
        const float mg = x0 + xq;
        const float rg = x0 - xq;
        const float gm = 1.0f / (1.0f + __builtin_fabsf(rg / a.p.delta1));
        f += tb1 * w[cnt + k] * gm;
        gg -= tb1 * w[cnt + k] * gm * mg;
      }
      const float G = gg + d;
      const float F = f;
      const float xpass = a.compat ? xp[jp - NN] : x0;
      const float nm = a.num_direct ? z[j] : -(xpass * z[j]);
      val = (-G + __builtin_sqrtf(G * G - 8.0f * nm * F)) / (4.0f * F);
      xnp[jp] = val;
    } else if (pi == 0 || pi == S + 1) {
      val = xp[(int64_t)pi * NN + pix];
    }
    v[q] = val;
  }
}

// Fused APMLR stencil (apmlr.cc:919-982 with FGInner/FGTop/FGBottom) and
// update (apmlr.cc:3-26) over ghost-padded images [S+2][N][N]; real slice s
// lives at padded index s+1.  The ray pass of slice s used padded slice s
// (compat, apmlr.cc:1082-1084) or s+1, which also supplies num's x factor.
// Thread (pixel, slot batch b): slot t of the batched copy holds padded
// slice t + off (off = 0 compat, 1 otherwise); real slices are updated,
// ghosts (padded 0 and S+1, constant) are copied, slots past S+1 hold 0.
// The grid covers every real slice even when its slot lies past nslots.
template <int SB>
__global__ void __launch_bounds__(kPixBlock)
k_update_apmlr(const float *__restrict__ xp, float *__restrict__ xnp,
               const float *__restrict__ z, const float *__restrict__ den,
               int S, int N, ApmlrArgs a, float *__restrict__ xb) {
  const int m = blockIdx.x * kPixBlock + threadIdx.x;
  const int n = blockIdx.y;
  const int b = blockIdx.z;
  if (m >= N) return;
  const int64_t NN = (int64_t)N * N;
  const int64_t pix = (int64_t)n * N + m;
  const int cls = pixel_class(m, n, N);
  const float d = den[pix];
  float v[SB];
  if (cls == 0)
    apmlr_pixel<SB, true>(xp, xnp, z, S, N, NN, pix, cls, d, a, b, v);
  else
    apmlr_pixel<SB, false>(xp, xnp, z, S, N, NN, pix, cls, d, a, b, v);
  if (b * SB < a.nslots) store_slices<SB>(xb + ((int64_t)b * NN + pix) * SB, v);
}

// ------------------------------------------------------------- batched --
//
// Persistent kernels over the work queues of a batch of slices with
// different geometries (BatchQueues).  Each workgroup takes an item, finds
// its slice by binary search in the phase's start table and runs the same
// per-item code as the single-geometry kernels (one slice per geometry).

__device__ __forceinline__ int batch_next_item(unsigned *counter) {
  __shared__ unsigned s_item;
  __syncthreads();  // every lane has read the previous item
  if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane((int)s_item);
}

__device__ __forceinline__ int batch_find(const int *__restrict__ start, int n, int item) {
  int lo = 0, hi = n;  // start[lo] <= item < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void k_batch_reset(unsigned *__restrict__ counters, unsigned *__restrict__ maxw,
                              int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3) counters[i] = 0u;
  if (i < n) maxw[i] = 0u;
}

template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_forward_batch(const BatchSliceDev *__restrict__ d, BatchQueues q, int it, int ba_log2) {
  const int *start = q.start[0];
  const int total = start[q.n];
  for (;;) {
    const int item = batch_next_item(q.counters + 0);
    if (item >= total) break;
    const int k = batch_find(start, q.n, item);
    const BatchSliceDev &sd = d[__builtin_amdgcn_readfirstlane(q.order[0][k])];
    if (it >= sd.iters) continue;
    const int64_t bundle = (int64_t)(item - start[k]) * (kRayBlock / 64) + (threadIdx.x >> 6);
    if (sd.alg == 0)
      forward_bundle<1, kFwdSIRT, kExact>(sd.g, sd.ang, sd.rays, sd.P, 1, ba_log2, sd.x,
                                          sd.sino, sd.w, sd.maxw, bundle, 0);
    else
      forward_bundle<1, kFwdEM, kExact>(sd.g, sd.ang, sd.rays, sd.P, 1, ba_log2, sd.x, sd.sino,
                                        sd.w, sd.maxw, bundle, 0);
  }
}

template <int TW, int TH, bool kExact>
__global__ void __launch_bounds__(kTileBlock)
k_bp_batch(const BatchSliceDev *__restrict__ d, BatchQueues q, int it) {
  extern __shared__ unsigned long long acc[];  // [TH][TW + 1], then int count[P]
  const int *start = q.start[1];
  const int total = start[q.n];
  for (;;) {
    const int item = batch_next_item(q.counters + 1);
    if (item >= total) break;
    const int k = batch_find(start, q.n, item);
    const BatchSliceDev &sd = d[__builtin_amdgcn_readfirstlane(q.order[1][k])];
    if (it >= sd.iters) continue;
    TileArgs a;
    a.g = sd.g;
    a.ang = sd.ang;
    a.rays = sd.rays;
    a.count = sd.count;
    a.P = sd.P;
    a.S = 1;
    a.count_lds = 1;
    a.yi0 = sd.yi0;
    a.w = sd.w;
    a.maxw = sd.maxw;
    a.maxlen = sd.maxlen;
    a.denmax = sd.denmax;
    a.z = sd.z;
    a.wraps = sd.wraps;
    bp_tile_body<TW, TH, 1, kBpSino, kExact>(a, item - start[k], 0, acc);
  }
}

// SIRT (sirt.cc:21-33) / MLEM (mlem.cc:21-34) update, kBatchUpdateBlocks
// blocks of kPixBlock pixels per item (one item per block made the shared
// item counter, ~1.2 million atomics per iteration at 200 uneven slices,
// cost 14 ms where the pixels take 1.6 ms).
__global__ void __launch_bounds__(kPixBlock)
k_update_batch(const BatchSliceDev *__restrict__ d, BatchQueues q, int it) {
  const int *start = q.start[2];
  const int total = start[q.n];
  for (;;) {
    const int item = batch_next_item(q.counters + 2);
    if (item >= total) break;
    const int k = batch_find(start, q.n, item);
    const BatchSliceDev &sd = d[__builtin_amdgcn_readfirstlane(q.order[2][k])];
    if (it >= sd.iters) continue;
    const int64_t NN = (int64_t)sd.g.N * sd.g.N;
    const int64_t j0 = (int64_t)(item - start[k]) * kPixBlock * kBatchUpdateBlocks + threadIdx.x;
#pragma unroll 4
    for (int u = 0; u < kBatchUpdateBlocks; ++u) {
      const int64_t j = j0 + (int64_t)u * kPixBlock;
      if (j < NN) {
        float xv = sd.x[j];
        if (sd.alg == 0) xv += sd.z[j] / sd.den[j];
        else xv *= sd.z[j] / sd.den[j];
        sd.x[j] = xv;
      }
    }
  }
}

// ------------------------------------------------- reference / debugging --

// Whole-ray scatter with global atomics (z must be zeroed): the plain
// ray-driven adjoint, kept as an independent check of the tile kernel.
template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_adjoint_atomic(GridInfo g, const AngleInfo *__restrict__ ang,
                 const RayInfo *__restrict__ rays, int P, int S, int mode,
                 const float *__restrict__ r, float *__restrict__ z) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  if (t >= (int64_t)P * C) return;
  const int p = (int)(t / C);
  const int s = blockIdx.y;
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  float w = 1.f;
  if (mode == kBpSino) {
    w = r[t * S + s];  // [P][C][S]
    if (w == 0.f) return;
  }
  float *zs = z + (int64_t)s * NN;
  traverse_full<kExact>(g, rays[t], ang[p].quadrant, [&](int ix, int iy, float l, float) {
    const unsigned idx = (unsigned)(ix + iy * g.N);
    if (idx < NN) atomic_add_f32(zs + idx, mode == kBpCount ? 1.0f : (mode == kBpOnes ? l : l * w));
  });
}

template <bool kExact>
__global__ void k_dump_rays(GridInfo g, const AngleInfo *__restrict__ ang,
                            const RayInfo *__restrict__ rays, int n,
                            const int *__restrict__ proj, const int *__restrict__ col,
                            int max_seg, int *__restrict__ counts, int *__restrict__ idx,
                            float *__restrict__ len, float *__restrict__ len2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int k = 0;
  const int64_t base = (int64_t)t * max_seg;
  const int64_t ray = (int64_t)proj[t] * g.C + col[t];
  const int cnt = traverse_full<kExact>(g, rays[ray], ang[proj[t]].quadrant,
                                        [&](int ix, int iy, float l, float l2) {
                                          if (k < max_seg) {
                                            idx[base + k] = ix + iy * g.N;
                                            len[base + k] = l;
                                            len2[base + k] = l2;
                                          }
                                          ++k;
                                        });
  counts[t] = cnt;
}

// Per-ray sums over all segments, in the walk order: len_sum = sum len
// (row sums of A) and len2_sum = sum len^2 (the SIRT normaliser a2 of
// sirt.cc:47-50, accumulated from the squared lengths as there).  Output
// [P][C]; either pointer may be null.
template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_ray_sums(GridInfo g, const AngleInfo *__restrict__ ang, const RayInfo *__restrict__ rays,
           int P, float *__restrict__ len_sum, float *__restrict__ len2_sum) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  if (t >= (int64_t)P * g.C) return;
  const int p = (int)(t / g.C);
  float a1 = 0.f, a2 = 0.f;
  traverse_full<kExact>(g, rays[t], ang[p].quadrant, [&](int, int, float l, float l2) {
    a1 += l;
    a2 += l2;
  });
  if (len_sum) len_sum[t] = a1;
  if (len2_sum) len2_sum[t] = a2;
}

// Segment statistics of one slice: total count, count outside the image,
// and the bits of the longest segment under either length rounding.
__global__ void __launch_bounds__(kRayBlock)
k_count_segments(GridInfo g, const AngleInfo *__restrict__ ang,
                 const RayInfo *__restrict__ rays, int P,
                 unsigned long long *total, unsigned long long *oor, unsigned *maxlen,
                 unsigned long long *wraps) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  unsigned cnt = 0, bad = 0, wr = 0;
  float lmax = 0.f;
  if (t < (int64_t)P * C) {
    const int p = (int)(t / C);
    const unsigned NN = (unsigned)g.N * (unsigned)g.N;
    const RayInfo r = rays[t];
    wr = (r.alo & kRayWraps) ? 1u : 0u;
    cnt = traverse_full<true>(g, r, ang[p].quadrant, [&](int ix, int iy, float l, float l2) {
      bad += ((unsigned)(ix + iy * g.N) >= NN) ? 1u : 0u;
      lmax = fmaxf(lmax, fmaxf(l, __builtin_amdgcn_sqrtf(l2)));
    });
  }
  unsigned lb = __float_as_uint(lmax);
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_down(cnt, off, 64);
    bad += __shfl_down(bad, off, 64);
    wr += __shfl_down(wr, off, 64);
    lb = max(lb, (unsigned)__shfl_down((int)lb, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(total, (unsigned long long)cnt);
    if (bad) atomicAdd(oor, (unsigned long long)bad);
    if (wr && wraps) atomicAdd(wraps, (unsigned long long)wr);
    atomicMax(maxlen, lb);
  }
}

__global__ void k_check_finite(const float *__restrict__ v, int64_t n,
                               int *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kPixBlock;
  int bad = 0;
  for (; i < n; i += stride) bad |= !__builtin_isfinite(v[i]);
  if (bad) atomicOr(flag, 1);
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------ launch-shape dispatch --

template <int SB, bool E, int LPR>
hipError_t fwd_mode(int mode, const Geometry &g, const RayInfo *rays, int ba_log2,
                    const float *xb, const float *sino, float *out, unsigned *maxw,
                    hipStream_t st) {
  const int bc = (64 / LPR) >> ba_log2, ba = 1 << ba_log2;
  if (bc < 1) return hipErrorInvalidValue;
  const int64_t bundles = (int64_t)((g.P + ba - 1) / ba) * ((g.C + bc - 1) / bc);
  dim3 grid(blocks_for(bundles, kRayBlock / 64), (g.S + SB - 1) / SB), block(kRayBlock);
  if (mode == kFwdY)
    hipLaunchKernelGGL((k_forward<SB, kFwdY, E, LPR>), grid, block, 0, st, g.grid, g.d_ang,
                       rays, g.P, g.S, ba_log2, xb, sino, out, maxw);
  else if (mode == kFwdSIRT)
    hipLaunchKernelGGL((k_forward<SB, kFwdSIRT, E, LPR>), grid, block, 0, st, g.grid, g.d_ang,
                       rays, g.P, g.S, ba_log2, xb, sino, out, maxw);
  else
    hipLaunchKernelGGL((k_forward<SB, kFwdEM, E, LPR>), grid, block, 0, st, g.grid, g.d_ang,
                       rays, g.P, g.S, ba_log2, xb, sino, out, maxw);
  return hipGetLastError();
}

template <int SB, int LPR>
hipError_t fwd_exact(bool exact, int mode, const Geometry &g, const RayInfo *rays, int ba_log2,
                     const float *xb, const float *sino, float *out, unsigned *maxw,
                     hipStream_t st) {
  return exact ? fwd_mode<SB, true, LPR>(mode, g, rays, ba_log2, xb, sino, out, maxw, st)
               : fwd_mode<SB, false, LPR>(mode, g, rays, ba_log2, xb, sino, out, maxw, st);
}

constexpr size_t kLdsBytes = 160 * 1024;

template <int TW, int TH, int SB, int MODE, bool E>
hipError_t bp_launch(TileArgs a, int nbatch, hipStream_t st) {
  const int N = a.g.N;
  const int tiles = ((N + TW - 1) / TW) * ((N + TH - 1) / TH);
  size_t lds = (size_t)bp_acc_words<TW, TH, SB, MODE>() * sizeof(BpWord<MODE>);
  // stage the per-angle column counts in LDS when they fit
  a.count_lds = (lds + (size_t)a.P * sizeof(int) <= kLdsBytes / 2) ? 1 : 0;
  if (a.count_lds) lds += (size_t)a.P * sizeof(int);
  auto kern = k_bp_tile<TW, TH, SB, MODE, E>;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(tiles, nbatch), dim3(kTileBlock), lds, st, a);
  return hipGetLastError();
}

template <int TW, int TH, int SB>
hipError_t bp_mode(int mode, bool exact, const TileArgs &a, int nbatch, hipStream_t st) {
  if (mode == kBpSino)
    return exact ? bp_launch<TW, TH, SB, kBpSino, true>(a, nbatch, st)
                 : bp_launch<TW, TH, SB, kBpSino, false>(a, nbatch, st);
  if (mode == kBpOnes)
    return exact ? bp_launch<TW, TH, 1, kBpOnes, true>(a, 1, st)
                 : bp_launch<TW, TH, 1, kBpOnes, false>(a, 1, st);
  return bp_launch<TW, TH, 1, kBpCount, true>(a, 1, st);
}

}  // namespace

// ------------------------------------------------------------ launchers --

int tile_count_bound(float sinq, float cosq, int TW, int TH) {
  // columns in an interval of length TW|sin| + TH|cos| + 2 kColumnMargin
  // (plus rounding of the float projections, far below the 2 x 0.05 pixel
  // that kCountMargin adds): at most floor(len) + 1.
  // At most ceil(sqrt(TW^2 + TH^2)) + 4, the column ranks bp_tile_loop
  // enumerates for one-slice tiles (ensure_count checks it).
  const double len =
      TW * std::fabs((double)sinq) + TH * std::fabs((double)cosq) + 2.0 * kCountMargin;
  return (int)std::floor(len) + 1;
}

hipError_t launch_ray_info(const Geometry &g, RayInfo *rays, hipStream_t st, bool flags) {
  const int64_t n = g.rays_per_slice();
  hipLaunchKernelGGL(k_ray_info, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0, st,
                     g.grid, g.d_ang, g.d_yi, g.P, flags ? 1 : 0, rays);
  return hipGetLastError();
}

hipError_t launch_forward(const Geometry &g, const RayInfo *rays, int sb, int bundle_angles,
                          int lanes_per_ray, int mode, bool exact, const float *xb,
                          const float *sino, float *out, unsigned *maxw, hipStream_t st) {
  int ba_log2 = 0;
  while ((1 << ba_log2) < bundle_angles && ba_log2 < 6) ++ba_log2;
  if ((1 << ba_log2) != bundle_angles) return hipErrorInvalidValue;
#define TRACE_FWD_CASE(SB_, LPR_)                                                    \
  if (sb == SB_ && lanes_per_ray == LPR_)                                            \
    return fwd_exact<SB_, LPR_>(exact, mode, g, rays, ba_log2, xb, sino, out, maxw, st);
  TRACE_FWD_SHAPES(TRACE_FWD_CASE)
#undef TRACE_FWD_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_backproject(const Geometry &g, const RayInfo *rays, const int *count,
                              const TileShape &ts, int mode, bool exact, const float *w,
                              const unsigned *maxw, float maxlen, float denmax, float *z,
                              bool wraps, hipStream_t st) {
  TileArgs a;
  a.wraps = wraps ? 1 : 0;
  a.maxw = maxw;
  a.maxlen = maxlen;
  a.denmax = denmax;
  a.g = g.grid;
  a.ang = g.d_ang;
  a.rays = rays;
  a.count = count;
  a.P = g.P;
  a.S = g.S;
  a.yi0 = g.yi0;
  a.w = w;
  a.z = z;
  const int nb = (g.S + ts.sb - 1) / ts.sb;
#define TRACE_BP_CASE(TW_, TH_, SB_)                              \
  if (ts.tw == TW_ && ts.th == TH_ && ts.sb == SB_)               \
    return bp_mode<TW_, TH_, SB_>(mode, exact, a, nb, st);
  TRACE_BP_SHAPES(TRACE_BP_CASE)
#undef TRACE_BP_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_pack(const float *src, int nsrc, int off, float fill, int64_t NN, int nslots,
                       int sb, float *xb, hipStream_t st) {
  const int nb = (nslots + sb - 1) / sb;
  dim3 grid(blocks_for(NN, kPixBlock), nb), block(kPixBlock);
  switch (sb) {
    case 1: hipLaunchKernelGGL(k_pack<1>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    case 2: hipLaunchKernelGGL(k_pack<2>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    case 4: hipLaunchKernelGGL(k_pack<4>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    case 8: hipLaunchKernelGGL(k_pack<8>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    case 16: hipLaunchKernelGGL(k_pack<16>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    case 32: hipLaunchKernelGGL(k_pack<32>, grid, block, 0, st, src, nsrc, off, fill, NN, nslots, xb); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_update_em(bool sirt, float *x, const float *z, const float *den, int S,
                            int64_t NN, int sb, float *xb, hipStream_t st) {
  const int nb = (S + sb - 1) / sb;
  dim3 grid(blocks_for(NN, kPixBlock), nb), block(kPixBlock);
#define TRACE_UPD_CASE(SB_)                                                                   \
  case SB_:                                                                                   \
    if (sirt)                                                                                 \
      hipLaunchKernelGGL((k_update_em<SB_, true>), grid, block, 0, st, x, z, den, S, NN, xb); \
    else                                                                                      \
      hipLaunchKernelGGL((k_update_em<SB_, false>), grid, block, 0, st, x, z, den, S, NN, xb); \
    break;
  switch (sb) {
    TRACE_UPD_CASE(1)
    TRACE_UPD_CASE(2)
    TRACE_UPD_CASE(4)
    TRACE_UPD_CASE(8)
    TRACE_UPD_CASE(16)
    TRACE_UPD_CASE(32)
    default: return hipErrorInvalidValue;
  }
#undef TRACE_UPD_CASE
  return hipGetLastError();
}

hipError_t launch_fill(float *p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, kPixBlock)), dim3(kPixBlock), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_update_pml(const float *x, float *xn, const float *z, const float *den,
                             int S, int N, float beta, bool num_direct, int sb, int nslots,
                             float *xb, hipStream_t st) {
  dim3 grid(blocks_for(N, kPixBlock), N, nslots / sb);
#define TRACE_PML_CASE(SB_)                                                              \
  case SB_:                                                                              \
    hipLaunchKernelGGL(k_update_pml<SB_>, grid, dim3(kPixBlock), 0, st, x, xn, z, den, S, \
                       N, beta, num_direct ? 1 : 0, xb);                                  \
    break;
  switch (sb) {
    TRACE_PML_CASE(1)
    TRACE_PML_CASE(2)
    TRACE_PML_CASE(4)
    TRACE_PML_CASE(8)
    TRACE_PML_CASE(16)
    TRACE_PML_CASE(32)
    default: return hipErrorInvalidValue;
  }
#undef TRACE_PML_CASE
  return hipGetLastError();
}

hipError_t launch_update_apmlr(const float *xpad, float *xnpad, const float *z,
                               const float *den, int S, int N, const RegParams &p,
                               const StencilWeights &w, int compat, bool num_direct,
                               int sb, int nslots, float *xb, hipStream_t st) {
  // batches: every slot of the copy and every real slice's slot
  const int off = compat ? 0 : 1;
  const int need = (nslots > S + 1 - off) ? nslots : S + 1 - off;
  dim3 grid(blocks_for(N, kPixBlock), N, (need + sb - 1) / sb);
  ApmlrArgs a{w, p, compat, nslots, num_direct ? 1 : 0};
#define TRACE_APMLR_CASE(SB_)                                                         \
  case SB_:                                                                           \
    hipLaunchKernelGGL(k_update_apmlr<SB_>, grid, dim3(kPixBlock), 0, st, xpad, xnpad, z, \
                       den, S, N, a, xb);                                              \
    break;
  switch (sb) {
    TRACE_APMLR_CASE(1)
    TRACE_APMLR_CASE(2)
    TRACE_APMLR_CASE(4)
    TRACE_APMLR_CASE(8)
    TRACE_APMLR_CASE(16)
    TRACE_APMLR_CASE(32)
    default: return hipErrorInvalidValue;
  }
#undef TRACE_APMLR_CASE
  return hipGetLastError();
}

hipError_t launch_batch_reset(const BatchQueues &q, hipStream_t st) {
  const int n = q.n > 3 ? q.n : 3;
  hipLaunchKernelGGL(k_batch_reset, dim3((n + 255) / 256), dim3(256), 0, st, q.counters,
                     q.maxw_all, q.n);
  return hipGetLastError();
}

hipError_t launch_batch_forward(const BatchSliceDev *d, const BatchQueues &q, int it,
                                int bundle_angles, bool exact, int grid, hipStream_t st) {
  int ba_log2 = 0;
  while ((1 << ba_log2) < bundle_angles && ba_log2 < 6) ++ba_log2;
  if ((1 << ba_log2) != bundle_angles || grid <= 0) return hipErrorInvalidValue;
  if (exact)
    hipLaunchKernelGGL(k_forward_batch<true>, dim3(grid), dim3(kRayBlock), 0, st, d, q, it,
                       ba_log2);
  else
    hipLaunchKernelGGL(k_forward_batch<false>, dim3(grid), dim3(kRayBlock), 0, st, d, q, it,
                       ba_log2);
  return hipGetLastError();
}

hipError_t launch_batch_backproject(const BatchSliceDev *d, const BatchQueues &q, int it,
                                    int tw, int th, bool exact, int grid, size_t lds,
                                    hipStream_t st) {
  if (grid <= 0) return hipErrorInvalidValue;
#define TRACE_BATCH_BP(TW_, TH_)                                                             \
  if (tw == TW_ && th == TH_) {                                                              \
    if (exact)                                                                               \
      hipLaunchKernelGGL((k_bp_batch<TW_, TH_, true>), dim3(grid), dim3(kTileBlock), lds, st, \
                         d, q, it);                                                          \
    else                                                                                     \
      hipLaunchKernelGGL((k_bp_batch<TW_, TH_, false>), dim3(grid), dim3(kTileBlock), lds,    \
                         st, d, q, it);                                                      \
    return hipGetLastError();                                                                \
  }
  TRACE_BATCH_TILES(TRACE_BATCH_BP)
#undef TRACE_BATCH_BP
  return hipErrorInvalidValue;
}

hipError_t launch_batch_update(const BatchSliceDev *d, const BatchQueues &q, int it, int grid,
                               hipStream_t st) {
  if (grid <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_update_batch, dim3(grid), dim3(kPixBlock), 0, st, d, q, it);
  return hipGetLastError();
}

bool batch_tile_supported(int tw, int th) {
#define TRACE_BATCH_MATCH(TW_, TH_) \
  if (tw == TW_ && th == TH_) return true;
  TRACE_BATCH_TILES(TRACE_BATCH_MATCH)
#undef TRACE_BATCH_MATCH
  return false;
}

int batch_blocks_per_cu(int phase, int tw, int th, bool exact, size_t lds) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (phase == 0) {
    e = exact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward_batch<true>,
                                                             kRayBlock, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_forward_batch<false>,
                                                             kRayBlock, 0);
  } else if (phase == 2) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_update_batch, kPixBlock, 0);
  } else {
#define TRACE_BATCH_OCC(TW_, TH_)                                                            \
  if (tw == TW_ && th == TH_)                                                                \
    e = exact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bp_batch<TW_, TH_, true>, \
                                                             kTileBlock, lds)                \
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(                                \
                    &nb, k_bp_batch<TW_, TH_, false>, kTileBlock, lds);
    TRACE_BATCH_TILES(TRACE_BATCH_OCC)
#undef TRACE_BATCH_OCC
  }
  return e == hipSuccess ? nb : 0;
}

hipError_t launch_adjoint_atomic(const Geometry &g, const RayInfo *rays, int mode, bool exact,
                                 const float *r, float *z, int nslices, hipStream_t st) {
  const int64_t n = g.rays_per_slice();
  dim3 grid(blocks_for(n, kRayBlock), nslices), block(kRayBlock);
  if (exact)
    hipLaunchKernelGGL(k_adjoint_atomic<true>, grid, block, 0, st, g.grid, g.d_ang, rays, g.P,
                       g.S, mode, r, z);
  else
    hipLaunchKernelGGL(k_adjoint_atomic<false>, grid, block, 0, st, g.grid, g.d_ang, rays, g.P,
                       g.S, mode, r, z);
  return hipGetLastError();
}

hipError_t launch_dump_rays(const Geometry &g, const RayInfo *rays, int n, const int *proj,
                            const int *col, int max_seg, int *counts, int *idx, float *len,
                            float *len2, hipStream_t st, bool exact) {
  if (n <= 0) return hipSuccess;
  if (exact)
    hipLaunchKernelGGL(k_dump_rays<true>, dim3(blocks_for(n, 64)), dim3(64), 0, st, g.grid,
                       g.d_ang, rays, n, proj, col, max_seg, counts, idx, len, len2);
  else
    hipLaunchKernelGGL(k_dump_rays<false>, dim3(blocks_for(n, 64)), dim3(64), 0, st, g.grid,
                       g.d_ang, rays, n, proj, col, max_seg, counts, idx, len, len2);
  return hipGetLastError();
}

hipError_t launch_ray_sums(const Geometry &g, const RayInfo *rays, bool exact, float *len_sum,
                           float *len2_sum, hipStream_t st) {
  const int64_t n = g.rays_per_slice();
  if (exact)
    hipLaunchKernelGGL(k_ray_sums<true>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0, st,
                       g.grid, g.d_ang, rays, g.P, len_sum, len2_sum);
  else
    hipLaunchKernelGGL(k_ray_sums<false>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0, st,
                       g.grid, g.d_ang, rays, g.P, len_sum, len2_sum);
  return hipGetLastError();
}

hipError_t launch_count_segments(const Geometry &g, const RayInfo *rays,
                                 unsigned long long *total, unsigned long long *oor,
                                 unsigned *maxlen, unsigned long long *wraps, hipStream_t st) {
  const int64_t n = g.rays_per_slice();
  hipLaunchKernelGGL(k_count_segments, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0, st,
                     g.grid, g.d_ang, rays, g.P, total, oor, maxlen, wraps);
  return hipGetLastError();
}

hipError_t launch_to_pcs(const float *r, int P, int S, int C, float *w, unsigned *maxw,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_to_pcs, dim3((unsigned)((int64_t)P * S)), dim3(256), 0, st, r, P, S, C,
                     w, maxw);
  return hipGetLastError();
}

hipError_t launch_absmax_slices(const float *w, int P, int S, int C, unsigned *maxw,
                                hipStream_t st) {
  int64_t n = (int64_t)P * C;
  unsigned bx = blocks_for(n, kPixBlock);
  if (bx > 512) bx = 512;
  hipLaunchKernelGGL(k_absmax_slices, dim3(bx, S), dim3(kPixBlock), 0, st, w, P, S, C, maxw);
  return hipGetLastError();
}

hipError_t launch_check_finite(const float *v, int64_t n, int *flag, hipStream_t st) {
  unsigned blocks = blocks_for(n, kPixBlock);
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_finite, dim3(blocks), dim3(kPixBlock), 0, st, v, n, flag);
  return hipGetLastError();
}

}  // namespace trace_hip
