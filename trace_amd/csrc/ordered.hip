// ordered.hip -- backprojection summed in the reference's order.
//
// The reference accumulates each pixel's backprojected sums in float, one
// segment at a time, in the order its one-thread engine meets the rays:
// projections ascending, detector columns ascending, segments along the ray
// (rows of NextMirroredRegion, data_region_base.h:98-111; the per-segment
// updates of UpdateReconReplica):
//
//   den_j += len                      sirt.cc:62, mlem.cc:57, pml.cc:244
//   num_j += len * upd                sirt.cc:61, mlem.cc:56
//   num_j -= (x_j * len) * upd        pml.cc:243, apmlr.cc:1024
//
// A float sum depends on its order.  For PML and APMLR it matters: the
// closed-form update (pml.cc:18-20) evaluates -G + sqrt(G^2 - 8 num F) with
// G ~ den ~ P, which cancels, so one ulp of den or num in G^2 - 8 num F
// moves the updated pixel by up to ~1e-4 relative.  The tile backprojector
// (kernels.hip) sums in fixed point, exactly but in no particular order;
// this one reproduces the reference's float sums bit for bit.
//
// Gather form: each thread owns a strip of PW x PH pixels and loops over
// the projections in order; per projection it walks the columns whose ray
// can touch the strip in ascending order (tile_columns), each over the
// window of the ray around the strip (traverse_window, the same walk the
// tile kernels use), and adds the segments falling in the strip to its
// register accumulators.  Every pixel's terms therefore arrive in the
// reference's (projection, column, segment) order.  No atomics, no LDS.
//
// Wrapped segments (ix == N: the reference's truncation sends them to pixel
// 0 of the next row, trace_utils.cc:177-185) belong to strips at column 0.
// Such strips walk the flagged rays (kRayWraps) crossing the strip left of
// the right edge whole, so that their segments keep the ray's order.
#include <hip/hip_runtime.h>

#include "internal.h"

#pragma clang fp contract(off)

namespace trace_hip {

namespace {

constexpr int kOrdBlock = 256;
constexpr int kOrdStrips = 16;  // a block is 16 x 16 strips

struct OrdArgs {
  GridInfo g;
  const AngleInfo *ang;
  const RayInfo *rays;
  int P, S;
  float yi0;
  const float *w;   // [P][C][S] per-ray weights (kOrdNum, kOrdEM)
  const float *xb;  // [S/xsb][N*N][xsb] image the ray pass read (kOrdNum)
  int xsb;          // slices per slot batch of xb
  float *out;       // kOrdDen: [N][N]; kOrdNum, kOrdEM: [S][N][N]
  int wraps;        // some ray has a segment with ix == N
};

template <int PW, int PH, int SB, int MODE, bool kExact>
__global__ void __launch_bounds__(kOrdBlock)
k_bp_ordered(OrdArgs a) {
  constexpr int NP = PW * PH;
  const int N = a.g.N, C = a.g.C;
  const int sxn = (N + PW - 1) / PW, syn = (N + PH - 1) / PH;
  const int bxn = (sxn + kOrdStrips - 1) / kOrdStrips;
  const int sx = (int)(blockIdx.x % bxn) * kOrdStrips + (threadIdx.x % kOrdStrips);
  const int sy = (int)(blockIdx.x / bxn) * kOrdStrips + (threadIdx.x / kOrdStrips);
  if (sx >= sxn || sy >= syn) return;  // no barrier in this kernel
  const int tx0 = sx * PW, ty0 = sy * PH;
  const int tx1 = min(tx0 + PW, N), ty1 = min(ty0 + PH, N);
  const int b = blockIdx.y;
  const int64_t NN = (int64_t)N * N;

  float acc[NP][SB];
  float xv[NP][SB];
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int q = 0; q < SB; ++q) acc[i][q] = 0.f, xv[i][q] = 0.f;
  if (MODE == kOrdNum) {
    // x of the ray pass, slots b*SB .. b*SB+SB-1 (SB divides xsb)
    const int s0 = b * SB;
    const int xb_b = s0 / a.xsb, xb_q = s0 % a.xsb;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int px = tx0 + i % PW, py = ty0 + i / PW;
      if (px < N && py < N) {
        const float *src = a.xb + ((int64_t)xb_b * NN + (int64_t)py * N + px) * a.xsb + xb_q;
#pragma unroll
        for (int q = 0; q < SB; ++q) xv[i][q] = src[q];
      }
    }
  }
  const float x0 = grid_line(N, tx0), x1 = grid_line(N, tx1);
  const float y0 = grid_line(N, ty0), y1 = grid_line(N, ty1);
  // strip just left of the right edge, rows ty0-1 .. ty1: the rays whose
  // wrapped segments can land in a column-0 strip
  const bool edge = a.wraps && tx0 == 0;
  const float ex0 = grid_line(N, N - 1), ex1 = grid_line(N, N);
  const float ey0 = grid_line(N, ty0 - 1), ey1 = y1;

  for (int p = 0; p < a.P; ++p) {
    const AngleInfo ang = a.ang[p];
    int cmin, cmax;
    tile_columns(ang, x0, x1, y0, y1, a.yi0, cmin, cmax);
    int emin = 1, emax = 0;
    if (edge) tile_columns(ang, ex0, ex1, ey0, ey1, a.yi0, emin, emax);
    const int lo = max(0, edge ? min(cmin, emin) : cmin);
    const int hi = min(C - 1, edge ? max(cmax, emax) : cmax);
    for (int c = lo; c <= hi; ++c) {
      const bool in_main = (c >= cmin && c <= cmax);
      const bool in_erange = edge && c >= emin && c <= emax;
      if (!in_main && !in_erange) {
        // c lies in the gap between the two ranges: go to the upper one
        c = (c > cmax ? emin : cmin) - 1;
        continue;
      }
      const int64_t t = (int64_t)p * C + c;
      const RayInfo ray = a.rays[t];
      const bool in_edge = in_erange && (ray.alo & kRayWraps);
      if (!in_main && !in_edge) continue;
      float wv[SB];
      if (MODE != kOrdDen) {
        const float *wr = a.w + t * a.S + b * SB;
#pragma unroll
        for (int q = 0; q < SB; ++q) wv[q] = (b * SB + q < a.S) ? wr[q] : 0.f;
      }
      auto visit = [&](int ix, int iy, float l, float) {
        const bool wrap = (ix == N);
        if (wrap && !in_edge) return;
        ix = wrap ? 0 : ix;
        iy = wrap ? iy + 1 : iy;
        const unsigned lx = (unsigned)(ix - tx0), ly = (unsigned)(iy - ty0);
        if (lx < (unsigned)(tx1 - tx0) && ly < (unsigned)(ty1 - ty0)) {
          const int pix = (int)(ly * PW + lx);
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            if (NP == 1 || pix == i) {
#pragma unroll
              for (int q = 0; q < SB; ++q) {
                if (MODE == kOrdDen)
                  acc[i][q] = acc[i][q] + l;
                else if (MODE == kOrdEM)
                  acc[i][q] = acc[i][q] + l * wv[q];
                else
                  acc[i][q] = acc[i][q] - (xv[i][q] * l) * wv[q];
              }
            }
          }
        }
      };
      if (in_edge)
        traverse_full<kExact>(a.g, ray, ang.quadrant, visit);
      else
        traverse_window<kExact, true>(a.g, ray, ang.quadrant, tx0, tx1, ty0, ty1, visit);
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const int px = tx0 + i % PW, py = ty0 + i / PW;
    if (px >= N || py >= N) continue;
    const int64_t j = (int64_t)py * N + px;
#pragma unroll
    for (int q = 0; q < SB; ++q) {
      const int s = b * SB + q;
      if (MODE == kOrdDen)
        a.out[j] = acc[i][0];
      else if (s < a.S)
        a.out[(int64_t)s * NN + j] = acc[i][q];
    }
  }
}

template <int PW, int PH, int SB, int MODE>
hipError_t ord_launch(const OrdArgs &a, bool exact, hipStream_t st) {
  const int N = a.g.N;
  const int sxn = (N + PW - 1) / PW, syn = (N + PH - 1) / PH;
  const int blocks = ((sxn + kOrdStrips - 1) / kOrdStrips) * ((syn + kOrdStrips - 1) / kOrdStrips);
  const int nb = MODE == kOrdDen ? 1 : (a.S + SB - 1) / SB;
  if (exact)
    hipLaunchKernelGGL((k_bp_ordered<PW, PH, SB, MODE, true>), dim3(blocks, nb),
                       dim3(kOrdBlock), 0, st, a);
  else
    hipLaunchKernelGGL((k_bp_ordered<PW, PH, SB, MODE, false>), dim3(blocks, nb),
                       dim3(kOrdBlock), 0, st, a);
  return hipGetLastError();
}

}  // namespace

bool ordered_sb_supported(int mode, int sb, int xsb) {
  if (mode == kOrdDen) return true;
  if (sb != 1 && sb != 2 && sb != 4 && sb != 8) return false;
  return mode == kOrdEM || (xsb >= sb && xsb % sb == 0);
}

hipError_t launch_bp_ordered(const Geometry &g, const RayInfo *rays, int mode, bool exact,
                             const float *w, const float *xb, int xsb, int sb, float *out,
                             bool wraps, hipStream_t st) {
  OrdArgs a;
  a.g = g.grid;
  a.ang = g.d_ang;
  a.rays = rays;
  a.P = g.P;
  a.S = g.S;
  a.yi0 = g.yi0;
  a.w = w;
  a.xb = xb;
  a.xsb = xsb;
  a.out = out;
  a.wraps = wraps ? 1 : 0;
  if (!ordered_sb_supported(mode, sb, xsb)) return hipErrorInvalidValue;
  if (mode == kOrdDen) return ord_launch<2, 2, 1, kOrdDen>(a, exact, st);
#define TRACE_ORD_CASE(SB_)                                                              \
  case SB_:                                                                              \
    return mode == kOrdNum ? ord_launch<1, 1, SB_, kOrdNum>(a, exact, st)                \
                           : ord_launch<1, 1, SB_, kOrdEM>(a, exact, st);
  switch (sb) {
    TRACE_ORD_CASE(1)
    TRACE_ORD_CASE(2)
    TRACE_ORD_CASE(4)
    TRACE_ORD_CASE(8)
    default: return hipErrorInvalidValue;
  }
#undef TRACE_ORD_CASE
}

}  // namespace trace_hip
