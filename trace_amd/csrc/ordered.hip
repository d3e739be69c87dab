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
//
// Tile form (k_bp_ordered_tile, the default where it applies): a wavefront
// owns an 8 x 8 tile and takes the projections four at a time.  Its 64
// lanes first walk the four projections' columns through the tile at once
// (one (projection, column) pair per lane, as the LDS tile backprojector
// does) and drop every segment's length into an LDS slot keyed by (pixel,
// projection, column - the pixel's first candidate column, segment number
// of that ray in that pixel); then each lane sums its own pixel's slots in
// (projection, column, segment) order.  A pixel sees at most three
// candidate columns per projection (its projected width is at most sqrt 2
// plus twice the 0.45-pixel margin), and a ray puts at most two segments in
// one pixel on the geometries measured (two where it grazes a corner); a
// geometry that would need more slots sets a flag when den is computed, and
// the handle then keeps the per-pixel form.  Geometries with wrapped
// segments use the per-pixel form.
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


// Tile form: see the header.  One wavefront per (8 x 8 tile, slice batch).
constexpr int kOT = 8;       // tile side
constexpr int kOA = 4;       // projections per round
constexpr int kORanks = 16;  // column slots per projection (>= the tile's count bound)
constexpr int kOSlots = 3;   // candidate columns per pixel and projection
constexpr int kODups = 2;    // segments per ray and pixel
constexpr int kOWave = 64;

// projection slot (0 .. kOA-1) of a lane's phase-1 pair
__device__ __forceinline__ int ai1_of(int lane) { return lane / kORanks; }

template <int SB>
struct OrdTileLds {
  static constexpr int kWPitch = SB >= 4 ? SB + 4 : SB;  // row pitch of s_w (bank spread)
};

template <int SB, int MODE, bool kExact>
__global__ void __launch_bounds__(kOWave)
k_bp_ordered_tile(OrdArgs a, unsigned *__restrict__ overflow) {
  static_assert(SB <= kWeightPad, "a slice batch of weights past the last row must be allocated");
  constexpr int WP = OrdTileLds<SB>::kWPitch;
  __shared__ unsigned s_cnt[kOA][kOWave];                  // byte counter per column slot
  __shared__ float s_len[kOA][kOSlots * kODups][kOWave];
  __shared__ float s_w[MODE == kOrdDen ? 1 : kOA * kORanks * WP];
  __shared__ int s_cmin[kOA];
  // first candidate column of each pixel for the round's projections,
  // written by the pixel's lane one round ahead (double-buffered)
  __shared__ int s_pc0[2][kOA][kOWave];
  const int N = a.g.N, C = a.g.C;
  const int tiles_x = (N + kOT - 1) / kOT;
  const int tile = (int)xcd_chunked(blockIdx.x, gridDim.x);
  const int tx0 = (tile % tiles_x) * kOT, ty0 = (tile / tiles_x) * kOT;
  const int tx1 = min(tx0 + kOT, N), ty1 = min(ty0 + kOT, N);
  const int b = blockIdx.y;
  const int lane = threadIdx.x;
  const int64_t NN = (int64_t)N * N;
  // the pixel this lane sums
  const int px = tx0 + (lane % kOT), py = ty0 + (lane / kOT);
  const bool pix_ok = px < N && py < N;
  const float pcx0 = grid_line(N, px), pcx1 = grid_line(N, px + 1);
  const float pcy0 = grid_line(N, py), pcy1 = grid_line(N, py + 1);
  float acc[SB], xv[SB];
#pragma unroll
  for (int q = 0; q < SB; ++q) acc[q] = 0.f, xv[q] = 0.f;
  if (MODE == kOrdNum && pix_ok) {
    const int s0 = b * SB;
    const float *src =
        a.xb + ((int64_t)(s0 / a.xsb) * NN + (int64_t)py * N + px) * a.xsb + s0 % a.xsb;
#pragma unroll
    for (int q = 0; q < SB; ++q) xv[q] = src[q];
  }
  // Wait for the pixel's values here, once.  Otherwise the compiler waits
  // for them where phase 2 first reads them, inside the loop, and its
  // in-order counter then also waits for the next round's prefetch issued
  // just before, which serialises every round on that fetch.
  wait_vector_loads();
  const float x0 = grid_line(N, tx0), x1 = grid_line(N, tx1);
  const float y0 = grid_line(N, ty0), y1 = grid_line(N, ty1);
  // The angle records of the next round are loaded at the start of a round
  // (nxt: the four projections, for the pixels' columns; fang: this lane's
  // phase-1 projection, for its pair), so that they have arrived by the
  // time they are used and their wait never covers the pair prefetch.
  AngleInfo nxt[kOA], fang;
  auto load_angles = [&](int p) {
#pragma unroll
    for (int ai = 0; ai < kOA; ++ai) nxt[ai] = a.ang[min(p + ai, a.P - 1)];
    fang = a.ang[min(p + ai1_of(lane), a.P - 1)];
  };
  // this lane's pixel's first candidate column for projections p .. p+kOA-1
  auto pixel_columns = [&](int p, int buf) {
#pragma unroll
    for (int ai = 0; ai < kOA; ++ai) {
      int pc0 = 0, pc1;
      if (p + ai < a.P) tile_columns(nxt[ai], pcx0, pcx1, pcy0, pcy1, a.yi0, pc0, pc1);
      s_pc0[buf][ai][lane] = pc0;
    }
  };
#pragma unroll
  for (int ai = 0; ai < kOA; ++ai) s_cnt[ai][lane] = 0;
  load_angles(0);
  pixel_columns(0, 0);
  __syncthreads();
  const int ai1 = ai1_of(lane), k1 = lane % kORanks;  // phase-1 pair of this lane
  // The lane's phase-1 pair of a round, fetched one round ahead: its ray
  // table entry and weights mostly miss L2, so the next round's loads are
  // issued before this round's phase 2 and arrive while it runs.
  struct Pair {
    int c, cmin, quadrant;
    bool ok;
    RayInfo ray;
    float w[SB];
  };
  // Every lane issues the same loads on every path (row 0 when it has no
  // pair; a whole slice batch of weights, which the kWeightPad floats after
  // the weights keep in bounds for the last row): with one path, the
  // compiler's wait for a load covers that load and the older ones only, so
  // the prefetch is never waited for before its use.
  auto fetch = [&](int p, const AngleInfo &ang, Pair &q) {
    int cmin, cmax;
    tile_columns(ang, x0, x1, y0, y1, a.yi0, cmin, cmax);
    q.cmin = cmin;
    q.quadrant = ang.quadrant;
    q.c = cmin + k1;
    q.ok = p < a.P && q.c <= cmax && q.c >= 0 && q.c < C;
    const int64_t t = q.ok ? (int64_t)p * C + q.c : 0;
    q.ray = a.rays[t];
    if (MODE != kOrdDen) {
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
  Pair cur;
  fetch(ai1, fang, cur);
  for (int p0 = 0, buf = 0; p0 < a.P; p0 += kOA, buf ^= 1) {
    load_angles(p0 + kOA);
    // phase 1: walk the pairs, file each segment of the tile in its slot
    const int p = p0 + ai1;
    if (p < a.P) {
      if (k1 == 0) s_cmin[ai1] = cur.cmin;
      if (cur.ok) {
        const int c = cur.c;
        if (MODE != kOrdDen) {
          // slots past the stack's last slice (read from the next row) add 0
          float *dst = s_w + (ai1 * kORanks + k1) * WP;
#pragma unroll
          for (int q = 0; q < SB; ++q) dst[q] = b * SB + q < a.S ? cur.w[q] : 0.f;
        }
        traverse_window<kExact, true>(a.g, cur.ray, cur.quadrant, tx0, tx1, ty0, ty1,
                                      [&](int ix, int iy, float l, float) {
          const unsigned lx = (unsigned)(ix - tx0), ly = (unsigned)(iy - ty0);
          if (lx < (unsigned)(tx1 - tx0) && ly < (unsigned)(ty1 - ty0)) {
            const int lp = (int)(ly * kOT + lx);
            const int slot = c - s_pc0[buf][ai1][lp];
            if ((unsigned)slot >= (unsigned)kOSlots) {
              atomicOr(overflow, 1u);
              return;
            }
            const unsigned old = atomicAdd(&s_cnt[ai1][lp], 1u << (8 * slot));
            const int dup = (int)((old >> (8 * slot)) & 0xffu);
            if (dup >= kODups) {
              atomicOr(overflow, 2u);
              return;
            }
            s_len[ai1][slot * kODups + dup][lp] = l;
          }
        });
      }
    }
    __syncthreads();
    fetch(p0 + kOA + ai1, fang, cur);  // next round's pair, in flight during phase 2
    // phase 2: each lane adds its pixel's segments in order, (projection,
    // column slot, segment) ascending.  With many slices per segment the
    // lane's segments of the round are walked by one flat loop, so that the
    // wavefront runs as many iterations as its busiest pixel has segments
    // rather than one per slot that any lane uses; with few slices the
    // decoding would cost more than it saves, and the loops nest.
    auto add = [&](int ai, int kbase, int slot, int d) {
      const float l = s_len[ai][slot * kODups + d][lane];
      if (MODE == kOrdDen) {
        acc[0] = acc[0] + l;
      } else {
        const float *wr = s_w + (ai * kORanks + kbase + slot) * WP;
#pragma unroll
        for (int q = 0; q < SB; ++q) {
          if (MODE == kOrdEM)
            acc[q] = acc[q] + l * wr[q];
          else
            acc[q] = acc[q] - (xv[q] * l) * wr[q];
        }
      }
    };
    if (pix_ok && SB >= 8) {
      static_assert(kOSlots == 3 && kODups <= 3, "two-bit counts of three slots");
      unsigned cnt[kOA];  // n0 | n1 << 2 | n2 << 4 per projection
      int kb[kOA], end[kOA];
      int total = 0;
#pragma unroll
      for (int ai = 0; ai < kOA; ++ai) {
        const unsigned word = s_cnt[ai][lane];
        s_cnt[ai][lane] = 0;
        const unsigned n0 = min(word & 0xffu, (unsigned)kODups);
        const unsigned n1 = min((word >> 8) & 0xffu, (unsigned)kODups);
        const unsigned n2 = min((word >> 16) & 0xffu, (unsigned)kODups);
        cnt[ai] = n0 | (n1 << 2) | (n2 << 4);
        total += (int)(n0 + n1 + n2);
        end[ai] = total;
        kb[ai] = s_pc0[buf][ai][lane] - s_cmin[ai];
      }
      for (int k = 0; k < total; ++k) {
        // projection of the k-th segment, and its rank r within it
        int ai = 0, base = 0;
#pragma unroll
        for (int i = 0; i < kOA - 1; ++i)
          if (k >= end[i]) ai = i + 1, base = end[i];
        unsigned c = cnt[0];
        int kbase = kb[0];
#pragma unroll
        for (int i = 1; i < kOA; ++i)
          if (ai == i) c = cnt[i], kbase = kb[i];
        const int r = k - base;
        const int n0 = (int)(c & 3u), n01 = n0 + (int)((c >> 2) & 3u);
        const int slot = (r >= n0) + (r >= n01);
        add(ai, kbase, slot, r - (slot == 0 ? 0 : (slot == 1 ? n0 : n01)));
      }
    } else if (pix_ok) {
#pragma unroll
      for (int ai = 0; ai < kOA; ++ai) {
        const unsigned word = s_cnt[ai][lane];
        if (word == 0) continue;
        s_cnt[ai][lane] = 0;
        const int kbase = s_pc0[buf][ai][lane] - s_cmin[ai];
#pragma unroll
        for (int slot = 0; slot < kOSlots; ++slot) {
          const int n = min((int)((word >> (8 * slot)) & 0xffu), kODups);
          for (int d = 0; d < n; ++d) add(ai, kbase, slot, d);
        }
      }
    }
    pixel_columns(p0 + kOA, buf ^ 1);
    __syncthreads();
  }
  if (!pix_ok) return;
  const int64_t j = (int64_t)py * N + px;
  if (MODE == kOrdDen) {
    a.out[j] = acc[0];
    return;
  }
#pragma unroll
  for (int q = 0; q < SB; ++q)
    if (b * SB + q < a.S) a.out[(int64_t)(b * SB + q) * NN + j] = acc[q];
}

template <int SB, int MODE>
hipError_t ord_tile_launch(const OrdArgs &a, bool exact, unsigned *overflow, hipStream_t st) {
  const int N = a.g.N;
  const int tiles = ((N + kOT - 1) / kOT) * ((N + kOT - 1) / kOT);
  const int nb = MODE == kOrdDen ? 1 : (a.S + SB - 1) / SB;
  if (exact)
    hipLaunchKernelGGL((k_bp_ordered_tile<SB, MODE, true>), dim3(tiles, nb), dim3(kOWave), 0, st,
                       a, overflow);
  else
    hipLaunchKernelGGL((k_bp_ordered_tile<SB, MODE, false>), dim3(tiles, nb), dim3(kOWave), 0,
                       st, a, overflow);
  return hipGetLastError();
}

}  // namespace

bool ordered_sb_supported(int mode, int sb, int xsb) {
  if (mode == kOrdDen) return true;
  if (sb != 1 && sb != 2 && sb != 4 && sb != 8 && sb != 16 && sb != 32) return false;
  return mode == kOrdEM || (xsb >= sb && xsb % sb == 0);
}

hipError_t launch_bp_ordered(const Geometry &g, const RayInfo *rays, int mode, bool exact,
                             const float *w, const float *xb, int xsb, int sb, float *out,
                             bool wraps, unsigned *tile_overflow, hipStream_t st) {
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
  if (tile_overflow) {
    // tile form (no wrapped segments)
    if (wraps) return hipErrorInvalidValue;
    if (mode == kOrdDen) return ord_tile_launch<1, kOrdDen>(a, exact, tile_overflow, st);
#define TRACE_ORDT_CASE(SB_)                                                               \
  case SB_:                                                                                \
    return mode == kOrdNum ? ord_tile_launch<SB_, kOrdNum>(a, exact, tile_overflow, st)    \
                           : ord_tile_launch<SB_, kOrdEM>(a, exact, tile_overflow, st);
    switch (sb) {
      TRACE_ORDT_CASE(1)
      TRACE_ORDT_CASE(2)
      TRACE_ORDT_CASE(4)
      TRACE_ORDT_CASE(8)
      TRACE_ORDT_CASE(16)
      TRACE_ORDT_CASE(32)
      default: return hipErrorInvalidValue;
    }
#undef TRACE_ORDT_CASE
  }
  if (sb > 8) return hipErrorInvalidValue;
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
