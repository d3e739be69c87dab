// staged.hip -- forward projector with the image staged in LDS, for
// 32-slice batches (one 128-byte line of xb per pixel).
//
// The whole-ray form (kernels.hip) reads every segment's pixel line from
// L2; rays share lines only when they happen to meet in time, so each line
// comes from memory hundreds of times per launch.  Here one workgroup holds
// a bundle of kStG adjacent angles x kStW adjacent columns (one wavefront
// per angle, one lane per ray) and walks all of its rays together through
// the image in thin slabs across the walk direction:
//
//   * shallow rays (|slope| <= 1) sweep slabs of kStH pixel columns left to
//     right (the walk is sorted by x, so the pixel column never decreases);
//   * steep rays sweep slabs of kStH pixel rows, upwards or downwards with
//     the sign of the slope.
//
// For each slab the workgroup stages the pixels its rays cross (a window of
// at most kStEmax pixels across the slab, planned once per geometry from the
// bundle's outermost rays) in LDS, double-buffered, while the lanes consume
// the previous slab.  Each lane still walks its own ray from start to end in
// the reference's order and adds x[idx]*len per slice exactly as the
// whole-ray form does (sirt.cc:4-19), so y is bit-identical; a segment whose
// pixel is not in the staged window (outside the plan's margin, a row wrap
// at ix == N, a window clipped to kStEmax) reads its line from HBM instead.
// The staging plan therefore affects only speed, never the sums.
#include <hip/hip_runtime.h>

#include <cmath>

#include "staged.h"

#pragma clang fp contract(off)

namespace trace_hip {

namespace {

constexpr int kSB = kStSlices;
// Four floats as a native vector: staged values live in registers across
// the slab loop (an array of the float4 class does not reliably).
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kBlock = kStG * 64;              // one wavefront per angle
// A pixel's 128-byte line takes 144 bytes of LDS (nine float4 slots): the
// lines of neighbouring pixels then start 4 banks apart, so the lanes of a
// wavefront reading one group of neighbouring pixels hit different banks,
// and the eight groups of a line sit at fixed offsets from its start.
constexpr int kPix4 = 9;
constexpr int kBuf4 = kStH * kStEmax * kPix4;  // float4 slots per slab buffer
constexpr int kStage4 = (kStH * kStEmax * 8 + kBlock - 1) / kBlock;  // float4 staged per thread
constexpr size_t kStLds = 2 * (size_t)kBuf4 * sizeof(f4);
static_assert(kStW == 64, "one lane per column of the bundle");
static_assert(kStH == 4, "the staging order below takes the slab position from 2 bits");
static_assert(kStLds <= 160 * 1024, "two slab buffers must fit the 160 KB of LDS");

// Slab coordinates of pixel (row, col): u along the sweep, v across it.
__device__ __forceinline__ void slab_uv(int axis, int N, int row, int col, int &u, int &v) {
  u = axis == kStAxisX ? col : (axis == kStAxisYUp ? row : N - 1 - row);
  v = axis == kStAxisX ? row : col;
}

// Plan of one bundle: its sweep axis, and per slab the first pixel and the
// extent of the window across the slab (extent 0: nothing to stage).  One
// block per bundle; the window of slab k holds the pixels, with one pixel
// of margin, that the lines of the bundle's outermost rays (first and last
// column of every angle: at a fixed position along the sweep the rays of an
// angle are parallel lines ordered by column) cross within the slab.
__global__ void __launch_bounds__(256)
k_staged_plan(GridInfo g, const RayInfo *__restrict__ rays, int P, int nslab, int ngc,
              int *__restrict__ axis_out, int2 *__restrict__ range) {
  const int wg = blockIdx.x;
  const int ag = wg / ngc, cb = wg % ngc;
  const int p0 = ag * kStG, c0 = cb * kStW;
  const int pn = min(kStG, P - p0), cn = min(kStW, g.C - c0);
  const int N = g.N;
  const RayInfo &mid = rays[(int64_t)(p0 + pn / 2) * g.C + c0 + cn / 2];
  const int axis = __builtin_fabsf(mid.slope) <= 1.0f
                       ? kStAxisX
                       : (mid.slope > 0.0f ? kStAxisYUp : kStAxisYDown);
  if (threadIdx.x == 0) axis_out[wg] = axis;
  const double mg = 0.5 * N;
  for (int k = threadIdx.x; k < nslab; k += blockDim.x) {
    // grid-coordinate interval of the slab along the sweep, half a pixel of
    // margin on either side
    double ua, ub;
    if (axis == kStAxisYDown) {
      ua = (double)(N - k * kStH - kStH) - mg - 0.5;
      ub = (double)(N - k * kStH) - mg + 0.5;
    } else {
      ua = (double)(k * kStH) - mg - 0.5;
      ub = (double)(k * kStH + kStH) - mg + 0.5;
    }
    double lo = 1e30, hi = -1e30;
    bool any = false, wide = false;
    for (int e = 0; e < 2 * pn; ++e) {
      const int p = p0 + (e >> 1);
      const int c = (e & 1) ? c0 + cn - 1 : c0;
      const RayInfo r = rays[(int64_t)p * g.C + c];
      double va, vb;
      if (axis == kStAxisX) {  // y on the ray at x = ua, ub
        va = (double)r.slope * (ua - (double)r.srcx) + (double)r.srcy;
        vb = (double)r.slope * (ub - (double)r.srcx) + (double)r.srcy;
      } else {                 // x on the ray at y = ua, ub
        va = (double)r.islope * (ua - (double)r.srcy) + (double)r.srcx;
        vb = (double)r.islope * (ub - (double)r.srcy) + (double)r.srcx;
      }
      if (!std::isfinite(va) || !std::isfinite(vb)) {
        wide = true;
        continue;
      }
      lo = fmin(lo, fmin(va, vb));
      hi = fmax(hi, fmax(va, vb));
      any = true;
    }
    int first = 0, last = -1;
    if (wide) {
      first = 0;
      last = N - 1;
    } else if (any) {
      const double fl = floor(lo + mg) - 1.0, fh = floor(hi + mg) + 1.0;
      first = (int)fmax(fl, 0.0);
      last = (int)fmin(fh, (double)(N - 1));
    }
    int extent = last >= first ? last - first + 1 : 0;
    if (extent > kStEmax) {  // keep the middle of the window
      first += (extent - kStEmax) / 2;
      extent = kStEmax;
    }
    range[(int64_t)wg * nslab + k] = make_int2(first, extent);
  }
}

// The walk of one lane, one segment ahead: the pending segment is computed
// but not yet added (idx, len, len2; u, v its slab coordinates; key its
// slab, -1 for a pixel outside the image, taken at once and adding nothing,
// and kStDone once the walk is over).
constexpr int kStDone = 1 << 30;
struct StagedWalk {
  Walker w;
  float px, py;
  int rem;
  unsigned idx;
  float len, len2;
  int key, u, v;
};

template <bool kExact, int AXIS>
__device__ __forceinline__ void staged_advance(StagedWalk &r, const GridInfo &g, unsigned NN) {
  if (r.rem <= 0) {
    r.key = kStDone;
    return;
  }
  const int N = g.N;
  float qx, qy;
  r.w.next(qx, qy);
  int ix, iy;
  segment<kExact>(r.px, r.py, qx, qy, g.mg, ix, iy, r.len, r.len2);
  r.px = qx;
  r.py = qy;
  --r.rem;
  r.idx = (unsigned)(ix + iy * N);
  const bool in = r.idx < NN;
  int row = iy, col = ix;
  if ((unsigned)ix >= (unsigned)N) {  // a wrap onto the next row
    row = (int)(r.idx / (unsigned)N);
    col = (int)(r.idx - (unsigned)row * (unsigned)N);
  }
  slab_uv(AXIS, N, row, col, r.u, r.v);
  r.key = in ? (int)((unsigned)r.u / kStH) : -1;
}

// Global offset (in floats) of staged float4 f of slab k, or -1 if it holds
// no pixel.  Float4 f is group f & 7 of pixel f >> 3; pixels in the order
// (v, u) with u fastest.
__device__ __forceinline__ int64_t stage_src(int f, int k, int lo, int E, int N, int axis) {
  const int qg = f >> 3, du = qg & (kStH - 1), dv = qg >> 2;
  const int u = k * kStH + du;
  if (f >= kStH * E * 8 || u >= N) return -1;
  const int v = lo + dv;
  const int row = axis == kStAxisX ? v : (axis == kStAxisYUp ? u : N - 1 - u);
  const int col = axis == kStAxisX ? u : v;
  return ((int64_t)row * N + col) * kSB + 4 * (f & 7);
}

// LDS slot (in float4) of staged float4 f: pixel (du, dv) at q = du * E + dv.
__device__ __forceinline__ int stage_dst(int f, int E) {
  const int qg = f >> 3, du = qg & (kStH - 1), dv = qg >> 2;
  return (du * E + dv) * kPix4 + (f & 7);
}

template <int MODE, bool kExact, int AXIS>
__device__ __forceinline__ void forward_staged(
    f4 *slab, int wg, int b, const GridInfo &g, const AngleInfo *__restrict__ ang,
    const RayInfo *__restrict__ rays, int P, int S, int nslab, int ngc,
    const int2 *__restrict__ range, const float *__restrict__ xb,
    const float *__restrict__ sino, float *__restrict__ out, unsigned *__restrict__ maxw) {
  constexpr int axis = AXIS;
  const int ag = wg / ngc, cb = wg % ngc;
  const int lane = threadIdx.x & 63;
  const int p = ag * kStG + (threadIdx.x >> 6), c = cb * kStW + lane;
  const int C = g.C, N = g.N;
  const bool valid = p < P && c < C;
  const unsigned NN = (unsigned)N * (unsigned)N;
  const float *xbb = xb + (int64_t)b * NN * kSB;
  const int2 *rng = range + (int64_t)wg * nslab;

  float acc[kSB];
#pragma unroll
  for (int s = 0; s < kSB; ++s) acc[s] = 0.f;
  float a2 = 0.f;
  bool hit = false;  // the ray has a segment in the image

  StagedWalk r;
  r.px = r.py = 0.f;
  r.rem = 0;
  r.idx = 0;
  r.len = r.len2 = 0.f;
  r.key = kStDone;
  r.u = r.v = 0;
  const int64_t t = (int64_t)p * C + c;
  if (valid) {
    r.w.init(rays[t], N, ang[p].quadrant);
    r.w.seek(0, 0);
    const int n = r.w.alen + r.w.blen;
    if (n >= 2) {
      r.w.next(r.px, r.py);
      r.rem = n - 1;
      staged_advance<kExact, AXIS>(r, g, NN);
    }
  }

  f4 stg[kStage4];
  int2 cur = rng[0];
#pragma unroll
  for (int j = 0; j < kStage4; ++j) {
    const int f = threadIdx.x + j * kBlock;
    const int64_t o = stage_src(f, 0, cur.x, cur.y, N, axis);
    if (o >= 0) slab[stage_dst(f, cur.y)] = *reinterpret_cast<const f4 *>(xbb + o);
  }
  __syncthreads();
  // Slab k: stage slab k + 1 into the other buffer while the lanes add their
  // segments up to slab k (and those of earlier slabs, from HBM: they were
  // not staged).  The pass k == nslab, past the last slab (every segment in
  // the image has a slab below nslab), adds what is left, from HBM (nothing
  // for a ray inside the grid).
  for (int k = 0; k <= nslab; ++k) {
    const bool more = k + 1 < nslab;
    const int2 nxt = more ? rng[k + 1] : make_int2(0, 0);
    int64_t so[kStage4];
#pragma unroll
    for (int j = 0; j < kStage4; ++j) {
      so[j] = more ? stage_src(threadIdx.x + j * kBlock, k + 1, nxt.x, nxt.y, N, axis) : -1;
      if (so[j] >= 0) stg[j] = *reinterpret_cast<const f4 *>(xbb + so[j]);
    }
    const int kk = k;
    const int lo = cur.x, E = k < nslab ? cur.y : 0;
    const int base = (k & 1) * kBuf4;
    while (r.key <= kk) {
      if (r.key >= 0) {
        const int dv = r.v - lo;
        const bool staged = r.key == kk && (unsigned)dv < (unsigned)E;
        const int q = (r.u - kk * kStH) * E + dv;
        const int lq = base + q * kPix4;
        if (staged) {
          // two halves of the line, so that at most 16 loaded values are live
#pragma unroll
          for (int hlf = 0; hlf < 2; ++hlf) {
            f4 f[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) f[j] = slab[lq + 4 * hlf + j];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              float *a = acc + 16 * hlf + 4 * j;
              a[0] += f[j].x * r.len; a[1] += f[j].y * r.len;
              a[2] += f[j].z * r.len; a[3] += f[j].w * r.len;
            }
          }
        } else {
          // from HBM; the explicit wait keeps the merge point free of
          // waits, which would otherwise also wait for the staging loads
          const f4 *gsrc = reinterpret_cast<const f4 *>(xbb + (int64_t)r.idx * kSB);
#pragma unroll
          for (int hlf = 0; hlf < 2; ++hlf) {
            f4 f[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) f[j] = gsrc[4 * hlf + j];
            wait_vector_loads();
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              float *a = acc + 16 * hlf + 4 * j;
              a[0] += f[j].x * r.len; a[1] += f[j].y * r.len;
              a[2] += f[j].z * r.len; a[3] += f[j].w * r.len;
            }
          }
        }
      }
      hit |= r.key >= 0;
      if (MODE == kFwdOutSIRT) a2 += r.len2;
      staged_advance<kExact, AXIS>(r, g, NN);
    }
    if (more) {
      const int nbase = ((k + 1) & 1) * kBuf4;
#pragma unroll
      for (int j = 0; j < kStage4; ++j)
        if (so[j] >= 0) slab[nbase + stage_dst(threadIdx.x + j * kBlock, nxt.y)] = stg[j];
    }
    // every staging load has been waited for by its store; saying so here
    // keeps the compiler from waiting again before the next slab's loads
    wait_vector_loads();
    if (k < nslab) __syncthreads();
    cur = nxt;
  }

#pragma unroll
  for (int k = 0; k < kSB; ++k) {
    const int s = b * kSB + k;
    const bool live = valid && (s < S);
    unsigned m = 0;
    if (live) {
      const int64_t o = ((int64_t)p * S + s) * C + c;
      if (MODE == kFwdOutY) {
        out[o] = acc[k];
      } else {
        const float wv = (MODE == kFwdOutSIRT) ? (sino[o] - acc[k]) / a2 : sino[o] / acc[k];
        out[t * S + s] = wv;
        // a ray with no segment in the image does not set the scale (as in
        // the whole-ray form)
        if (hit) m = __float_as_uint(__builtin_fabsf(wv));
      }
    }
    if (MODE != kFwdOutY) {
      for (int off = 32; off >= 1; off >>= 1)
        m = max(m, (unsigned)__shfl_down((int)m, off, 64));
      if (lane == 0 && s < S) atomicMax(maxw + s, m);
    }
  }
}

template <int MODE, bool kExact>
__global__ void __launch_bounds__(kBlock)
k_forward_staged(GridInfo g, const AngleInfo *__restrict__ ang,
                 const RayInfo *__restrict__ rays, int P, int S, int nslab, int ngc,
                 const int *__restrict__ axis_of, const int2 *__restrict__ range,
                 const float *__restrict__ xb, const float *__restrict__ sino,
                 float *__restrict__ out, unsigned *__restrict__ maxw) {
  extern __shared__ f4 slab[];  // [2][kBuf4]
  const int wg = (int)xcd_chunked(blockIdx.x, gridDim.x);
  const int b = blockIdx.y;
  const int axis = axis_of[wg];  // uniform: one specialised body per sweep axis
  if (axis == kStAxisX)
    forward_staged<MODE, kExact, kStAxisX>(slab, wg, b, g, ang, rays, P, S, nslab, ngc, range,
                                           xb, sino, out, maxw);
  else if (axis == kStAxisYUp)
    forward_staged<MODE, kExact, kStAxisYUp>(slab, wg, b, g, ang, rays, P, S, nslab, ngc,
                                             range, xb, sino, out, maxw);
  else
    forward_staged<MODE, kExact, kStAxisYDown>(slab, wg, b, g, ang, rays, P, S, nslab, ngc,
                                               range, xb, sino, out, maxw);
}

template <int MODE, bool E>
hipError_t staged_launch(const Geometry &g, const RayInfo *rays, const StagedPlan &pl,
                         const float *xb, const float *sino, float *out, unsigned *maxw,
                         hipStream_t st) {
  auto kern = k_forward_staged<MODE, E>;
  hipError_t e = hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStLds);
  if (e != hipSuccess) return e;
  const int64_t nwg = (int64_t)pl.ngc * ((g.P + kStG - 1) / kStG);
  dim3 grid((unsigned)nwg, (g.S + kSB - 1) / kSB), block(kBlock);
  hipLaunchKernelGGL(kern, grid, block, kStLds, st, g.grid, g.d_ang, rays, g.P, g.S, pl.nslab,
                     pl.ngc, pl.axis, pl.range, xb, sino, out, maxw);
  return hipGetLastError();
}

}  // namespace

void staged_plan_size(const Geometry &g, int &nwg, int &nslab, int &ngc) {
  ngc = (g.C + kStW - 1) / kStW;
  nwg = ngc * ((g.P + kStG - 1) / kStG);
  nslab = (g.N + kStH - 1) / kStH;
}

hipError_t launch_staged_plan(const Geometry &g, const RayInfo *rays, StagedPlan &pl,
                              hipStream_t st) {
  int nwg = 0;
  staged_plan_size(g, nwg, pl.nslab, pl.ngc);
  hipLaunchKernelGGL(k_staged_plan, dim3(nwg), dim3(256), 0, st, g.grid, rays, g.P, pl.nslab,
                     pl.ngc, pl.axis, pl.range);
  return hipGetLastError();
}

hipError_t launch_forward_staged(const Geometry &g, const RayInfo *rays, const StagedPlan &pl,
                                 int mode, bool exact, const float *xb, const float *sino,
                                 float *out, unsigned *maxw, hipStream_t st) {
  if (mode == kFwdOutY)
    return exact ? staged_launch<kFwdOutY, true>(g, rays, pl, xb, sino, out, maxw, st)
                 : staged_launch<kFwdOutY, false>(g, rays, pl, xb, sino, out, maxw, st);
  if (mode == kFwdOutSIRT)
    return exact ? staged_launch<kFwdOutSIRT, true>(g, rays, pl, xb, sino, out, maxw, st)
                 : staged_launch<kFwdOutSIRT, false>(g, rays, pl, xb, sino, out, maxw, st);
  return exact ? staged_launch<kFwdOutEM, true>(g, rays, pl, xb, sino, out, maxw, st)
               : staged_launch<kFwdOutEM, false>(g, rays, pl, xb, sino, out, maxw, st);
}

}  // namespace trace_hip
