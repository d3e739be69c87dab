// internal.h -- host-side structures shared by the C-ABI and the kernel
// launchers.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "traverse.h"

namespace trace_hip {

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b to XCD b % 8),
// and each XCD has its own L2.  Renumbering so that XCD x runs a contiguous
// range of logical blocks puts neighbouring work (adjacent columns of one
// angle group, adjacent tiles) behind one L2: measured 2% on the forward
// projector at 2048^2 / 1500 angles.  A bijection of [0, nb).
__device__ __forceinline__ int64_t xcd_chunked(int64_t b, int64_t nb) {
  constexpr int kXcd = 8;
  const int64_t q = nb / kXcd, r = nb % kXcd;
  const int64_t x = b % kXcd, j = b / kXcd;
  return x * q + (x < r ? x : r) + j;
}

// s_waitcnt vmcnt(0): waits until every vector memory load of the wave has
// returned (gfx9 encoding: vmcnt in bits 3:0 and 15:14; expcnt, bits 6:4,
// and lgkmcnt, bits 11:8, at their maxima, i.e. not waited for).
__device__ __forceinline__ void wait_vector_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Floats allocated after the [P][C][S] weights: the ordered tile kernel
// reads a whole slice batch (up to 32 floats) from every row, so the last
// row's partial batch stays inside the allocation.
constexpr int kWeightPad = 32;

// Pixel blocks (of 256 pixels) per work item of the batched image update.
constexpr int kBatchUpdateBlocks = 16;

enum Alg { kSIRT = 0, kMLEM = 1, kPML = 2, kAPMLR = 3 };
enum FwdModeHost { kFwdOutY = 0, kFwdOutSIRT = 1, kFwdOutEM = 2 };
enum BpModeHost { kBpOutSino = 0, kBpOutOnes = 1, kBpOutCount = 2 };
// Backprojections summed in the reference's order (ordered.hip):
//   kOrdDen  den_j = sum len                 (one slice, [N][N])
//   kOrdNum  num_j = -sum (x_j * len) * w    (PML/APMLR, [S][N][N])
//   kOrdEM   num_j = sum len * w             (SIRT/MLEM, [S][N][N])
enum OrdMode { kOrdDen = 0, kOrdNum = 1, kOrdEM = 2 };

// One slice stack's geometry on the device.
struct Geometry {
  int P = 0, S = 0, C = 0, N = 0;
  float center = 0.f;  // as given by the caller
  float mov = 0.f;     // trace_data.h:97-101
  float yi0 = 0.f;     // detector coordinate of column 0 (sirt.cc:129)
  GridInfo grid{};
  AngleInfo *d_ang = nullptr;  // [P]
  float *d_yi = nullptr;       // [C]
  int64_t NN() const { return (int64_t)N * N; }
  int64_t rays_per_slice() const { return (int64_t)P * C; }
};

// Launch shape of the tile backprojector: tile of tw x th pixels, sb slices
// per workgroup (LDS: tw*th*sb floats).
struct TileShape {
  int tw = 16, th = 16, sb = 16;
};

// Regularisation weights for the PML/APMLR stencils, computed on the host
// with the reference's float/double promotion order.
struct StencilWeights {
  // [z mode][pixel class][neighbour]: z mode 0 = two z neighbours,
  // 1 = first slice (+1 only), 2 = last slice (-1 only); classes 0 inner,
  // 1 top, 2 bottom, 3 left, 4 right, 5..8 corners TL, TR, BL, BR;
  // neighbours in-plane first, then z.
  float w[3][9][10];
};

struct RegParams {
  float beta, beta1, delta, delta1, regw;
};

// Upper bound of the detector columns whose ray can touch one tile at an
// angle (host side, for the flattened (angle, column) work list).
int tile_count_bound(float sinq, float cosq, int TW, int TH);

// Launchers (kernels.hip).  All are asynchronous on `st`.
// Ray table; `flags` adds the window flags (kRayWhole, kRayWraps, kRay1D)
// from one walk per ray, which the tile kernels need and whole-ray walks
// do not.
hipError_t launch_ray_info(const Geometry &g, RayInfo *rays, hipStream_t st, bool flags = true);
hipError_t launch_forward(const Geometry &g, const RayInfo *rays, int sb, int bundle_angles,
                          int lanes_per_ray, int mode, bool exact, const float *xb,
                          const float *sino, float *out, unsigned *maxw, hipStream_t st);
// One slice of a batched reconstruction: its own geometry (P, C = N,
// centre), tables and buffers; one slice per geometry.
struct BatchSliceDev {
  GridInfo g;
  const AngleInfo *ang;   // [P]
  const RayInfo *rays;    // [P][C]
  const int *count;       // [P] columns-per-tile bound for the batch tile
  int P;
  int alg;                // 0 SIRT, 1 MLEM
  int iters;
  float yi0;
  float maxlen;           // longest segment of the geometry
  float denmax;           // bound on the sum of lengths per pixel (max of den)
  const float *sino;      // [P][C]
  float *x;               // [N][N] image (updated in place)
  float *w;               // [P][C] per-ray weights
  float *z;               // [N][N] A^T w
  const float *den;       // [N][N] A^T 1
  unsigned *maxw;         // bits of max |w| of this slice
  int wraps;              // some ray has a segment with ix == N (kRayWraps)
};

// Work queues of one batch: phase f's items of slice order_f[k] are
// [start_f[k], start_f[k+1]); workgroups of a persistent kernel take items
// from counters[f] until they run out.  Slices are ordered by decreasing
// cost per item, so the tail of each phase is filled with short items.
struct BatchQueues {
  const int *order[3];    // forward, backprojector, update
  const int *start[3];    // [n + 1] each
  int n;
  unsigned *counters;     // [3]
  unsigned *maxw_all;     // [n]
};

// Compiled tile shapes of the batched backprojector (one slice per tile).
#define TRACE_BATCH_TILES(X) X(16, 16) X(32, 32) X(64, 64)

// Forward projector shapes compiled into the library: (slices per lane
// batch, lanes per ray).
#define TRACE_FWD_SHAPES(X)                                                \
  X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(16, 1) X(32, 1) X(8, 2) X(16, 2) X(32, 2) \
  X(16, 4) X(32, 4)

// Tile shapes (tile_w, tile_h, slices per workgroup) compiled into the
// backprojector; the C-ABI accepts exactly these.
#define TRACE_BP_SHAPES(X)                                                 \
  X(16, 16, 8) X(16, 16, 16) X(16, 16, 4) X(24, 24, 8) X(32, 16, 8)          \
  X(16, 32, 8) X(32, 32, 4) X(32, 32, 8) X(8, 8, 16) X(64, 64, 2)           \
  X(32, 32, 1) X(64, 64, 1) X(32, 32, 2) X(16, 16, 2) X(16, 16, 1)          \
  X(16, 16, 32) X(32, 32, 16) X(32, 16, 16) X(24, 24, 16)

// Backprojection in the reference's summation order (ordered.hip): w
// [P][C][S] weights, xb the slice-batched image the ray pass read (xsb
// slices per slot batch; kOrdNum only), sb slices per thread (a divisor of
// xsb for kOrdNum; at most 8 in the per-pixel form).  With tile_overflow
// non-null the tile form runs (geometries without wrapped segments) and
// ORs a nonzero value into *tile_overflow when a pixel needs more LDS
// slots than it has; without, the per-pixel form.
hipError_t launch_bp_ordered(const Geometry &g, const RayInfo *rays, int mode, bool exact,
                             const float *w, const float *xb, int xsb, int sb, float *out,
                             bool wraps, unsigned *tile_overflow, hipStream_t st);
bool ordered_sb_supported(int mode, int sb, int xsb);

hipError_t launch_backproject(const Geometry &g, const RayInfo *rays, const int *count,
                              const TileShape &ts, int mode, bool exact, const float *w,
                              const unsigned *maxw, float maxlen, float denmax, float *z,
                              bool wraps, hipStream_t st);
hipError_t launch_absmax_slices(const float *w, int P, int S, int C, unsigned *maxw,
                                hipStream_t st);
hipError_t launch_to_pcs(const float *r, int P, int S, int C, float *w, unsigned *maxw,
                         hipStream_t st);
hipError_t launch_pack(const float *src, int nsrc, int off, float fill, int64_t NN,
                       int nslots, int sb, float *xb, hipStream_t st);
hipError_t launch_update_em(bool sirt, float *x, const float *z, const float *den, int S,
                            int64_t NN, int sb, float *xb, hipStream_t st);
hipError_t launch_fill(float *p, int64_t n, float v, hipStream_t st);
// num_direct: z holds num summed as the reference does (launch_bp_ordered,
// kOrdNum); otherwise z = A^T w and num = -x * z.
hipError_t launch_update_pml(const float *x, float *xn, const float *z, const float *den,
                             int S, int N, float beta, bool num_direct, int sb, int nslots,
                             float *xb, hipStream_t st);
hipError_t launch_update_apmlr(const float *xpad, float *xnpad, const float *z,
                               const float *den, int S, int N, const RegParams &p,
                               const StencilWeights &w, int compat, bool num_direct,
                               int sb, int nslots, float *xb, hipStream_t st);
hipError_t launch_batch_reset(const BatchQueues &q, hipStream_t st);
hipError_t launch_batch_forward(const BatchSliceDev *d, const BatchQueues &q, int it,
                                int bundle_angles, bool exact, int grid, hipStream_t st);
hipError_t launch_batch_backproject(const BatchSliceDev *d, const BatchQueues &q, int it,
                                    int tw, int th, bool exact, int grid, size_t lds,
                                    hipStream_t st);
hipError_t launch_batch_update(const BatchSliceDev *d, const BatchQueues &q, int it, int grid,
                               hipStream_t st);
// Resident workgroups per CU of the batched kernels (0 if not compiled).
int batch_blocks_per_cu(int phase, int tw, int th, bool exact, size_t lds);
bool batch_tile_supported(int tw, int th);
hipError_t launch_adjoint_atomic(const Geometry &g, const RayInfo *rays, int mode, bool exact,
                                 const float *r, float *z, int nslices, hipStream_t st);
hipError_t launch_dump_rays(const Geometry &g, const RayInfo *rays, int n, const int *proj,
                            const int *col, int max_seg, int *counts, int *idx, float *len,
                            float *len2, hipStream_t st, bool exact);
hipError_t launch_ray_sums(const Geometry &g, const RayInfo *rays, bool exact, float *len_sum,
                           float *len2_sum, hipStream_t st);
hipError_t launch_count_segments(const Geometry &g, const RayInfo *rays,
                                 unsigned long long *total, unsigned long long *oor,
                                 unsigned *maxlen, unsigned long long *wraps, hipStream_t st);
hipError_t launch_check_finite(const float *v, int64_t n, int *flag, hipStream_t st);

}  // namespace trace_hip
