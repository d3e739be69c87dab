// capi.hip -- implementation of the C-ABI declared in include/trace_hip.h.
//
// The host side does what the reference's TraceMetadata constructor and the
// per-row prologue of Reduce do (angle tables with the host libm, detector
// coordinates, centre/mov rules) and drives the per-iteration kernel
// sequence that replaces DISPEngineReduction + ReconSpace:
//
//   forward projector   y = A x, per-ray weight w        (k_forward)
//   tile backprojector  z = A^T w                        (k_bp_tile)
//   image update        x' from x, z, den [, stencil]    (k_update_*)
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/trace_hip.h"
#include "handle.h"
#include "internal.h"

using namespace trace_hip;


namespace trace_hip {

namespace {
thread_local std::string g_err;
std::atomic<long long> g_live_allocs{0};
std::atomic<int> g_fail_after{-1};  // successful allocations before one fails
}  // namespace

hipError_t device_alloc(void **p, size_t bytes) {
  *p = nullptr;
  int k = g_fail_after.load();
  while (k >= 0) {
    if (g_fail_after.compare_exchange_weak(k, k - 1)) {
      if (k == 0) return hipErrorOutOfMemory;  // the injected failure (then off)
      break;
    }
  }
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess && *p) g_live_allocs.fetch_add(1);
  return e;
}

void dev_free(void *p) {
  if (p && hipFree(p) == hipSuccess) g_live_allocs.fetch_sub(1);
}

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int nbatch(int S, int sb) { return (S + sb - 1) / sb; }

const char *last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

// Ray table and segment statistics (count, longest segment) of the
// geometry; one synchronous read-back per handle.
int ensure_rays(trace_handle *h) {
  if (h->rays_ready) return TRACE_OK;
  if (!h->rays) HIP_TRY(dev_alloc(&h->rays, (size_t)h->geo.rays_per_slice()));
  TIMED(TRACE_PHASE_SETUP, launch_ray_info(h->geo, h->rays, h->stream));
  HIP_TRY(hipMemsetAsync(h->counters, 0, 4 * sizeof(unsigned long long), h->stream));
  TIMED(TRACE_PHASE_SETUP, launch_count_segments(h->geo, h->rays, h->counters, h->counters + 1,
                                                 reinterpret_cast<unsigned *>(h->counters + 2),
                                                 h->counters + 3, h->stream));
  unsigned long long c[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(c, h->counters, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->nnz = (int64_t)c[0];
  h->nnz_oor = (int64_t)c[1];
  h->wraps = c[3] > 0;
  const unsigned bits = (unsigned)c[2];
  std::memcpy(&h->maxlen, &bits, sizeof(float));
  h->rays_ready = true;
  return TRACE_OK;
}

// Slab plan of the staged forward projector (once per geometry).
int ensure_staged(trace_handle *h) {
  if (h->staged_ready) return TRACE_OK;
  if (int rc = ensure_rays(h)) return rc;
  int nwg = 0, nslab = 0, ngc = 0;
  staged_plan_size(h->geo, nwg, nslab, ngc);
  if (!h->staged.axis) HIP_TRY(dev_alloc(&h->staged.axis, (size_t)nwg));
  if (!h->staged.range) HIP_TRY(dev_alloc(&h->staged.range, (size_t)nwg * nslab));
  TIMED(TRACE_PHASE_SETUP, launch_staged_plan(h->geo, h->rays, h->staged, h->stream));
  h->staged_ready = true;
  return TRACE_OK;
}

int ensure_count(trace_handle *h) {
  const TileShape &t = h->tile;
  if (h->count && h->count_shape.tw == t.tw && h->count_shape.th == t.th) return TRACE_OK;
  // the one-slice tile kernel enumerates ceil(sqrt(tw^2 + th^2)) + 4 column
  // ranks per angle (bp_tile_loop): every count must fit
  int ranks = 0;
  while (ranks * ranks < t.tw * t.tw + t.th * t.th) ++ranks;
  ranks += 4;
  std::vector<int> cnt(h->geo.P);
  for (int p = 0; p < h->geo.P; ++p) {
    cnt[p] = tile_count_bound(h->ang[p].sinq, h->ang[p].cosq, t.tw, t.th);
    if (cnt[p] > ranks) return fail(TRACE_E_UNSUPPORTED, "tile column bound exceeds the rank count");
  }
  if (!h->count) HIP_TRY(dev_alloc(&h->count, (size_t)h->geo.P));
  HIP_TRY(hipMemcpyAsync(h->count, cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));  // cnt leaves scope
  h->count_shape = t;
  return TRACE_OK;
}

// A^T applied with the configured backprojector.  mode: kBpOutSino (w
// [P][C][S] -> z [S][N][N], with h->maxw holding max |w| per slice, as left
// by the forward projector or launch_to_pcs), kBpOutOnes / kBpOutCount
// (-> z [N][N]).
int backproject(trace_handle *h, int mode, const float *w, float *z, int phase) {
  if (int rc = ensure_rays(h)) return rc;
  if (h->backprojector == 0) {
    if (int rc = ensure_count(h)) return rc;
    TIMED(phase, launch_backproject(h->geo, h->rays, h->count, h->tile, mode, h->exact, w,
                                    h->maxw, h->maxlen, h->den_max, z, h->wraps, h->stream));
  } else {
    const int ns = (mode == kBpOutSino) ? h->geo.S : 1;
    TIMED(phase, launch_fill(z, (int64_t)ns * h->geo.NN(), 0.f, h->stream));
    TIMED(phase, launch_adjoint_atomic(h->geo, h->rays, mode, h->exact, w, z, ns, h->stream));
  }
  return TRACE_OK;
}

int ensure_den(trace_handle *h) {
  const bool ord = h->ordered >= 1;
  if (h->den_ready && h->den_exact == h->exact && h->den_ordered == ord) return TRACE_OK;
  // the den the updates divide by: the reference's float sum in its order,
  // or the exact fixed-point sum rounded once
  if (ord) {
    if (int rc = ordered_den(h, h->den)) return rc;
  } else {
    if (int rc = backproject(h, kBpOutOnes, nullptr, h->den, TRACE_PHASE_SETUP)) return rc;
  }
  // largest column sum: bounds every pixel's sum of segment lengths, which
  // sets the 32-bit fixed-point scale of the weights backprojector
  unsigned *bits = reinterpret_cast<unsigned *>(h->flag);
  TIMED(TRACE_PHASE_SETUP, launch_fill(reinterpret_cast<float *>(bits), 1, 0.f, h->stream));
  TIMED(TRACE_PHASE_SETUP,
        launch_absmax_slices(h->den, 1, 1, (int)h->geo.NN(), bits, h->stream));
  unsigned b = 0;
  HIP_TRY(hipMemcpyAsync(&b, bits, sizeof(b), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  float m = 0.f;
  std::memcpy(&m, &b, sizeof(m));
  // The fixed-point den is exact to far below 2^-20 relative.  The float
  // sum in order is within (n - 1) * 2^-24 of the exact one for n terms,
  // and a pixel takes at most 6 segments per projection (3 candidate
  // columns, 2 segments each), so the margin covers 6 P terms.
  const double margin = 1.0 + 0x1p-16 + (ord ? 6.0 * h->geo.P * 0x1p-24 : 0.0);
  h->den_max = std::isfinite(m) && m > 0.f ? (float)(m * margin) : 1.0f;
  h->den_ready = true;
  h->den_exact = h->exact;
  h->den_ordered = ord;
  return TRACE_OK;
}

// Slices per thread of the ordered backprojector: up to 32 in the tile form
// and 8 in the per-pixel form, a divisor of the slice-batched copy's batch
// for kOrdNum.
static int ordered_batch(int mode, int S, int xsb, bool tile) {
  int sb = tile ? 32 : 8;
  while (sb > 1 && sb / 2 >= S) sb /= 2;
  if (mode == kOrdNum)
    while (sb > 1 && (xsb % sb) != 0) sb /= 2;
  return sb;
}

int ordered_den(trace_handle *h, float *out) {
  if (int rc = ensure_rays(h)) return rc;
  if (h->ord_tile != 0 && !h->wraps) {
    unsigned *ovf = reinterpret_cast<unsigned *>(h->flag);
    HIP_TRY(hipMemsetAsync(ovf, 0, sizeof(unsigned), h->stream));
    TIMED(TRACE_PHASE_SETUP, launch_bp_ordered(h->geo, h->rays, kOrdDen, h->exact, nullptr,
                                               nullptr, 1, 1, out, false, ovf, h->stream));
    unsigned f = 0;
    HIP_TRY(hipMemcpyAsync(&f, ovf, sizeof(f), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->ord_tile = f ? 0 : 1;
    if (!f) return TRACE_OK;
  } else {
    h->ord_tile = 0;
  }
  TIMED(TRACE_PHASE_SETUP, launch_bp_ordered(h->geo, h->rays, kOrdDen, h->exact, nullptr, nullptr,
                                             1, 1, out, h->wraps, nullptr, h->stream));
  return TRACE_OK;
}

int backproject_ordered(trace_handle *h, int mode, const float *w, float *z, int phase) {
  if (int rc = ensure_rays(h)) return rc;
  if (h->ord_tile < 0)  // decided by the den pass (same walks, same slots)
    if (int rc = ordered_den(h, h->z)) return rc;
  if (mode == kOrdNum && !h->xb) return fail(TRACE_E_INVALID, "no slice-batched image copy");
  const bool tile = h->ord_tile == 1;
  const int sb = ordered_batch(mode, h->geo.S, h->xb_sb, tile);
  // the tile form's overflow flag was checked on this geometry by the den pass
  unsigned *ovf = tile ? reinterpret_cast<unsigned *>(h->flag) : nullptr;
  TIMED(phase, launch_bp_ordered(h->geo, h->rays, mode, h->exact, w, h->xb, h->xb_sb, sb, z,
                                 h->wraps, ovf, h->stream));
  return TRACE_OK;
}

int ensure_work(trace_handle *h, bool images, bool batched_copy) {
  const size_t nn = (size_t)h->geo.NN();
  const size_t S = (size_t)h->geo.S;
  if (h->xb && h->xb_sb != h->fwd_sb) {
    dev_free(h->xb);
    h->xb = nullptr;
  }
  if (!h->xb && batched_copy) {
    HIP_TRY(dev_alloc(&h->xb, (size_t)nbatch((int)S, h->fwd_sb) * h->fwd_sb * nn));
    h->xb_sb = h->fwd_sb;
  }
  if (!h->w) HIP_TRY(dev_alloc(&h->w, (size_t)h->geo.P * S * h->geo.C + kWeightPad));
  if (!h->maxw) HIP_TRY(dev_alloc(&h->maxw, S));
  if (!h->z) HIP_TRY(dev_alloc(&h->z, S * nn));
  if (images) {
    if (!h->xa) HIP_TRY(dev_alloc(&h->xa, (S + 2) * nn));
    if (!h->xc) HIP_TRY(dev_alloc(&h->xc, (S + 2) * nn));
  }
  return TRACE_OK;
}

int check_finite(trace_handle *h, const float *v, int64_t n) {
  HIP_TRY(hipMemsetAsync(h->flag, 0, sizeof(int), h->stream));
  HIP_TRY(launch_check_finite(v, n, h->flag, h->stream));
  int flag = 0;
  HIP_TRY(hipMemcpyAsync(&flag, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (flag) return fail(TRACE_E_NONFINITE, "sinogram holds NaN or Inf values");
  return TRACE_OK;
}

}  // namespace trace_hip

// Scratch of trace_forward_noalloc: the ray table and the slice-batched
// image copy of one call, in device memory the library never allocates at
// run time.  Every GPU the library runs on holds the array until the
// process ends; calls on one GPU share it.  In host code `scratch` names a
// host-side stand-in, never the device memory: its device address comes
// from hipGetSymbolAddress, once per device (scratch_address).
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up on the
// device's first call.
int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(TRACE_E_INVALID, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TRACE_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return TRACE_OK;
}

}  // namespace

namespace {

constexpr float kPI = 3.14159265358979f;  // trace_utils.h:8

// trace_utils.cc:41-46
int quadrant_of(float q) {
  return ((q >= 0 && q < kPI / 2) || (q >= kPI && q < 3 * kPI / 2)) ? 1 : 0;
}

// trace_data.h:97-101: default centre, mov and its unconditional nudge.
float mov_of(int C, float center) {
  float c = center;
  if (c <= 0.) c = (float)((double)(float)C / 2. + 1.);
  float mov = (float)((double)(float)C / 2. - (double)c);
  if ((double)(mov - std::ceil(mov)) < 1e-6) mov = (float)((double)mov + 1e-6);
  return mov;
}

// Launch shapes compiled into the library (kernels.hip, launch_backproject).
bool fwd_shape_supported(int sb, int lanes) {
#define TRACE_FWD_MATCH(SB_, LPR_) \
  if (sb == SB_ && lanes == LPR_) return true;
  TRACE_FWD_SHAPES(TRACE_FWD_MATCH)
#undef TRACE_FWD_MATCH
  return false;
}

bool tile_shape_supported(int tw, int th, int sb) {
#define TRACE_SHAPE_MATCH(TW_, TH_, SB_) \
  if (tw == TW_ && th == TH_ && sb == SB_) return true;
  TRACE_BP_SHAPES(TRACE_SHAPE_MATCH)
#undef TRACE_SHAPE_MATCH
  return false;
}


// Forward projection of xb into weights w (mode kFwdOutSIRT/EM) or y
// (kFwdOutY), with max |w| per slice folded into h->maxw.
int forward(trace_handle *h, int mode, const float *sino, float *out) {
  if (mode != kFwdOutY)
    TIMED(TRACE_PHASE_SETUP,
          launch_fill(reinterpret_cast<float *>(h->maxw), h->geo.S, 0.f, h->stream));
  if (h->fwd_staged && h->fwd_sb == kStSlices) {
    if (int rc = ensure_staged(h)) return rc;
    TIMED(TRACE_PHASE_FORWARD, launch_forward_staged(h->geo, h->rays, h->staged, mode, h->exact,
                                                     h->xb, sino, out, h->maxw, h->stream));
    return TRACE_OK;
  }
  TIMED(TRACE_PHASE_FORWARD, launch_forward(h->geo, h->rays, h->fwd_sb, h->fwd_bundle,
                                            h->fwd_lanes, mode, h->exact, h->xb, sino, out,
                                            h->maxw, h->stream));
  return TRACE_OK;
}

// APMLR weights (apmlr.cc:40-51, 94-101, 245-251 and the one-sided variants
// in FGTop/FGBottom): float total from a double expression, axial 1/total in
// float, diagonal (1/sqrt 2)/total in double, z regw/total in float.
void apmlr_weights(float regw, StencilWeights &w) {
  static const int nin[9] = {8, 5, 5, 5, 5, 3, 3, 3, 3};
  static const int naxial[9] = {4, 3, 3, 3, 3, 2, 2, 2, 2};
  std::memset(&w, 0, sizeof(w));
  for (int zm = 0; zm < 3; ++zm) {
    const int nz = zm == 0 ? 2 : 1;
    for (int cls = 0; cls < 9; ++cls) {
      const double base = cls == 0 ? 4 + 4 / std::sqrt(2.)
                                   : (cls <= 4 ? 3 + 2 / std::sqrt(2.) : 2 + 1 / std::sqrt(2.));
      const double zterm = nz == 2 ? (double)(2 * regw) : (double)regw;
      const float total = (float)(base + zterm);
      for (int q = 0; q < nin[cls]; ++q)
        w.w[zm][cls][q] = q < naxial[cls] ? 1 / total : (float)(1 / std::sqrt(2.) / total);
      for (int q = 0; q < nz; ++q) w.w[zm][cls][nin[cls] + q] = regw / total;
    }
  }
}

int validate_params(trace_alg_t alg, const trace_params_t *prm) {
  if (alg == TRACE_PML || alg == TRACE_APMLR) {
    if (!prm) return fail(TRACE_E_INVALID, "PML/APMLR need parameters");
    if (!(prm->beta > 0.f) || !std::isfinite(prm->beta))
      return fail(TRACE_E_INVALID, "beta must be finite and > 0 (F = 0 otherwise)");
    if (alg == TRACE_APMLR) {
      if (!(prm->delta > 0.f) || !(prm->delta1 > 0.f) || !std::isfinite(prm->delta) ||
          !std::isfinite(prm->delta1))
        return fail(TRACE_E_INVALID, "delta and delta1 must be finite and > 0");
      if (!std::isfinite(prm->beta1) || !std::isfinite(prm->regw) || prm->beta1 < 0.f ||
          prm->regw < 0.f)
        return fail(TRACE_E_INVALID, "beta1 and regw must be finite and >= 0");
      if (prm->apmlr_compat != 0 && prm->apmlr_compat != 1)
        return fail(TRACE_E_INVALID, "apmlr_compat must be 0 or 1");
    }
  }
  return TRACE_OK;
}

}  // namespace

extern "C" {

const char *trace_last_error(void) { return g_err.c_str(); }

const char *trace_version(void) { return "trace-hip 0.3 (gfx950)"; }

void trace_debug_fail_alloc(int after) { g_fail_after.store(after < 0 ? -1 : after); }

int64_t trace_debug_live_allocations(void) { return (int64_t)g_live_allocs.load(); }

void trace_deg2rad(const float *deg, float *rad, int n) {
  // trace_utils.cc:22: float product, double division, float result
  if (!deg || !rad) return;
  for (int i = 0; i < n; ++i) rad[i] = (float)((double)(deg[i] * kPI) / 180.0);
}

void trace_absolute(float *v, int64_t n) {
  if (!v) return;
  for (int64_t i = 0; i < n; ++i) v[i] = std::fabs(v[i]);
}

int trace_create(trace_handle_t **out, const trace_geom_t *g, int device,
                 void *hip_stream) {
  g_err.clear();
  if (!out) return fail(TRACE_E_INVALID, "handle pointer is null");
  *out = nullptr;
  if (!g) return fail(TRACE_E_INVALID, "geometry is null");
  if (!g->theta) return fail(TRACE_E_INVALID, "theta is null");
  if (g->num_projs <= 0 || g->num_slices <= 0 || g->num_cols <= 1 || g->num_grids <= 1)
    return fail(TRACE_E_INVALID, "P, S must be >= 1 and C, N >= 2");
  if (g->num_cols > 32768 || g->num_grids > 32768)
    return fail(TRACE_E_INVALID, "num_cols or num_grids above 32768");
  if (g->num_slices > 65535) return fail(TRACE_E_INVALID, "num_slices above 65535");
  if (!std::isfinite(g->center)) return fail(TRACE_E_INVALID, "center is not finite");
  if (device < 0) return fail(TRACE_E_INVALID, "device index is negative");
  for (int p = 0; p < g->num_projs; ++p)
    if (!std::isfinite(g->theta[p])) return fail(TRACE_E_INVALID, "theta holds NaN/Inf");

  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device >= ndev) return fail(TRACE_E_INVALID, "device index out of range");
  HIP_TRY(hipSetDevice(device));

  auto *h = new trace_handle();
  h->device = device;
  h->stream = (hipStream_t)hip_stream;
  // default launch shape by stack depth (the fastest measured at 2048^2,
  // 1500 angles; trace_amd.autotune refines it for a given geometry)
  const int S = g->num_slices;
  if (S >= 32) {
    h->fwd_sb = 32, h->fwd_lanes = 4, h->tile = TileShape{16, 16, 32};
  } else if (S >= 16) {
    h->fwd_sb = 16, h->fwd_lanes = 2, h->tile = TileShape{16, 16, 16};
  } else if (S >= 8) {
    h->fwd_sb = 8, h->fwd_lanes = 2, h->tile = TileShape{16, 16, 8};
  } else if (S >= 3) {
    h->fwd_sb = 4, h->fwd_lanes = 1, h->tile = TileShape{32, 32, 4};
  } else if (S == 2) {
    h->fwd_sb = 2, h->fwd_lanes = 1, h->tile = TileShape{32, 32, 2};
  } else {
    // one slice: one atomic per segment, so larger tiles pay off as the
    // image grows (the autotuner's picks at 1500 angles: 16x16 up to 512
    // pixels, 32x32 at 1024 and 1536, 64x64 at 2048)
    const int n = g->num_grids > 0 ? g->num_grids : g->num_cols;
    const int side = n >= 1792 ? 64 : (n >= 768 ? 32 : 16);
    h->fwd_sb = 1, h->fwd_lanes = 1, h->tile = TileShape{side, side, 1};
  }
  // 16 angles x 1 column per wavefront at four lanes per ray (measured at
  // 2048^2 / 1500 angles: 110 vs 116 ms for 8 angles x 2 columns), 8 x 8/LPR
  // otherwise
  h->fwd_bundle = (h->fwd_lanes == 4) ? 16 : 8;
  Geometry &geo = h->geo;
  geo.P = g->num_projs;
  geo.S = g->num_slices;
  geo.C = g->num_cols;
  geo.N = g->num_grids;
  geo.center = g->center;
  geo.mov = mov_of(geo.C, g->center);
  geo.yi0 = (float)((double)(1 - geo.C) / 2. + (double)geo.mov);
  geo.grid.N = geo.N;
  geo.grid.C = geo.C;
  geo.grid.gx0 = (float)(-(double)geo.N / 2.);
  geo.grid.gxN = (float)(-(double)geo.N / 2. + geo.N);
  geo.grid.mg = (float)((double)geo.N / 2.);
  h->theta.assign(g->theta, g->theta + geo.P);

  h->ang.resize(geo.P);
  for (int p = 0; p < geo.P; ++p) {
    const float q = h->theta[p];
    h->ang[p].sinq = sinf(q);  // host libm, as sirt.cc:119-120
    h->ang[p].cosq = cosf(q);
    h->ang[p].quadrant = quadrant_of(q);
    h->ang[p].pad = 0;
  }
  std::vector<float> yi(geo.C);
  for (int c = 0; c < geo.C; ++c)  // sirt.cc:129
    yi[c] = (float)((double)(1 - geo.C) / 2. + (double)c + (double)geo.mov);

  auto cleanup_fail = [&](int code) {
    trace_destroy(h);
    return code;
  };
  hipError_t e;
  if ((e = dev_alloc(&geo.d_ang, (size_t)geo.P)) != hipSuccess ||
      (e = dev_alloc(&geo.d_yi, (size_t)geo.C)) != hipSuccess ||
      (e = dev_alloc(&h->den, (size_t)geo.NN())) != hipSuccess ||
      (e = dev_alloc(&h->flag, 1)) != hipSuccess ||
      (e = dev_alloc(&h->counters, 4)) != hipSuccess)
    return cleanup_fail(fail(e == hipErrorOutOfMemory ? TRACE_E_NOMEM : TRACE_E_HIP,
                             std::string("allocation: ") + hipGetErrorString(e)));
  if ((e = hipMemcpy(geo.d_ang, h->ang.data(), h->ang.size() * sizeof(AngleInfo),
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(geo.d_yi, yi.data(), yi.size() * sizeof(float),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup_fail(fail(TRACE_E_HIP, std::string("upload: ") + hipGetErrorString(e)));
  *out = h;
  return TRACE_OK;
}

void trace_destroy(trace_handle_t *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  else (void)hipDeviceSynchronize();
  h->timer.release();
  for (void *p : {(void *)h->geo.d_ang, (void *)h->geo.d_yi, (void *)h->rays, (void *)h->count,
                  (void *)h->den, (void *)h->xb, (void *)h->w, (void *)h->z, (void *)h->xa,
                  (void *)h->xc, (void *)h->flag, (void *)h->counters, (void *)h->maxw,
                  (void *)h->staged.axis, (void *)h->staged.range})
    dev_free(p);
  delete h;
}

int trace_set_option(trace_handle_t *h, int option, int value) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  switch (option) {
    case TRACE_OPT_EXACT_SQRT:
      if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
      h->exact = value != 0;
      return TRACE_OK;
    case TRACE_OPT_PHASE_TIMING:
      if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
      h->timer.on = value != 0;
      return TRACE_OK;
    case TRACE_OPT_BACKPROJECTOR:
      if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
      h->backprojector = value;
      return TRACE_OK;
    case TRACE_OPT_FWD_BUNDLE:
      if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
        return fail(TRACE_E_INVALID, "angles per bundle must be 1, 2, 4, 8 or 16");
      if (value * h->fwd_lanes > 64)
        return fail(TRACE_E_INVALID, "angles per bundle x lanes per ray exceeds 64");
      h->fwd_bundle = value;
      return TRACE_OK;
    case TRACE_OPT_ORDERED_SUMS:
      if (value < 0 || value > 2) return fail(TRACE_E_INVALID, "value must be 0, 1 or 2");
      h->ordered = value;
      return TRACE_OK;
    case TRACE_OPT_ORDERED_TILE:
      if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
      h->ord_tile = value ? -1 : 0;
      h->den_ready = false;  // den is recomputed (same bits) by the chosen form
      return TRACE_OK;
    case TRACE_OPT_FWD_LANES:
      if (value != 1 && value != 2 && value != 4)
        return fail(TRACE_E_INVALID, "lanes per ray must be 1, 2 or 4");
      if (!fwd_shape_supported(h->fwd_sb, value))
        return fail(TRACE_E_UNSUPPORTED, "slice batch and lanes per ray not compiled");
      h->fwd_lanes = value;
      return TRACE_OK;
    case TRACE_OPT_FWD_STAGED:
      if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
      h->fwd_staged = value;
      return TRACE_OK;
    default:
      return fail(TRACE_E_UNSUPPORTED, "unknown option");
  }
}

int trace_get_option(trace_handle_t *h, int option, int *value) {
  clear_error();
  if (!h || !value) return fail(TRACE_E_INVALID, "null argument");
  switch (option) {
    case TRACE_OPT_EXACT_SQRT: *value = h->exact ? 1 : 0; return TRACE_OK;
    case TRACE_OPT_PHASE_TIMING: *value = h->timer.on ? 1 : 0; return TRACE_OK;
    case TRACE_OPT_BACKPROJECTOR: *value = h->backprojector; return TRACE_OK;
    case TRACE_OPT_FWD_BUNDLE: *value = h->fwd_bundle; return TRACE_OK;
    case TRACE_OPT_FWD_LANES: *value = h->fwd_lanes; return TRACE_OK;
    case TRACE_OPT_ORDERED_SUMS: *value = h->ordered; return TRACE_OK;
    case TRACE_OPT_ORDERED_TILE: *value = h->ord_tile != 0 ? 1 : 0; return TRACE_OK;
    case TRACE_OPT_FWD_STAGED: *value = h->fwd_staged; return TRACE_OK;
    default: return fail(TRACE_E_UNSUPPORTED, "unknown option");
  }
}

int trace_set_launch_shape(trace_handle_t *h, int fwd_batch, int tile_w, int tile_h,
                           int tile_batch) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (fwd_batch != 1 && fwd_batch != 2 && fwd_batch != 4 && fwd_batch != 8 &&
      fwd_batch != 16 && fwd_batch != 32)
    return fail(TRACE_E_UNSUPPORTED, "forward slice batch must be 1, 2, 4, 8, 16 or 32");
  if (!fwd_shape_supported(fwd_batch, 1))
    return fail(TRACE_E_UNSUPPORTED, "forward slice batch not compiled");
  if (!tile_shape_supported(tile_w, tile_h, tile_batch))
    return fail(TRACE_E_UNSUPPORTED, "tile shape not compiled into this library");
  h->fwd_sb = fwd_batch;
  if (!fwd_shape_supported(fwd_batch, h->fwd_lanes)) h->fwd_lanes = 1;
  h->tile = TileShape{tile_w, tile_h, tile_batch};
  return TRACE_OK;
}

int trace_get_launch_shape(trace_handle_t *h, int *fwd_batch, int *tile_w, int *tile_h,
                           int *tile_batch) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (fwd_batch) *fwd_batch = h->fwd_sb;
  if (tile_w) *tile_w = h->tile.tw;
  if (tile_h) *tile_h = h->tile.th;
  if (tile_batch) *tile_batch = h->tile.sb;
  return TRACE_OK;
}

int trace_reconstruct(trace_handle_t *h, trace_alg_t alg, int iters,
                      const float *sino, float *recon, const trace_params_t *prm) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!sino || !recon) return fail(TRACE_E_INVALID, "null buffer");
  if (alg < TRACE_SIRT || alg > TRACE_APMLR) return fail(TRACE_E_UNSUPPORTED, "unknown algorithm");
  if (iters < 0) return fail(TRACE_E_INVALID, "iters must be >= 0");
  if (int rc = validate_params(alg, prm)) return rc;
  HIP_TRY(hipSetDevice(h->device));
  const Geometry &g = h->geo;
  const int64_t NN = g.NN();
  const int S = g.S;
  if (int rc = check_finite(h, sino, (int64_t)g.P * S * g.C)) return rc;
  if (iters == 0) return TRACE_OK;
  if (int rc = ensure_rays(h)) return rc;
  if (int rc = ensure_den(h)) return rc;
  const bool reg = (alg == TRACE_PML || alg == TRACE_APMLR);
  if (int rc = ensure_work(h, reg)) return rc;
  hipStream_t st = h->stream;
  const int sb = h->fwd_sb;
  const int nslots = nbatch(S, sb) * sb;
  const int fmode = (alg == TRACE_SIRT) ? kFwdOutSIRT : kFwdOutEM;

  // sums in the reference's order: PML/APMLR num unless ordered == 0, SIRT/
  // MLEM num at ordered == 2 (den: ensure_den)
  const bool ord_num = reg ? h->ordered >= 1 : h->ordered >= 2;
  auto back = [&]() {
    return ord_num ? backproject_ordered(h, reg ? kOrdNum : kOrdEM, h->w, h->z, TRACE_PHASE_BACK)
                   : backproject(h, kBpOutSino, h->w, h->z, TRACE_PHASE_BACK);
  };
  if (alg == TRACE_SIRT || alg == TRACE_MLEM) {
    TIMED(TRACE_PHASE_SETUP, launch_pack(recon, S, 0, 0.f, NN, nslots, sb, h->xb, st));
    for (int it = 0; it < iters; ++it) {
      if (int rc = forward(h, fmode, sino, h->w)) return rc;
      if (int rc = back()) return rc;
      TIMED(TRACE_PHASE_UPDATE,
            launch_update_em(alg == TRACE_SIRT, recon, h->z, h->den, S, NN, sb, h->xb, st));
    }
    return TRACE_OK;
  }
  if (alg == TRACE_PML) {
    float *cur = recon, *alt = h->xa;
    TIMED(TRACE_PHASE_SETUP, launch_pack(cur, S, 0, 0.f, NN, nslots, sb, h->xb, st));
    for (int it = 0; it < iters; ++it) {
      if (int rc = forward(h, fmode, sino, h->w)) return rc;
      if (int rc = back()) return rc;
      TIMED(TRACE_PHASE_UPDATE, launch_update_pml(cur, alt, h->z, h->den, S, g.N, prm->beta,
                                                  ord_num, sb, nslots, h->xb, st));
      std::swap(cur, alt);
    }
    if (cur != recon)
      HIP_TRY(hipMemcpyAsync(recon, cur, sizeof(float) * S * NN, hipMemcpyDeviceToDevice, st));
    return TRACE_OK;
  }
  // APMLR on ghost-padded images [S+2][N][N]; ghosts hold the reference's
  // initial value 1 (apmlr_main.cc:170) and are never updated at one rank.
  // The ray pass of slice s reads padded slice s (compat) or s+1.
  StencilWeights w;
  apmlr_weights(prm->regw, w);
  RegParams rp{prm->beta, prm->beta1, prm->delta, prm->delta1, prm->regw};
  const int off = prm->apmlr_compat ? 0 : 1;
  float *pa = h->xa, *pb = h->xc;
  TIMED(TRACE_PHASE_SETUP, launch_fill(pa, (S + 2) * NN, 1.f, st));
  TIMED(TRACE_PHASE_SETUP, launch_fill(pb, (S + 2) * NN, 1.f, st));
  HIP_TRY(hipMemcpyAsync(pa + NN, recon, sizeof(float) * S * NN, hipMemcpyDeviceToDevice, st));
  TIMED(TRACE_PHASE_SETUP, launch_pack(pa, S + 2, off, 0.f, NN, nslots, sb, h->xb, st));
  for (int it = 0; it < iters; ++it) {
    if (int rc = forward(h, kFwdOutEM, sino, h->w)) return rc;
    if (int rc = back()) return rc;
    TIMED(TRACE_PHASE_UPDATE, launch_update_apmlr(pa, pb, h->z, h->den, S, g.N, rp, w,
                                                  prm->apmlr_compat, ord_num, sb, nslots,
                                                  h->xb, st));
    std::swap(pa, pb);
  }
  HIP_TRY(hipMemcpyAsync(recon, pa + NN, sizeof(float) * S * NN, hipMemcpyDeviceToDevice, st));
  return TRACE_OK;
}

int trace_reconstruct_host(trace_handle_t *h, trace_alg_t alg, int iters,
                           const float *sino, float *recon, const trace_params_t *prm) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!sino || !recon) return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  const Geometry &g = h->geo;
  const size_t nsino = (size_t)g.P * g.S * g.C, nrec = (size_t)g.S * g.NN();
  float *ds = nullptr, *dr = nullptr;
  hipError_t e = dev_alloc(&ds, nsino);
  if (e == hipSuccess) e = dev_alloc(&dr, nrec);
  if (e != hipSuccess) {
    dev_free(ds);
    return fail(e == hipErrorOutOfMemory ? TRACE_E_NOMEM : TRACE_E_HIP,
                std::string("allocation: ") + hipGetErrorString(e));
  }
  int rc = TRACE_OK;
  if ((e = hipMemcpyAsync(ds, sino, nsino * 4, hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(dr, recon, nrec * 4, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
    rc = fail(TRACE_E_HIP, std::string("upload: ") + hipGetErrorString(e));
  if (rc == TRACE_OK) rc = trace_reconstruct(h, alg, iters, ds, dr, prm);
  if (rc == TRACE_OK) {
    if ((e = hipMemcpyAsync(recon, dr, nrec * 4, hipMemcpyDeviceToHost, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
      rc = fail(TRACE_E_HIP, std::string("download: ") + hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(h->stream);
  }
  dev_free(ds);
  dev_free(dr);
  return rc;
}

int trace_forward(trace_handle_t *h, const float *x, float *y) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!x || !y) return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  if (int rc = ensure_rays(h)) return rc;
  if (int rc = ensure_work(h, false)) return rc;
  const Geometry &g = h->geo;
  const int sb = h->fwd_sb;
  TIMED(TRACE_PHASE_SETUP,
        launch_pack(x, g.S, 0, 0.f, g.NN(), nbatch(g.S, sb) * sb, sb, h->xb, h->stream));
  return forward(h, kFwdOutY, nullptr, y);
}

int trace_forward_noalloc(trace_handle_t *h, const float *x, float *y) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!x || !y) return fail(TRACE_E_INVALID, "null buffer");
  const Geometry &g = h->geo;
  const int sb = h->fwd_sb;
  const size_t rays_b = ((size_t)g.rays_per_slice() * sizeof(RayInfo) + 255) & ~size_t(255);
  const size_t xb_b = (size_t)nbatch(g.S, sb) * sb * (size_t)g.NN() * sizeof(float);
  if (rays_b + xb_b > kScratchBytes)
    return fail(TRACE_E_INVALID, "this geometry needs " +
                                     std::to_string((rays_b + xb_b + (1 << 20) - 1) >> 20) +
                                     " MiB of scratch; trace_forward_noalloc has 128 MiB");
  HIP_TRY(hipSetDevice(h->device));
  void *base = nullptr;
  if (int rc = scratch_address(h->device, &base)) return rc;
  RayInfo *rays = static_cast<RayInfo *>(base);
  float *xb = reinterpret_cast<float *>(static_cast<unsigned char *>(base) + rays_b);
  // every scratch byte read below is written first on the same stream
  HIP_TRY(launch_ray_info(g, rays, h->stream, false));
  HIP_TRY(launch_pack(x, g.S, 0, 0.f, g.NN(), nbatch(g.S, sb) * sb, sb, xb, h->stream));
  HIP_TRY(launch_forward(g, rays, sb, h->fwd_bundle, h->fwd_lanes, kFwdOutY, h->exact, xb,
                         nullptr, y, nullptr, h->stream));
  return TRACE_OK;
}

int trace_adjoint(trace_handle_t *h, const float *r, float *z) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!r || !z) return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  if (int rc = ensure_work(h, false)) return rc;
  if (int rc = ensure_den(h)) return rc;  // fixed-point scale of A^T
  TIMED(TRACE_PHASE_SETUP,
        launch_fill(reinterpret_cast<float *>(h->maxw), h->geo.S, 0.f, h->stream));
  TIMED(TRACE_PHASE_SETUP,
        launch_to_pcs(r, h->geo.P, h->geo.S, h->geo.C, h->w, h->maxw, h->stream));
  return backproject(h, kBpOutSino, h->w, z, TRACE_PHASE_BACK);
}

int trace_column_sums(trace_handle_t *h, float *den) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!den) return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  if (h->ordered == 0) return backproject(h, kBpOutOnes, nullptr, den, TRACE_PHASE_SETUP);
  return ordered_den(h, den);
}

int trace_segment_counts(trace_handle_t *h, float *counts) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!counts) return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  return backproject(h, kBpOutCount, nullptr, counts, TRACE_PHASE_SETUP);
}

int trace_ray_sums(trace_handle_t *h, float *len_sum, float *len2_sum) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (!len_sum && !len2_sum) return TRACE_OK;
  HIP_TRY(hipSetDevice(h->device));
  if (int rc = ensure_rays(h)) return rc;
  TIMED(TRACE_PHASE_SETUP,
        launch_ray_sums(h->geo, h->rays, h->exact, len_sum, len2_sum, h->stream));
  return TRACE_OK;
}

int trace_dump_rays(trace_handle_t *h, int n, const int *proj, const int *col,
                    int max_seg, int *counts, int *idx, float *len, float *len2) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  if (n < 0 || max_seg < 0) return fail(TRACE_E_INVALID, "negative size");
  if (n == 0) return TRACE_OK;
  if (!proj || !col || !counts || !idx || !len || !len2)
    return fail(TRACE_E_INVALID, "null buffer");
  HIP_TRY(hipSetDevice(h->device));
  // the ray ids index device tables: check them on the host first
  std::vector<int> hp(n), hc(n);
  HIP_TRY(hipMemcpy(hp.data(), proj, n * sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hc.data(), col, n * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i)
    if (hp[i] < 0 || hp[i] >= h->geo.P || hc[i] < 0 || hc[i] >= h->geo.C)
      return fail(TRACE_E_INVALID, "ray id out of range");
  if (int rc = ensure_rays(h)) return rc;
  HIP_TRY(launch_dump_rays(h->geo, h->rays, n, proj, col, max_seg, counts, idx, len, len2,
                           h->stream, h->exact));
  return TRACE_OK;
}

int trace_phase_times(trace_handle_t *h, double *ms, int64_t *launches, int reset) {
  g_err.clear();
  if (!h) return fail(TRACE_E_INVALID, "handle is null");
  HIP_TRY(hipSetDevice(h->device));
  h->timer.collect();
  for (int i = 0; i < TRACE_NUM_PHASES; ++i) {
    if (ms) ms[i] = h->timer.ms[i];
    if (launches) launches[i] = h->timer.count[i];
    if (reset) {
      h->timer.ms[i] = 0.0;
      h->timer.count[i] = 0;
    }
  }
  return TRACE_OK;
}

int trace_count_segments(trace_handle_t *h, int64_t *nnz, int64_t *out_of_image) {
  g_err.clear();
  if (!h || !nnz) return fail(TRACE_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  if (int rc = ensure_rays(h)) return rc;
  *nnz = h->nnz;
  if (out_of_image) *out_of_image = h->nnz_oor;
  return TRACE_OK;
}

}  // extern "C"
