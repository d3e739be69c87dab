// handle.h -- host-side state of a C-ABI handle and the helpers shared by
// the single-stack entry points (capi.hip) and the batched ones (batch.hip).
// Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/trace_hip.h"
#include "internal.h"
#include "staged.h"

namespace trace_hip {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string &msg);
const char *last_error();
void clear_error();

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      return fail(e_ == hipErrorOutOfMemory ? TRACE_E_NOMEM : TRACE_E_HIP,   \
                  std::string(#expr) + ": " + hipGetErrorString(e_));        \
    }                                                                        \
  } while (0)

// Every device allocation of the library goes through dev_alloc and is
// released by dev_free: they keep a process-wide count of live allocations
// and can be told to fail once (trace_debug_fail_alloc), so that tests reach
// the out-of-memory paths from the host.
hipError_t device_alloc(void **p, size_t bytes);
void dev_free(void *p);

template <typename T>
hipError_t dev_alloc(T **p, size_t count) {
  return device_alloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

// HIP-event brackets around kernel launches, per phase.
struct PhaseTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double ms[TRACE_NUM_PHASES] = {0, 0, 0, 0};
  int64_t count[TRACE_NUM_PHASES] = {0, 0, 0, 0};
  hipEvent_t open_ev = nullptr;
  int open_phase = -1;

  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  void begin(int phase, hipStream_t st) {
    if (!on) return;
    open_ev = take();
    open_phase = phase;
    if (open_ev) (void)hipEventRecord(open_ev, st);
  }
  void end(hipStream_t st) {
    if (!on || !open_ev) return;
    hipEvent_t e = take();
    if (!e) return;
    (void)hipEventRecord(e, st);
    pending.push_back({open_phase, {open_ev, e}});
    open_ev = nullptr;
  }
  void collect() {
    for (auto &p : pending) {
      float t = 0.f;
      if (hipEventSynchronize(p.second.second) == hipSuccess &&
          hipEventElapsedTime(&t, p.second.first, p.second.second) == hipSuccess) {
        ms[p.first] += t;
        count[p.first] += 1;
      }
      pool.push_back(p.second.first);
      pool.push_back(p.second.second);
    }
    pending.clear();
  }
  void release() {
    collect();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace trace_hip

struct trace_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  trace_hip::Geometry geo;
  std::vector<float> theta;
  std::vector<trace_hip::AngleInfo> ang;
  bool exact = true;
  int backprojector = 0;  // 0 tile (LDS), 1 whole-ray global atomics
  int fwd_sb = 16;        // slices per forward lane (xb layout)
  int fwd_bundle = 8;     // adjacent angles per forward wavefront
  int fwd_lanes = 1;      // lanes per ray of the forward projector
  int fwd_staged = 1;     // TRACE_OPT_FWD_STAGED
  trace_hip::StagedPlan staged;  // slab plan of the staged forward projector
  bool staged_ready = false;
  trace_hip::TileShape tile{16, 16, 16};
  trace_hip::RayInfo *rays = nullptr;     // [P][C]
  bool rays_ready = false;
  int *count = nullptr;        // [P] columns-per-tile bound for count_shape
  trace_hip::TileShape count_shape{0, 0, 0};
  float *den = nullptr;        // [N][N]
  bool den_ready = false;
  bool den_exact = true;
  bool den_ordered = false;    // den summed in the reference's order
  int ordered = 1;             // TRACE_OPT_ORDERED_SUMS
  int ord_tile = -1;           // tile form of the ordered sums: -1 untried, 1 usable, 0 not
  float *xb = nullptr;         // [ceil(S/fwd_sb)][N][N][fwd_sb]
  int xb_sb = 0;
  float *w = nullptr;          // [P][C][S] per-ray weights
  float *z = nullptr;          // [S][N][N]
  float *xa = nullptr;         // [S+2][N][N] work image (PML/APMLR)
  float *xc = nullptr;         // [S+2][N][N]
  int *flag = nullptr;
  unsigned long long *counters = nullptr;  // [4]: segments, outside image, max length bits,
                                           // rays with a segment at ix == N
  unsigned *maxw = nullptr;    // [S] bits of max |w| per slice
  float maxlen = 0.f;          // longest segment of the geometry
  float den_max = 0.f;         // bound on the sum of segment lengths per pixel
  int64_t nnz = 0, nnz_oor = 0;
  bool wraps = false;  // some ray has a segment with ix == N (the tile pass for it runs)
  trace_hip::PhaseTimer timer;
};

// Launch bracketed by the phase timer.
#define TIMED(phase, expr)              \
  do {                                  \
    h->timer.begin((phase), h->stream); \
    HIP_TRY(expr);                      \
    h->timer.end(h->stream);            \
  } while (0)

namespace trace_hip {

// Number of batches of sb slices covering S slices.
int nbatch(int S, int sb);

// Ray table and segment statistics (count, longest segment); one
// synchronous read-back per handle.
int ensure_rays(trace_handle *h);
// Per-angle bound of the columns touching a tile of the handle's tile shape.
int ensure_count(trace_handle *h);
// den = A^T 1 of one slice.
int ensure_den(trace_handle *h);
// Work buffers: weights, A^T w, max |w|, the slice-batched image copy
// (unless `batched_copy` is false) and, with `images`, the two padded
// work images of PML/APMLR.
int ensure_work(trace_handle *h, bool images, bool batched_copy = true);
// A^T applied with the configured backprojector (see capi.hip).
int backproject(trace_handle *h, int mode, const float *w, float *z, int phase);
// den = A^T 1 summed in the reference's order into out [N][N]: the tile
// form where the geometry allows it (decided once, synchronously), else
// the per-pixel form.
int ordered_den(trace_handle *h, float *out);
// Backprojected sums of w in the reference's order (kOrdNum: PML/APMLR num
// with the x factor from the slice-batched copy xb; kOrdEM: SIRT/MLEM num).
int backproject_ordered(trace_handle *h, int mode, const float *w, float *z, int phase);
// TRACE_E_NONFINITE if v[0..n) holds NaN or Inf (synchronous).
int check_finite(trace_handle *h, const float *v, int64_t n);

}  // namespace trace_hip
