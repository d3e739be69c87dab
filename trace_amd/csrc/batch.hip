// batch.hip -- batched reconstruction of many slices with different
// geometries (detector widths, angle counts, algorithms) in one captured
// HIP graph.
//
// Each slice keeps the single-geometry tables of a one-slice handle (angle
// table, ray table, per-angle tile column bound, den = A^T 1), built once.
// An iteration of the whole batch is four launches on every slice at once:
//
//   k_batch_reset      zero the work counters and each slice's max |w|
//   k_forward_batch    weights w = (p - A x)/a2 (SIRT) or p/(A x) (MLEM)
//   k_bp_batch         z = A^T w, LDS-tiled, fixed-point accumulation
//   k_update_batch     x += z/den (SIRT) or x *= z/den (MLEM)
//
// The three main kernels are persistent: as many workgroups as fit on the
// GPU take (slice, item) work from per-phase counters, slices ordered by
// decreasing cost per item so that short items fill the tail.  All
// iterations are captured once into a HIP graph on a private stream and
// replayed with one hipGraphLaunch per run, ordered after the caller's
// stream and before its later work by events.  The graph holds kernel
// nodes only (the counters are cleared by k_batch_reset).
//
// Per slice the arithmetic is that of trace_reconstruct with one slice: the
// same forward projector, the same deterministic fixed-point backprojector
// and the same update expressions (sirt.cc:21-64, mlem.cc:21-61), so every
// slice's result equals its own single-slice reconstruction bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/trace_hip.h"
#include "handle.h"
#include "internal.h"

using namespace trace_hip;

struct trace_batch {
  int device = 0;
  hipStream_t user = nullptr;  // the caller's stream
  hipStream_t cap = nullptr;   // private stream: capture and graph launches
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<trace_handle_t *> h;
  std::vector<trace_batch_slice_t> s;
  std::vector<int64_t> nnz;
  BatchSliceDev *d_desc = nullptr;  // [n]
  int *d_tables = nullptr;          // order[3][n], start[3][n + 1]
  unsigned *counters = nullptr;     // [3]
  unsigned *maxw = nullptr;         // [n]
  int *flag = nullptr;
  // the batch autotuner's usual pick on uneven slices of 256-2048 pixels
  int tw = 64, th = 64, bundle = 8;
  bool exact = true;
  bool prepared = false;
  int grid[3] = {0, 0, 0};
  size_t lds = 0;
  int64_t items[3] = {0, 0, 0};
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;  // by iteration count
};

namespace {

void drop_graphs(trace_batch *b) {
  for (auto &g : b->graphs) {
    (void)hipGraphExecDestroy(g.second.second);
    (void)hipGraphDestroy(g.second.first);
  }
  b->graphs.clear();
}

// Image side of a slice: num_grids, or C when it is 0.
int slice_grids(const trace_batch_slice_t &s) { return s.num_grids > 0 ? s.num_grids : s.num_cols; }

// Work items of one slice per phase for the batch's launch shape.
void slice_items(const trace_batch *b, int i, int64_t out[3]) {
  const int P = b->s[i].num_projs, C = b->s[i].num_cols, N = slice_grids(b->s[i]);
  const int ba = b->bundle, bc = 64 / ba;
  const int64_t bundles = (int64_t)((P + ba - 1) / ba) * ((C + bc - 1) / bc);
  out[0] = (bundles + 3) / 4;  // four wavefront bundles per 256-lane workgroup
  out[1] = (int64_t)((N + b->tw - 1) / b->tw) * ((N + b->th - 1) / b->th);
  out[2] = ((int64_t)N * N + 256 * kBatchUpdateBlocks - 1) / (256 * kBatchUpdateBlocks);
}

// Tables, descriptors and launch grids for the current launch shape.
int prepare(trace_batch *b) {
  if (b->prepared) return TRACE_OK;
  drop_graphs(b);
  const int n = (int)b->s.size();
  int maxP = 0;
  for (int i = 0; i < n; ++i) {
    trace_handle_t *h = b->h[i];
    h->tile = TileShape{b->tw, b->th, 1};
    h->exact = b->exact;
    if (int rc = ensure_count(h)) return rc;
    maxP = std::max(maxP, h->geo.P);
  }
  // 32-bit accumulators of one slice, then the per-angle column counts
  b->lds = (size_t)b->th * (b->tw + 1) * sizeof(unsigned) + (size_t)maxP * sizeof(int);
  if (b->lds > 64 * 1024)
    return fail(TRACE_E_UNSUPPORTED, "tile and angle count exceed the batched kernel's LDS");
  // per-phase order (decreasing cost per item) and start tables
  std::vector<int> tables(3 * (size_t)n + 3 * ((size_t)n + 1));
  for (int f = 0; f < 3; ++f) {
    std::vector<int64_t> it(n);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
      int64_t c[3];
      slice_items(b, i, c);
      it[i] = c[f];
      cost[i] = (f == 2) ? (double)b->s[i].num_cols : (double)b->nnz[i] / (double)c[f];
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return cost[a] > cost[c]; });
    int *ord = tables.data() + (size_t)f * n;
    int *st = tables.data() + 3 * (size_t)n + (size_t)f * (n + 1);
    int64_t acc = 0;
    for (int k = 0; k < n; ++k) {
      ord[k] = order[k];
      st[k] = (int)acc;
      acc += it[order[k]];
    }
    st[n] = (int)acc;
    if (acc > 0x7fffffff) return fail(TRACE_E_UNSUPPORTED, "batch has too many work items");
    b->items[f] = acc;
  }
  std::vector<BatchSliceDev> desc(n);
  for (int i = 0; i < n; ++i) {
    const trace_handle_t *h = b->h[i];
    BatchSliceDev &d = desc[i];
    d.g = h->geo.grid;
    d.ang = h->geo.d_ang;
    d.rays = h->rays;
    d.count = h->count;
    d.P = h->geo.P;
    d.alg = b->s[i].alg == TRACE_SIRT ? 0 : 1;
    d.iters = b->s[i].iters;
    d.yi0 = h->geo.yi0;
    d.maxlen = h->maxlen;
    d.denmax = h->den_max;
    d.sino = b->s[i].sino;
    d.x = b->s[i].recon;
    d.w = h->w;
    d.z = h->z;
    d.den = h->den;
    d.maxw = b->maxw + i;
    d.wraps = h->wraps ? 1 : 0;
  }
  HIP_TRY(hipMemcpyAsync(b->d_desc, desc.data(), desc.size() * sizeof(BatchSliceDev),
                         hipMemcpyHostToDevice, b->user));
  HIP_TRY(hipMemcpyAsync(b->d_tables, tables.data(), tables.size() * sizeof(int),
                         hipMemcpyHostToDevice, b->user));
  HIP_TRY(hipStreamSynchronize(b->user));  // host tables leave scope
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  for (int f = 0; f < 3; ++f) {
    const int per_cu = batch_blocks_per_cu(f, b->tw, b->th, b->exact, b->lds);
    if (per_cu <= 0) return fail(TRACE_E_HIP, "batched kernel cannot be resident");
    const int64_t g = (int64_t)per_cu * cus;
    b->grid[f] = (int)std::max<int64_t>(1, std::min<int64_t>(g, b->items[f]));
  }
  b->prepared = true;
  return TRACE_OK;
}

BatchQueues queues(const trace_batch *b) {
  const int n = (int)b->s.size();
  BatchQueues q;
  for (int f = 0; f < 3; ++f) {
    q.order[f] = b->d_tables + (size_t)f * n;
    q.start[f] = b->d_tables + 3 * (size_t)n + (size_t)f * (n + 1);
  }
  q.n = n;
  q.counters = b->counters;
  q.maxw_all = b->maxw;
  return q;
}

// Graph of `iters` batch iterations (captured on the private stream).
int graph_for(trace_batch *b, int iters, hipGraphExec_t *out) {
  auto found = b->graphs.find(iters);
  if (found != b->graphs.end()) {
    *out = found->second.second;
    return TRACE_OK;
  }
  const BatchQueues q = queues(b);
  HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int it = 0; it < iters && e == hipSuccess; ++it) {
    e = launch_batch_reset(q, b->cap);
    if (e == hipSuccess)
      e = launch_batch_forward(b->d_desc, q, it, b->bundle, b->exact, b->grid[0], b->cap);
    if (e == hipSuccess)
      e = launch_batch_backproject(b->d_desc, q, it, b->tw, b->th, b->exact, b->grid[1], b->lds,
                                   b->cap);
    if (e == hipSuccess) e = launch_batch_update(b->d_desc, q, it, b->grid[2], b->cap);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e_end = hipStreamEndCapture(b->cap, &graph);
  if (e != hipSuccess || e_end != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(TRACE_E_HIP, std::string("graph capture: ") +
                                 hipGetErrorString(e != hipSuccess ? e : e_end));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(TRACE_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  b->graphs[iters] = {graph, exec};
  *out = exec;
  return TRACE_OK;
}

}  // namespace

extern "C" {

void trace_batch_destroy(trace_batch_t *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->cap) (void)hipStreamSynchronize(b->cap);
  if (b->user) (void)hipStreamSynchronize(b->user);
  else (void)hipDeviceSynchronize();
  drop_graphs(b);
  for (trace_handle_t *h : b->h) trace_destroy(h);
  for (void *p : {(void *)b->d_desc, (void *)b->d_tables, (void *)b->counters, (void *)b->maxw,
                  (void *)b->flag})
    dev_free(p);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

int trace_batch_create(trace_batch_t **out, const trace_batch_slice_t *slices, int n,
                       int device, void *hip_stream) {
  clear_error();
  if (!out) return fail(TRACE_E_INVALID, "batch pointer is null");
  *out = nullptr;
  if (n <= 0 || !slices) return fail(TRACE_E_INVALID, "a batch needs at least one slice");
  for (int i = 0; i < n; ++i) {
    const trace_batch_slice_t &s = slices[i];
    if (s.alg != TRACE_SIRT && s.alg != TRACE_MLEM)
      return fail(TRACE_E_UNSUPPORTED, "batched slices run SIRT or MLEM");
    if (s.iters < 0) return fail(TRACE_E_INVALID, "iters must be >= 0");
    if (s.num_grids < 0) return fail(TRACE_E_INVALID, "num_grids must be >= 0");
    if (!s.sino || !s.recon || !s.theta) return fail(TRACE_E_INVALID, "null slice buffer");
  }
  HIP_TRY(hipSetDevice(device < 0 ? 0 : device));
  auto *b = new trace_batch();
  b->device = device;
  b->user = (hipStream_t)hip_stream;
  b->s.assign(slices, slices + n);
  auto bail = [&](int rc) {
    trace_batch_destroy(b);
    return rc;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming)) != hipSuccess)
    return bail(fail(TRACE_E_HIP, std::string("stream/event: ") + hipGetErrorString(e)));
  b->h.reserve(n);
  b->nnz.resize(n);
  for (int i = 0; i < n; ++i) {
    const trace_batch_slice_t &s = slices[i];
    trace_geom_t g{s.num_projs, 1, s.num_cols, slice_grids(s), s.center, s.theta};
    trace_handle_t *h = nullptr;
    if (int rc = trace_create(&h, &g, device, hip_stream)) return bail(rc);
    b->h.push_back(h);
    h->fwd_sb = 1;
    h->tile = TileShape{b->tw, b->th, 1};
    if (int rc = ensure_rays(h)) return bail(rc);
    if (int rc = ensure_den(h)) return bail(rc);
    if (int rc = ensure_work(h, false, false)) return bail(rc);
    b->nnz[i] = h->nnz;
  }
  if ((e = dev_alloc(&b->d_desc, (size_t)n)) != hipSuccess ||
      (e = dev_alloc(&b->d_tables, 3 * (size_t)n + 3 * ((size_t)n + 1))) != hipSuccess ||
      (e = dev_alloc(&b->counters, 3)) != hipSuccess ||
      (e = dev_alloc(&b->maxw, (size_t)n)) != hipSuccess ||
      (e = dev_alloc(&b->flag, 1)) != hipSuccess)
    return bail(fail(e == hipErrorOutOfMemory ? TRACE_E_NOMEM : TRACE_E_HIP,
                     std::string("allocation: ") + hipGetErrorString(e)));
  if (int rc = prepare(b)) return bail(rc);
  *out = b;
  return TRACE_OK;
}

int trace_batch_set_launch_shape(trace_batch_t *b, int bundle_angles, int tile_w, int tile_h) {
  clear_error();
  if (!b) return fail(TRACE_E_INVALID, "batch is null");
  if (bundle_angles != 1 && bundle_angles != 2 && bundle_angles != 4 && bundle_angles != 8 &&
      bundle_angles != 16)
    return fail(TRACE_E_INVALID, "angles per bundle must be 1, 2, 4, 8 or 16");
  if (!batch_tile_supported(tile_w, tile_h))
    return fail(TRACE_E_UNSUPPORTED, "tile shape not compiled into the batched backprojector");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->cap));  // no replay may still read the old tables
  b->bundle = bundle_angles;
  b->tw = tile_w;
  b->th = tile_h;
  b->prepared = false;
  return prepare(b);
}

int trace_batch_get_launch_shape(trace_batch_t *b, int *bundle_angles, int *tile_w,
                                 int *tile_h) {
  clear_error();
  if (!b) return fail(TRACE_E_INVALID, "batch is null");
  if (bundle_angles) *bundle_angles = b->bundle;
  if (tile_w) *tile_w = b->tw;
  if (tile_h) *tile_h = b->th;
  return TRACE_OK;
}

int trace_batch_set_option(trace_batch_t *b, int option, int value) {
  clear_error();
  if (!b) return fail(TRACE_E_INVALID, "batch is null");
  if (option != TRACE_OPT_EXACT_SQRT) return fail(TRACE_E_UNSUPPORTED, "unknown batch option");
  if (value != 0 && value != 1) return fail(TRACE_E_INVALID, "value must be 0 or 1");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->cap));
  if (b->exact != (value != 0)) {
    b->exact = value != 0;
    for (trace_handle_t *h : b->h) {
      h->exact = b->exact;
      if (int rc = ensure_den(h)) return rc;  // den follows the length rounding
    }
    b->prepared = false;
  }
  return prepare(b);
}

namespace {

int run_iterations(const trace_batch *b, int max_iters) {
  int iters = 0;
  for (const trace_batch_slice_t &s : b->s)
    iters = std::max(iters, max_iters > 0 ? std::min(s.iters, max_iters) : s.iters);
  return iters;
}

}  // namespace

int trace_batch_prepare(trace_batch_t *b, int max_iters) {
  clear_error();
  if (!b) return fail(TRACE_E_INVALID, "batch is null");
  HIP_TRY(hipSetDevice(b->device));
  if (int rc = prepare(b)) return rc;
  const int iters = run_iterations(b, max_iters);
  if (iters == 0) return TRACE_OK;
  hipGraphExec_t exec = nullptr;
  return graph_for(b, iters, &exec);
}

int trace_batch_run(trace_batch_t *b, int max_iters) {
  clear_error();
  if (!b) return fail(TRACE_E_INVALID, "batch is null");
  HIP_TRY(hipSetDevice(b->device));
  if (int rc = prepare(b)) return rc;
  // the sinograms may have changed since the last run: check them all
  HIP_TRY(hipMemsetAsync(b->flag, 0, sizeof(int), b->user));
  for (const trace_batch_slice_t &s : b->s)
    HIP_TRY(launch_check_finite(s.sino, (int64_t)s.num_projs * s.num_cols, b->flag, b->user));
  int flag = 0;
  HIP_TRY(hipMemcpyAsync(&flag, b->flag, sizeof(int), hipMemcpyDeviceToHost, b->user));
  HIP_TRY(hipStreamSynchronize(b->user));
  if (flag) return fail(TRACE_E_NONFINITE, "a sinogram of the batch holds NaN or Inf values");
  const int iters = run_iterations(b, max_iters);
  if (iters == 0) return TRACE_OK;
  hipGraphExec_t exec = nullptr;
  if (int rc = graph_for(b, iters, &exec)) return rc;
  HIP_TRY(hipEventRecord(b->ev_in, b->user));
  HIP_TRY(hipStreamWaitEvent(b->cap, b->ev_in, 0));
  HIP_TRY(hipGraphLaunch(exec, b->cap));
  HIP_TRY(hipEventRecord(b->ev_out, b->cap));
  HIP_TRY(hipStreamWaitEvent(b->user, b->ev_out, 0));
  return TRACE_OK;
}

int trace_batch_reconstruct(const trace_batch_slice_t *slices, int n, int device,
                            void *hip_stream) {
  trace_batch_t *b = nullptr;
  if (int rc = trace_batch_create(&b, slices, n, device, hip_stream)) return rc;
  int rc = trace_batch_run(b, 0);
  if (rc == TRACE_OK) {
    const hipError_t e = hipStreamSynchronize(b->user);
    if (e != hipSuccess) rc = fail(TRACE_E_HIP, std::string("run: ") + hipGetErrorString(e));
  }
  trace_batch_destroy(b);
  return rc;
}

}  // extern "C"
