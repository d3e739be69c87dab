// Single-ray traversal and forward projection through the reference's own
// functions (trace_utils.cc:41-215, sirt.cc:4-19), geometry from
// TraceMetadata (trace_data.h:53-109) and the per-row prologue of Reduce
// (sirt.cc:111-129).  Test infrastructure only.
#include <cmath>
#include "harness.h"
#include "sirt.h"
#include "trace_utils.h"

namespace {
// CalculateSimdata is protected; a derived probe exposes it unchanged.
struct SimdataProbe : public SIRTReconSpace {
  SimdataProbe() : SIRTReconSpace(1, 2) {}
  float Sim(float *recon, int len, int *indi, float *leng) {
    return CalculateSimdata(recon, len, indi, leng);
  }
};

struct RayBufs {
  int n;
  float *coordx, *coordy, *ax, *ay, *bx, *by, *coorx, *coory;
  explicit RayBufs(int N) : n(N) {
    coordx = new float[N + 1]; coordy = new float[N + 1];
    ax = new float[N + 1]; ay = new float[N + 1];
    bx = new float[N + 1]; by = new float[N + 1];
    coorx = new float[2 * N + 2]; coory = new float[2 * N + 2];
  }
  ~RayBufs() {
    delete[] coordx; delete[] coordy; delete[] ax; delete[] ay;
    delete[] bx; delete[] by; delete[] coorx; delete[] coory;
  }
};

int trace(const TraceMetadata &md, RayBufs &b, float theta_q, int col,
          int *idx, float *len, float *len2) {
  int num_cols = md.num_cols();
  int num_grids = md.num_grids();
  int quadrant = trace_utils::CalculateQuadrant(theta_q);
  float sinq = sinf(theta_q);
  float cosq = cosf(theta_q);
  float xi = -1e6;
  float yi = (1 - num_cols) / 2. + col + md.mov();
  trace_utils::CalculateCoordinates(num_grids, xi, yi, sinq, cosq, md.gridx(),
                                    md.gridy(), b.coordx, b.coordy);
  int alen, blen;
  trace_utils::MergeTrimCoordinates(num_grids, b.coordx, b.coordy, md.gridx(),
                                    md.gridy(), &alen, &blen, b.ax, b.ay, b.bx,
                                    b.by);
  trace_utils::SortIntersectionPoints(quadrant, alen, blen, b.ax, b.ay, b.bx,
                                      b.by, b.coorx, b.coory);
  int l = alen + blen;
  if (len2)
    trace_utils::CalculateDistanceLengths(l, num_grids, b.coorx, b.coory, len,
                                          len2, idx);
  else
    trace_utils::CalculateDistanceLengths(l, num_grids, b.coorx, b.coory, len,
                                          idx);
  return l > 1 ? l - 1 : 0;
}
}  // namespace

extern "C" {

// Segments of ray (theta_q [radians], col) on an N = C grid.  Buffers hold
// 2C+2 entries.  Returns the segment count.
int ref_trace_ray(int C, float center, float theta_q, int col, int *idx,
                  float *len, float *len2) {
  float *th = new float[1];
  th[0] = theta_q;
  TraceMetadata md(th, 0, 0, 0, 1, 1, 1, C, C, center);
  RayBufs b(C);
  return trace(md, b, theta_q, col, idx, len, len2);
}

// Segments of ray (theta_q, col) with C detector columns on an N x N grid:
// the reference's TraceMetadata and traversal functions take num_cols and
// num_grids separately (trace_data.h:53-109, trace_utils.cc:63-215) although
// its Reduce passes num_grids = num_cols (sirt.cc:107-108).  Buffers hold
// 2N+2 entries.
int ref_trace_ray_n(int N, int C, float center, float theta_q, int col, int *idx,
                    float *len, float *len2) {
  float *th = new float[1];
  th[0] = theta_q;
  TraceMetadata md(th, 0, 0, 0, 1, 1, 1, C, N, center);
  RayBufs b(N);
  return trace(md, b, theta_q, col, idx, len, len2);
}

// y[p][s][c] = CalculateSimdata over each ray of slice s of x [S][C][C].
int ref_forward(int P, int S, int C, float center, const float *theta,
                const float *x, float *y) {
  float *th = ref_harness::new_copy(theta, (size_t)P);
  TraceMetadata md(th, 0, 0, 0, S, P, S, C, C, center);
  RayBufs b(C);
  int *idx = new int[2 * C + 2];
  float *len = new float[2 * C + 2];
  SimdataProbe probe;
  float *xc = ref_harness::new_copy(x, (size_t)S * C * C);
  for (int p = 0; p < P; ++p)
    for (int c = 0; c < C; ++c) {
      int ns = trace(md, b, theta[p], c, idx, len, nullptr);
      for (int s = 0; s < S; ++s)
        y[((size_t)p * S + s) * C + c] =
            probe.Sim(xc + (size_t)s * C * C, ns + 1, idx, len);
    }
  delete[] idx; delete[] len; delete[] xc;
  return 0;
}

// Theta as the apps use it after preprocessing (trace_utils.cc:10-24).
void ref_deg2rad(const float *deg, float *rad, int n) {
  ref_harness::ThetaBox tb(deg, n);
  trace_utils::DegreeToRadian(tb.data);
  std::memcpy(rad, tb.data.data, sizeof(float) * (size_t)n);
  delete[] static_cast<float *>(tb.data.data);
}

int ref_quadrant(float q) { return trace_utils::CalculateQuadrant(q); }

float ref_mov(int C, float center) {
  float *th = new float[1];
  th[0] = 0.f;
  TraceMetadata md(th, 0, 0, 0, 1, 1, 1, C, C, center);
  return md.mov();
}

// Compiler that built this library (reported beside CPU timings).
const char *ref_compiler(void) { return "g++ " __VERSION__; }

}  // extern "C"
