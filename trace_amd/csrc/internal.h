// internal.h -- host-side structures shared by the C-ABI and the kernel
// launchers.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "traverse.h"

namespace trace_hip {

enum Alg { kSIRT = 0, kMLEM = 1, kPML = 2, kAPMLR = 3 };

// One slice stack's geometry on the device.
struct Geometry {
  int P = 0, S = 0, C = 0, N = 0;
  float center = 0.f;  // as given by the caller
  float mov = 0.f;     // trace_data.h:97-101
  GridInfo grid{};
  AngleInfo *d_ang = nullptr;  // [P]
  float *d_yi = nullptr;       // [C]
  int64_t NN() const { return (int64_t)N * N; }
  int64_t rays_per_slice() const { return (int64_t)P * C; }
};

// Regularisation weights for the PML/APMLR stencils, computed on the host
// with the reference's float/double promotion order.
struct StencilWeights {
  // [z mode][pixel class][neighbour]: z mode 0 = two z neighbours,
  // 1 = first slice (+1 only), 2 = last slice (-1 only); classes as in
  // traverse-free stencil code (0 inner, 1 top, 2 bottom, 3 left, 4 right,
  // 5..8 corners TL, TR, BL, BR); neighbours in-plane first, then z.
  float w[3][9][10];
};

struct RegParams {
  float beta, beta1, delta, delta1, regw;
};

// Launchers (kernels.hip).  All are asynchronous on `st`.
hipError_t launch_den(const Geometry &g, float *den, hipStream_t st, bool exact);
hipError_t launch_forward(const Geometry &g, const float *x, int64_t x_stride,
                          float *y, hipStream_t st, bool exact);
hipError_t launch_adjoint(const Geometry &g, const float *r, float *z,
                          hipStream_t st, bool exact);
hipError_t launch_ray_pass(int alg, const Geometry &g, const float *x,
                           int64_t x_stride, const float *sino, float *num,
                           hipStream_t st, bool exact);
hipError_t launch_fill(float *p, int64_t n, float v, hipStream_t st);
hipError_t launch_update_sirt(float *x, const float *num, const float *den,
                              int S, int64_t NN, hipStream_t st);
hipError_t launch_update_mlem(float *x, const float *num, const float *den,
                              int S, int64_t NN, hipStream_t st);
hipError_t launch_update_pml(const float *x, float *xn, const float *num,
                             const float *den, int S, int N, float beta,
                             hipStream_t st, float *F_out, float *G_out);
hipError_t launch_update_apmlr(const float *xpad, float *xnpad,
                               const float *num, const float *den, int S,
                               int N, const RegParams &p,
                               const StencilWeights &w, hipStream_t st,
                               float *F_out, float *G_out);
hipError_t launch_dump_rays(const Geometry &g, int n, const int *proj,
                            const int *col, int max_seg, int *counts, int *idx,
                            float *len, float *len2, hipStream_t st, bool exact);
hipError_t launch_count_segments(const Geometry &g, unsigned long long *total,
                                 unsigned long long *oor, hipStream_t st);
hipError_t launch_check_finite(const float *v, int64_t n, int *flag,
                               hipStream_t st);

}  // namespace trace_hip
