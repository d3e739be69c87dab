// kernels.hip -- gfx950 kernels of the reconstruction hot path.
//
// Ray kernels: one lane per ray; the 64 lanes of a wavefront are 64 adjacent
// detector columns of one (projection, slice) row, so at every step they
// touch neighbouring pixels and their gathers/atomics share cache lines.
// Segments are generated on the fly by traverse_ray (traverse.h).
//
// Image kernels: one lane per pixel, rows of pixels across a wavefront for
// coalesced HBM traffic.
#include <hip/hip_runtime.h>

#include "internal.h"

#pragma clang fp contract(off)

namespace trace_hip {

namespace {

constexpr int kRayBlock = 256;
constexpr int kPixBlock = 256;

__device__ __forceinline__ void atomic_add_f32(float *p, float v) {
  // hardware global float atomic (no CAS loop); relaxed ordering, device scope
  unsafeAtomicAdd(p, v);
}

// ray id t -> (p, s, c) in sinogram order [P][S][C]
__device__ __forceinline__ void ray_coords(int64_t t, int S, int C, int &p,
                                           int &s, int &c) {
  c = (int)(t % C);
  int64_t ps = t / C;
  s = (int)(ps % S);
  p = (int)(ps / S);
}

template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_den(GridInfo g, const AngleInfo *__restrict__ ang, const float *__restrict__ yi,
      int P, float *__restrict__ den) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  if (t >= (int64_t)P * C) return;
  const int c = (int)(t % C), p = (int)(t / C);
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  traverse_ray<kExact>(g, yi[c], ang[p], [&](int idx, float l, float) {
    if ((unsigned)idx < NN) atomic_add_f32(den + idx, l);
  });
}

// y[p][s][c] = sum_k x_s[idx_k] * len_k in segment order (sirt.cc:4-19).
template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_forward(GridInfo g, const AngleInfo *__restrict__ ang,
          const float *__restrict__ yi, int P, int S,
          const float *__restrict__ x, int64_t x_stride, float *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  if (t >= (int64_t)P * S * C) return;
  int p, s, c;
  ray_coords(t, S, C, p, s, c);
  const float *xs = x + (int64_t)s * x_stride;
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  float acc = 0.f;
  traverse_ray<kExact>(g, yi[c], ang[p], [&](int idx, float l, float) {
    if ((unsigned)idx < NN) acc += xs[idx] * l;
  });
  y[t] = acc;
}

// z_s[idx] += len * r[p][s][c] (the adjoint; z must be zeroed).
template <bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_adjoint(GridInfo g, const AngleInfo *__restrict__ ang,
          const float *__restrict__ yi, int P, int S,
          const float *__restrict__ r, float *__restrict__ z) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  if (t >= (int64_t)P * S * C) return;
  int p, s, c;
  ray_coords(t, S, C, p, s, c);
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  float *zs = z + (int64_t)s * NN;
  const float w = r[t];
  if (w == 0.f) return;
  traverse_ray<kExact>(g, yi[c], ang[p], [&](int idx, float l, float) {
    if ((unsigned)idx < NN) atomic_add_f32(zs + idx, l * w);
  });
}

// One iteration's ray work for one ray (Reduce + UpdateReconReplica):
// simdata (and SIRT's a2) in a first walk, the per-ray weight, then the
// scatter of the numerator in a second walk.  den is iteration-invariant and
// precomputed by k_den.
//   SIRT  w = (p - sim)/a2;  num += len*w               (sirt.cc:35-64)
//   MLEM  w = p/sim;         num += len*w               (mlem.cc:37-61)
//   PML/APMLR w = p/sim;     num -= x*len*w             (pml.cc:222-246)
template <int ALG, bool kExact>
__global__ void __launch_bounds__(kRayBlock)
k_ray_pass(GridInfo g, const AngleInfo *__restrict__ ang,
           const float *__restrict__ yi, int P, int S,
           const float *__restrict__ x, int64_t x_stride,
           const float *__restrict__ sino, float *__restrict__ num) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  if (t >= (int64_t)P * S * C) return;
  int p, s, c;
  ray_coords(t, S, C, p, s, c);
  const float *xs = x + (int64_t)s * x_stride;
  const unsigned NN = (unsigned)g.N * (unsigned)g.N;
  float *nums = num + (int64_t)s * NN;
  const float yic = yi[c];
  const AngleInfo a = ang[p];
  float sim = 0.f, a2 = 0.f;
  traverse_ray<kExact>(g, yic, a, [&](int idx, float l, float l2) {
    if ((unsigned)idx < NN) sim += xs[idx] * l;
    if (ALG == kSIRT) a2 += l2;
  });
  const float ray = sino[t];
  float w;
  if (ALG == kSIRT)
    w = (ray - sim) / a2;
  else
    w = ray / sim;
  traverse_ray<kExact>(g, yic, a, [&](int idx, float l, float) {
    if ((unsigned)idx < NN) {
      if (ALG == kSIRT || ALG == kMLEM)
        atomic_add_f32(nums + idx, l * w);
      else
        atomic_add_f32(nums + idx, -(xs[idx] * l * w));
    }
  });
}

__global__ void k_fill(float *__restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// sirt.cc:21-33
__global__ void k_update_sirt(float *__restrict__ x, const float *__restrict__ num,
                              const float *__restrict__ den, int64_t n, int64_t NN) {
  const int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  if (i < n) x[i] += num[i] / den[i % NN];
}

// mlem.cc:21-34
__global__ void k_update_mlem(float *__restrict__ x, const float *__restrict__ num,
                              const float *__restrict__ den, int64_t n, int64_t NN) {
  const int64_t i = (int64_t)blockIdx.x * kPixBlock + threadIdx.x;
  if (i < n) x[i] *= num[i] / den[i % NN];
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

// Fused PML stencil (pml.cc:25-201, F and G from zero) and closed-form
// update (pml.cc:3-23):  x' = (-G + sqrt(G^2 - 8 num F)) / (4F).
__global__ void __launch_bounds__(kPixBlock)
k_update_pml(const float *__restrict__ x, float *__restrict__ xn,
             const float *__restrict__ num, const float *__restrict__ den,
             int S, int N, float beta, float *__restrict__ F_out,
             float *__restrict__ G_out) {
  const int m = blockIdx.x * kPixBlock + threadIdx.x;
  const int n = blockIdx.y;
  const int s = blockIdx.z;
  if (m >= N) return;
  const int64_t NN = (int64_t)N * N;
  const int64_t j = (int64_t)s * NN + (int64_t)n * N + m;
  const int cls = pixel_class(m, n, N);
  const float *w = c_pml_w[cls == 0 ? 0 : (cls <= 4 ? 1 : 2)];
  const float x0 = x[j];
  float f = 0.f, gg = 0.f;
  const int cnt = c_nbn[cls];
  for (int q = 0; q < cnt; ++q) {
    const float xq = x[j + c_nb[cls][q][0] + (int64_t)c_nb[cls][q][1] * N];
    f += 2.0f * beta * w[q];
    gg -= 2.0f * beta * w[q] * (x0 + xq);
  }
  const float G = gg + den[(int64_t)n * N + m];
  const float F = f;
  const float nm = num[j];
  if (F_out) { F_out[j] = F; G_out[j] = G; }
  xn[j] = (-G + __builtin_sqrtf(G * G - 8.0f * nm * F)) / (4.0f * F);
}

struct ApmlrArgs {
  StencilWeights w;
  RegParams p;
};

// Fused APMLR stencil (apmlr.cc:919-982 with FGInner/FGTop/FGBottom) and
// update (apmlr.cc:3-26) over ghost-padded images [S+2][N][N]; real slice s
// lives at padded index s+1.
__global__ void __launch_bounds__(kPixBlock)
k_update_apmlr(const float *__restrict__ xp, float *__restrict__ xnp,
               const float *__restrict__ num, const float *__restrict__ den,
               int S, int N, ApmlrArgs a, float *__restrict__ F_out,
               float *__restrict__ G_out) {
  const int m = blockIdx.x * kPixBlock + threadIdx.x;
  const int n = blockIdx.y;
  const int s = blockIdx.z;
  if (m >= N) return;
  const int64_t NN = (int64_t)N * N;
  const int64_t jp = (int64_t)(s + 1) * NN + (int64_t)n * N + m;  // padded
  const int64_t j = (int64_t)s * NN + (int64_t)n * N + m;
  const bool top = (s == 0), bottom = (s == S - 1 && S - 1 != 0);
  const int zmode = top ? 1 : (bottom ? 2 : 0);
  const int nz = zmode == 0 ? 2 : 1;
  const int cls = pixel_class(m, n, N);
  const float *w = a.w.w[zmode][cls];
  const int cnt = c_nbn[cls];
  const float x0 = xp[jp];
  float f = 0.f, gg = 0.f;
  const float tb = 2.0f * a.p.beta;
  for (int q = 0; q < cnt; ++q) {
    const float xq = xp[jp + c_nb[cls][q][0] + (int64_t)c_nb[cls][q][1] * N];
    const float mg = x0 + xq;
    const float rg = x0 - xq;
    const float gm = 1.0f / (1.0f + __builtin_fabsf(rg / a.p.delta));
    f += tb * w[q] * gm;
    gg -= tb * w[q] * gm * mg;
  }
  const float tb1 = 2.0f * a.p.beta1;
  for (int q = 0; q < nz; ++q) {
    const int64_t dz = (zmode == 2 || q == 1) ? -NN : NN;
    const float xq = xp[jp + dz];
    const float mg = x0 + xq;
    const float rg = x0 - xq;
    const float gm = 1.0f / (1.0f + __builtin_fabsf(rg / a.p.delta1));
    f += tb1 * w[cnt + q] * gm;
    gg -= tb1 * w[cnt + q] * gm * mg;
  }
  const float G = gg + den[(int64_t)n * N + m];
  const float F = f;
  const float nm = num[j];
  if (F_out) { F_out[j] = F; G_out[j] = G; }
  xnp[jp] = (-G + __builtin_sqrtf(G * G - 8.0f * nm * F)) / (4.0f * F);
}

template <bool kExact>
__global__ void k_dump_rays(GridInfo g, const AngleInfo *__restrict__ ang,
                            const float *__restrict__ yi, int n,
                            const int *__restrict__ proj,
                            const int *__restrict__ col, int max_seg,
                            int *__restrict__ counts, int *__restrict__ idx,
                            float *__restrict__ len, float *__restrict__ len2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int k = 0;
  const int64_t base = (int64_t)t * max_seg;
  int cnt = traverse_ray<kExact>(g, yi[col[t]], ang[proj[t]],
                                 [&](int i, float l, float l2) {
                                   if (k < max_seg) {
                                     idx[base + k] = i;
                                     len[base + k] = l;
                                     len2[base + k] = l2;
                                   }
                                   ++k;
                                 });
  counts[t] = cnt;
}

__global__ void __launch_bounds__(kRayBlock)
k_count_segments(GridInfo g, const AngleInfo *__restrict__ ang,
                 const float *__restrict__ yi, int P,
                 unsigned long long *total, unsigned long long *oor) {
  const int64_t t = (int64_t)blockIdx.x * kRayBlock + threadIdx.x;
  const int C = g.C;
  unsigned cnt = 0, bad = 0;
  if (t < (int64_t)P * C) {
    const int c = (int)(t % C), p = (int)(t / C);
    const unsigned NN = (unsigned)g.N * (unsigned)g.N;
    cnt = traverse_ray<true>(g, yi[c], ang[p], [&](int idx, float, float) {
      bad += ((unsigned)idx >= NN) ? 1u : 0u;
    });
  }
  // wavefront reduction, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_down(cnt, off, 64);
    bad += __shfl_down(bad, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(total, (unsigned long long)cnt);
    if (bad) atomicAdd(oor, (unsigned long long)bad);
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

}  // namespace

hipError_t launch_den(const Geometry &g, float *den, hipStream_t st, bool exact) {
  const int64_t n = g.rays_per_slice();
  if (exact)
    hipLaunchKernelGGL(k_den<true>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0,
                       st, g.grid, g.d_ang, g.d_yi, g.P, den);
  else
    hipLaunchKernelGGL(k_den<false>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock), 0,
                       st, g.grid, g.d_ang, g.d_yi, g.P, den);
  return hipGetLastError();
}

hipError_t launch_forward(const Geometry &g, const float *x, int64_t x_stride,
                          float *y, hipStream_t st, bool exact) {
  const int64_t n = (int64_t)g.P * g.S * g.C;
  if (exact)
    hipLaunchKernelGGL(k_forward<true>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock),
                       0, st, g.grid, g.d_ang, g.d_yi, g.P, g.S, x, x_stride, y);
  else
    hipLaunchKernelGGL(k_forward<false>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock),
                       0, st, g.grid, g.d_ang, g.d_yi, g.P, g.S, x, x_stride, y);
  return hipGetLastError();
}

hipError_t launch_adjoint(const Geometry &g, const float *r, float *z,
                          hipStream_t st, bool exact) {
  const int64_t n = (int64_t)g.P * g.S * g.C;
  if (exact)
    hipLaunchKernelGGL(k_adjoint<true>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock),
                       0, st, g.grid, g.d_ang, g.d_yi, g.P, g.S, r, z);
  else
    hipLaunchKernelGGL(k_adjoint<false>, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock),
                       0, st, g.grid, g.d_ang, g.d_yi, g.P, g.S, r, z);
  return hipGetLastError();
}

template <int ALG>
static void ray_pass_alg(const Geometry &g, const float *x, int64_t x_stride,
                         const float *sino, float *num, hipStream_t st, bool exact) {
  const int64_t n = (int64_t)g.P * g.S * g.C;
  dim3 grid(blocks_for(n, kRayBlock)), block(kRayBlock);
  if (exact)
    hipLaunchKernelGGL((k_ray_pass<ALG, true>), grid, block, 0, st, g.grid, g.d_ang,
                       g.d_yi, g.P, g.S, x, x_stride, sino, num);
  else
    hipLaunchKernelGGL((k_ray_pass<ALG, false>), grid, block, 0, st, g.grid, g.d_ang,
                       g.d_yi, g.P, g.S, x, x_stride, sino, num);
}

hipError_t launch_ray_pass(int alg, const Geometry &g, const float *x,
                           int64_t x_stride, const float *sino, float *num,
                           hipStream_t st, bool exact) {
  switch (alg) {
    case kSIRT: ray_pass_alg<kSIRT>(g, x, x_stride, sino, num, st, exact); break;
    case kMLEM: ray_pass_alg<kMLEM>(g, x, x_stride, sino, num, st, exact); break;
    default: ray_pass_alg<kPML>(g, x, x_stride, sino, num, st, exact); break;
  }
  return hipGetLastError();
}

hipError_t launch_fill(float *p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, kPixBlock)), dim3(kPixBlock), 0, st, p,
                     n, v);
  return hipGetLastError();
}

hipError_t launch_update_sirt(float *x, const float *num, const float *den, int S,
                              int64_t NN, hipStream_t st) {
  const int64_t n = (int64_t)S * NN;
  hipLaunchKernelGGL(k_update_sirt, dim3(blocks_for(n, kPixBlock)), dim3(kPixBlock), 0,
                     st, x, num, den, n, NN);
  return hipGetLastError();
}

hipError_t launch_update_mlem(float *x, const float *num, const float *den, int S,
                              int64_t NN, hipStream_t st) {
  const int64_t n = (int64_t)S * NN;
  hipLaunchKernelGGL(k_update_mlem, dim3(blocks_for(n, kPixBlock)), dim3(kPixBlock), 0,
                     st, x, num, den, n, NN);
  return hipGetLastError();
}

hipError_t launch_update_pml(const float *x, float *xn, const float *num,
                             const float *den, int S, int N, float beta,
                             hipStream_t st, float *F_out, float *G_out) {
  dim3 grid(blocks_for(N, kPixBlock), N, S);
  hipLaunchKernelGGL(k_update_pml, grid, dim3(kPixBlock), 0, st, x, xn, num, den, S,
                     N, beta, F_out, G_out);
  return hipGetLastError();
}

hipError_t launch_update_apmlr(const float *xpad, float *xnpad, const float *num,
                               const float *den, int S, int N, const RegParams &p,
                               const StencilWeights &w, hipStream_t st,
                               float *F_out, float *G_out) {
  dim3 grid(blocks_for(N, kPixBlock), N, S);
  ApmlrArgs a{w, p};
  hipLaunchKernelGGL(k_update_apmlr, grid, dim3(kPixBlock), 0, st, xpad, xnpad, num,
                     den, S, N, a, F_out, G_out);
  return hipGetLastError();
}

hipError_t launch_dump_rays(const Geometry &g, int n, const int *proj, const int *col,
                            int max_seg, int *counts, int *idx, float *len,
                            float *len2, hipStream_t st, bool exact) {
  if (n <= 0) return hipSuccess;
  if (exact)
    hipLaunchKernelGGL(k_dump_rays<true>, dim3(blocks_for(n, 64)), dim3(64), 0, st,
                       g.grid, g.d_ang, g.d_yi, n, proj, col, max_seg, counts, idx,
                       len, len2);
  else
    hipLaunchKernelGGL(k_dump_rays<false>, dim3(blocks_for(n, 64)), dim3(64), 0, st,
                       g.grid, g.d_ang, g.d_yi, n, proj, col, max_seg, counts, idx,
                       len, len2);
  return hipGetLastError();
}

hipError_t launch_count_segments(const Geometry &g, unsigned long long *total,
                                 unsigned long long *oor, hipStream_t st) {
  const int64_t n = g.rays_per_slice();
  hipLaunchKernelGGL(k_count_segments, dim3(blocks_for(n, kRayBlock)), dim3(kRayBlock),
                     0, st, g.grid, g.d_ang, g.d_yi, g.P, total, oor);
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
