/*
 * trace_oracle.c -- CPU restatement of Trace's iterative-reconstruction hot
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * path is compared against.  It is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, never by
 * the product (trace_amd/).
 *
 * Every function states the reference code it restates (file:line relative
 * to the reference repository root).  Floating-point semantics follow the
 * reference's SIRT build (-O3 -ffp-contract=off, no fast-math): this file is
 * compiled with -ffp-contract=off and every expression keeps the reference's
 * float/double promotion order, so at one thread the fp32 entry points are
 * bit-identical to the compiled reference.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   or_trace_ray          bit-exact vs the compiled reference (oracle/_ref)
 *                         and tests/golden fixtures; closed forms at theta=0,
 *                         brute-force fp64 Siddon lengths at small N.
 *   or_forward/or_adjoint dot-product identity, A*1 = chord, vs _ref simdata.
 *   or_reconstruct        bit-exact vs _ref (T=1) for SIRT, MLEM, PML, APMLR
 *                         (both APMLR slice-offset modes are pinned: compat
 *                         against _ref, fixed against the invariants).
 *   or_reconstruct_f64    fp64 twin; pinned by agreement with the fp32 path to
 *                         fp32 rounding level and by the same invariants.
 *
 * Image side N and detector columns C are separate parameters.  The
 * reference's Reduce sets num_grids = num_cols (sirt.cc:107-108); its
 * traversal functions (trace_utils.cc) and TraceMetadata (trace_data.h:
 * 53-109) take them separately, and the N != C traversal is pinned against
 * those functions (oracle/ref/harness_trace.cc: ref_trace_ray_n).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_KPI 3.14159265358979f /* trace_utils.h:8 */

enum { OR_SIRT = 0, OR_MLEM = 1, OR_PML = 2, OR_APMLR = 3 };

typedef struct {
  float beta;   /* PML beta / APMLR in-plane beta0 */
  float beta1;  /* APMLR inter-slice beta */
  float delta;  /* APMLR in-plane delta0 */
  float delta1; /* APMLR inter-slice delta */
  float regw;   /* APMLR inter-slice weight */
  int apmlr_compat; /* 1: forward/backprojection read the ghost-padded recon
                       without the ghost offset, as apmlr.cc:1082-1084 does */
} or_params;

/* ------------------------------------------------------------------ */
/* Host preprocessing                                                  */
/* ------------------------------------------------------------------ */

/* trace_utils.cc:10-24 (DegreeToRadian): float product, then a double
 * division by 180, rounded back to float. */
void or_deg2rad(const float *deg, float *rad, int n) {
  for (int i = 0; i < n; ++i) rad[i] = (float)((double)(deg[i] * OR_KPI) / 180.0);
}

/* trace_utils.cc:4-8 (Absolute). */
void or_absolute(float *v, int64_t n) {
  for (int64_t i = 0; i < n; ++i) v[i] = fabsf(v[i]);
}

/* trace_utils.cc:41-46 (CalculateQuadrant), evaluated in float. */
int or_quadrant(float q) {
  return ((q >= 0 && q < OR_KPI / 2) || (q >= OR_KPI && q < 3 * OR_KPI / 2)) ? 1 : 0;
}

/* trace_data.h:97-101: default centre C/2+1 when centre <= 0, then
 * mov = C/2 - centre and an unconditional +1e-6 nudge (the guard
 * mov - ceil(mov) < 1e-6 can never be false). */
float or_mov(int C, float center) {
  float c = center;
  if (c <= 0.) c = (float)((double)(float)C / 2. + 1.);
  float mov = (float)((double)(float)C / 2. - (double)c);
  if ((double)(mov - ceilf(mov)) < 1e-6) mov = (float)((double)mov + 1e-6);
  return mov;
}

/* trace_data.h:103-108: grid line coordinates -N/2 + i, i = 0..N. */
static void or_grid(int N, float *g) {
  for (int i = 0; i <= N; ++i) g[i] = (float)(-(double)N / 2. + i);
}

/* sirt.cc:128-129: detector coordinate of column `col`. */
float or_yi(int C, int col, float mov) {
  return (float)((double)(1 - C) / 2. + (double)col + (double)mov);
}

/* ------------------------------------------------------------------ */
/* Ray traversal                                                       */
/* ------------------------------------------------------------------ */

typedef struct {
  int N;
  float *gx, *gy;            /* [N+1] */
  float *cx, *cy;            /* [N+1] crossings */
  float *ax, *ay, *bx, *by;  /* [N+1] trimmed lists */
  float *px, *py;            /* [2N+2] merged points */
} or_ws;

static int ws_init(or_ws *w, int N) {
  w->N = N;
  size_t n1 = (size_t)N + 1, n2 = 2 * (size_t)N + 2;
  w->gx = malloc(n1 * sizeof(float));
  w->gy = malloc(n1 * sizeof(float));
  w->cx = malloc(n1 * sizeof(float));
  w->cy = malloc(n1 * sizeof(float));
  w->ax = malloc(n1 * sizeof(float));
  w->ay = malloc(n1 * sizeof(float));
  w->bx = malloc(n1 * sizeof(float));
  w->by = malloc(n1 * sizeof(float));
  w->px = malloc(n2 * sizeof(float));
  w->py = malloc(n2 * sizeof(float));
  if (!w->gx || !w->gy || !w->cx || !w->cy || !w->ax || !w->ay || !w->bx ||
      !w->by || !w->px || !w->py)
    return -1;
  or_grid(N, w->gx);
  or_grid(N, w->gy);
  return 0;
}

static void ws_free(or_ws *w) {
  free(w->gx); free(w->gy); free(w->cx); free(w->cy); free(w->ax);
  free(w->ay); free(w->bx); free(w->by); free(w->px); free(w->py);
}

/* One ray: trace_utils.cc:48-80 (crossings of the source->detector line with
 * every grid line), :82-111 (strict trim to the open grid square),
 * :113-158 (two-list merge; A walked forward iff quadrant==1, ties go to the
 * B list), :160-187 (segment length, truncating pixel index).
 * Writes nseg = len-1 segments; returns nseg. */
static int trace_one(or_ws *w, float yi, float sinq, float cosq, int quadrant,
                     int *idx, float *len, float *len2) {
  const int N = w->N;
  const float xi = -1e6f;
  float srcx = xi * cosq - yi * sinq;
  float srcy = xi * sinq + yi * cosq;
  float detx = -1 * (xi * cosq + yi * sinq);
  float dety = -xi * sinq + yi * cosq;
  float slope = (srcy - dety) / (srcx - detx);
  float islope = 1 / slope;
  for (int n = 0; n <= N; ++n) {
    w->cx[n] = islope * (w->gy[n] - srcy) + srcx;
    w->cy[n] = slope * (w->gx[n] - srcx) + srcy;
  }
  int alen = 0, blen = 0;
  for (int i = 0; i <= N; ++i) {
    if (w->cx[i] > w->gx[0] && w->cx[i] < w->gx[N]) {
      w->ax[alen] = w->cx[i];
      w->ay[alen] = w->gy[i];
      alen++;
    }
    if (w->cy[i] > w->gy[0] && w->cy[i] < w->gy[N]) {
      w->bx[blen] = w->gx[i];
      w->by[blen] = w->cy[i];
      blen++;
    }
  }
  int i = 0, j = 0, k = 0;
  while (i < alen && j < blen) {
    int a = quadrant ? i : (alen - 1 - i);
    if (w->ax[a] < w->bx[j]) {
      w->px[k] = w->ax[a]; w->py[k] = w->ay[a]; i++;
    } else {
      w->px[k] = w->bx[j]; w->py[k] = w->by[j]; j++;
    }
    k++;
  }
  for (; i < alen; ++i, ++k) {
    int a = quadrant ? i : (alen - 1 - i);
    w->px[k] = w->ax[a]; w->py[k] = w->ay[a];
  }
  for (; j < blen; ++j, ++k) {
    w->px[k] = w->bx[j]; w->py[k] = w->by[j];
  }
  int total = alen + blen;
  float mg = (float)((double)N / 2.);
  for (int s = 0; s < total - 1; ++s) {
    float dx = w->px[s + 1] - w->px[s];
    float mx = (float)((double)(w->px[s + 1] + w->px[s]) / 2.);
    float dy = w->py[s + 1] - w->py[s];
    float my = (float)((double)(w->py[s + 1] + w->py[s]) / 2.);
    float l2 = dx * dx + dy * dy;
    if (len2) len2[s] = l2;
    len[s] = sqrtf(l2);
    int ix = (int)(mx + mg);
    int iy = (int)(my + mg);
    idx[s] = ix + iy * N;
  }
  return total > 1 ? total - 1 : 0;
}

/* Public single-ray entry: geometry from (C, centre) exactly as
 * TraceMetadata (trace_data.h:53-109) and Reduce (sirt.cc:116-129) set it up.
 * Buffers must hold 2N+2 entries.  Returns the segment count or -1. */
int or_trace_ray(int N, int C, float center, float theta, int col, int *idx,
                 float *len, float *len2) {
  or_ws w;
  if (ws_init(&w, N)) { ws_free(&w); return -1; }
  float mov = or_mov(C, center);
  float yi = or_yi(C, col, mov);
  int n = trace_one(&w, yi, sinf(theta), cosf(theta), or_quadrant(theta), idx,
                    len, len2);
  ws_free(&w);
  return n;
}

/* Batch of single rays (theta[i], col[i]): segments concatenated into
 * idx/len, ray i owning [offsets[i], offsets[i+1]).  Same per-ray code path
 * as or_trace_ray (trace_utils.cc:48-187 with sirt.cc:116-129's set-up).
 * Returns the total segment count, or -1 when it would exceed `cap`. */
int64_t or_trace_rays(int N, int C, float center, int n, const float *theta,
                      const int *col, int64_t *offsets, int *idx, float *len,
                      int64_t cap) {
  or_ws w;
  if (ws_init(&w, N)) { ws_free(&w); return -1; }
  int *ti = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *tl = malloc((2 * (size_t)N + 2) * sizeof(float));
  const float mov = or_mov(C, center);
  int64_t total = 0;
  offsets[0] = 0;
  for (int r = 0; r < n; ++r) {
    const float q = theta[r];
    int k = trace_one(&w, or_yi(C, col[r], mov), sinf(q), cosf(q), or_quadrant(q),
                      ti, tl, NULL);
    if (k < 0 || total + k > cap) { total = -1; break; }
    memcpy(idx + total, ti, (size_t)k * sizeof(int));
    memcpy(len + total, tl, (size_t)k * sizeof(float));
    total += k;
    offsets[r + 1] = total;
  }
  free(ti); free(tl); ws_free(&w);
  return total;
}

/* Merged intersection points of one ray (debugging aid; same code path). */
int or_ray_points(int N, int C, float center, float theta, int col, float *px,
                  float *py) {
  or_ws w;
  if (ws_init(&w, N)) { ws_free(&w); return -1; }
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  float yi = or_yi(C, col, or_mov(C, center));
  int n = trace_one(&w, yi, sinf(theta), cosf(theta), or_quadrant(theta), idx,
                    len, NULL);
  int npts = n > 0 ? n + 1 : 0;
  /* total points may be 1 when n == 0; report the merged list length */
  memcpy(px, w.px, (size_t)(npts) * sizeof(float));
  memcpy(py, w.py, (size_t)(npts) * sizeof(float));
  free(idx); free(len); ws_free(&w);
  return npts;
}

/* ------------------------------------------------------------------ */
/* Operators                                                           */
/* ------------------------------------------------------------------ */

/* y[p][s][c] = sum_k x_s[idx_k] * len_k, k ascending, fp32
 * (CalculateSimdata, sirt.cc:4-19).  Segments whose index falls outside the
 * slice are skipped (the reference would read past the image there). */
int or_forward(int P, int S, int C, int N, float center, const float *theta,
               const float *x, float *y) {
  const int64_t NN = (int64_t)N * N;
  or_ws w;
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  if (ws_init(&w, N) || !idx || !len) { ws_free(&w); free(idx); free(len); return -1; }
  float mov = or_mov(C, center);
  for (int p = 0; p < P; ++p) {
    float q = theta[p];
    int quad = or_quadrant(q);
    float sq = sinf(q), cq = cosf(q);
    for (int c = 0; c < C; ++c) {
      int ns = trace_one(&w, or_yi(C, c, mov), sq, cq, quad, idx, len, NULL);
      for (int s = 0; s < S; ++s) {
        const float *xs = x + (int64_t)s * NN;
        float acc = 0.f;
        for (int k = 0; k < ns; ++k)
          if (idx[k] >= 0 && idx[k] < NN) acc += xs[idx[k]] * len[k];
        y[((int64_t)p * S + s) * C + c] = acc;
      }
    }
  }
  ws_free(&w); free(idx); free(len);
  return 0;
}

/* z_s[j] = sum over rays of len * r[p][s][c] (the adjoint A^T r), rays in
 * sinogram order; fp64 accumulation so the result is order-insensitive. */
int or_adjoint(int P, int S, int C, int N, float center, const float *theta,
               const float *r, double *z) {
  const int64_t NN = (int64_t)N * N;
  or_ws w;
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  if (ws_init(&w, N) || !idx || !len) { ws_free(&w); free(idx); free(len); return -1; }
  memset(z, 0, (size_t)(S * NN) * sizeof(double));
  float mov = or_mov(C, center);
  for (int p = 0; p < P; ++p) {
    float q = theta[p];
    int quad = or_quadrant(q);
    float sq = sinf(q), cq = cosf(q);
    for (int c = 0; c < C; ++c) {
      int ns = trace_one(&w, or_yi(C, c, mov), sq, cq, quad, idx, len, NULL);
      for (int s = 0; s < S; ++s) {
        double v = r[((int64_t)p * S + s) * C + c];
        double *zs = z + (int64_t)s * NN;
        for (int k = 0; k < ns; ++k)
          if (idx[k] >= 0 && idx[k] < NN) zs[idx[k]] += (double)len[k] * v;
      }
    }
  }
  ws_free(&w); free(idx); free(len);
  return 0;
}

/* Total segment count sum_rays (len-1) for one slice; also the number of
 * segments whose pixel index is out of the image (reported separately). */
int64_t or_count_segments(int P, int C, int N, float center, const float *theta,
                          int64_t *out_of_range) {
  or_ws w;
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  if (ws_init(&w, N) || !idx || !len) { ws_free(&w); free(idx); free(len); return -1; }
  float mov = or_mov(C, center);
  int64_t tot = 0, oor = 0;
  for (int p = 0; p < P; ++p) {
    float q = theta[p];
    int quad = or_quadrant(q);
    float sq = sinf(q), cq = cosf(q);
    for (int c = 0; c < C; ++c) {
      int ns = trace_one(&w, or_yi(C, c, mov), sq, cq, quad, idx, len, NULL);
      tot += ns;
      for (int k = 0; k < ns; ++k)
        if (idx[k] < 0 || idx[k] >= (int64_t)N * N) oor++;
    }
  }
  if (out_of_range) *out_of_range = oor;
  ws_free(&w); free(idx); free(len);
  return tot;
}

/* ------------------------------------------------------------------ */
/* Regularisation stencils                                            */
/* ------------------------------------------------------------------ */

/* PML weights, pml.h:22-29 (the float values the reference stores). */
static const float kPmlIn[8] = {
    0.1464466094f, 0.1464466094f, 0.1464466094f, 0.1464466094f,
    0.10355339059f, 0.10355339059f, 0.10355339059f, 0.10355339059f};
static const float kPmlEdge[5] = {0.226540919667f, 0.226540919667f,
                                  0.226540919667f, 0.1601886205f, 0.1601886205f};
static const float kPmlCorner[3] = {0.36939806251f, 0.36939806251f,
                                    0.26120387496f};

/* Neighbour offsets (dx, dy) per pixel class, in the reference's q order.
 * Class: 0 inner, 1 top row, 2 bottom row, 3 left column, 4 right column,
 * 5..8 corners top-left, top-right, bottom-left, bottom-right.  Orders follow
 * pml.cc:45-53, 69-74, 87-92, 105-110, 123-128, 141-144, 156-159, 171-174,
 * 187-190 (the same orders are used by apmlr.cc for the in-plane part). */
static const int kNb[9][8][2] = {
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
static const int kNbCount[9] = {8, 5, 5, 5, 5, 3, 3, 3, 3};

static int pixel_class(int m, int n, int N) {
  int top = (n == 0), bot = (n == N - 1), left = (m == 0), right = (m == N - 1);
  if (N == 1) return -1;
  if (top && left) return 5;
  if (top && right) return 6;
  if (bot && left) return 7;
  if (bot && right) return 8;
  if (top) return 1;
  if (bot) return 2;
  if (left) return 3;
  if (right) return 4;
  return 0;
}

/* pml.cc:25-201 (CalculateFG) for one slice, starting from F = G = 0
 * (SetFG(0), pml.h:49-57), then G += den (pml.cc:194-200). */
static void pml_fg_slice(const float *x, const float *den, float *F, float *G,
                         int N, float beta) {
  for (int n = 0; n < N; ++n)
    for (int m = 0; m < N; ++m) {
      int cls = pixel_class(m, n, N);
      int64_t j = (int64_t)n * N + m;
      float f = 0.f, g = 0.f;
      if (cls >= 0) {
        const float *wg = cls == 0 ? kPmlIn : (cls <= 4 ? kPmlEdge : kPmlCorner);
        for (int q = 0; q < kNbCount[cls]; ++q) {
          int64_t jq = j + kNb[cls][q][0] + (int64_t)kNb[cls][q][1] * N;
          f += 2 * beta * wg[q];
          g -= 2 * beta * wg[q] * (x[j] + x[jq]);
        }
      }
      F[j] = f;
      G[j] = g + den[j];
    }
}

/* APMLR weights for one pixel class: apmlr.cc:40-51 (inner), 94-101 (edges),
 * 245-251 (corners) with both z neighbours; the one-sided variants
 * (apmlr.cc:361-370 and the Bottom twin) add regw once instead of twice. */
static void apmlr_weights(int cls, float regw, int nz, float *wg) {
  float total;
  int nin = kNbCount[cls];
  if (cls == 0)
    total = (float)(4 + 4 / sqrt(2.) + (nz == 2 ? (double)(2 * regw) : (double)regw));
  else if (cls <= 4)
    total = (float)(3 + 2 / sqrt(2.) + (nz == 2 ? (double)(2 * regw) : (double)regw));
  else
    total = (float)(2 + 1 / sqrt(2.) + (nz == 2 ? (double)(2 * regw) : (double)regw));
  /* axial neighbours come first in every class, diagonal ones last */
  int naxial = cls == 0 ? 4 : (cls <= 4 ? 3 : 2);
  for (int q = 0; q < nin; ++q)
    wg[q] = q < naxial ? 1 / total : (float)(1 / sqrt(2.) / total);
  for (int q = 0; q < nz; ++q) wg[nin + q] = regw / total;
}

/* apmlr.cc:919-982 (CalculateFG), with CalculateFGInner (:28-347), FGTop
 * (:349-630) and FGBottom (:632-915).  xr points at the first real slice of
 * the ghost-padded recon (padded + N*N); slices 0 and S-1 (when S > 1) use
 * the one-sided z stencil, others the two-sided one.  G += den at the end. */
static void apmlr_fg(const float *xr, const float *den, float *F, float *G, int S,
                     int N, float beta, float beta1, float delta, float delta1,
                     float regw) {
  const int64_t NN = (int64_t)N * N;
  for (int s = 0; s < S; ++s) {
    int top = (s == 0), bottom = (s == S - 1 && S - 1 != 0);
    int nz;
    int zoff[2];
    if (top) { nz = 1; zoff[0] = 1; }
    else if (bottom) { nz = 1; zoff[0] = -1; }
    else { nz = 2; zoff[0] = 1; zoff[1] = -1; }
    float wcls[9][10];
    for (int cls = 0; cls < 9; ++cls) apmlr_weights(cls, regw, nz, wcls[cls]);
    const float *x = xr + s * NN;
    for (int n = 0; n < N; ++n)
      for (int m = 0; m < N; ++m) {
        int cls = pixel_class(m, n, N);
        int64_t j = (int64_t)n * N + m;
        float f = 0.f, g = 0.f;
        if (cls >= 0) {
          const float *wg = wcls[cls];
          int nin = kNbCount[cls];
          for (int q = 0; q < nin; ++q) {
            int64_t jq = j + kNb[cls][q][0] + (int64_t)kNb[cls][q][1] * N;
            float mg = x[j] + x[jq];
            float rg = x[j] - x[jq];
            float gm = 1 / (1 + fabsf(rg / delta));
            f += 2 * beta * wg[q] * gm;
            g -= 2 * beta * wg[q] * gm * mg;
          }
          for (int q = 0; q < nz; ++q) {
            float xq = x[j + zoff[q] * NN];
            float mg = x[j] + xq;
            float rg = x[j] - xq;
            float gm = 1 / (1 + fabsf(rg / delta1));
            f += 2 * beta1 * wg[nin + q] * gm;
            g -= 2 * beta1 * wg[nin + q] * gm * mg;
          }
        }
        F[s * NN + j] = f;
        G[s * NN + j] = g + den[s * NN + j];
      }
  }
}

/* ------------------------------------------------------------------ */
/* Reconstruction driver (one thread, reference row order)            */
/* ------------------------------------------------------------------ */

/* One full run of `iters` iterations, restating the mains' loop
 * (sirt_main.cc:177-202, mlem_main.cc:171-182, pml_main.cc:178-193,
 * apmlr_main.cc:209-244) at T = 1: rows (p, s) in sinogram order, each row's
 * C rays traced and accumulated into num/den of slice s
 * (UpdateReconReplica: sirt.cc:35-64, mlem.cc:37-61, pml.cc:222-246,
 * apmlr.cc:1003-1027), then the per-algorithm update (sirt.cc:21-33,
 * mlem.cc:21-34, pml.cc:3-23 / :25-201, apmlr.cc:3-26 / :919-982).
 *
 * theta is in radians; sino [P][S][C] is used as given (callers apply the
 * reference's per-algorithm preprocessing).  recon [S][N][N] holds the
 * initial image on entry and the result on return.  Optional outputs of the
 * last iteration: num_out/den_out [S][N][N] (before the update), F_out/G_out
 * [S][N][N] (PML/APMLR). */
int or_reconstruct(int alg, int P, int S, int C, int N, const float *theta,
                   float center, int iters, const float *sino, float *recon,
                   const or_params *prm, float *num_out, float *den_out,
                   float *F_out, float *G_out) {
  const int64_t NN = (int64_t)N * N;
  const int ghost = (alg == OR_APMLR) ? 1 : 0;
  const int64_t SP = (int64_t)S + 2 * ghost;
  float *x = malloc((size_t)(SP * NN) * sizeof(float));
  float *num = malloc((size_t)(S * NN) * sizeof(float));
  float *den = malloc((size_t)(S * NN) * sizeof(float));
  float *F = NULL, *G = NULL;
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  float *len2 = malloc((2 * (size_t)N + 2) * sizeof(float));
  or_ws w;
  int rc = 0;
  if (alg == OR_PML || alg == OR_APMLR) {
    F = malloc((size_t)(S * NN) * sizeof(float));
    G = malloc((size_t)(S * NN) * sizeof(float));
  }
  if (ws_init(&w, N) || !x || !num || !den || !idx || !len || !len2 ||
      ((alg == OR_PML || alg == OR_APMLR) && (!F || !G))) {
    rc = -1;
    goto done;
  }
  /* ghost slices hold the reference's recon init value 1 (apmlr_main.cc:170) */
  for (int64_t i = 0; i < SP * NN; ++i) x[i] = 1.f;
  memcpy(x + ghost * NN, recon, (size_t)(S * NN) * sizeof(float));
  const float mov = or_mov(C, center);
  const int compat = prm ? prm->apmlr_compat : 1;

  for (int it = 0; it < iters; ++it) {
    memset(num, 0, (size_t)(S * NN) * sizeof(float));
    memset(den, 0, (size_t)(S * NN) * sizeof(float));
    for (int p = 0; p < P; ++p) {
      float q = theta[p];
      int quad = or_quadrant(q);
      float sq = sinf(q), cq = cosf(q);
      for (int s = 0; s < S; ++s) {
        /* slice image seen by the ray pass: APMLR compat reads padded slice s
         * (= real slice s-1), fixed mode reads the real slice s */
        const int64_t xoff = (alg == OR_APMLR && !compat) ? (s + 1) * NN : s * NN;
        const float *xs = x + xoff;
        float *nums = num + s * NN, *dens = den + s * NN;
        for (int c = 0; c < C; ++c) {
          int ns = trace_one(&w, or_yi(C, c, mov), sq, cq, quad, idx, len, len2);
          float ray = sino[((int64_t)p * S + s) * C + c];
          float sim = 0.f;
          for (int k = 0; k < ns; ++k)
            if (idx[k] >= 0 && idx[k] < NN) sim += xs[idx[k]] * len[k];
          if (alg == OR_SIRT) {
            float a2 = 0.f;
            for (int k = 0; k < ns; ++k) a2 += len2[k];
            float upd = (ray - sim) / a2;
            for (int k = 0; k < ns; ++k) {
              if (idx[k] < 0 || idx[k] >= NN) continue;
              nums[idx[k]] += len[k] * upd;
              dens[idx[k]] += len[k];
            }
          } else if (alg == OR_MLEM) {
            float upd = ray / sim;
            for (int k = 0; k < ns; ++k) {
              if (idx[k] < 0 || idx[k] >= NN) continue;
              nums[idx[k]] += len[k] * upd;
              dens[idx[k]] += len[k];
            }
          } else {
            float upd = ray / sim;
            for (int k = 0; k < ns; ++k) {
              if (idx[k] < 0 || idx[k] >= NN) continue;
              nums[idx[k]] -= xs[idx[k]] * len[k] * upd;
              dens[idx[k]] += len[k];
            }
          }
        }
      }
    }
    if (it == iters - 1) {
      if (num_out) memcpy(num_out, num, (size_t)(S * NN) * sizeof(float));
      if (den_out) memcpy(den_out, den, (size_t)(S * NN) * sizeof(float));
    }
    float *xr = x + ghost * NN;
    if (alg == OR_SIRT) {
      for (int64_t i = 0; i < S * NN; ++i) xr[i] += num[i] / den[i];
    } else if (alg == OR_MLEM) {
      for (int64_t i = 0; i < S * NN; ++i) xr[i] *= num[i] / den[i];
    } else {
      if (alg == OR_PML) {
        for (int s = 0; s < S; ++s)
          pml_fg_slice(xr + s * NN, den + s * NN, F + s * NN, G + s * NN, N,
                       prm->beta);
      } else {
        apmlr_fg(xr, den, F, G, S, N, prm->beta, prm->beta1, prm->delta,
                 prm->delta1, prm->regw);
      }
      if (it == iters - 1) {
        if (F_out) memcpy(F_out, F, (size_t)(S * NN) * sizeof(float));
        if (G_out) memcpy(G_out, G, (size_t)(S * NN) * sizeof(float));
      }
      for (int64_t i = 0; i < S * NN; ++i)
        xr[i] = (-G[i] + sqrtf(G[i] * G[i] - 8 * num[i] * F[i])) / (4 * F[i]);
    }
  }
  memcpy(recon, x + ghost * NN, (size_t)(S * NN) * sizeof(float));
done:
  ws_free(&w);
  free(x); free(num); free(den); free(F); free(G);
  free(idx); free(len); free(len2);
  return rc;
}

/* ------------------------------------------------------------------ */
/* FP64 twin of or_reconstruct                                         */
/* ------------------------------------------------------------------ */

/* Same algorithm and row order as or_reconstruct, with the system matrix
 * taken from the fp32 traversal (the reference's A) and every accumulation,
 * weight and update carried in double.  The recon is double in and out. */
int or_reconstruct_f64(int alg, int P, int S, int C, int N, const float *theta,
                       float center, int iters, const float *sino,
                       double *recon, const or_params *prm) {
  const int64_t NN = (int64_t)N * N;
  const int ghost = (alg == OR_APMLR) ? 1 : 0;
  const int64_t SP = (int64_t)S + 2 * ghost;
  double *x = malloc((size_t)(SP * NN) * sizeof(double));
  double *num = malloc((size_t)(S * NN) * sizeof(double));
  double *den = malloc((size_t)(S * NN) * sizeof(double));
  int *idx = malloc((2 * (size_t)N + 2) * sizeof(int));
  float *len = malloc((2 * (size_t)N + 2) * sizeof(float));
  float *len2 = malloc((2 * (size_t)N + 2) * sizeof(float));
  or_ws w;
  int rc = 0;
  if (ws_init(&w, N) || !x || !num || !den || !idx || !len || !len2) {
    rc = -1;
    goto done;
  }
  for (int64_t i = 0; i < SP * NN; ++i) x[i] = 1.0;
  memcpy(x + ghost * NN, recon, (size_t)(S * NN) * sizeof(double));
  const float mov = or_mov(C, center);
  const int compat = prm ? prm->apmlr_compat : 1;
  for (int it = 0; it < iters; ++it) {
    memset(num, 0, (size_t)(S * NN) * sizeof(double));
    memset(den, 0, (size_t)(S * NN) * sizeof(double));
    for (int p = 0; p < P; ++p) {
      float q = theta[p];
      int quad = or_quadrant(q);
      float sq = sinf(q), cq = cosf(q);
      for (int s = 0; s < S; ++s) {
        const int64_t xoff = (alg == OR_APMLR && !compat) ? (s + 1) * NN : s * NN;
        const double *xs = x + xoff;
        double *nums = num + s * NN, *dens = den + s * NN;
        for (int c = 0; c < C; ++c) {
          int ns = trace_one(&w, or_yi(C, c, mov), sq, cq, quad, idx, len, len2);
          double ray = sino[((int64_t)p * S + s) * C + c];
          double sim = 0.0, a2 = 0.0;
          for (int k = 0; k < ns; ++k) {
            if (idx[k] >= 0 && idx[k] < NN) sim += xs[idx[k]] * (double)len[k];
            a2 += (double)len[k] * (double)len[k];
          }
          double upd = alg == OR_SIRT ? (ray - sim) / a2 : ray / sim;
          for (int k = 0; k < ns; ++k) {
            if (idx[k] < 0 || idx[k] >= NN) continue;
            if (alg == OR_SIRT || alg == OR_MLEM)
              nums[idx[k]] += (double)len[k] * upd;
            else
              nums[idx[k]] -= xs[idx[k]] * (double)len[k] * upd;
            dens[idx[k]] += (double)len[k];
          }
        }
      }
    }
    double *xr = x + ghost * NN;
    if (alg == OR_SIRT) {
      for (int64_t i = 0; i < S * NN; ++i) xr[i] += num[i] / den[i];
    } else if (alg == OR_MLEM) {
      for (int64_t i = 0; i < S * NN; ++i) xr[i] *= num[i] / den[i];
    } else {
      double *F = malloc((size_t)(S * NN) * sizeof(double));
      double *G = malloc((size_t)(S * NN) * sizeof(double));
      if (!F || !G) { free(F); free(G); rc = -1; goto done; }
      for (int s = 0; s < S; ++s) {
        int top = (s == 0), bottom = (s == S - 1 && S - 1 != 0);
        int nz = (alg == OR_PML) ? 0 : ((top || bottom) ? 1 : 2);
        int zoff[2] = {top ? 1 : -1, -1};
        if (nz == 2) { zoff[0] = 1; zoff[1] = -1; }
        const double *xs = xr + s * NN;
        for (int n = 0; n < N; ++n)
          for (int m = 0; m < N; ++m) {
            int cls = pixel_class(m, n, N);
            int64_t j = (int64_t)n * N + m;
            double f = 0.0, g = 0.0;
            if (cls >= 0) {
              int nin = kNbCount[cls];
              double wg[10];
              if (alg == OR_PML) {
                const float *wf = cls == 0 ? kPmlIn : (cls <= 4 ? kPmlEdge : kPmlCorner);
                for (int q = 0; q < nin; ++q) wg[q] = wf[q];
              } else {
                int naxial = cls == 0 ? 4 : (cls <= 4 ? 3 : 2);
                double total = (cls == 0 ? 4 + 4 / sqrt(2.) : cls <= 4 ? 3 + 2 / sqrt(2.)
                                                                      : 2 + 1 / sqrt(2.)) +
                               nz * (double)prm->regw;
                for (int q = 0; q < nin; ++q)
                  wg[q] = q < naxial ? 1 / total : 1 / sqrt(2.) / total;
                for (int q = 0; q < nz; ++q) wg[nin + q] = (double)prm->regw / total;
              }
              for (int q = 0; q < nin; ++q) {
                int64_t jq = j + kNb[cls][q][0] + (int64_t)kNb[cls][q][1] * N;
                double gm = 1.0;
                if (alg == OR_APMLR) gm = 1 / (1 + fabs((xs[j] - xs[jq]) / prm->delta));
                f += 2 * (double)prm->beta * wg[q] * gm;
                g -= 2 * (double)prm->beta * wg[q] * gm * (xs[j] + xs[jq]);
              }
              for (int q = 0; q < nz; ++q) {
                double xq = xs[j + zoff[q] * NN];
                double gm = 1 / (1 + fabs((xs[j] - xq) / prm->delta1));
                f += 2 * (double)prm->beta1 * wg[nin + q] * gm;
                g -= 2 * (double)prm->beta1 * wg[nin + q] * gm * (xs[j] + xq);
              }
            }
            F[s * NN + j] = f;
            G[s * NN + j] = g + den[s * NN + j];
          }
      }
      for (int64_t i = 0; i < S * NN; ++i)
        xr[i] = (-G[i] + sqrt(G[i] * G[i] - 8 * num[i] * F[i])) / (4 * F[i]);
      free(F); free(G);
    }
  }
  memcpy(recon, x + ghost * NN, (size_t)(S * NN) * sizeof(double));
done:
  ws_free(&w);
  free(x); free(num); free(den); free(idx); free(len); free(len2);
  return rc;
}
