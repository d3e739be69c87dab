/*
 * trace_hip.h -- C-ABI of the MI355X (gfx950) iterative-reconstruction hot
 * path.  It replaces, for one process and one GPU, the reference's
 * reconstruction engine loop:
 *
 *   TraceMetadata set-up            src/common/trace_data.h:53-109
 *   DISPEngineReduction + the       src/common/DISP/core/disp_engine_reduction.h:169-202
 *   per-algorithm ReconSpace        src/sirt/sirt.h:16-71, src/mlem/mlem.h:16-68,
 *                                   src/pml/pml.h:63-122, src/apmlr/apmlr.h:55-130
 *   iteration loop of the mains     src/sirt/sirt_main.cc:177-202,
 *                                   src/mlem/mlem_main.cc:171-182,
 *                                   src/pml/pml_main.cc:178-193,
 *                                   src/apmlr/apmlr_main.cc:209-244
 *
 * Conventions
 * - Plain C types only.  The caller owns every buffer; nothing is freed or
 *   adopted by the library (the reference's TraceMetadata deletes the theta
 *   array it is given, trace_data.h:129; this interface copies instead).
 * - Every entry returns TRACE_OK (0) or a negative TRACE_E_* code; the text
 *   of the last error of the calling thread is in trace_last_error().
 * - A handle is bound to one device and one HIP stream; calls on a handle are
 *   asynchronous on that stream unless stated otherwise, and a handle must
 *   not be used from two host threads at once.  Distinct handles are
 *   independent.
 * - Layouts are the reference's: sinogram [P][S][C] (trace_data.h:139-149,
 *   trace_h5io.cc:157-205), image [S][N][N] with pixel index ix + iy*N.
 *   N and C may differ.  The reference's Reduce hard-wires num_grids =
 *   num_cols (sirt.cc:107-108, sirt_main.cc:135), while its TraceMetadata
 *   and traversal functions take them separately (trace_data.h:53-109,
 *   trace_utils.cc:63-215); N != C follows those functions.
 * - Theta is in radians.  The reference's per-application preprocessing
 *   (degree conversion for SIRT/MLEM/APMLR, |sinogram| for MLEM/PML/APMLR,
 *   initial image 0 for SIRT and 1 otherwise) is the caller's, exactly as in
 *   the mains; trace_deg2rad() and trace_absolute() reproduce the reference
 *   helpers bit for bit.
 */
#ifndef TRACE_HIP_H_
#define TRACE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TRACE_OK 0
#define TRACE_E_INVALID (-1)     /* bad argument (null, size, value) */
#define TRACE_E_UNSUPPORTED (-2) /* unknown algorithm or option */
#define TRACE_E_HIP (-3)         /* HIP runtime error (text in trace_last_error) */
#define TRACE_E_NOMEM (-4)       /* device or host allocation failed */
#define TRACE_E_NONFINITE (-5)   /* NaN or Inf in an input sinogram */

/* Algorithm selection; values follow the draft unified CLI's list
 * (src/all/trace.cc:112-118). */
typedef enum {
  TRACE_SIRT = 0,
  TRACE_MLEM = 1,
  TRACE_PML = 2,
  TRACE_APMLR = 3
} trace_alg_t;

/* Geometry of one slice stack: the arguments the mains pass to
 * TraceMetadata (sirt_main.cc:126-136). */
typedef struct {
  int num_projs;       /* P: projections (angles) */
  int num_slices;      /* S: slices */
  int num_cols;        /* C: detector columns */
  int num_grids;       /* N: image side (the reference's mains pass N = C) */
  float center;        /* rotation centre; <= 0 selects C/2+1 (trace_data.h:97) */
  const float *theta;  /* [P] projection angles in radians, host memory (copied) */
} trace_geom_t;

/* Regularisation parameters.  Defaults of the reference CLIs: PML beta 10
 * (pml_main.cc:48-49); APMLR beta0 10, beta1 1, delta0 1, delta1 1, regw 1
 * (apmlr_main.cc:22-26). */
typedef struct {
  float beta;    /* PML beta, APMLR in-plane beta0 (> 0) */
  float beta1;   /* APMLR inter-slice beta1 */
  float delta;   /* APMLR in-plane delta0 (> 0) */
  float delta1;  /* APMLR inter-slice delta1 (> 0) */
  float regw;    /* APMLR inter-slice neighbour weight */
  int apmlr_compat; /* 1: slice s is forward/back-projected with real slice
                       s-1 (slice 0 with a constant-1 ghost), as the reference
                       does (apmlr.cc:1082-1084); 0: with slice s */
} trace_params_t;

typedef struct trace_handle trace_handle_t;

/* Options of a handle (trace_set_option). */
#define TRACE_OPT_EXACT_SQRT 1 /* 1 (default): correctly rounded segment
                                  lengths, bit-identical to the reference;
                                  0: hardware v_sqrt_f32 (<= 1 ulp) */
#define TRACE_OPT_PHASE_TIMING 2 /* 1: bracket every kernel launch with HIP
                                    events on the handle's stream
                                    (trace_phase_times) */
#define TRACE_OPT_BACKPROJECTOR 3 /* 0 (default): LDS-tiled backprojector;
                                     1: whole-ray scatter with global float
                                     atomics (an independent check) */
#define TRACE_OPT_FWD_BUNDLE 4 /* adjacent angles per forward-projector
                                  wavefront (1, 2, 4, 8, 16; the 64 lanes are
                                  angles x 64/(angles x lanes) adjacent
                                  columns x lanes per ray); default 16 at
                                  four lanes per ray, else 8 */
#define TRACE_OPT_FWD_LANES 5 /* lanes per ray of the forward projector
                                 (1, 2, 4; default by stack depth: 4 at 32 or
                                 more slices, 2 at 8-31, else 1): the lanes
                                 of a ray split its slice batch; compiled for
                                 batches 8, 16 and 32 (2 lanes) and 16, 32
                                 (4 lanes) */

#define TRACE_OPT_ORDERED_SUMS 6 /* order of the backprojected float sums:
                                    0: every sum in 32-bit fixed point
                                       (exact to ~1e-8 of a pixel's sum,
                                       rounded once);
                                    1 (default): den = A^T 1 and the PML/
                                       APMLR num summed in the reference's
                                       order (projection, column, segment:
                                       sirt.cc:62, pml.cc:243-244), bit for
                                       bit; SIRT/MLEM num in fixed point;
                                    2: SIRT/MLEM num in the reference's order
                                       too (sirt.cc:61, mlem.cc:56) */
#define TRACE_OPT_ORDERED_TILE 7 /* form of the ordered sums (same bits either
                                    way): 1 (default) LDS-slotted 8x8 tiles
                                    where the geometry allows (no segment
                                    wrapping to the next row, at most two
                                    segments per ray and pixel; checked once
                                    per geometry), else per pixel; 0 always
                                    per pixel */
#define TRACE_OPT_FWD_STAGED 8 /* form of the forward projector at 32-slice
                                  batches (same bits either way): 1 (default)
                                  bundles of 16 angles x 64 columns sweep the
                                  image in slabs staged in LDS; 0: every ray
                                  reads its pixels from HBM (the form used at
                                  other slice batches) */

/* Kernel phases reported by trace_phase_times. */
#define TRACE_PHASE_FORWARD 0 /* forward projection and per-ray weights */
#define TRACE_PHASE_BACK 1    /* backprojection A^T w */
#define TRACE_PHASE_UPDATE 2  /* image update (with the PML/APMLR stencil) */
#define TRACE_PHASE_SETUP 3   /* ray table, column sums, packing, checks */
#define TRACE_NUM_PHASES 4

/* Creates a handle for geometry *g on HIP device `device`, working on
 * `hip_stream` (a hipStream_t; NULL = the device's null stream).  Replaces
 * the TraceMetadata constructor (trace_data.h:53-109) and the per-row
 * prologue of Reduce (sirt.cc:111-124): per-angle sinf/cosf/quadrant and
 * per-column detector coordinates are computed here on the host. */
int trace_create(trace_handle_t **h, const trace_geom_t *g, int device,
                 void *hip_stream);

/* Releases every resource of the handle (NULL is accepted). */
void trace_destroy(trace_handle_t *h);

int trace_set_option(trace_handle_t *h, int option, int value);
/* Current value of an option (TRACE_OPT_*). */
int trace_get_option(trace_handle_t *h, int option, int *value);

/* Launch shape: slices per forward-projector ray (1, 2, 4, 8, 16 or 32; the
 * lanes per ray are kept when the batch has that form, else reset to 1) and the
 * backprojector's image tile (tile_w x tile_h pixels, tile_batch slices per
 * workgroup).  Only the shapes compiled into the library are accepted
 * (TRACE_E_UNSUPPORTED otherwise); trace_amd.autotune picks among them.
 * No reference counterpart (the reference's work unit is fixed at one
 * sinogram row, sirt_main.cc:172). */
int trace_set_launch_shape(trace_handle_t *h, int fwd_batch, int tile_w, int tile_h,
                           int tile_batch);
int trace_get_launch_shape(trace_handle_t *h, int *fwd_batch, int *tile_w, int *tile_h,
                           int *tile_batch);

/* Runs `iters` iterations of `alg` on device buffers: sino [P][S][C] and
 * recon [S][N][N] (in: initial image; out: result).  Replaces the mains'
 * iteration loops (sirt_main.cc:177-202, mlem_main.cc:171-182,
 * pml_main.cc:178-193, apmlr_main.cc:209-244).  prm may be NULL for SIRT and
 * MLEM.  Asynchronous except for the input validation, which reads back one
 * flag (non-finite sinogram values -> TRACE_E_NONFINITE). */
int trace_reconstruct(trace_handle_t *h, trace_alg_t alg, int iters,
                      const float *sino, float *recon,
                      const trace_params_t *prm);

/* As trace_reconstruct with host buffers; copies in and out over PCIe and
 * returns after the result is in `recon`.  This is the call a reference main
 * makes in place of its engine loop (sirt_main.cc:145-202). */
int trace_reconstruct_host(trace_handle_t *h, trace_alg_t alg, int iters,
                           const float *sino, float *recon,
                           const trace_params_t *prm);

/* y = A x: y [P][S][C] from x [S][N][N] (device buffers).  Per ray the sum
 * runs in the reference's segment order (CalculateSimdata, sirt.cc:4-19). */
int trace_forward(trace_handle_t *h, const float *x, float *y);

/* y = A x as trace_forward, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid it).  Its
 * scratch -- the ray table and the slice-batched image copy, P*C*32 bytes
 * (rounded up to 256) plus ceil(S/b)*b*N*N*4 bytes, b the forward slice
 * batch of the launch shape -- lives in a static 128 MiB device array that
 * every GPU the library runs on holds until the process ends; a geometry
 * that needs more gets TRACE_E_INVALID before anything is launched.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other
 * streams, other host threads) race.  Same result as trace_forward, bit for
 * bit. */
int trace_forward_noalloc(trace_handle_t *h, const float *x, float *y);

/* z = A^T r: z [S][N][N] (overwritten) from r [P][S][C] (device buffers).
 * The scatter half of UpdateReconReplica (sirt.cc:35-64) with a free
 * per-ray weight. */
int trace_adjoint(trace_handle_t *h, const float *r, float *z);

/* den = A^T 1 for one slice of this geometry into den [N][N] (device): the
 * column sums the reference accumulates as the second replica entry
 * (sirt.cc:62), summed in its order (bit for bit) unless
 * TRACE_OPT_ORDERED_SUMS is 0. */
int trace_column_sums(trace_handle_t *h, float *den);

/* Per-ray sums of one slice's rays into [P][C] device arrays (either may be
 * null): len_sum = sum of segment lengths (row sums of A) and len2_sum = sum
 * of squared segment lengths, the SIRT normaliser a2 of UpdateReconReplica
 * (sirt.cc:47-50), both accumulated in segment order. */
int trace_ray_sums(trace_handle_t *h, float *len_sum, float *len2_sum);

/* Parity entry: number of segments per pixel of one slice, counts [N][N]
 * (device, float), through the configured backprojector.  Integer-exact, so
 * the tile and whole-ray backprojectors must agree bit for bit. */
int trace_segment_counts(trace_handle_t *h, float *counts);

/* Debug/parity entry: segments of n rays (proj[i], col[i]) (device int
 * arrays), written to idx/len/len2 [n][max_seg] and counts [n] (device).  The
 * per-ray pipeline trace_utils.cc:48-187 as executed by the kernels. */
int trace_dump_rays(trace_handle_t *h, int n, const int *proj, const int *col,
                    int max_seg, int *counts, int *idx, float *len,
                    float *len2);

/* Number of segments (nonzeros of A) of one slice, and how many of them fall
 * outside the image (skipped).  Synchronous. */
int trace_count_segments(trace_handle_t *h, int64_t *nnz, int64_t *out_of_image);

/* Device time in milliseconds and launch count of each phase, summed over
 * the launches recorded since the last reset (TRACE_OPT_PHASE_TIMING).
 * Synchronises with the recorded events.  The reference's -DTIMERON phase
 * timers (sirt_main.cc:98-101, 174-197) are the counterpart. */
int trace_phase_times(trace_handle_t *h, double *ms, int64_t *launches, int reset);

/* Host helpers identical to trace_utils::DegreeToRadian
 * (trace_utils.cc:10-24) and trace_utils::Absolute (trace_utils.cc:4-8). */
void trace_deg2rad(const float *deg, float *rad, int n);
void trace_absolute(float *v, int64_t n);

/* ---- Batched reconstruction of slices with different geometries ------
 *
 * The reference reconstructs the slices of one dataset (one geometry) per
 * run (sirt_main.cc:145-202, mlem_main.cc:139-182).  A batch takes many
 * slices, each with its own detector width C (image C x C), angle set,
 * centre, algorithm (SIRT or MLEM, with the preprocessing and initial
 * image of its main already applied) and iteration count, and runs them
 * together: one HIP-graph launch per trace_batch_run, whose persistent
 * kernels take (slice, work item) pairs from device work queues.  Each
 * slice's result is the one trace_reconstruct gives for it alone.
 * Buffers stay the caller's; their addresses are fixed at creation. */
typedef struct trace_batch trace_batch_t;

typedef struct {
  int num_projs;        /* P */
  int num_cols;         /* C: detector columns */
  float center;         /* rotation axis; <= 0 -> C/2 + 1 (trace_data.h:97) */
  const float *theta;   /* [P] radians (host; copied at creation) */
  trace_alg_t alg;      /* TRACE_SIRT or TRACE_MLEM */
  int iters;            /* iterations of this slice */
  const float *sino;    /* device [P][C] */
  float *recon;         /* device [N][N]: initial image in, result out */
  int num_grids;        /* N: image side; 0 selects N = C (the mains' choice) */
} trace_batch_slice_t;

/* Builds the per-slice tables (synchronous) on `device`, ordered on
 * `hip_stream`. */
int trace_batch_create(trace_batch_t **b, const trace_batch_slice_t *slices, int n,
                       int device, void *hip_stream);

/* Ray-bundle width (adjacent angles per forward wavefront: 1, 2, 4, 8, 16)
 * and backprojector tile (16x16, 32x32 or 64x64); trace_amd.autotune picks
 * them.  Default 8 and 32x32. */
int trace_batch_set_launch_shape(trace_batch_t *b, int bundle_angles, int tile_w, int tile_h);
int trace_batch_get_launch_shape(trace_batch_t *b, int *bundle_angles, int *tile_w,
                                 int *tile_h);

/* TRACE_OPT_EXACT_SQRT as for a handle. */
int trace_batch_set_option(trace_batch_t *b, int option, int value);

/* Captures and instantiates the graph trace_batch_run(b, max_iters) will
 * launch, without launching it (graphs are cached per iteration count and
 * dropped when the launch shape or options change). */
int trace_batch_prepare(trace_batch_t *b, int max_iters);

/* Runs every slice's iterations (at most max_iters each when max_iters > 0)
 * as one graph launch, asynchronous on the batch's stream after checking
 * that no sinogram holds NaN/Inf (TRACE_E_NONFINITE, synchronous). */
int trace_batch_run(trace_batch_t *b, int max_iters);

/* Create, run, synchronise, destroy. */
int trace_batch_reconstruct(const trace_batch_slice_t *slices, int n, int device,
                            void *hip_stream);

void trace_batch_destroy(trace_batch_t *b);

/* ---- Testing hooks (no reference counterpart) ------------------------
 * The reference ignores allocation failures (new throws, hipMalloc does not
 * exist there); these hooks let a test force the library's out-of-memory
 * paths from the host.  trace_debug_fail_alloc(k): the library's device
 * allocations succeed k more times, then one fails with hipErrorOutOfMemory
 * (reported as TRACE_E_NOMEM) and the hook switches itself off; k < 0
 * switches it off.  Process-wide.  trace_debug_live_allocations(): device
 * allocations of the library currently live, process-wide. */
void trace_debug_fail_alloc(int after);
int64_t trace_debug_live_allocations(void);

/* Thread-local text of the last error ("" if none). */
const char *trace_last_error(void);

/* Library version string. */
const char *trace_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TRACE_HIP_H_ */
