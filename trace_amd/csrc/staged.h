// staged.h -- the slab-staged forward projector (staged.hip): host-side
// declarations.  Not part of the public interface.
#pragma once
#include "internal.h"

namespace trace_hip {

// Slab-staged forward projector (staged.hip), for 32-slice batches: a
// workgroup holds kStG adjacent angles x kStW adjacent columns, one lane per
// ray, and sweeps the image in slabs of kStH pixels, staging a window of at
// most kStEmax pixels across each slab in LDS.
constexpr int kStSlices = 32;
constexpr int kStG = 16;
constexpr int kStW = 64;
constexpr int kStH = 4;
constexpr int kStEmax = 128;
enum StagedAxis { kStAxisX = 0, kStAxisYUp = 1, kStAxisYDown = 2 };
// Per-geometry plan: per bundle its sweep axis, per (bundle, slab) the
// first pixel across the slab and the window's extent (0: empty).
struct StagedPlan {
  int *axis = nullptr;    // [nwg]
  int2 *range = nullptr;  // [nwg][nslab]
  int nslab = 0, ngc = 0;
};
void staged_plan_size(const Geometry &g, int &nwg, int &nslab, int &ngc);
hipError_t launch_staged_plan(const Geometry &g, const RayInfo *rays, StagedPlan &pl,
                              hipStream_t st);
hipError_t launch_forward_staged(const Geometry &g, const RayInfo *rays, const StagedPlan &pl,
                                 int mode, bool exact, const float *xb, const float *sino,
                                 float *out, unsigned *maxw, hipStream_t st);

}  // namespace trace_hip
