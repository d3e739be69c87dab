// SIRT run through the reference's own engine and reconstruction space,
// mirroring sirt_main.cc:120-202 (input reading and output writing replaced
// by caller buffers).  Test infrastructure only.
#include "harness.h"
#include "data_region_base.h"
#include "disp_engine_reduction.h"
#include "sirt.h"

extern "C" int ref_sirt(int P, int S, int C, const float *theta_deg,
                        float center, int iters, int threads,
                        const float *sino, float *recon_out, float *num_out,
                        float *den_out) {
  ref_harness::ThetaBox tb(theta_deg, P);
  trace_utils::DegreeToRadian(tb.data);                 // sirt_main.cc:121
  TraceMetadata md(static_cast<float *>(tb.data.data), 0, 0, 0, S, P, S, C, C,
                   center);                             // :126-136
  auto *slices = new DataRegionBase<float, TraceMetadata>(
      ref_harness::new_copy(sino, (size_t)P * S * C), md.count(), &md);
  auto *space = new SIRTReconSpace(S, 2 * C * C);       // :151-157
  space->Initialize(md.num_grids());
  auto &replica = space->reduction_objects();
  float init_val = 0.;
  replica.ResetAllItems(init_val);
  DISPEngineBase<SIRTReconSpace, float> *engine =
      new DISPEngineReduction<SIRTReconSpace, float>(nullptr, space, threads);
  int64_t req_number = md.num_cols();
  for (int i = 0; i < iters; ++i) {                     // :177-202
    engine->RunParallelReduction(*slices, req_number);
    engine->ParInPlaceLocalSynchWrapper();
    if (i == iters - 1) {
      for (int s = 0; s < S; ++s)
        for (int j = 0; j < C * C; ++j) {
          if (num_out) num_out[(size_t)s * C * C + j] = replica[s][2 * j];
          if (den_out) den_out[(size_t)s * C * C + j] = replica[s][2 * j + 1];
        }
    }
    space->UpdateRecon(md.recon(), replica);
    engine->ResetReductionSpaces(init_val);
    slices->ResetMirroredRegionIter();
  }
  std::memcpy(recon_out, &md.recon()[0], sizeof(float) * (size_t)S * C * C);
  delete slices;
  delete engine;
  return 0;
}
