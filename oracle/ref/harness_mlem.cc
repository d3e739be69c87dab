// MLEM run through the reference's own engine and reconstruction space,
// mirroring mlem_main.cc:113-182.  Test infrastructure only.
#include "harness.h"
#include "data_region_base.h"
#include "disp_engine_reduction.h"
#include "mlem.h"

extern "C" int ref_mlem(int P, int S, int C, const float *theta_deg,
                        float center, int iters, int threads,
                        const float *sino, float *recon_out, float *num_out,
                        float *den_out) {
  ref_harness::ThetaBox tb(theta_deg, P);
  trace_utils::DegreeToRadian(tb.data);                 // mlem_main.cc:114
  float *data = ref_harness::new_copy(sino, (size_t)P * S * C);
  trace_utils::Absolute(data, (size_t)P * S * C);       // :115-117
  TraceMetadata md(static_cast<float *>(tb.data.data), 0, 0, 0, S, P, S, C, C,
                   center, 0, 1.);                      // :121-134
  auto *slices = new DataRegionBase<float, TraceMetadata>(data, md.count(), &md);
  auto *space = new MLEMReconSpace(S, 2 * C * C);
  space->Initialize(md.num_grids());
  auto &replica = space->reduction_objects();
  float init_val = 0.;
  replica.ResetAllItems(init_val);
  DISPEngineBase<MLEMReconSpace, float> *engine =
      new DISPEngineReduction<MLEMReconSpace, float>(nullptr, space, threads);
  int64_t req_number = md.num_cols();
  for (int i = 0; i < iters; ++i) {                     // :171-182
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
