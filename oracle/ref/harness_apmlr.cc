// APMLR run through the reference's own engine, data region and
// reconstruction space, mirroring apmlr_main.cc:150-244 at one rank (the
// neighbour exchange trace_comm.cc:4-67 returns at once for one rank and is
// not called).  recon_out receives the S real slices (ghosts stripped).
// Test infrastructure only.
#include "harness.h"
#include "data_region_base.h"
#include "disp_engine_reduction.h"
#include "apmlr.h"

extern "C" int ref_apmlr(int P, int S, int C, const float *theta_deg,
                         float center, int iters, int threads,
                         const float *sino, float *recon_out, float beta0,
                         float beta1, float delta0, float delta1, float regw,
                         float *num_out, float *den_out, float *F_out,
                         float *G_out) {
  ref_harness::ThetaBox tb(theta_deg, P);
  trace_utils::DegreeToRadian(tb.data);                 // apmlr_main.cc:151
  float *data = ref_harness::new_copy(sino, (size_t)P * S * C);
  trace_utils::Absolute(data, (size_t)P * S * C);       // :152-154
  TraceMetadata md(static_cast<float *>(tb.data.data), 0, 0, 0, S, P, S, C, C,
                   center, 1, 1.);                      // :158-171
  auto *slices = new APMLRDataRegion(data, md.count(), &md);
  auto *space = new APMLRReconSpace(S, 2 * C * C);
  space->Initialize(md.num_grids());
  auto &replica = space->reduction_objects();
  float init_val = 0.;
  replica.ResetAllItems(init_val);
  DISPEngineBase<APMLRReconSpace, float> *engine =
      new DISPEngineReduction<APMLRReconSpace, float>(nullptr, space, threads);
  int64_t req_number = md.num_cols();
  for (int i = 0; i < iters; ++i) {                     // :209-244
    engine->RunParallelReduction(*slices, req_number);
    engine->ParInPlaceLocalSynchWrapper();
    slices->SetFG(0.);
    space->CalculateFG(*slices, beta0, beta1, delta0, delta1, regw);
    if (i == iters - 1) {
      size_t nn = (size_t)C * C;
      for (int s = 0; s < S; ++s)
        for (size_t j = 0; j < nn; ++j) {
          if (num_out) num_out[s * nn + j] = replica[s][2 * j];
          if (den_out) den_out[s * nn + j] = replica[s][2 * j + 1];
          if (F_out) F_out[s * nn + j] = slices->F()[s * nn + j];
          if (G_out) G_out[s * nn + j] = slices->G()[s * nn + j];
        }
    }
    space->UpdateRecon(*slices, replica);
    engine->ResetReductionSpaces(init_val);
    slices->ResetMirroredRegionIter();
  }
  std::memcpy(recon_out, &md.recon()[(size_t)C * C],
              sizeof(float) * (size_t)S * C * C);
  delete slices;
  delete engine;
  return 0;
}
