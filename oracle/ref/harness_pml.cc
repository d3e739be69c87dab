// PML run through the reference's own engine, data region and
// reconstruction space, mirroring pml_main.cc:120-193 (theta is used as
// read, without degree conversion: pml_main.cc:121).  Test infrastructure
// only.
#include "harness.h"
#include "data_region_base.h"
#include "disp_engine_reduction.h"
#include "pml.h"

extern "C" int ref_pml(int P, int S, int C, const float *theta, float center,
                       int iters, int threads, const float *sino,
                       float *recon_out, float beta, float *num_out,
                       float *den_out, float *F_out, float *G_out) {
  float *th = ref_harness::new_copy(theta, (size_t)P);
  float *data = ref_harness::new_copy(sino, (size_t)P * S * C);
  trace_utils::Absolute(data, (size_t)P * S * C);       // pml_main.cc:122-124
  TraceMetadata md(th, 0, 0, 0, S, P, S, C, C, center, 0, 1.);  // :128-141
  auto *slices = new PMLDataRegion(data, md.count(), &md);
  auto *space = new PMLReconSpace(S, 2 * C * C);
  space->Initialize(md.num_grids());
  auto &replica = space->reduction_objects();
  float init_val = 0.;
  replica.ResetAllItems(init_val);
  DISPEngineBase<PMLReconSpace, float> *engine =
      new DISPEngineReduction<PMLReconSpace, float>(nullptr, space, threads);
  int64_t req_number = md.num_cols();
  for (int i = 0; i < iters; ++i) {                     // :178-193
    engine->RunParallelReduction(*slices, req_number);
    engine->ParInPlaceLocalSynchWrapper();
    slices->SetFG(0.);
    space->CalculateFG(*slices, beta);
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
  std::memcpy(recon_out, &md.recon()[0], sizeof(float) * (size_t)S * C * C);
  delete slices;
  delete engine;
  return 0;
}
