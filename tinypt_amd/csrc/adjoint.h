// adjoint.h -- path-replay adjoint of the FP32 wavefront render (k_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tinypt_gpu.h"
#include "tpt_kernels.h"
#include "tpt_trace32.h"

namespace tpt {

// Device memory of one adjoint call, owned by the caller (a context keeps it between calls, so
// the optimisation loop allocates nothing per step).
struct AdjScratch {
  void* dldp;    // d loss / d pixel: pixels * 12 bytes
  void* grad;    // FP64 gradient: materials * 32 bytes
  void* stats;   // 64 bytes of counters
  void* apm;     // perturbed roughness pairs: materials * 8 bytes
  void* gstack;  // traversal stack overflow: gstack_bytes (null when none is needed)
};
enum { ADJ_SCRATCH_DLDP, ADJ_SCRATCH_GRAD, ADJ_SCRATCH_STATS, ADJ_SCRATCH_APM, ADJ_SCRATCH_GSTACK, ADJ_SCRATCH_N };
// Bytes of each buffer of AdjScratch for a render of cfg over n_mats materials.
void adjoint_scratch_bytes(const tpt_render_config* cfg, int n_mats, int stack_need, size_t bytes[ADJ_SCRATCH_N]);

int adjoint_render(hipStream_t s, const SceneF32& S, const SceneF64& S64, const EnvDev& env, const LightsDev& lights, const MatsDev& mats,
                   const CamF32& cam, const double* mat_param, int stack_need, const tpt_render_config* cfg,
                   const float* d_loss_d_pixel, double* grad_params, tpt_stats* stats, const AdjScratch& scratch,
                   std::string* err);

}  // namespace tpt
