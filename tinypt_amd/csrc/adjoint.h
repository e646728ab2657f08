// adjoint.h -- path-replay adjoint of the FP32 wavefront render (k_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tinypt_gpu.h"
#include "tpt_kernels.h"
#include "tpt_trace32.h"

namespace tpt {

int adjoint_render(hipStream_t s, const SceneF32& S, const SceneF64& S64, const EnvDev& env, const LightsDev& lights, const MatsDev& mats,
                   const CamF32& cam, const double* mat_param, int stack_need, const tpt_render_config* cfg,
                   const float* d_loss_d_pixel, double* grad_params, tpt_stats* stats, std::string* err);

}  // namespace tpt
