// k_adjoint.hip -- path-replay adjoint (placeholder until the replay kernels land).
#include "adjoint.h"

namespace tpt {

int adjoint_render(hipStream_t, const SceneF32&, const EnvDev&, const LightsDev&, const MatsDev&, const CamF32&, int,
                   const tpt_render_config*, const float*, double*, tpt_stats*, std::string* err) {
  *err = "adjoint not built yet";
  return TPT_ERR_UNSUPPORTED;
}

}  // namespace tpt
