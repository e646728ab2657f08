// plain64.h -- parameters of the plain FP64 megakernel (k_plain64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tpt_kernels.h"

namespace tpt {

#define TPT_PLAIN_BLOCK 64

struct Plain64Params {
  SceneF64 S;
  EnvDev env;
  LightsDev lights;
  MatsDev mats;
  CameraF64 cam;
  int32_t width, height, spp, max_bounces, rr_start, adaptive, min_spp;
  int32_t n_pix;            // pixels of the domain; domain pixel q is image pixel pix_map[q] (q if null)
  const int32_t* pix_map;
  int32_t s_begin, s_end;  // samples of this launch
  double threshold;
  double inv_w, inv_h;
  uint64_t seed;
  // per-pixel state, persistent across launches
  uint64_t* rng;
  double* acc;      // 3 per pixel
  double* acc_sq;   // 3 per pixel (optional)
  double* acc_w;
  uint32_t* nsamp;
  double* welford;  // mean, m2
  uint8_t* done;
  uint64_t* stats;  // [0] extension rays, [1] shadow rays, [2] paths
};

hipError_t launch_plain64(const Plain64Params& P, hipStream_t s);

}  // namespace tpt
