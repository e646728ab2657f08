// tpt_kernels.h -- kernel-side scene view and launch declarations shared by api.hip and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tinypt_gpu.h"
#include "tpt_layout.h"

namespace tpt {

// Environment (env.rs) as the kernels see it.
struct EnvDev {
  int32_t kind;  // TPT_ENV_SKY / TPT_ENV_IMAGE
  int32_t w, h;
  int32_t black;  // image environment whose texels are all zero: it adds nothing, NEE included
  double total;   // total_weight (env.rs:32)
  // FP32 side
  const float4* __restrict__ rgb32;       // w*h texels
  const float* __restrict__ alias_prob;   // w*h
  const int32_t* __restrict__ alias_idx;  // w*h
  const float* __restrict__ cell_pdf;     // w*h: (weight/total) W H / (2 pi^2)
  // FP64 side (reference layout)
  const double* __restrict__ rgb64;    // 3*w*h
  const double* __restrict__ row_cdf;  // h+1
  const double* __restrict__ col_cdf;  // h*(w+1)
};

struct LightsDev {
  int32_t n;
  int32_t pad;
  double total;
  float total32;
  float pad2;
  const LightF32* __restrict__ l32;
  const float* __restrict__ cdf32;
  const LightF64* __restrict__ l64;
  const double* __restrict__ cdf64;
};

struct CamF32 {
  float o[3], u[3], v[3], w[3];
  float half_h, half_w, lens_radius, focus_dist;
};

struct MatsDev {
  int32_t n;
  const int32_t* __restrict__ kind;
  const float4* __restrict__ p32;   // rgb, scalar
  const double* __restrict__ p64;   // 4 per material
};

// Plain FP64 scene (reference layout).
struct SceneF64 {
  const RefNodeDev* __restrict__ nodes;
  const int32_t* __restrict__ indices;
  const double* __restrict__ tris;  // 9 per triangle, original order
  const int32_t* __restrict__ tri_mat;
  const MeshDev* __restrict__ meshes;
  const InstF64* __restrict__ inst;
  const double* __restrict__ sph;  // 4 per sphere
  const int32_t* __restrict__ sph_mat;
  int32_t n_inst, n_meshes, n_spheres, pad;
};

}  // namespace tpt
