// wavefront.h -- parameters of the FP32 wavefront kernels (k_wavefront.hip) and their launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tpt_kernels.h"
#include "tpt_trace32.h"

namespace tpt {

#define TPT_WF_BLOCK 256
#define TPT_WF_MAX_BLOCKS 2048
static_assert(TPT_WF_BLOCK % 64 == 0 && TPT_WF_BLOCK / 64 <= 7, "block_slot keeps per-wave counts in 7 LDS words");
#define TPT_WF_STACK 24      // LDS stack entries per lane (24 KB per 256-lane block)
#define TPT_WF_STACK_ADJ 16  // the adjoint kernel keeps 16 in LDS (plus 4 KB of gradient slots)

#define TPT_CNT_Q0 0
#define TPT_CNT_Q1 1
#define TPT_CNT_SH 2  // environment shadow queue
#define TPT_CNT_ACTIVE 3
#define TPT_CNT_SHL 4  // area-light shadow queue
#define TPT_NUM_CNT 8

// Path state, structure of arrays.  The live paths of a bounce sit densely at queue positions
// 0..n-1 of one of two buffer sets; shading writes the survivors densely into the other set, so
// every kernel reads its inputs coalesced.  A path carries its id (the slot of its (pixel,
// sample) in the pass) and gathers radiance in L[id].
struct PathState {
  float4* o[2];     // origin xyz, pdf of the last BSDF sample
  float4* d[2];     // direction xyz, flags (bits 0-7 bounce, bit 8 last bounce non-delta)
  float4* beta[2];  // throughput xyz, path id (int bits)
  uint64_t* rng[2];
  int4* hit;        // closest hit of queue position i: t bits, instance code, triangle slot
  float4* L;        // radiance gathered so far, by path id
  // shadow rays, one dense queue per kind: [0] environment, [1] area light
  float4* sh_o[2];  // origin, tmax
  float4* sh_d[2];  // direction, path id (int bits)
  float4* sh_c[2];  // contribution if unoccluded; w = 1, set to 0 when the ray is dropped
  float4* sh_l;     // light queue: shading point xyz, light-sample u
  float2* sh_lv;    // light queue: light-sample v, light index (int bits)
  uint32_t* counters;
  uint64_t* stats;  // [0] extension rays, [1] shadow rays, [2] paths
};

struct WfParams {
  SceneF32 S;
  SceneF64 S64;  // reference-layout scene, for the light self-occlusion decision
  EnvDev env;
  LightsDev lights;
  MatsDev mats;
  CamF32 cam;
  PathState ps;
  int32_t* gstack;  // per-thread overflow stacks (may be null)
  int32_t gstack_cap;
  int32_t width, height;
  int32_t max_bounces, rr_start;
  float inv_w, inv_h;
  uint64_t seed;
};

struct RaygenArgs {
  uint32_t n_paths;   // pixels * k
  uint32_t n_pixels;  // pixels in this pass
  uint32_t pixel0;    // first pixel (non-adaptive)
  uint32_t s0;        // first sample index (non-adaptive)
  const int32_t* active;  // adaptive: active pixel list (else null)
  const uint32_t* nsamp;  // adaptive: samples taken per pixel
  const int32_t* pix_map;  // image pixel of each domain pixel (null: the domain is the image)
};

struct AccumArgs {
  uint32_t n_pixels, pixel0, k;
  double* acc;
  double* acc_w;
  double* acc_sq;
};

struct AdaptiveArgs {
  uint32_t n_active, k, min_spp, max_spp;
  double threshold;
  const int32_t* active_in;
  int32_t* active_out;
  uint32_t* n_active_out;
  uint32_t* nsamp;
  double* welford;  // mean, m2 per pixel
  double* acc;
  double* acc_w;
  double* acc_sq;
};

hipError_t wf_launch_raygen(const WfParams& W, const RaygenArgs& A, hipStream_t s);
hipError_t wf_launch_extend(const WfParams& W, int cur, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_shade(const WfParams& W, int cur, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_light_selfocc(const WfParams& W, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_shadow(const WfParams& W, int queue, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_accumulate(const WfParams& W, const AccumArgs& A, hipStream_t s);
hipError_t wf_launch_accumulate_adaptive(const WfParams& W, const AdaptiveArgs& A, hipStream_t s);
uint32_t wf_grid_threads(uint32_t n);

// trace_closest entry (k_trace.hip)
hipError_t launch_trace_closest(const SceneF32& S, int64_t n, const float* rays, float* t, int32_t* kind, int32_t* inst,
                                int32_t* prim, int32_t* gstack, int32_t gcap, hipStream_t s);
uint32_t trace_grid_threads(int64_t n);

}  // namespace tpt
