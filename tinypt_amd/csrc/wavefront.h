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

// Queues are split into TPT_WF_SEGS segments.  A block reserves its entries in segment
// blockIdx % TPT_WF_SEGS -- with grids a multiple of TPT_WF_SEGS, the blocks of one XCD -- with
// one atomic on that segment's counter.  Reservations that all hit one counter (or one cache
// line) serialise at the memory side across the 8 XCDs and throttled the shading kernel; each
// counter therefore sits on a 256 B line of its own.  Consumers see a queue as dense positions
// 0..n-1 running through segment 0, then 1, ... (QueueMap).
#define TPT_WF_SEGS 8
#define TPT_CNT_STRIDE 64  // uint32 words between counters (256 B)
#define TPT_CNT_Q0 0
#define TPT_CNT_Q1 1
#define TPT_CNT_SH 2   // environment shadow queue
#define TPT_CNT_SHL 3  // area-light shadow queue
#define TPT_CNT_SHK 4  // slots of the area-light queue entries that the light does not shadow itself
#define TPT_CNT_IDX(q, g) (((q) * TPT_WF_SEGS + (g)) * TPT_CNT_STRIDE)
#define TPT_CNT_ACTIVE (5 * TPT_WF_SEGS * TPT_CNT_STRIDE)    // adaptive active-pixel list (one dense queue)
#define TPT_CNT_OVERFLOW (TPT_CNT_ACTIVE + TPT_CNT_STRIDE)  // set if a segment would overflow (never by construction)
#define TPT_CNT_FETCH (TPT_CNT_OVERFLOW + TPT_CNT_STRIDE)    // next ray of the refilling extension kernel
#define TPT_CNT_FETCH_SH(q) (TPT_CNT_FETCH + (1 + (q)) * TPT_CNT_STRIDE)  // next ray of shadow queue q (refilling)
#define TPT_NUM_CNT (TPT_CNT_FETCH + 3 * TPT_CNT_STRIDE)
#define TPT_STAT_STRIDE 32  // uint64 words of statistics per segment (256 B)
#define TPT_STAT_WORDS (TPT_WF_SEGS * TPT_STAT_STRIDE)
static_assert(TPT_WF_MAX_BLOCKS % TPT_WF_SEGS == 0, "grids are multiples of TPT_WF_SEGS");

// Entries per queue segment for queues of at most P entries.  The shading kernel's block b
// handles input chunks c (TPT_WF_BLOCK positions each) with c = b (mod grid), so with a grid that
// is a multiple of TPT_WF_SEGS a segment receives the outputs of at most ceil(chunks / SEGS)
// chunks.
inline uint32_t wf_seg_cap(uint32_t P) {
  const uint32_t chunks = (P + TPT_WF_BLOCK - 1) / TPT_WF_BLOCK;
  return (chunks + TPT_WF_SEGS - 1) / TPT_WF_SEGS * TPT_WF_BLOCK;
}

// Closest hit of one extension ray: t (float bits), instance code, triangle slot (12 B).
struct HitRec {
  int32_t t, inst, prim;
};

// Path state, structure of arrays, in two buffer sets: the live paths of a bounce sit in the
// segmented queue of one set; shading writes the survivors into the other set, so every kernel
// reads its inputs coalesced.  A path carries its id (the slot of its (pixel, sample) in the
// pass) and gathers radiance in L[id].  hit is indexed by dense queue position.
struct PathState {
  float4* o[2];     // origin xyz, pdf of the last BSDF sample
  float4* d[2];     // direction xyz, flags (bits 0-7 bounce, bit 8 last bounce non-delta)
  float4* beta[2];  // throughput xyz, path id (int bits)
  uint64_t* rng[2];
  HitRec* hit;      // closest hit of queue position i
  float4* L;        // radiance gathered so far, by path id
  // shadow rays, one segmented queue per kind: [0] environment, [1] area light
  float4* sh_o[2];  // environment queue: origin, tmax ([1] unused)
  float4* sh_d[2];  // direction, path id (int bits)
  float4* sh_c[2];  // contribution if unoccluded; w: 1 (environment) or the distance to the light
                    // sample (light; origin p + eps wi and tmax follow from it)
  float4* sh_l;     // light queue: shading point xyz, dense position of the path in its queue (int bits)
  uint64_t* sh_lv;  // light queue: random stream state before the light sample (its light and
                    // 53-bit uniforms are drawn again from it for the self-occlusion test)
  uint32_t* shl_keep;  // segmented queue of light-queue slots whose ray is traced (not self-occluded)
  uint32_t* counters;
  uint64_t* stats;  // per segment g: [g * TPT_STAT_STRIDE + k], k = 0 extension rays, 1 shadow rays, 2 paths,
                    // 3 area-light samples, 4 of them shadowed by their own light
  uint32_t seg_cap; // entries per queue segment (wf_seg_cap of the allocated path count)
  uint32_t pad;
};

// Dense queue position -> slot of a segmented queue.
struct QueueMap {
  uint32_t pre[TPT_WF_SEGS];  // entries before segment g
  uint32_t n;                 // entries in the queue
  uint32_t cap;
  __device__ __forceinline__ void load(const uint32_t* counters, int q, uint32_t seg_cap) {
    uint32_t s = 0;
#pragma unroll
    for (int g = 0; g < TPT_WF_SEGS; g++) {
      pre[g] = s;
      s += counters[TPT_CNT_IDX(q, g)];
    }
    n = s;
    cap = seg_cap;
  }
  __device__ __forceinline__ uint32_t slot(uint32_t i) const {
    uint32_t base = 0, off = i;
#pragma unroll
    for (int g = 1; g < TPT_WF_SEGS; g++) {
      if (i >= pre[g]) {
        base = (uint32_t)g * cap;
        off = i - pre[g];
      }
    }
    return base + off;
  }
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
  int32_t refill;        // extension waves take new rays once fewer than this many lanes trace (0: one ray per lane)
  int32_t refill_first;  // the same for the first bounce's extension (its camera rays)
  int32_t refill_sh;     // the same for the shadow kernels
  int32_t pad;
};

// The paths of one pass: path i is pixel (i mod n_pixels) of the pass's pixel set, sample
// (i / n_pixels) of the samples the pass adds to it.
struct PassArgs {
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

hipError_t wf_launch_clear_radiance(const WfParams& W, uint32_t n, hipStream_t s);
// gen: the pass's paths, for the first bounce (its kernels generate the camera rays); null after
// ev0 / ev1 (optional): events the dispatch itself records at the kernel's start and end (no
// separate marker packets between kernels)
hipError_t wf_launch_extend(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t wf_launch_shade(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s);
// cur: the queue of paths shaded this bounce; gen: the pass's paths on the first bounce (their
// rays are generated again from their ids), null after
hipError_t wf_launch_light_selfocc(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_shadow(const WfParams& W, int queue, uint32_t max_n, hipStream_t s);
hipError_t wf_launch_accumulate(const WfParams& W, const AccumArgs& A, hipStream_t s);
hipError_t wf_launch_accumulate_adaptive(const WfParams& W, const AdaptiveArgs& A, hipStream_t s);
uint32_t wf_grid_threads(uint32_t n);

// trace_closest entry (k_trace.hip)
// max_blocks > 0 caps the grid (each thread then traces several rays); gstack holds gcap
// overflow entries per thread of the launched grid
hipError_t launch_trace_closest(const SceneF32& S, int64_t n, const float* rays, float* t, int32_t* kind, int32_t* inst,
                                int32_t* prim, int32_t* gstack, int32_t gcap, hipStream_t s, uint32_t max_blocks = 0);
uint32_t trace_grid_threads(int64_t n);  // threads of the default grid for n rays
uint32_t trace_block_threads();          // threads per block

}  // namespace tpt
