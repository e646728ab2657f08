// k_trace.hip -- FP32 closest hit for a caller-supplied ray batch (tpt_trace_closest).
#include "tpt_trace32.h"
#include "wavefront.h"

namespace tpt {

#define TPT_TR_BLOCK 256

__global__ void __launch_bounds__(TPT_TR_BLOCK) k_trace_closest(SceneF32 S, int64_t n, const float4* __restrict__ rays,
                                                                 float* __restrict__ t_out, int32_t* __restrict__ kind,
                                                                 int32_t* __restrict__ inst, int32_t* __restrict__ prim,
                                                                 int32_t* gstack, int32_t gcap) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_TR_BLOCK];
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{s_stack + threadIdx.x, gstack ? gstack + gid * gcap : nullptr, TPT_TR_BLOCK, TPT_WF_STACK};
  for (int64_t i = gid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float4 a = rays[2 * i], b = rays[2 * i + 1];  // o.xyz, d.x | d.yz, tmin, tmax
    HitF32 h;
    bool found = trace_f32<false>(S, mk(a.x, a.y, a.z), mk(a.w, b.x, b.y), b.z, b.w, &h, st);
    if (!found) {
      t_out[i] = __int_as_float(0x7f800000);
      kind[i] = TPT_HIT_NONE;
      inst[i] = -1;
      prim[i] = -1;
    } else if ((uint32_t)h.inst & TPT_ITEM_SPHERE) {
      t_out[i] = h.t;
      kind[i] = TPT_HIT_SPHERE;
      inst[i] = (int32_t)((uint32_t)h.inst & ~TPT_ITEM_SPHERE);
      prim[i] = -1;
    } else {
      t_out[i] = h.t;
      kind[i] = TPT_HIT_TRI;
      inst[i] = (int32_t)((uint32_t)h.inst & ~TPT_HIT_FLAT);
      prim[i] = S.tri_orig[h.prim];
    }
  }
}

static inline uint32_t tr_grid(int64_t n) {
  int64_t b = (n + TPT_TR_BLOCK - 1) / TPT_TR_BLOCK;
  return (uint32_t)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}
uint32_t trace_grid_threads(int64_t n) { return tr_grid(n) * TPT_TR_BLOCK; }
uint32_t trace_block_threads() { return TPT_TR_BLOCK; }

hipError_t launch_trace_closest(const SceneF32& S, int64_t n, const float* rays, float* t, int32_t* kind, int32_t* inst,
                                int32_t* prim, int32_t* gstack, int32_t gcap, hipStream_t s, uint32_t max_blocks) {
  uint32_t blocks = tr_grid(n);
  if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks;  // the kernel loops over its rays
  hipLaunchKernelGGL(k_trace_closest, dim3(blocks), dim3(TPT_TR_BLOCK), 0, s, S, n, (const float4*)rays, t, kind, inst,
                     prim, gstack, gcap);
  return hipGetLastError();
}

}  // namespace tpt
