// k_wavefront.hip -- FP32 wavefront path tracer (radiance(), integrator.rs:53-135, split into
// kernels over SoA path queues).
//
// One pass renders `k` samples of a set of pixels: P = pixels * k paths, slot i -> pixel
// (i mod pixels), sample (s0 + i / pixels), so neighbouring slots are neighbouring pixels.
// Per bounce:  extend (closest hit) -> shade (miss / RR / emission / NEE sampling / BSDF
// sampling; survivors compacted into the next queue) -> light self-occlusion -> shadow (any hit
// for the <= 2 NEE rays of each path, added to that path's radiance by the thread that owns the
// path).  The first bounce's extend and shade generate the camera rays themselves.
// After the last bounce, accumulate adds each pixel's k path radiances into FP64 sums.
// Queue compaction: a ballot per wave and one atomic per block on the block's queue segment.
// Random numbers: one PCG32 stream per (pixel, sample): seed_for(x, y, s) ^ seed.
#include "tpt_bsdf.h"
#include "tpt_kernels.h"
#include "tpt_sample32.h"
#include "tpt_selfocc64.h"
#include "tpt_trace32.h"
#include "wavefront.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <mutex>
#include <unordered_map>

namespace tpt {

// ---------------------------------------------------------------- queue compaction
// Position of this lane's entry in a queue when `pred` holds: a ballot per wave, the waves'
// counts combined in LDS, and one atomic per block, lanes ordered by block thread id.  Every
// thread of the block must call it (it synchronises the block); `sh` is 8 words of LDS.
__device__ __forceinline__ uint32_t block_slot(bool pred, uint32_t* __restrict__ count, uint32_t* sh) {
  const uint64_t mask = __ballot(pred);
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < TPT_WF_BLOCK / 64; w++) tot += sh[w];
    sh[7] = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = sh[7];
  for (uint32_t w = 0; w < wave; w++) base += sh[w];
  const uint32_t pos = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  __syncthreads();  // sh is reused by the next call
  return pos;
}

// Slot of this lane's entry in queue q (segment blockIdx % TPT_WF_SEGS) when `pred` holds, or
// UINT32_MAX (and the overflow flag raised) if the segment were full, which the segment size
// rules out.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t seg_slot(bool pred, const PathState& ps, int q, uint32_t* sh) {
  const uint32_t g = blockIdx.x & (TPT_WF_SEGS - 1);
  const uint32_t pos = block_slot(pred, ps.counters + TPT_CNT_IDX(q, g), sh);
  if (pred && pos >= ps.seg_cap) {
    ps.counters[TPT_CNT_OVERFLOW] = 1u;
    return UINT32_MAX;
  }
  return g * ps.seg_cap + pos;
}

// ---------------------------------------------------------------- kernels
__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Path i of a pass (render.rs:142-157): its pixel and sample, its random stream
// seed_for(x, y, s) ^ seed, and its camera ray.  The first bounce's extension and shading
// kernels call it for their own paths, so no path state is written for bounce 0.
__device__ __forceinline__ void gen_path(const WfParams& W, const PassArgs& A, uint32_t i, RngF32* rng, f3* o, f3* d) {
  uint32_t pix, s;
  if (A.active) {
    const uint32_t j = i % A.n_pixels;
    pix = (uint32_t)A.active[j];
    s = A.nsamp[pix] + i / A.n_pixels;
  } else {
    pix = A.pixel0 + i % A.n_pixels;
    s = A.s0 + i / A.n_pixels;
  }
  if (A.pix_map) pix = (uint32_t)A.pix_map[pix];
  const uint32_t x = pix % (uint32_t)W.width, y = pix / (uint32_t)W.width;
  *rng = RngF32(seed_for(x, y, s) ^ W.seed);
  camera_ray32(W.cam, x, y, W.inv_w, W.inv_h, *rng, o, d);
}

// A pass without bounces (max_bounces 0) gathers no radiance: clear its paths' sums.
__global__ void __launch_bounds__(TPT_WF_BLOCK) k_clear_radiance(WfParams W, uint32_t n) {
  for (uint32_t i = gtid(); i < n; i += gridDim.x * blockDim.x) W.ps.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// OVF: the scene's stack bound exceeds the LDS part, so the traversal stack may overflow to the
// per-thread global area (W.gstack); otherwise that path is compiled out.  GEN: the first
// bounce, whose paths are the pass's camera rays, generated here (path i at position i).  Six
// waves per SIMD fit the later bounces in 80 VGPRs without spills, five the first (its camera
// code needs more); the traversal is VALU-bound, and more waves hide its load latency.
template <bool OVF, bool GEN>
__global__ void __launch_bounds__(TPT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(GEN ? 5 : 6)))
k_extend(WfParams W, int cur, PassArgs A) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  QueueMap qm;
  uint32_t n;
  if constexpr (GEN) {
    n = A.n_paths;
  } else {
    qm.load(W.ps.counters, cur ? TPT_CNT_Q1 : TPT_CNT_Q0, W.ps.seg_cap);
    n = qm.n;
  }
  const float4* __restrict__ po = W.ps.o[cur];
  const float4* __restrict__ pd = W.ps.d[cur];
  LaneStackT<OVF> st{s_stack + threadIdx.x, OVF ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
                     TPT_WF_STACK};
  if (gtid() < TPT_WF_SEGS) {  // clear the queues the shading kernel fills next
    W.ps.counters[TPT_CNT_IDX(cur ? TPT_CNT_Q0 : TPT_CNT_Q1, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SH, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SHL, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SHK, gtid())] = 0;
  }
  if (gtid() < 2) W.ps.counters[TPT_CNT_FETCH_SH(gtid())] = 0;  // cursors of this bounce's shadow kernels
  if (gtid() == 0) atomicAdd((unsigned long long*)&W.ps.stats[0], (unsigned long long)n);
  for (uint32_t i = gtid(); i < n; i += gridDim.x * blockDim.x) {
    f3 o, d;
    if constexpr (GEN) {
      RngF32 rng(1);
      gen_path(W, A, i, &rng, &o, &d);
    } else {
      const uint32_t slot = qm.slot(i);
      o = f3of(po[slot]);
      d = f3of(pd[slot]);
    }
    HitF32 h;
    trace_f32<false>(W.S, o, d, (float)TPT_RAY_EPSILON, 1e30f, &h, st);
    W.ps.hit[i] = HitRec{__float_as_int(h.t), h.inst, h.prim};
  }
}

// Persistent form of k_extend for deep trees, where the rays of one wave need very different
// numbers of traversal steps: a wave keeps its lanes' traversals resumable (TraceF32 step by
// step) and, once fewer than T lanes are still tracing (W.refill, or W.refill_first on the first
// bounce), hands its idle lanes the next rays of the queue (one atomic per refill on a global
// cursor; dynamic ray fetch, Aila and Laine 2009).  The hits are those of k_extend: each ray's traversal is the same sequence of steps.
// The cursor (TPT_CNT_FETCH) is zero at launch: the render clears it, and every k_shade launch
// clears it for the next extension launch.  Every wave exits once the cursor has passed the
// queue's end and its lanes are idle.  With the overflow path the resumable state needs more
// registers: five waves per SIMD (96 VGPRs) for the later bounces and four (117) for the first
// keep it out of scratch, where six and five spilled 44 and 64 bytes per lane that the loop
// reloaded at every step (measured: the 1.31M-triangle mesh 4% and the 4096 instances 11%
// faster).
template <bool OVF, bool GEN>
__global__ void __launch_bounds__(TPT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(GEN && OVF ? 4 : GEN || OVF ? 5 : 6)))
k_extend_refill(WfParams W, int cur, PassArgs A) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  QueueMap qm;
  uint32_t n;
  if constexpr (GEN) {
    n = A.n_paths;
  } else {
    qm.load(W.ps.counters, cur ? TPT_CNT_Q1 : TPT_CNT_Q0, W.ps.seg_cap);
    n = qm.n;
  }
  const float4* __restrict__ po = W.ps.o[cur];
  const float4* __restrict__ pd = W.ps.d[cur];
  LaneStackT<OVF> st{s_stack + threadIdx.x, OVF ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
                     TPT_WF_STACK};
  if (gtid() < TPT_WF_SEGS) {  // clear the queues the shading kernel fills next
    W.ps.counters[TPT_CNT_IDX(cur ? TPT_CNT_Q0 : TPT_CNT_Q1, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SH, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SHL, gtid())] = 0;
    W.ps.counters[TPT_CNT_IDX(TPT_CNT_SHK, gtid())] = 0;
  }
  if (gtid() < 2) W.ps.counters[TPT_CNT_FETCH_SH(gtid())] = 0;  // cursors of this bounce's shadow kernels
  if (gtid() == 0) atomicAdd((unsigned long long*)&W.ps.stats[0], (unsigned long long)n);
  uint32_t* cursor = W.ps.counters + TPT_CNT_FETCH;
  const uint32_t lane = __lane_id();
  TraceF32<false> tr;
  bool live = false, exhausted = false;
  uint32_t my = 0;
  // nl < threshold <= 64: every refill moves the cursor
  const int threshold = min(GEN ? W.refill_first : W.refill, 64);
  for (;;) {
    const uint64_t mask = __ballot(live);
    const int nl = __popcll(mask);
    if (!exhausted && nl < threshold) {  // wave-uniform: refill the idle lanes
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(cursor, (uint32_t)(64 - nl));
      base = (uint32_t)__shfl((int)base, 0);
      if (base >= n) exhausted = true;
      if (!live) {
        my = base + (uint32_t)__popcll(~mask & ((1ull << lane) - 1ull));
        if (my < n) {
          f3 o, d;
          if constexpr (GEN) {
            RngF32 rng(1);
            gen_path(W, A, my, &rng, &o, &d);
          } else {
            const uint32_t slot = qm.slot(my);
            o = f3of(po[slot]);
            d = f3of(pd[slot]);
          }
          live = tr.init(W.S, o, d, (float)TPT_RAY_EPSILON, 1e30f);
          if (!live) W.ps.hit[my] = HitRec{__float_as_int(tr.hit.t), tr.hit.inst, tr.hit.prim};
        }
      }
      continue;
    }
    if (nl == 0) break;  // the queue is taken and every lane is idle
    if (live && !tr.step(W.S, st)) {
      W.ps.hit[my] = HitRec{__float_as_int(tr.hit.t), tr.hit.inst, tr.hit.prim};
      live = false;
    }
  }
}

__device__ __forceinline__ void add_radiance(float4* L, int32_t id, f3 c) {
  float4 l = L[id];
  L[id] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
}

// ENV_NEE: the environment is an image with a non-zero texel, so environment NEE runs (the
// variant without it needs fewer registers).  GEN: the first bounce; path i (position i) is
// regenerated from its pixel and sample, as k_extend traced it, and its radiance sum L[i] is
// written rather than added to.
template <bool ENV_NEE, bool GEN>
__global__ void __launch_bounds__(TPT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_shade(WfParams W, int cur, PassArgs A) {
  __shared__ uint32_t s_slot[8];
  QueueMap qm;
  uint32_t n;
  if constexpr (GEN) {
    n = A.n_paths;
  } else {
    qm.load(W.ps.counters, cur ? TPT_CNT_Q1 : TPT_CNT_Q0, W.ps.seg_cap);
    n = qm.n;
  }
  const int nxt = cur ^ 1;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (gtid() == 0) W.ps.counters[TPT_CNT_FETCH] = 0;  // for the next extension launch (k_extend_refill)
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    bool alive = false, has_env = false, has_light = false;
    // path state (valid when i < n)
    f3 o = mk(0.f, 0.f, 0.f), d = o, beta = o;
    float last_pdf = 0.f;
    uint32_t flags = 0;
    int32_t id = 0;
    RngF32 rng(1);
    int4 h4 = make_int4(0, -1, 0, 0);
    if (i < n) {
      if constexpr (GEN) {
        gen_path(W, A, i, &rng, &o, &d);
        beta = mk(1.f, 1.f, 1.f);
        id = (int32_t)i;
      } else {
        const uint32_t slot = qm.slot(i);
        float4 o4 = W.ps.o[cur][slot], d4 = W.ps.d[cur][slot], b4 = W.ps.beta[cur][slot];
        o = f3of(o4);
        d = f3of(d4);
        beta = f3of(b4);
        id = __float_as_int(b4.w);
        last_pdf = o4.w;
        flags = __float_as_uint(d4.w);
        rng.p.state = W.ps.rng[cur][slot];
      }
      const HitRec hr = W.ps.hit[i];
      h4 = make_int4(hr.t, hr.inst, hr.prim, 0);
    }
    // radiance of this bounce: added to L[id], or (first bounce) gathered and written once
    f3 l0 = mk(0.f, 0.f, 0.f);
    auto gather = [&](f3 c) {
      if constexpr (GEN) l0 = l0 + c;
      else add_radiance(W.ps.L, id, c);
    };
    const int bounce = (int)(flags & 0xffu);
    const bool last_nd = (flags & 0x100u) != 0;
    // ---- phase 1: miss, Russian roulette, emission, next-event estimation
    float4 sh_c0 = make_float4(0.f, 0.f, 0.f, 0.f), sh_c1 = sh_c0;
    float4 sh_o0 = sh_c0, sh_d0 = sh_c0, sh_d1 = sh_c0, sh_l = sh_c0;  // 0: environment, 1: light
    uint64_t sh_lv = 0;
    bool go = false;  // hit a scattering surface: sample the BSDF in phase 2
    f3 p = mk(0.f, 0.f, 0.f), ng = p;
    int mat = 0, kind = 0;
    float4 mp = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
      if (h4.y == -1) {
        // miss (integrator.rs:71-83); a black map adds nothing, so its paths just end
        if (!W.env.black) {
          f3 bg = W.env.kind == TPT_ENV_IMAGE ? env_lookup32(W.env, d) : sky32(d);
          f3 c = had(beta, bg);
          if (W.env.kind == TPT_ENV_IMAGE && last_nd) c = c * mis_weight(last_pdf, env_pdf32(W.env, d));
          gather(clamp_luminance(c, (float)TPT_FIREFLY_CLAMP));
        }
      } else {
        go = true;
        if (bounce >= W.rr_start) {  // integrator.rs:89-95
          float q = fmaxf(fminf(fmaxf(fmaxf(beta.x, beta.y), beta.z), 0.95f), 0.05f);
          if (rng.uni() > q) go = false;
          else beta = beta / q;
        }
        if (go) {
          const float t = __int_as_float(h4.x);
          hit_geometry32(W.S, h4.y, h4.z, t, o, d, &p, &ng, &mat);
          kind = W.mats.kind[mat];
          mp = W.mats.p32[mat];
          const f3 mc = f3of(mp);
          if (kind == TPT_MAT_DIFFUSE_LIGHT) {  // integrator.rs:100-104
            gather(clamp_luminance(had(beta, mc), (float)TPT_FIREFLY_CLAMP));
            go = false;
          } else if (!mat_is_delta(kind)) {
            f3 n_o = dot(ng, d) < 0.f ? ng : -ng;  // oriented_normal (integrator.rs:138-140)
            f3 wo = vnorm(-d);
            // nee_environment (integrator.rs:148-172); a black map contributes nothing, so the
            // FP32 form skips it (its own random streams need not follow the reference's draws)
            if (ENV_NEE) {
              f3 wi, li;
              float pdf_env;
              env_sample32(W.env, rng, &wi, &li, &pdf_env);
              float cs = fmaxf(dot(n_o, wi), 0.f);
              if (cs > 0.f && pdf_env > 0.f && (li.x != 0.f || li.y != 0.f || li.z != 0.f)) {
                f3 f;
                float pdf_b;
                bsdf_eval(kind, mc, mp.w, wo, wi, n_o, &f, &pdf_b);
                f3 c = had(had(beta, f), li) * (cs * mis_weight(pdf_env, pdf_b) / pdf_env);
                if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                  f3 so = p + (float)TPT_RAY_EPSILON * wi;
                  sh_o0 = make_float4(so.x, so.y, so.z, 1e30f);
                  sh_d0 = make_float4(wi.x, wi.y, wi.z, __int_as_float(id));
                  sh_c0 = make_float4(c.x, c.y, c.z, 1.f);
                  has_env = true;
                }
              }
            }
            LightSample32 ls;  // nee_area_light (integrator.rs:178-213)
            const uint64_t light_state = rng.p.state;
            if (sample_light32(W.lights, rng, p, &ls)) {
              f3 to_light = ls.pos - p;
              float dist = sqrtf(dot(to_light, to_light));
              if (dist > 1e-6f) {
                f3 wi = to_light / dist;
                float cs = fmaxf(dot(n_o, wi), 0.f);
                if (cs > 0.f) {
                  f3 f;
                  float pdf_b;
                  bsdf_eval(kind, mc, mp.w, wo, wi, n_o, &f, &pdf_b);
                  f3 c = had(had(beta, f), ls.emit) * (cs * mis_weight(ls.pdf, pdf_b) / ls.pdf);
                  if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                    // the shadow kernel rebuilds origin p + eps wi and tmax from p and dist
                    sh_d1 = make_float4(wi.x, wi.y, wi.z, __int_as_float(id));
                    sh_c1 = make_float4(c.x, c.y, c.z, dist);
                    sh_l = make_float4(p.x, p.y, p.z, __int_as_float((int32_t)i));  // also for k_light_selfocc
                    sh_lv = light_state;
                    has_light = true;
                  }
                }
              }
            }
          }
        }
      }
    }
    if (GEN && i < n) W.ps.L[id] = make_float4(l0.x, l0.y, l0.z, 0.f);
    const uint32_t je = ENV_NEE ? seg_slot(has_env, W.ps, TPT_CNT_SH, s_slot) : UINT32_MAX;
    if (ENV_NEE && has_env && je != UINT32_MAX) {
      W.ps.sh_o[0][je] = sh_o0;
      W.ps.sh_d[0][je] = sh_d0;
      W.ps.sh_c[0][je] = sh_c0;
    }
    const uint32_t jl = seg_slot(has_light, W.ps, TPT_CNT_SHL, s_slot);
    if (has_light && jl != UINT32_MAX) {
      W.ps.sh_d[1][jl] = sh_d1;
      W.ps.sh_c[1][jl] = sh_c1;
      W.ps.sh_l[jl] = sh_l;
      W.ps.sh_lv[jl] = sh_lv;
    }
    // ---- phase 2: BSDF sampling; survivors go densely into the next buffer set
    bool bsdf_nd = false;
    if (go && bounce + 1 < W.max_bounces) {  // the scattered ray of the last bounce is never traced
      BsdfSample<float> bs;
      if (bsdf_sample(kind, f3of(mp), mp.w, d, p, ng, __int_as_float(h4.x), rng, &bs)) {
        beta = had(beta, bs.weight);
        o = bs.o;
        d = bs.d;
        last_pdf = bs.pdf;
        bsdf_nd = !bs.is_delta;
        alive = true;
      }
    }
    const uint32_t k = seg_slot(alive, W.ps, cur ? TPT_CNT_Q0 : TPT_CNT_Q1, s_slot);
    if (alive && k != UINT32_MAX) {
      uint32_t nflags = (uint32_t)(bounce + 1) | (bsdf_nd ? 0x100u : 0u);
      W.ps.o[nxt][k] = make_float4(o.x, o.y, o.z, last_pdf);
      W.ps.d[nxt][k] = make_float4(d.x, d.y, d.z, __uint_as_float(nflags));
      W.ps.beta[nxt][k] = make_float4(beta.x, beta.y, beta.z, __int_as_float(id));
      W.ps.rng[nxt][k] = rng.p.state;
    }
  }
}

// The reference's light shadow ray can be blocked by the light itself; that decision is taken
// in FP64 (tpt_selfocc64.h) here, in a kernel of its own so the FP64 registers do not weigh on
// k_shade.  Its inputs are rebuilt with the reference's precision: the shading point by the
// FP64 World::hit of the path's FP32 ray on the primitive it hit (the ray from the path queue,
// or, on the first bounce, the camera ray generated again), the light and its 53-bit uniforms
// drawn again from the stream state saved before the light sample.  The slots of the rays that
// remain go densely into the SHK queue, so k_shadow's waves carry no lanes for dropped rays
// (about 30% on a rectangle light).
template <bool GEN>
__global__ void __launch_bounds__(TPT_WF_BLOCK) k_light_selfocc(WfParams W, int cur, PassArgs A) {
  __shared__ uint32_t s_slot[8];
  QueueMap qm, qp;
  qm.load(W.ps.counters, TPT_CNT_SHL, W.ps.seg_cap);
  if constexpr (!GEN) qp.load(W.ps.counters, cur ? TPT_CNT_Q1 : TPT_CNT_Q0, W.ps.seg_cap);
  const uint32_t n = qm.n;
  uint32_t dropped = 0;  // self-occluded samples this thread saw
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    uint32_t slot = 0;
    bool keep = false;
    if (j < n) {
      slot = qm.slot(j);
      const float4 l = W.ps.sh_l[slot];
      // the light and its uniforms first: their loads overlap the hit point's chain below
      RngF32 rng(1);
      rng.p.state = W.ps.sh_lv[slot];
      const int idx = light_pick32(W.lights, rng);
      const double u = rng.uni53(), v = rng.uni53();
      const uint32_t pos = (uint32_t)__float_as_int(l.w);  // the path's dense position this bounce
      f3 o, d;
      if constexpr (GEN) {
        RngF32 cam(1);
        gen_path(W, A, pos, &cam, &o, &d);
      } else {
        const uint32_t ps = qp.slot(pos);
        o = f3of(W.ps.o[cur][ps]);
        d = f3of(W.ps.d[cur][ps]);
      }
      const HitRec hr = W.ps.hit[pos];
      const D3 p64 = hit_point64(W.S64, W.S.tri_orig, hr.inst, hr.prim, o, d, f3of(l));
      keep = !light_self_occluded64(W.S64, W.lights, idx, p64, u, v);
    }
    const uint32_t k = seg_slot(keep, W.ps, TPT_CNT_SHK, s_slot);  // every thread of the block
    if (keep && k != UINT32_MAX) W.ps.shl_keep[k] = slot;
    if (j < n && !keep) dropped++;
  }
  for (int off = 32; off > 0; off >>= 1) dropped += (uint32_t)__shfl_down((int)dropped, off);
  uint64_t* st = W.ps.stats + (blockIdx.x & (TPT_WF_SEGS - 1)) * TPT_STAT_STRIDE;
  if (__lane_id() == 0 && dropped) atomicAdd((unsigned long long*)&st[4], (unsigned long long)dropped);
  if (gtid() == 0 && n) atomicAdd((unsigned long long*)&W.ps.stats[3], (unsigned long long)n);
}

// Any hit for one shadow queue (0: environment, 1: area light); an unoccluded ray adds its
// contribution to its path's radiance.  A path has at most one ray per queue, so the
// read-modify-write of L[id] has a single writer per launch.
template <bool OVF>
__global__ void __launch_bounds__(TPT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) k_shadow(WfParams W, int queue) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  QueueMap qm;
  qm.load(W.ps.counters, queue ? TPT_CNT_SHK : TPT_CNT_SH, W.ps.seg_cap);
  const float4* __restrict__ qo = queue ? W.ps.sh_l : W.ps.sh_o[0];
  const float4* __restrict__ qd = W.ps.sh_d[queue];
  const float4* __restrict__ qc = W.ps.sh_c[queue];
  LaneStackT<OVF> st{s_stack + threadIdx.x, OVF ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
                     TPT_WF_STACK};
  uint32_t traced = 0;
  for (uint32_t j = gtid(); j < qm.n; j += gridDim.x * blockDim.x) {
    const uint32_t slot = queue ? W.ps.shl_keep[qm.slot(j)] : qm.slot(j);
    const float4 o4 = qo[slot], d4 = qd[slot];
    f3 o = f3of(o4);
    float tmax = o4.w;
    if (queue) {  // light queue: o4 is the shading point, sh_c's w the distance to the light sample
      o = o + (float)TPT_RAY_EPSILON * f3of(d4);
      tmax = light_shadow_tmax32(qc[slot].w);
    }
    HitF32 h;
    traced++;
    // the contribution is read only for an unoccluded ray, after the traversal
    if (!trace_f32<true>(W.S, o, f3of(d4), (float)TPT_RAY_EPSILON, tmax, &h, st))
      add_radiance(W.ps.L, __float_as_int(d4.w), f3of(qc[slot]));
  }
  // one atomic per wave for the shadow-ray count
  uint32_t tot = traced;
  for (int off = 32; off > 0; off >>= 1) tot += (uint32_t)__shfl_down((int)tot, off);
  if (__lane_id() == 0 && tot)
    atomicAdd((unsigned long long*)&W.ps.stats[(blockIdx.x & (TPT_WF_SEGS - 1)) * TPT_STAT_STRIDE + 1],
              (unsigned long long)tot);
}

// Non-adaptive: pixel p accumulates the k paths of this pass (slots p + j * pixels).
__global__ void __launch_bounds__(TPT_WF_BLOCK) k_accumulate(WfParams W, AccumArgs A) {
  for (uint32_t p = gtid(); p < A.n_pixels; p += gridDim.x * blockDim.x) {
    double r = 0, g = 0, b = 0, r2 = 0, g2 = 0, b2 = 0;
    for (uint32_t j = 0; j < A.k; j++) {
      float4 l = W.ps.L[p + j * A.n_pixels];
      double x = l.x, y = l.y, z = l.z;
      r += x; g += y; b += z;
      r2 += x * x; g2 += y * y; b2 += z * z;
    }
    uint32_t g_pix = A.pixel0 + p;
    A.acc[3 * g_pix] += r; A.acc[3 * g_pix + 1] += g; A.acc[3 * g_pix + 2] += b;
    A.acc_w[g_pix] += (double)A.k;
    if (A.acc_sq) { A.acc_sq[3 * g_pix] += r2; A.acc_sq[3 * g_pix + 1] += g2; A.acc_sq[3 * g_pix + 2] += b2; }
  }
  if (gtid() == 0) atomicAdd((unsigned long long*)&W.ps.stats[2], (unsigned long long)A.n_pixels * A.k);
}

// Adaptive (render.rs:143-179): for each active pixel take its k new samples in sample order,
// Welford on luminance, stop at the reference's criterion; samples past the stop are dropped.
// Survivors are compacted into `active_out`.
__global__ void __launch_bounds__(TPT_WF_BLOCK) k_accumulate_adaptive(WfParams W, AdaptiveArgs A) {
  __shared__ uint32_t s_slot[8];
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t kept = 0;  // samples this thread added to its pixels (the paths the render counts)
  for (uint32_t base = blockIdx.x * blockDim.x; base < A.n_active; base += stride) {
    uint32_t j = base + threadIdx.x;
    bool keep = false;
    int32_t pix = 0;
    if (j < A.n_active) {
      pix = A.active_in[j];
      uint32_t n = A.nsamp[pix];
      const uint32_t n0 = n;
      double mean = A.welford[2 * pix], m2 = A.welford[2 * pix + 1];
      double r = 0, g = 0, b = 0, r2 = 0, g2 = 0, b2 = 0;
      bool stop = false;
      for (uint32_t k = 0; k < A.k && !stop; k++) {
        float4 l = W.ps.L[j + k * A.n_active];
        double x = l.x, y = l.y, z = l.z;
        r += x; g += y; b += z;
        r2 += x * x; g2 += y * y; b2 += z * z;
        n += 1;
        double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
        double delta = lum - mean;
        mean += delta / (double)n;
        double delta2 = lum - mean;
        m2 += delta * delta2;
        if (n >= A.min_spp) {
          double var = n > 1 ? m2 / (double)(n - 1) : __longlong_as_double(0x7ff0000000000000ll);
          double rel = sqrt(var) / fmax(fabs(mean), 1e-4);
          if (rel < A.threshold) stop = true;
        }
        if (n >= A.max_spp) stop = true;
      }
      A.acc[3 * pix] += r; A.acc[3 * pix + 1] += g; A.acc[3 * pix + 2] += b;
      if (A.acc_sq) { A.acc_sq[3 * pix] += r2; A.acc_sq[3 * pix + 1] += g2; A.acc_sq[3 * pix + 2] += b2; }
      A.acc_w[pix] = (double)n;
      A.nsamp[pix] = n;
      kept += n - n0;
      A.welford[2 * pix] = mean;
      A.welford[2 * pix + 1] = m2;
      keep = !stop;
    }
    const uint32_t q = block_slot(keep, A.n_active_out, s_slot);
    if (keep) A.active_out[q] = pix;
  }
  for (int off = 32; off > 0; off >>= 1) kept += (uint32_t)__shfl_down((int)kept, off);
  if (__lane_id() == 0 && kept)
    atomicAdd((unsigned long long*)&W.ps.stats[(blockIdx.x & (TPT_WF_SEGS - 1)) * TPT_STAT_STRIDE + 2],
              (unsigned long long)kept);
}

// Persistent form of k_shadow for deep trees (see k_extend_refill): lanes whose shadow ray is
// done take the queue's next rays once fewer than W.refill_sh lanes are still tracing.  The
// cursor TPT_CNT_FETCH_SH(queue) is cleared by the extension kernel of the same bounce.
template <bool OVF>
__global__ void __launch_bounds__(TPT_WF_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_shadow_refill(WfParams W, int queue) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  QueueMap qm;
  qm.load(W.ps.counters, queue ? TPT_CNT_SHK : TPT_CNT_SH, W.ps.seg_cap);
  const uint32_t n = qm.n;
  const float4* __restrict__ qo = queue ? W.ps.sh_l : W.ps.sh_o[0];
  const float4* __restrict__ qd = W.ps.sh_d[queue];
  const float4* __restrict__ qc = W.ps.sh_c[queue];
  LaneStackT<OVF> st{s_stack + threadIdx.x, OVF ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
                     TPT_WF_STACK};
  uint32_t* cursor = W.ps.counters + TPT_CNT_FETCH_SH(queue);
  const uint32_t lane = __lane_id();
  TraceF32<true> tr;
  bool live = false, exhausted = false;
  uint32_t slot = 0, traced = 0;
  int32_t id = 0;
  const int threshold = min(W.refill_sh, 64);  // nl < threshold <= 64: every refill moves the cursor
  for (;;) {
    const uint64_t mask = __ballot(live);
    const int nl = __popcll(mask);
    if (!exhausted && nl < threshold) {  // wave-uniform: refill the idle lanes
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(cursor, (uint32_t)(64 - nl));
      base = (uint32_t)__shfl((int)base, 0);
      if (base >= n) exhausted = true;
      if (!live) {
        const uint32_t j = base + (uint32_t)__popcll(~mask & ((1ull << lane) - 1ull));
        if (j < n) {
          slot = queue ? W.ps.shl_keep[qm.slot(j)] : qm.slot(j);
          const float4 o4 = qo[slot], d4 = qd[slot];
          f3 o = f3of(o4);
          float tmax = o4.w;
          if (queue) {  // light queue: o4 is the shading point, sh_c's w the distance to the light sample
            o = o + (float)TPT_RAY_EPSILON * f3of(d4);
            tmax = light_shadow_tmax32(qc[slot].w);
          }
          id = __float_as_int(d4.w);
          traced++;
          live = tr.init(W.S, o, f3of(d4), (float)TPT_RAY_EPSILON, tmax);
          if (!live && !tr.found) add_radiance(W.ps.L, id, f3of(qc[slot]));
        }
      }
      continue;
    }
    if (nl == 0) break;  // the queue is taken and every lane is idle
    if (live && !tr.step(W.S, st)) {
      if (!tr.found) add_radiance(W.ps.L, id, f3of(qc[slot]));  // unoccluded
      live = false;
    }
  }
  uint32_t tot = traced;
  for (int off = 32; off > 0; off >>= 1) tot += (uint32_t)__shfl_down((int)tot, off);
  if (__lane_id() == 0 && tot)
    atomicAdd((unsigned long long*)&W.ps.stats[(blockIdx.x & (TPT_WF_SEGS - 1)) * TPT_STAT_STRIDE + 1],
              (unsigned long long)tot);
}

// ---------------------------------------------------------------- host-side launchers
// Grid of at most TPT_WF_MAX_BLOCKS blocks: the kernels loop over their queues, so the grid only
// needs to fill the chip (256 CUs x 8 blocks); larger grids cost dispatch time per bounce,
// most of all for launches whose queues have shrunk to a few paths.
// The grid is a multiple of TPT_WF_SEGS (queue segments follow blockIdx % TPT_WF_SEGS).
static inline uint32_t grid_for(uint32_t n) {
  uint32_t b = (n + TPT_WF_BLOCK - 1) / TPT_WF_BLOCK;
  b = (b + TPT_WF_SEGS - 1) / TPT_WF_SEGS * TPT_WF_SEGS;
  return b < TPT_WF_SEGS ? TPT_WF_SEGS : (b > TPT_WF_MAX_BLOCKS ? TPT_WF_MAX_BLOCKS : b);
}

// Blocks of `kernel` that are resident on the whole device at once (CUs x blocks per CU at the
// kernel's register and LDS use), rounded down to a multiple of TPT_WF_SEGS.  A grid-stride
// kernel launched with more blocks than this runs its last blocks after the first ones finish,
// with the device partly idle (5 resident blocks per CU against 8 launched leaves a 60 %-full
// second round); capping the grid at the resident count gives every block the same share.
static uint32_t resident_blocks(const void* kernel) {
  static std::mutex mu;
  static std::unordered_map<const void*, uint32_t> cache;  // one device model per process
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int dev = 0, cus = 0, per_cu = 0;
  uint32_t r = TPT_WF_MAX_BLOCKS;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TPT_WF_BLOCK, 0) == hipSuccess && cus > 0 &&
      per_cu > 0) {
    r = (uint32_t)cus * (uint32_t)per_cu / TPT_WF_SEGS * TPT_WF_SEGS;
    r = std::max<uint32_t>(TPT_WF_SEGS, std::min<uint32_t>(r, TPT_WF_MAX_BLOCKS));
  }
  cache[kernel] = r;
  return r;
}

static inline dim3 grid_of(uint32_t n, const void* kernel) { return dim3(std::min(grid_for(n), resident_blocks(kernel))); }

hipError_t wf_launch_clear_radiance(const WfParams& W, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_radiance, dim3(grid_for(n)), dim3(TPT_WF_BLOCK), 0, s, W, n);
  return hipGetLastError();
}
// A launch, timed by events that the dispatch records itself when ev0 is given.
template <class K>
static void launch_timed(K kernel, dim3 grid, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, const WfParams& W, int cur,
                         const PassArgs& A) {
  if (ev0) hipExtLaunchKernelGGL(kernel, grid, dim3(TPT_WF_BLOCK), 0, s, ev0, ev1, 0, W, cur, A);
  else hipLaunchKernelGGL(kernel, grid, dim3(TPT_WF_BLOCK), 0, s, W, cur, A);
}
template <bool OVF, bool GEN>
static void launch_extend(const WfParams& W, int cur, const PassArgs& A, uint32_t max_n, hipStream_t s, hipEvent_t ev0,
                          hipEvent_t ev1) {
  launch_timed(k_extend<OVF, GEN>, grid_of(max_n, (const void*)k_extend<OVF, GEN>), s, ev0, ev1, W, cur, A);
}
template <bool OVF, bool GEN>
static void launch_extend_refill(const WfParams& W, int cur, const PassArgs& A, uint32_t max_n, hipStream_t s,
                                 hipEvent_t ev0, hipEvent_t ev1) {
  launch_timed(k_extend_refill<OVF, GEN>, grid_of(max_n, (const void*)k_extend_refill<OVF, GEN>), s, ev0, ev1, W, cur,
               A);
}
hipError_t wf_launch_extend(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s,
                            hipEvent_t ev0, hipEvent_t ev1) {
  const PassArgs A = gen ? *gen : PassArgs{};
  if (gen ? W.refill_first > 0 : W.refill > 0) {
    if (W.gstack) {
      if (gen) launch_extend_refill<true, true>(W, cur, A, max_n, s, ev0, ev1);
      else launch_extend_refill<true, false>(W, cur, A, max_n, s, ev0, ev1);
    } else {
      if (gen) launch_extend_refill<false, true>(W, cur, A, max_n, s, ev0, ev1);
      else launch_extend_refill<false, false>(W, cur, A, max_n, s, ev0, ev1);
    }
    return hipGetLastError();
  }
  if (W.gstack) {
    if (gen) launch_extend<true, true>(W, cur, A, max_n, s, ev0, ev1);
    else launch_extend<true, false>(W, cur, A, max_n, s, ev0, ev1);
  } else {
    if (gen) launch_extend<false, true>(W, cur, A, max_n, s, ev0, ev1);
    else launch_extend<false, false>(W, cur, A, max_n, s, ev0, ev1);
  }
  return hipGetLastError();
}
template <bool ENV_NEE, bool GEN>
static void launch_shade(const WfParams& W, int cur, const PassArgs& A, uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL((k_shade<ENV_NEE, GEN>), grid_of(max_n, (const void*)k_shade<ENV_NEE, GEN>), dim3(TPT_WF_BLOCK), 0,
                     s, W, cur, A);
}
hipError_t wf_launch_shade(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s) {
  const PassArgs A = gen ? *gen : PassArgs{};
  if (W.env.kind == TPT_ENV_IMAGE && !W.env.black) {
    if (gen) launch_shade<true, true>(W, cur, A, max_n, s);
    else launch_shade<true, false>(W, cur, A, max_n, s);
  } else {
    if (gen) launch_shade<false, true>(W, cur, A, max_n, s);
    else launch_shade<false, false>(W, cur, A, max_n, s);
  }
  return hipGetLastError();
}
hipError_t wf_launch_light_selfocc(const WfParams& W, int cur, const PassArgs* gen, uint32_t max_n, hipStream_t s) {
  const PassArgs A = gen ? *gen : PassArgs{};
  if (gen)
    hipLaunchKernelGGL(k_light_selfocc<true>, grid_of(max_n, (const void*)k_light_selfocc<true>), dim3(TPT_WF_BLOCK), 0, s,
                       W, cur, A);
  else
    hipLaunchKernelGGL(k_light_selfocc<false>, grid_of(max_n, (const void*)k_light_selfocc<false>), dim3(TPT_WF_BLOCK), 0,
                       s, W, cur, A);
  return hipGetLastError();
}
hipError_t wf_launch_shadow(const WfParams& W, int queue, uint32_t max_n, hipStream_t s) {
  if (W.refill_sh > 0) {
    if (W.gstack)
      hipLaunchKernelGGL(k_shadow_refill<true>, grid_of(max_n, (const void*)k_shadow_refill<true>), dim3(TPT_WF_BLOCK),
                         0, s, W, queue);
    else
      hipLaunchKernelGGL(k_shadow_refill<false>, grid_of(max_n, (const void*)k_shadow_refill<false>),
                         dim3(TPT_WF_BLOCK), 0, s, W, queue);
    return hipGetLastError();
  }
  if (W.gstack)
    hipLaunchKernelGGL(k_shadow<true>, grid_of(max_n, (const void*)k_shadow<true>), dim3(TPT_WF_BLOCK), 0, s, W, queue);
  else
    hipLaunchKernelGGL(k_shadow<false>, grid_of(max_n, (const void*)k_shadow<false>), dim3(TPT_WF_BLOCK), 0, s, W,
                       queue);
  return hipGetLastError();
}
hipError_t wf_launch_accumulate(const WfParams& W, const AccumArgs& A, hipStream_t s) {
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(A.n_pixels)), dim3(TPT_WF_BLOCK), 0, s, W, A);
  return hipGetLastError();
}
hipError_t wf_launch_accumulate_adaptive(const WfParams& W, const AdaptiveArgs& A, hipStream_t s) {
  hipLaunchKernelGGL(k_accumulate_adaptive, dim3(grid_for(A.n_active)), dim3(TPT_WF_BLOCK), 0, s, W, A);
  return hipGetLastError();
}
// Upper bound of the threads of any launch above over n entries (sizes per-thread areas).
uint32_t wf_grid_threads(uint32_t n) { return grid_for(n) * TPT_WF_BLOCK; }

}  // namespace tpt
