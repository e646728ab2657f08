// k_wavefront.hip -- FP32 wavefront path tracer (radiance(), integrator.rs:53-135, split into
// kernels over SoA path queues).
//
// One pass renders `k` samples of a set of pixels: P = pixels * k paths, slot i -> pixel
// (i mod pixels), sample (s0 + i / pixels), so neighbouring slots are neighbouring pixels.
// Per bounce:  extend (closest hit) -> shade (miss / RR / emission / NEE sampling / BSDF
// sampling; survivors compacted into the next queue) -> shadow (any hit for the <= 2 NEE rays
// of each path, added to that path's radiance by the thread that owns the path).
// After the last bounce, accumulate adds each pixel's k path radiances into FP64 sums.
// Queue compaction: one __ballot per wave, one atomicAdd per wave, lane prefix by popcount.
// Random numbers: one PCG32 stream per (pixel, sample): seed_for(x, y, s) ^ seed.
#include "tpt_bsdf.h"
#include "tpt_kernels.h"
#include "tpt_sample32.h"
#include "tpt_selfocc64.h"
#include "tpt_trace32.h"
#include "wavefront.h"

namespace tpt {

// ---------------------------------------------------------------- queue compaction
__device__ __forceinline__ void wave_enqueue(bool pred, int32_t val, int32_t* __restrict__ q, uint32_t* __restrict__ count) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  uint32_t lane = __lane_id();
  uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, (int)leader);
  if (pred) q[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = val;
}

// ---------------------------------------------------------------- kernels
__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

__global__ void __launch_bounds__(TPT_WF_BLOCK) k_raygen(WfParams W, RaygenArgs A) {
  const uint32_t P = A.n_paths;
  for (uint32_t i = gtid(); i < P; i += gridDim.x * blockDim.x) {
    uint32_t pix, s;
    if (A.active) {
      uint32_t j = i % A.n_pixels;
      pix = (uint32_t)A.active[j];
      s = A.nsamp[pix] + i / A.n_pixels;
    } else {
      pix = A.pixel0 + i % A.n_pixels;
      s = A.s0 + i / A.n_pixels;
    }
    uint32_t x = pix % (uint32_t)W.width, y = pix / (uint32_t)W.width;
    RngF32 rng(seed_for(x, y, s) ^ W.seed);
    f3 origin, dir;
    camera_ray32(W.cam, x, y, W.inv_w, W.inv_h, rng, &origin, &dir);
    W.ps.o[i] = make_float4(origin.x, origin.y, origin.z, 0.f);
    W.ps.d[i] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(0u));
    W.ps.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    W.ps.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    W.ps.rng[i] = rng.p.state;
    W.ps.pix[i] = pix;
    W.ps.q[0][i] = (int32_t)i;
  }
  if (gtid() == 0) {
    W.ps.counters[TPT_CNT_Q0] = P;
    W.ps.counters[TPT_CNT_Q1] = 0;
    W.ps.counters[TPT_CNT_SH] = 0;
  }
}

__global__ void __launch_bounds__(TPT_WF_BLOCK) k_extend(WfParams W, int cur) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  const uint32_t n = W.ps.counters[cur ? TPT_CNT_Q1 : TPT_CNT_Q0];
  const int32_t* __restrict__ q = W.ps.q[cur];
  LaneStack st{s_stack + threadIdx.x, W.gstack ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
               TPT_WF_STACK};
  if (gtid() == 0) {
    W.ps.counters[cur ? TPT_CNT_Q0 : TPT_CNT_Q1] = 0;  // next queue
    W.ps.counters[TPT_CNT_SH] = 0;
    atomicAdd((unsigned long long*)&W.ps.stats[0], (unsigned long long)n);
  }
  for (uint32_t i = gtid(); i < n; i += gridDim.x * blockDim.x) {
    int32_t slot = q[i];
    float4 o4 = W.ps.o[slot], d4 = W.ps.d[slot];
    HitF32 h;
    trace_f32<false>(W.S, f3of(o4), f3of(d4), (float)TPT_RAY_EPSILON, 1e30f, &h, st);
    W.ps.hit[slot] = make_int4(__float_as_int(h.t), h.inst, h.prim, 0);
  }
}

__global__ void __launch_bounds__(TPT_WF_BLOCK) k_shade(WfParams W, int cur) {
  const uint32_t n = W.ps.counters[cur ? TPT_CNT_Q1 : TPT_CNT_Q0];
  const int32_t* __restrict__ qin = W.ps.q[cur];
  int32_t* __restrict__ qout = W.ps.q[cur ^ 1];
  uint32_t* cout = &W.ps.counters[cur ? TPT_CNT_Q0 : TPT_CNT_Q1];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    bool alive = false, has_shadow = false;
    int32_t slot = 0;
    if (i < n) {
      slot = qin[i];
      float4 o4 = W.ps.o[slot], d4 = W.ps.d[slot], b4 = W.ps.beta[slot], l4 = W.ps.L[slot];
      int4 h4 = W.ps.hit[slot];
      f3 o = f3of(o4), d = f3of(d4), beta = f3of(b4), L = f3of(l4);
      float last_pdf = o4.w;
      uint32_t flags = __float_as_uint(d4.w);
      int bounce = (int)(flags & 0xffu);
      bool last_nd = (flags & 0x100u) != 0;
      RngF32 rng(1);
      rng.p.state = W.ps.rng[slot];
      float4 sh_c0 = make_float4(0.f, 0.f, 0.f, 0.f), sh_c1 = sh_c0;
      if (h4.y == -1) {
        // ---- miss (integrator.rs:71-83)
        f3 bg = W.env.kind == TPT_ENV_IMAGE ? env_lookup32(W.env, d) : sky32(d);
        f3 c = had(beta, bg);
        if (W.env.kind == TPT_ENV_IMAGE && last_nd) c = c * mis_weight(last_pdf, env_pdf32(W.env, d));
        L = L + clamp_luminance(c, (float)TPT_FIREFLY_CLAMP);
      } else {
        bool go = true;
        if (bounce >= W.rr_start) {  // integrator.rs:89-95
          float p = fmaxf(fminf(fmaxf(fmaxf(beta.x, beta.y), beta.z), 0.95f), 0.05f);
          if (rng.uni() > p) go = false;
          else beta = beta / p;
        }
        if (go) {
          const float t = __int_as_float(h4.x);
          f3 p, ng;
          int mat;
          hit_geometry32(W.S, h4.y, h4.z, t, o, d, &p, &ng, &mat);
          const int kind = W.mats.kind[mat];
          const float4 mp = W.mats.p32[mat];
          const f3 mc = f3of(mp);
          if (kind == TPT_MAT_DIFFUSE_LIGHT) {  // integrator.rs:100-104
            L = L + clamp_luminance(had(beta, mc), (float)TPT_FIREFLY_CLAMP);
          } else {
            f3 n = dot(ng, d) < 0.f ? ng : -ng;  // oriented_normal (integrator.rs:138-140)
            if (!mat_is_delta(kind)) {
              f3 wo = vnorm(-d);
              if (W.env.kind == TPT_ENV_IMAGE) {  // nee_environment (integrator.rs:148-172)
                f3 wi, li;
                float pdf_env;
                env_sample32(W.env, rng, &wi, &li, &pdf_env);
                float cs = fmaxf(dot(n, wi), 0.f);
                if (cs > 0.f && pdf_env > 0.f && (li.x != 0.f || li.y != 0.f || li.z != 0.f)) {
                  f3 f;
                  float pdf_b;
                  bsdf_eval(kind, mc, mp.w, wo, wi, n, &f, &pdf_b);
                  f3 c = had(had(beta, f), li) * (cs * mis_weight(pdf_env, pdf_b) / pdf_env);
                  if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                    f3 so = p + (float)TPT_RAY_EPSILON * wi;
                    W.ps.sh_o[2 * slot] = make_float4(so.x, so.y, so.z, 1e30f);
                    W.ps.sh_d[2 * slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
                    sh_c0 = make_float4(c.x, c.y, c.z, 1.f);
                    has_shadow = true;
                  }
                }
              }
              LightSample32 ls;  // nee_area_light (integrator.rs:178-213)
              if (sample_light32(W.lights, rng, p, &ls)) {
                f3 to_light = ls.pos - p;
                float dist = sqrtf(dot(to_light, to_light));
                if (dist > 1e-6f) {
                  f3 wi = to_light / dist;
                  float cs = fmaxf(dot(n, wi), 0.f);
                  if (cs > 0.f) {
                    f3 f;
                    float pdf_b;
                    bsdf_eval(kind, mc, mp.w, wo, wi, n, &f, &pdf_b);
                    f3 c = had(had(beta, f), ls.emit) * (cs * mis_weight(ls.pdf, pdf_b) / ls.pdf);
                    if ((c.x != 0.f || c.y != 0.f || c.z != 0.f) &&
                        !light_self_occluded64(W.S64, W.lights, ls.idx, p.x, p.y, p.z, ls.u, ls.v)) {
                      f3 so = p + (float)TPT_RAY_EPSILON * wi;
                      float tmax = light_shadow_tmax32(dist);
                      W.ps.sh_o[2 * slot + 1] = make_float4(so.x, so.y, so.z, tmax);
                      W.ps.sh_d[2 * slot + 1] = make_float4(wi.x, wi.y, wi.z, 0.f);
                      sh_c1 = make_float4(c.x, c.y, c.z, 1.f);
                      has_shadow = true;
                    }
                  }
                }
              }
            }
            if (bounce + 1 < W.max_bounces) {  // the scattered ray of the last bounce is never traced
              BsdfSample<float> bs;
              if (bsdf_sample(kind, mc, mp.w, d, p, ng, t, rng, &bs)) {
                beta = had(beta, bs.weight);
                o = bs.o;
                d = bs.d;
                last_pdf = bs.pdf;
                last_nd = !bs.is_delta;
                alive = true;
              }
            }
          }
        }
      }
      if (has_shadow) {
        W.ps.sh_c[2 * slot] = sh_c0;
        W.ps.sh_c[2 * slot + 1] = sh_c1;
      }
      W.ps.L[slot] = make_float4(L.x, L.y, L.z, 0.f);
      if (alive) {
        flags = (uint32_t)(bounce + 1) | (last_nd ? 0x100u : 0u);
        W.ps.o[slot] = make_float4(o.x, o.y, o.z, last_pdf);
        W.ps.d[slot] = make_float4(d.x, d.y, d.z, __uint_as_float(flags));
        W.ps.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
        W.ps.rng[slot] = rng.p.state;
      }
    }
    wave_enqueue(alive, slot, qout, cout);
    wave_enqueue(has_shadow, slot, W.ps.sq, &W.ps.counters[TPT_CNT_SH]);
  }
}

__global__ void __launch_bounds__(TPT_WF_BLOCK) k_shadow(WfParams W) {
  __shared__ int32_t s_stack[TPT_WF_STACK * TPT_WF_BLOCK];
  const uint32_t n = W.ps.counters[TPT_CNT_SH];
  LaneStack st{s_stack + threadIdx.x, W.gstack ? W.gstack + (size_t)gtid() * W.gstack_cap : nullptr, TPT_WF_BLOCK,
               TPT_WF_STACK};
  uint32_t traced = 0;
  for (uint32_t i = gtid(); i < n; i += gridDim.x * blockDim.x) {
    int32_t slot = W.ps.sq[i];
    f3 add = mk(0.f, 0.f, 0.f);
    for (int k = 0; k < 2; k++) {
      float4 c = W.ps.sh_c[2 * slot + k];
      if (c.w == 0.f) continue;
      float4 o4 = W.ps.sh_o[2 * slot + k], d4 = W.ps.sh_d[2 * slot + k];
      HitF32 h;
      traced++;
      if (!trace_f32<true>(W.S, f3of(o4), f3of(d4), (float)TPT_RAY_EPSILON, o4.w, &h, st)) add = add + f3of(c);
    }
    float4 l4 = W.ps.L[slot];
    W.ps.L[slot] = make_float4(l4.x + add.x, l4.y + add.y, l4.z + add.z, 0.f);
  }
  // one atomic per wave for the shadow-ray count
  uint32_t tot = traced;
  for (int off = 32; off > 0; off >>= 1) tot += (uint32_t)__shfl_down((int)tot, off);
  if (__lane_id() == 0 && tot) atomicAdd((unsigned long long*)&W.ps.stats[1], (unsigned long long)tot);
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
    wave_enqueue(keep, pix, A.active_out, A.n_active_out);
  }
  for (int off = 32; off > 0; off >>= 1) kept += (uint32_t)__shfl_down((int)kept, off);
  if (__lane_id() == 0 && kept) atomicAdd((unsigned long long*)&W.ps.stats[2], (unsigned long long)kept);
}

// ---------------------------------------------------------------- host-side launchers
static inline uint32_t grid_for(uint32_t n) {
  uint32_t b = (n + TPT_WF_BLOCK - 1) / TPT_WF_BLOCK;
  return b < 1 ? 1 : (b > 16384 ? 16384 : b);
}

hipError_t wf_launch_raygen(const WfParams& W, const RaygenArgs& A, hipStream_t s) {
  hipLaunchKernelGGL(k_raygen, dim3(grid_for(A.n_paths)), dim3(TPT_WF_BLOCK), 0, s, W, A);
  return hipGetLastError();
}
hipError_t wf_launch_extend(const WfParams& W, int cur, uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(grid_for(max_n)), dim3(TPT_WF_BLOCK), 0, s, W, cur);
  return hipGetLastError();
}
hipError_t wf_launch_shade(const WfParams& W, int cur, uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3(grid_for(max_n)), dim3(TPT_WF_BLOCK), 0, s, W, cur);
  return hipGetLastError();
}
hipError_t wf_launch_shadow(const WfParams& W, uint32_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(grid_for(max_n)), dim3(TPT_WF_BLOCK), 0, s, W);
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
uint32_t wf_grid_threads(uint32_t n) { return grid_for(n) * TPT_WF_BLOCK; }

}  // namespace tpt
