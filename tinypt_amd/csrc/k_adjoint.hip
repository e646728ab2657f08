// k_adjoint.hip -- path-replay adjoint of the FP32 render (gradients of a pixel loss with respect
// to material parameters).
//
// Every (pixel, sample) path is regenerated from its seed, seed_for(x, y, s) ^ seed, with the
// same random-number consumption as the forward wavefront kernels, so the adjoint sees exactly
// the paths the forward render summed.  Nothing is stored between paths; passes A and B below
// are one walk for paths short enough to log their BSDF vertices (ADJ_SUMGRAD), two otherwise.
//   Pass A sums, over the path's radiance terms c_k (misses, emitter hits, unoccluded NEE),
//     s_k = x_k o (delta^T J_k), where x_k is the term before the firefly clamp, J_k the clamp's
//     Jacobian (identity when not clamped) and delta = dLoss/dPixel / spp.
//   Pass B walks the path again.  A term's own parameters (BSDF value and MIS weight of its
//     vertex, emitter radiance) are differentiated where the term is formed; a BSDF sample's
//     weight w_j scales every later term, so its parameters receive
//     sum_c (dw_jc/dtheta / w_jc) * (S_total - S_upto_j)_c, the suffix of the pass-A sum.
// Sampling pdfs (BSDF, light selection, env) and Russian-roulette probabilities are held fixed:
// none of these parameters moves a sample, so differentiating the integrand alone gives the
// derivative of the expected image.  Parameters: per material 4 slots laid out like mat_param --
// Lambert / Metal / Subsurface albedo rgb, GGX F0 rgb, Dielectric absorption rgb, DiffuseLight
// emit rgb.
//
// GGX roughness is the exception.  The VNDF sampler of material.rs:327-347 does not draw
// directions with the density its pdf formula (material.rs:283-293) states, and the support
// of the density it does draw from has an edge that moves with alpha, so the expected image
// depends on alpha through the sampler in a way no per-sample derivative captures.  Its
// gradient is therefore taken by replay with common random numbers: every path that meets a
// GGX material is replayed with that material's alpha at +-TPT_ADJ_ALPHA_H and the central
// difference of its loss contribution is added.  The sum over paths is the central difference
// of the whole render's loss with the same seeds, computed in one pass.
#include <string>
#include <vector>

#include "adjoint.h"
#include "tpt_sample32.h"
#include "tpt_selfocc64.h"
#include "wavefront.h"

namespace tpt {

#define TPT_ADJ_BLOCK 256
#define TPT_ADJ_LDS_GRAD 1024  // per-block gradient slots in LDS (256 materials)
#define TPT_ADJ_ALPHA_H 1e-3  // roughness step of the per-path central difference
#define TPT_ADJ_MAX_GGX 16     // distinct GGX materials per path that receive a roughness gradient

// Replay modes: SUM returns S_total (and lists the path's GGX materials), GRAD accumulates the
// gradients, VALUE returns the path's loss contribution delta . L with one material's roughness
// shifted.  SUMGRAD does SUM and GRAD in one replay: a BSDF vertex's suffix term needs S_total,
// so the vertex's factors and S_upto are logged and the terms added once the path is done --
// possible when a path has at most TPT_ADJ_LOG BSDF vertices (max_bounces <= TPT_ADJ_LOG + 1).
#define ADJ_SUM 0
#define ADJ_GRAD 1
#define ADJ_VALUE 2
#define ADJ_SUMGRAD 3
#define TPT_ADJ_LOG 8

// BSDF vertices of one path whose weight gradient waits for S_total: the term for channel c is
// (S_total - S_upto)_c * num_c / den_c, skipped where den_c = 0.
struct VertexLog {
  int n;
  int mat[TPT_ADJ_LOG];
  f3 num[TPT_ADJ_LOG], den[TPT_ADJ_LOG], s_upto[TPT_ADJ_LOG];
};

struct GgxList {
  int n;
  int mat[TPT_ADJ_MAX_GGX];
  __device__ __forceinline__ void add(int m) {
    for (int i = 0; i < n; i++)
      if (mat[i] == m) return;
    if (n < TPT_ADJ_MAX_GGX) mat[n++] = m;
  }
};

struct AdjParams {
  SceneF32 S;
  SceneF64 S64;  // reference-layout scene, for the light self-occlusion decision
  EnvDev env;
  LightsDev lights;
  MatsDev mats;
  CamF32 cam;
  int32_t* gstack;
  int32_t gcap;
  int32_t width, height, spp, max_bounces, rr_start, n_mats;
  float inv_w, inv_h;
  uint64_t seed;
  uint64_t n_paths;
  const float* dldp;  // w*h*3
  const float* alpha_pm;  // n_mats*2: (float)(alpha + h), (float)(alpha - h)
  double* grad;       // n_mats*4
  uint64_t* stats;    // [0] extension rays, [1] shadow rays, [2] paths
};

// Firefly clamp (math.rs:42-49) and v^T J for its Jacobian.
__device__ __forceinline__ f3 clamp_vjp(f3 x, f3 v) {
  const float m = (float)TPT_FIREFLY_CLAMP;
  float l = luminance(x);
  if (!(l > m && l > 0.f)) return v;
  f3 lam = mk(0.2126f, 0.7152f, 0.0722f);
  float vx = dot(v, x);
  return (m / l) * v - (m * vx / (l * l)) * lam;
}

// d ggx_eval / dF0_c at (n, wo, wi) (material.rs:272-281): (1 - (1 - |wi.m|)^5) D G / denom,
// the same for every channel.
__device__ __forceinline__ float ggx_df_dF0(float s, f3 n, f3 wo, f3 wi) {
  float alpha = fmaxf(s, 1e-3f);
  float ci = dot(n, wo), co = dot(n, wi);
  if (ci <= 0.f || co <= 0.f) return 0.f;
  f3 m = vnorm(wo + wi);
  float ch = fmaxf(dot(n, m), 0.f);
  float D = ggx_distribution(alpha, ch);
  float G = ggx_smith_g1(alpha, ci) * ggx_smith_g1(alpha, co);
  float denom = 4.f * ci * co + 1e-6f;
  float p5 = pow5(1.f - fmaxf(fabsf(dot(wi, m)), 0.f));
  return (1.f - p5) * (D * G / denom);
}

struct GradSink {
  float* lds;  // block slots (or null)
  double* global;
  int n;
  __device__ __forceinline__ void add(int mat, int k, float v) {
    if (v == 0.f || !(v == v)) return;
    int i = 4 * mat + k;
    if (lds && i < TPT_ADJ_LDS_GRAD) atomicAdd(lds + i, v);
    else atomicAdd(global + i, (double)v);
  }
};

__device__ __forceinline__ void add_suffix_terms(GradSink& sink, int mat, f3 num, f3 den, f3 suf) {
  if (den.x != 0.f) sink.add(mat, 0, suf.x * num.x / den.x);
  if (den.y != 0.f) sink.add(mat, 1, suf.y * num.y / den.y);
  if (den.z != 0.f) sink.add(mat, 2, suf.z * num.z / den.z);
}

// One path (see the modes above).  VALUE: material pert_mat's roughness is replaced by
// pert_alpha (the FP32 rounding of the FP64 value +- h, as a material update would upload it).
// The mode is a run-time value, so the kernel holds one inlined copy of the replay: a copy per
// mode took 256 VGPRs plus 620 B of private memory per lane and ran one wave per SIMD; one copy
// runs three waves per SIMD (168 VGPRs, some spilled) and the kernel 2.3x faster.
__device__ __forceinline__ f3 replay(const int MODE, const AdjParams& A, uint32_t x, uint32_t y, uint32_t s, f3 delta,
                                     f3 S_total, GradSink& sink, GgxList* ggx, VertexLog* log, int pert_mat,
                                     float pert_alpha, const LaneStack& st, uint32_t* n_ext, uint32_t* n_sh) {
  const bool GRAD = MODE == ADJ_GRAD || MODE == ADJ_SUMGRAD;
  RngF32 rng(seed_for(x, y, s) ^ A.seed);
  f3 o, d;
  camera_ray32(A.cam, x, y, A.inv_w, A.inv_h, rng, &o, &d);
  f3 T = mk(1.f, 1.f, 1.f), S_upto = mk(0.f, 0.f, 0.f);
  float last_pdf = 0.f;
  bool last_nd = false;
  const bool env_img = A.env.kind == TPT_ENV_IMAGE && !A.env.black;  // as the forward kernels
  for (int bounce = 0; bounce < A.max_bounces; bounce++) {
    HitF32 h;
    (*n_ext)++;
    if (!trace_f32<false>(A.S, o, d, (float)TPT_RAY_EPSILON, 1e30f, &h, st)) {
      // miss (integrator.rs:71-83)
      f3 bg = A.env.black ? mk(0.f, 0.f, 0.f) : (env_img ? env_lookup32(A.env, d) : sky32(d));
      f3 x0 = had(T, bg);
      float w = 1.f, pe = 0.f;
      if (env_img && last_nd) {
        pe = env_pdf32(A.env, d);
        w = mis_weight(last_pdf, pe);
      }
      f3 xw = x0 * w;
      if (MODE == ADJ_VALUE) S_upto = S_upto + had(clamp_luminance(xw, (float)TPT_FIREFLY_CLAMP), delta);
      else S_upto = S_upto + had(xw, clamp_vjp(xw, delta));
      break;
    }
    if (bounce >= A.rr_start) {  // integrator.rs:89-95 (probability held fixed)
      float p = fmaxf(fminf(fmaxf(fmaxf(T.x, T.y), T.z), 0.95f), 0.05f);
      if (rng.uni() > p) break;
      T = T / p;
    }
    f3 p, ng;
    int mat;
    hit_geometry32(A.S, h.inst, h.prim, h.t, o, d, &p, &ng, &mat);
    const int kind = A.mats.kind[mat];
    float4 mp = A.mats.p32[mat];
    if (mat == pert_mat) mp.w = pert_alpha;
    const f3 mc = f3of(mp);
    if (kind == TPT_MAT_DIFFUSE_LIGHT) {  // integrator.rs:100-104
      f3 x0 = had(T, mc);
      f3 vj = clamp_vjp(x0, delta);
      if (MODE == ADJ_VALUE) S_upto = S_upto + had(clamp_luminance(x0, (float)TPT_FIREFLY_CLAMP), delta);
      else S_upto = S_upto + had(x0, vj);
      if (GRAD) {
        sink.add(mat, 0, vj.x * T.x);
        sink.add(mat, 1, vj.y * T.y);
        sink.add(mat, 2, vj.z * T.z);
      }
      break;
    }
    if (ggx && kind == TPT_MAT_GGX) ggx->add(mat);
    f3 n = dot(ng, d) < 0.f ? ng : -ng;
    f3 wo = vnorm(-d);
    if (!mat_is_delta(kind)) {
      if (env_img) {  // nee_environment (integrator.rs:148-172)
        f3 wi, li;
        float pe;
        env_sample32(A.env, rng, &wi, &li, &pe);
        float cs = fmaxf(dot(n, wi), 0.f);
        if (cs > 0.f && pe > 0.f && (li.x != 0.f || li.y != 0.f || li.z != 0.f)) {
          f3 f;
          float pb;
          bsdf_eval(kind, mc, mp.w, wo, wi, n, &f, &pb);
          float w = mis_weight(pe, pb);
          f3 c = had(had(T, f), li) * (cs * w / pe);
          if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
            HitF32 hs;
            (*n_sh)++;
            if (!trace_f32<true>(A.S, p + (float)TPT_RAY_EPSILON * wi, wi, (float)TPT_RAY_EPSILON, 1e30f, &hs, st)) {
              S_upto = S_upto + had(c, delta);
              if (GRAD) {
                f3 base = had(had(T, li), delta) * (cs / pe);  // dc_c/df_c * delta_c / w
                if (kind == TPT_MAT_LAMBERT || kind == TPT_MAT_SUBSURFACE) {
                  sink.add(mat, 0, base.x * w / (float)TPT_PI);
                  sink.add(mat, 1, base.y * w / (float)TPT_PI);
                  sink.add(mat, 2, base.z * w / (float)TPT_PI);
                } else if (kind == TPT_MAT_GGX) {
                  float g = ggx_df_dF0(mp.w, n, wo, wi) * w;
                  sink.add(mat, 0, base.x * g);
                  sink.add(mat, 1, base.y * g);
                  sink.add(mat, 2, base.z * g);
                }
              }
            }
          }
        }
      }
      LightSample32 ls;  // nee_area_light (integrator.rs:178-213)
      if (sample_light32(A.lights, rng, p, &ls)) {
        f3 to_light = ls.pos - p;
        float dist = sqrtf(dot(to_light, to_light));
        if (dist > 1e-6f) {
          f3 wi = to_light / dist;
          float cs = fmaxf(dot(n, wi), 0.f);
          if (cs > 0.f) {
            f3 f;
            float pb;
            bsdf_eval(kind, mc, mp.w, wo, wi, n, &f, &pb);
            float w = mis_weight(ls.pdf, pb);
            f3 c = had(had(T, f), ls.emit) * (cs * w / ls.pdf);
            if ((c.x != 0.f || c.y != 0.f || c.z != 0.f) &&
                !light_self_occluded64(A.S64, A.lights, ls.idx, hit_point64(A.S64, A.S.tri_orig, h.inst, h.prim, o, d, p),
                                       ls.u, ls.v)) {
              float tmax = light_shadow_tmax32(dist);
              HitF32 hs;
              (*n_sh)++;
              if (!trace_f32<true>(A.S, p + (float)TPT_RAY_EPSILON * wi, wi, (float)TPT_RAY_EPSILON, tmax, &hs, st)) {
                S_upto = S_upto + had(c, delta);
                if (GRAD) {
                  f3 base = had(had(T, ls.emit), delta) * (cs / ls.pdf);
                  if (kind == TPT_MAT_LAMBERT || kind == TPT_MAT_SUBSURFACE) {
                    sink.add(mat, 0, base.x * w / (float)TPT_PI);
                    sink.add(mat, 1, base.y * w / (float)TPT_PI);
                    sink.add(mat, 2, base.z * w / (float)TPT_PI);
                  } else if (kind == TPT_MAT_GGX) {
                    float g = ggx_df_dF0(mp.w, n, wo, wi) * w;
                    sink.add(mat, 0, base.x * g);
                    sink.add(mat, 1, base.y * g);
                    sink.add(mat, 2, base.z * g);
                  }
                  // emitter radiance of the sampled light
                  f3 e = had(had(T, f), delta) * (cs * w / ls.pdf);
                  sink.add(ls.mat, 0, e.x);
                  sink.add(ls.mat, 1, e.y);
                  sink.add(ls.mat, 2, e.z);
                }
              }
            }
          }
        }
      }
    }
    if (bounce + 1 >= A.max_bounces) break;
    BsdfSample<float> bs;
    if (!bsdf_sample(kind, mc, mp.w, d, p, ng, h.t, rng, &bs)) break;
    if (GRAD) {
      // the sampled weight scales every later term: d log w / d theta times the suffix sum,
      // (S_total - S_upto)_c * num_c / den_c
      f3 num = mk(0.f, 0.f, 0.f), den = num;
      if (kind == TPT_MAT_LAMBERT || kind == TPT_MAT_METAL || kind == TPT_MAT_SUBSURFACE) {
        num = mk(1.f, 1.f, 1.f);  // d log w_c / d albedo_c = 1 / albedo_c
        den = mk(mc.x > 0.f ? mc.x : 0.f, mc.y > 0.f ? mc.y : 0.f, mc.z > 0.f ? mc.z : 0.f);
      } else if (kind == TPT_MAT_GGX) {
        // w_c = F_c D G co / (denom pdf): d log w_c / dF0_c = (1 - (1 - |wi.m|)^5) / F_c
        f3 m = vnorm(wo + bs.d);
        float c = fabsf(dot(bs.d, m));
        f3 F = fresnel_schlick(c, mc);
        float dF = 1.f - pow5(1.f - fmaxf(c, 0.f));
        num = mk(dF, dF, dF);
        den = mk(F.x > 0.f ? F.x : 0.f, F.y > 0.f ? F.y : 0.f, F.z > 0.f ? F.z : 0.f);
      } else if (kind == TPT_MAT_DIELECTRIC && dot(ng, d) >= 0.f) {
        // exiting: weight carries exp(-sigma_c t) (material.rs:111-119)
        float dist = fmaxf(h.t, 0.f);
        num = mk(-dist, -dist, -dist);
        den = mk(1.f, 1.f, 1.f);
      }
      if (den.x != 0.f || den.y != 0.f || den.z != 0.f) {
        if (log) {
          const int j = log->n++;  // fits: a SUMGRAD path has at most TPT_ADJ_LOG BSDF vertices
          log->mat[j] = mat;
          log->num[j] = num;
          log->den[j] = den;
          log->s_upto[j] = S_upto;
        } else {
          add_suffix_terms(sink, mat, num, den, S_total - S_upto);
        }
      }
    }
    T = had(T, bs.weight);
    o = bs.o;
    d = bs.d;
    last_pdf = bs.pdf;
    last_nd = !bs.is_delta;
  }
  return S_upto;
}

__global__ void __launch_bounds__(TPT_ADJ_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) k_adjoint(AdjParams A) {
  __shared__ int32_t s_stack[TPT_WF_STACK_ADJ * TPT_ADJ_BLOCK];
  __shared__ float s_grad[TPT_ADJ_LDS_GRAD];
  for (int i = threadIdx.x; i < TPT_ADJ_LDS_GRAD; i += blockDim.x) s_grad[i] = 0.f;
  __syncthreads();
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  LaneStack st{s_stack + threadIdx.x, A.gstack ? A.gstack + gid * (uint64_t)A.gcap : nullptr, TPT_ADJ_BLOCK,
               TPT_WF_STACK_ADJ};
  GradSink sink{s_grad, A.grad, A.n_mats};
  const uint32_t npix = (uint32_t)A.width * (uint32_t)A.height;
  const bool one_pass = A.max_bounces <= TPT_ADJ_LOG + 1;  // at most max_bounces - 1 BSDF vertices
  uint32_t n_ext = 0, n_sh = 0, n_paths = 0;
  for (uint64_t i = gid; i < A.n_paths; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t pix = (uint32_t)(i % npix), s = (uint32_t)(i / npix);
    uint32_t x = pix % (uint32_t)A.width, y = pix / (uint32_t)A.width;
    f3 delta = mk(A.dldp[3 * pix], A.dldp[3 * pix + 1], A.dldp[3 * pix + 2]) * (1.f / (float)A.spp);
    if (delta.x == 0.f && delta.y == 0.f && delta.z == 0.f) continue;
    GgxList ggx;
    ggx.n = 0;
    VertexLog log;
    log.n = 0;
    const f3 zero = mk(0.f, 0.f, 0.f);
    // pass 0 sums the path's terms (and lists its GGX materials) -- and, when the path is short
    // enough to log its BSDF vertices, takes the gradients too; otherwise pass 1 takes them.
    // Then two passes per GGX material replay the path with its roughness at +h and -h.
    const int first_value = one_pass ? 1 : 2;
    f3 S_total = zero, dv = zero;
    for (int pass = 0; pass < first_value + 2 * ggx.n; pass++) {
      const int mode = pass == 0 ? (one_pass ? ADJ_SUMGRAD : ADJ_SUM) : pass < first_value ? ADJ_GRAD : ADJ_VALUE;
      const int vi = pass - first_value;
      const int m = vi >= 0 ? ggx.mat[vi >> 1] : -1;
      const float alpha = vi >= 0 ? A.alpha_pm[2 * m + (vi & 1)] : 0.f;
      const f3 v = replay(mode, A, x, y, s, delta, pass == 1 && !one_pass ? S_total : zero, sink,
                          pass == 0 ? &ggx : nullptr, pass == 0 && one_pass ? &log : nullptr, m, alpha, st, &n_ext,
                          &n_sh);
      if (pass == 0) {
        S_total = v;
        for (int j = 0; j < log.n; j++) add_suffix_terms(sink, log.mat[j], log.num[j], log.den[j], S_total - log.s_upto[j]);
      } else if (vi >= 0) {  // loss term at alpha + h minus the one at alpha - h
        dv = (vi & 1) ? dv - v : v;
        if (vi & 1) sink.add(m, 3, (dv.x + dv.y + dv.z) * (float)(0.5 / TPT_ADJ_ALPHA_H));
      }
    }
    n_paths++;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * A.n_mats && i < TPT_ADJ_LDS_GRAD; i += blockDim.x)
    if (s_grad[i] != 0.f) atomicAdd(A.grad + i, (double)s_grad[i]);
  unsigned long long e = n_ext, sh = n_sh, pa = n_paths;
  for (int off = 32; off > 0; off >>= 1) {
    e += (unsigned long long)__shfl_down((long long)e, off);
    sh += (unsigned long long)__shfl_down((long long)sh, off);
    pa += (unsigned long long)__shfl_down((long long)pa, off);
  }
  if (__lane_id() == 0) {
    if (e) atomicAdd((unsigned long long*)&A.stats[0], e);
    if (sh) atomicAdd((unsigned long long*)&A.stats[1], sh);
    if (pa) atomicAdd((unsigned long long*)&A.stats[2], pa);
  }
}

static int adjoint_blocks(uint64_t n_paths) {
  return (int)std::max<uint64_t>(std::min<uint64_t>((n_paths + TPT_ADJ_BLOCK - 1) / TPT_ADJ_BLOCK, 8192), 1);
}

void adjoint_scratch_bytes(const tpt_render_config* cfg, int n_mats, int stack_need, size_t bytes[ADJ_SCRATCH_N]) {
  const size_t npix = (size_t)cfg->width * cfg->height;
  const uint64_t n_paths = (uint64_t)npix * (uint64_t)std::max(cfg->spp, 0);
  const int gcap = stack_need > TPT_WF_STACK_ADJ ? stack_need - TPT_WF_STACK_ADJ : 0;
  bytes[ADJ_SCRATCH_DLDP] = npix * 12;
  bytes[ADJ_SCRATCH_GRAD] = (size_t)std::max(n_mats, 1) * 4 * 8;
  bytes[ADJ_SCRATCH_STATS] = 64;
  bytes[ADJ_SCRATCH_APM] = (size_t)std::max(n_mats, 1) * 2 * 4;
  bytes[ADJ_SCRATCH_GSTACK] = gcap > 0 ? (size_t)adjoint_blocks(n_paths) * TPT_ADJ_BLOCK * gcap * 4 : 0;
}

int adjoint_render(hipStream_t s, const SceneF32& S, const SceneF64& S64, const EnvDev& env, const LightsDev& lights, const MatsDev& mats,
                   const CamF32& cam, const double* mat_param, int stack_need, const tpt_render_config* cfg,
                   const float* d_loss_d_pixel, double* grad_params, tpt_stats* stats, const AdjScratch& scratch,
                   std::string* err) {
  auto fail = [&](int code, const std::string& m) {
    *err = m;
    return code;
  };
  if (cfg->adaptive_enabled) return fail(TPT_ERR_UNSUPPORTED, "the adjoint replays fixed-spp renders only");
  const size_t npix = (size_t)cfg->width * cfg->height;
  const uint64_t n_paths = (uint64_t)npix * (uint64_t)std::max(cfg->spp, 0);
  const int blocks = adjoint_blocks(n_paths);
  const int gcap = stack_need > TPT_WF_STACK_ADJ ? stack_need - TPT_WF_STACK_ADJ : 0;
  std::vector<float> apm(2 * (size_t)std::max(mats.n, 1));
  for (int m = 0; m < mats.n; m++) {
    apm[2 * m] = (float)(mat_param[4 * m + 3] + TPT_ADJ_ALPHA_H);
    apm[2 * m + 1] = (float)(mat_param[4 * m + 3] - TPT_ADJ_ALPHA_H);
  }
  void *d_dldp = scratch.dldp, *d_grad = scratch.grad, *d_stats = scratch.stats, *d_gs = scratch.gstack,
       *d_apm = scratch.apm;
  hipError_t e = hipMemcpyAsync(d_apm, apm.data(), apm.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_dldp, d_loss_d_pixel, npix * 12, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_grad, grad_params, (size_t)mats.n * 4 * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(d_stats, 0, 64, s);
  if (e != hipSuccess) return fail(TPT_ERR_DEVICE, std::string("adjoint setup: ") + hipGetErrorString(e));
  AdjParams A{};
  A.S = S;
  A.S64 = S64;
  A.env = env;
  A.lights = lights;
  A.mats = mats;
  A.cam = cam;
  A.gstack = (int32_t*)d_gs;
  A.gcap = gcap;
  A.width = cfg->width;
  A.height = cfg->height;
  A.spp = cfg->spp;
  A.max_bounces = cfg->max_bounces;
  A.rr_start = cfg->rr_start;
  A.n_mats = mats.n;
  A.inv_w = (float)(1.0 / (double)cfg->width);
  A.inv_h = (float)(1.0 / (double)cfg->height);
  A.seed = cfg->seed;
  A.n_paths = n_paths;
  A.dldp = (const float*)d_dldp;
  A.alpha_pm = (const float*)d_apm;
  A.grad = (double*)d_grad;
  A.stats = (uint64_t*)d_stats;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, s);
  if (n_paths > 0) hipLaunchKernelGGL(k_adjoint, dim3(blocks), dim3(TPT_ADJ_BLOCK), 0, s, A);
  e = hipGetLastError();
  (void)hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipMemcpyAsync(grad_params, d_grad, (size_t)mats.n * 4 * 8, hipMemcpyDeviceToHost, s);
  uint64_t h[8] = {0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, d_stats, 64, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(TPT_ERR_DEVICE, std::string("adjoint: ") + hipGetErrorString(e));
  if (stats) {
    stats->paths = h[2];
    stats->rays_extension = h[0];
    stats->rays_shadow = h[1];
    stats->elapsed_ms = ms;
    stats->trace_kernel_ms = ms;
    stats->trace_launches = 1;
    stats->trace_rays = h[0] + h[1];
    stats->light_samples = 0;
    stats->light_self_occluded = 0;
  }
  return TPT_OK;
}

}  // namespace tpt
