// tpt_bsdf.h -- the Material sample/eval contract (material.rs:60-347) on the device.
//
// Templated on the scalar T (float for the wavefront kernels, double for the plain kernel) and
// on the random source R (R::uni() returns a uniform in [0,1) of type T).  The draw order per
// material is the reference's: Lambert 2, Metal 0, Dielectric 1 only when refraction is
// possible, GGX 2 only when wo.n > 0, Subsurface 3.
#pragma once

#include "../../include/tinypt_gpu.h"
#include "tpt_math.h"

namespace tpt {

#define TPT_PI 3.14159265358979323846
#define TPT_TAU 6.28318530717958647692
#define TPT_RAY_EPSILON 1e-4
#define TPT_FIREFLY_CLAMP 50.0

template <class T>
struct BsdfSample {  // material.rs:49-58
  vec3<T> o, d;      // scattered ray (origin offset included)
  vec3<T> weight;    // f cos / pdf
  T pdf;
  bool is_delta;
};

template <class T> __device__ __forceinline__ T pow5(T x) {  // material.rs:259-262
  T x2 = x * x;
  return x2 * x2 * x;
}
template <class T> __device__ __forceinline__ vec3<T> fresnel_schlick(T cos_theta, vec3<T> f0) {  // material.rs:265-269
  vec3<T> one = mk(T(1), T(1), T(1));
  T x = T(1) - fmax(cos_theta, T(0));
  return f0 + (one - f0) * pow5(x);
}
template <class T> __device__ __forceinline__ T ggx_distribution(T alpha, T cos_h) {  // material.rs:272-279
  if (cos_h <= T(0)) return T(0);
  T a2 = alpha * alpha;
  T denom = cos_h * cos_h * (a2 - T(1)) + T(1);
  return a2 / (T(TPT_PI) * denom * denom);
}
template <class T> __device__ __forceinline__ T ggx_smith_g1(T alpha, T c) {  // material.rs:282-288
  if (c <= T(0)) return T(0);
  T a = alpha * sqrt(fmax(T(1) - c * c, T(0))) / fmax(c, T(1e-6));
  return T(2) / (T(1) + sqrt(T(1) + a * a));
}
template <class T> __device__ __forceinline__ T ggx_smith(T alpha, T ci, T co) { return ggx_smith_g1(alpha, ci) * ggx_smith_g1(alpha, co); }

template <class T>
__device__ __forceinline__ vec3<T> ggx_eval(vec3<T> albedo, T alpha, vec3<T> n, vec3<T> wo, vec3<T> wi) {  // material.rs:227-240
  T cos_i = dot(n, wo), cos_o = dot(n, wi);
  if (cos_i <= T(0) || cos_o <= T(0)) return mk(T(0), T(0), T(0));
  vec3<T> m = vnorm(wo + wi);
  T cos_h = fmax(dot(n, m), T(0));
  T d = ggx_distribution(alpha, cos_h);
  T g = ggx_smith(alpha, cos_i, cos_o);
  vec3<T> f = fresnel_schlick(fabs(dot(wi, m)), albedo);
  T denom = T(4) * cos_i * cos_o + T(1e-6);
  return f * (d * g / denom);
}
template <class T>
__device__ __forceinline__ T ggx_pdf(T alpha, vec3<T> n, vec3<T> wo, vec3<T> wi) {  // material.rs:243-256
  T cos_i = dot(n, wo), cos_o = dot(n, wi);
  if (cos_i <= T(0) || cos_o <= T(0)) return T(0);
  vec3<T> m = vnorm(wo + wi);
  T cos_h = fmax(dot(n, m), T(0));
  T d = ggx_distribution(alpha, cos_h);
  T g1 = ggx_smith_g1(alpha, cos_i);
  T pdf_m = d * g1 * cos_h / fmax(cos_i, T(1e-6));
  T denom = T(4) * fmax(fabs(dot(wi, m)), T(1e-6));
  return fmax(pdf_m / denom, T(0));
}
template <class T> __device__ __forceinline__ void tangent_frame(vec3<T> n, vec3<T>* t, vec3<T>* b) {  // material.rs:313-318
  vec3<T> a = fabs(n.x) > T(0.9) ? mk(T(0), T(1), T(0)) : mk(T(1), T(0), T(0));
  *t = vnorm(cross(n, a));
  *b = cross(*t, n);
}
template <class T, class R>
__device__ __forceinline__ vec3<T> sample_cosine_hemisphere(vec3<T> n, R& rng) {  // material.rs:299-310
  T u = rng.uni();
  T v = rng.uni();
  T r = sqrt(u);
  T phi = T(2) * T(TPT_PI) * v;
  T x = r * sample_cos(phi);
  T y = r * sample_sin(phi);
  T z = sqrt(T(1) - u);
  vec3<T> t, b;
  tangent_frame(n, &t, &b);
  return t * x + b * y + n * z;
}
template <class T>
__device__ __forceinline__ vec3<T> sample_ggx_vndf(vec3<T> wo, T alpha, T u1, T u2) {  // material.rs:327-347
  vec3<T> v = vnorm(mk(alpha * wo.x, alpha * wo.y, wo.z));
  vec3<T> t1, t2;
  if (v.z < T(0.9999)) {
    t1 = vnorm(mk(-v.y, v.x, T(0)));
    t2 = cross(v, t1);
  } else {
    t1 = mk(T(1), T(0), T(0));
    t2 = mk(T(0), T(1), T(0));
  }
  T r = sqrt(u1);
  T phi = T(TPT_TAU) * u2;
  T t1p = r * sample_cos(phi);
  T t2p_init = r * sample_sin(phi);
  T s = T(0.5) * (T(1) + v.z);
  T t2p = sqrt(T(1) - s) * t2p_init + s * sqrt(fmax(T(1) - t1p * t1p, T(0)));
  vec3<T> nh = vnorm(t1 * t1p + t2 * t2p + v * sqrt(fmax(T(1) - t1p * t1p - t2p * t2p, T(0))));
  return vnorm(mk(alpha * nh.x, alpha * nh.y, nh.z));
}

// Material::sample (material.rs:76-201).  kind/c/s: material kind, colour param, scalar param.
// hit_n is the geometric normal of the hit (not oriented), hit_t the hit distance.
template <class T, class R>
__device__ __forceinline__ bool bsdf_sample(int kind, vec3<T> c, T s, vec3<T> d_in, vec3<T> hit_p, vec3<T> hit_n, T hit_t,
                                            R& rng, BsdfSample<T>* out) {
  const T eps = T(TPT_RAY_EPSILON);
  bool entering = dot(hit_n, d_in) < T(0);
  vec3<T> n = entering ? hit_n : -hit_n;
  switch (kind) {
    case TPT_MAT_LAMBERT: {
      vec3<T> d = sample_cosine_hemisphere(n, rng);
      out->o = hit_p + eps * d;
      out->d = d;
      out->weight = c;
      out->pdf = fmax(dot(n, d), T(0)) / T(TPT_PI);
      out->is_delta = false;
      return true;
    }
    case TPT_MAT_METAL: {
      vec3<T> d = reflect(d_in, n);
      out->o = hit_p + eps * d;
      out->d = d;
      out->weight = c;
      out->pdf = T(0);
      out->is_delta = true;
      return true;
    }
    case TPT_MAT_DIELECTRIC: {
      T ior = s;
      T eta = entering ? div_approx(T(1), ior) : ior;
      T cos_i = fmax(dot(-d_in, n), T(0));
      T r0 = (ior - T(1)) / (ior + T(1));
      T f0 = r0 * r0;
      T fresnel = f0 + (T(1) - f0) * pow5(T(1) - cos_i);
      vec3<T> beta = mk(T(1), T(1), T(1));
      if (!entering) {
        T dist = fmax(hit_t, T(0));
        beta = had(beta, mk(exp(-c.x * dist), exp(-c.y * dist), exp(-c.z * dist)));
      }
      vec3<T> refl = reflect(d_in, n);
      vec3<T> refr;
      bool can_refr = refract(d_in, n, eta, &refr);
      bool choose_refl = !can_refr || rng.uni() < fresnel;
      vec3<T> d;
      if (choose_refl) {
        beta = beta * (fresnel / fmax(fresnel, T(1e-6)));
        d = refl;
      } else {
        T scale = (T(1) - fresnel) / fmax(T(1) - fresnel, T(1e-6));
        beta = beta * scale * (eta * eta);
        d = refr;
      }
      out->o = hit_p + eps * d;
      out->d = d;
      out->weight = beta;
      out->pdf = T(0);
      out->is_delta = true;
      return true;
    }
    case TPT_MAT_GGX: {
      T alpha = fmax(s, T(1e-3));
      vec3<T> t, b;
      tangent_frame(n, &t, &b);
      vec3<T> wo = vnorm(-d_in);
      vec3<T> wo_local = mk(dot(wo, t), dot(wo, b), dot(wo, n));
      if (wo_local.z <= T(0)) return false;
      T u1 = rng.uni();
      T u2 = rng.uni();
      vec3<T> ml = sample_ggx_vndf(wo_local, alpha, u1, u2);
      vec3<T> m = t * ml.x + b * ml.y + n * ml.z;
      vec3<T> d = reflect(d_in, m);
      T cos_o = dot(d, n);
      T cos_i = dot(-d_in, n);
      if (cos_o <= T(1e-6) || cos_i <= T(1e-6)) return false;
      T cos_h = fmax(dot(n, m), T(0));
      T dg = ggx_distribution(alpha, cos_h);
      T g = ggx_smith(alpha, cos_i, cos_o);
      vec3<T> f = fresnel_schlick(fabs(dot(d, m)), c);
      T denom = T(4) * cos_i * cos_o + T(1e-6);
      vec3<T> spec = f * (dg * g / denom);
      T pdf = ggx_pdf(alpha, n, wo, d);
      out->o = hit_p + eps * d;
      out->d = d;
      out->weight = spec * (cos_o / fmax(pdf, T(1e-6)));
      out->pdf = pdf;
      out->is_delta = false;
      return true;
    }
    case TPT_MAT_SUBSURFACE: {
      T scale = fmax(s, T(1e-4));
      T u = fmax(rng.uni(), T(1e-12));
      T dist = -log(u) * scale;
      vec3<T> normal_out = hit_n;
      vec3<T> d = sample_cosine_hemisphere(normal_out, rng);
      out->weight = c * exp(-dist / scale);
      out->o = hit_p - normal_out * dist + eps * d;
      out->d = d;
      out->pdf = fmax(dot(n, d), T(0)) / T(TPT_PI);
      out->is_delta = false;
      return true;
    }
    default:
      return false;
  }
}

// Material::eval (material.rs:207-223): f without the cosine, and the solid-angle pdf.
template <class T>
__device__ __forceinline__ void bsdf_eval(int kind, vec3<T> c, T s, vec3<T> wo, vec3<T> wi, vec3<T> n, vec3<T>* f, T* pdf) {
  if (kind == TPT_MAT_LAMBERT || kind == TPT_MAT_SUBSURFACE) {
    T cs = fmax(dot(n, wi), T(0));
    if (cs <= T(0)) {
      *f = mk(T(0), T(0), T(0));
      *pdf = T(0);
    } else {
      *f = c * (T(1) / T(TPT_PI));
      *pdf = cs / T(TPT_PI);
    }
  } else if (kind == TPT_MAT_GGX) {
    T a = fmax(s, T(1e-3));
    *f = ggx_eval(c, a, n, wo, wi);
    *pdf = ggx_pdf(a, n, wo, wi);
  } else {
    *f = mk(T(0), T(0), T(0));
    *pdf = T(0);
  }
}

__device__ __forceinline__ bool mat_is_delta(int kind) { return kind == TPT_MAT_METAL || kind == TPT_MAT_DIELECTRIC; }

template <class T> __device__ __forceinline__ T mis_weight(T a, T b) {  // integrator.rs:221-225
  T a2 = a * a, b2 = b * b;
  return a2 + b2 > T(0) ? div_approx(a2, a2 + b2) : T(0);
}

// cdf_search (world.rs:324-336, env.rs:185-197): largest lo with cdf[lo] <= x, starting at 0
// and never testing the last entry.
template <class T>
__device__ __forceinline__ int cdf_search(const T* cdf, int len, T x) {
  int lo = 0, hi = len > 0 ? len - 1 : 0;
  while (lo + 1 < hi) {
    int mid = (lo + hi) / 2;
    if (cdf[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace tpt
