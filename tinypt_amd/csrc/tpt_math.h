// tpt_math.h -- device vector math and random numbers shared by the HIP kernels.
//
// vec3<T> is instantiated for float (FP32 wavefront) and double (plain FP64 kernel).  Helpers
// mirror math.rs (norm with max(len, 1e-30), luminance, clamp_luminance, reflect, refract) in
// the same operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tpt_layout.h"

namespace tpt {

#define TPT_HD __host__ __device__ __forceinline__

template <class T>
struct vec3 {
  T x, y, z;
};
template <class T> TPT_HD vec3<T> mk(T x, T y, T z) { return vec3<T>{x, y, z}; }
template <class T> TPT_HD vec3<T> operator+(vec3<T> a, vec3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> TPT_HD vec3<T> operator-(vec3<T> a, vec3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> TPT_HD vec3<T> operator*(vec3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> TPT_HD vec3<T> operator*(T s, vec3<T> a) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> TPT_HD vec3<T> operator/(vec3<T> a, T s) { return mk(a.x / s, a.y / s, a.z / s); }
// FP32 division as the FP32 kernels compute it on the device: the hardware reciprocal (about
// 1 ulp) times the numerator, which is what their approximate-division build gives a / b.  Written
// out because the compiler drops that approximation from a division it rewrites (folding a
// negation or a select into it), leaving the correctly rounded dozen-instruction sequence there.
// The host build and FP64 divide exactly.
TPT_HD float div_approx(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;
#endif
}
TPT_HD double div_approx(double a, double b) { return a / b; }
TPT_HD vec3<float> operator/(vec3<float> a, float s) {
  const float r = div_approx(1.f, s);
  return mk(a.x * r, a.y * r, a.z * r);
}
template <class T> TPT_HD vec3<T> operator-(vec3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <class T> TPT_HD T dot(vec3<T> a, vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> TPT_HD vec3<T> cross(vec3<T> a, vec3<T> b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> TPT_HD vec3<T> had(vec3<T> a, vec3<T> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> TPT_HD T vlen(vec3<T> a) { return sqrt(dot(a, a)); }
template <class T> TPT_HD vec3<T> vnorm(vec3<T> a) {  // math.rs:69-72
  T l = fmax(vlen(a), T(1e-30));
  return a / l;
}
template <class T> TPT_HD T luminance(vec3<T> c) { return T(0.2126) * c.x + T(0.7152) * c.y + T(0.0722) * c.z; }
template <class T> TPT_HD vec3<T> clamp_luminance(vec3<T> c, T m) {  // math.rs:42-49
  T l = luminance(c);
  if (l > m && l > T(0)) return c * (m / l);
  return c;
}
template <class T> TPT_HD vec3<T> reflect(vec3<T> v, vec3<T> n) { return v - (T(2) * dot(v, n)) * n; }
template <class T> TPT_HD bool refract(vec3<T> v, vec3<T> n, T eta, vec3<T>* out) {  // math.rs:199-207
  T cosi = fmin(fmax(dot(-v, n), T(-1)), T(1));
  T sin2t = eta * eta * (T(1) - cosi * cosi);
  if (sin2t >= T(1)) return false;
  T cost = sqrt(T(1) - sin2t);
  *out = eta * v + (eta * cosi - cost) * n;
  return true;
}
template <class T> TPT_HD vec3<T> ld3(const T* p) { return mk(p[0], p[1], p[2]); }

// Sine and cosine of sampled angles.  The FP32 kernels take the hardware approximations
// (absolute error near 1e-6 over [0, 2 pi), far below the Monte Carlo noise, and the forward and
// adjoint kernels call the same code); FP64 keeps the libm-equivalent functions of the
// reference's arithmetic.
TPT_HD float sample_sin(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __sinf(x);
#else
  return sinf(x);
#endif
}
TPT_HD float sample_cos(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __cosf(x);
#else
  return cosf(x);
#endif
}
TPT_HD double sample_sin(double x) { return sin(x); }
TPT_HD double sample_cos(double x) { return cos(x); }

// ---------------------------------------------------------------- RNG (rng.rs:9-47)
struct Pcg32 {
  uint64_t state;
  TPT_HD explicit Pcg32(uint64_t seed) : state(seed == 0 ? 0x9E3779B97F4A7C15ull : seed) {}
  TPT_HD uint32_t next_u32() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
};
TPT_HD uint64_t seed_for(uint32_t px, uint32_t py, uint32_t s) {
  uint64_t h = 0xD1B54A32D192ED03ull;
  h ^= (uint64_t)px * 0x9E3779B185EBCA87ull;
  h ^= (uint64_t)py * 0xC2B2AE3D27D4EB4Full;
  h ^= (uint64_t)s * 0x165667B19E3779F9ull;
  return h ^ (h >> 32);
}
// Rng::next_f64: 53-bit uniform from two outputs (rng.rs:30-37).
struct RngF64 {
  Pcg32 p;
  TPT_HD explicit RngF64(uint64_t seed) : p(seed) {}
  TPT_HD double uni() {
    uint64_t hi = ((uint64_t)p.next_u32()) << 21;
    uint64_t lo = ((uint64_t)p.next_u32()) & ((1ull << 21) - 1);
    return (double)(hi | lo) * (1.0 / 9007199254740992.0);
  }
};
// FP32 kernels: one output per uniform, 24-bit mantissa in [0, 1).
struct RngF32 {
  Pcg32 p;
  TPT_HD explicit RngF32(uint64_t seed) : p(seed) {}
  TPT_HD float uni() { return (float)(p.next_u32() >> 8) * (1.0f / 16777216.0f); }
  // Rng::next_f64's 53-bit uniform (two outputs), where a decision taken in FP64 needs the
  // full-precision draw (the light self-occlusion test)
  TPT_HD double uni53() {
    uint64_t hi = ((uint64_t)p.next_u32()) << 21;
    uint64_t lo = ((uint64_t)p.next_u32()) & ((1ull << 21) - 1);
    return (double)(hi | lo) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace tpt
