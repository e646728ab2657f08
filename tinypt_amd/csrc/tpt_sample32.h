// tpt_sample32.h -- FP32 environment and light sampling shared by the wavefront render and
// the path-replay adjoint (env.rs, world.rs:171-272, integrator.rs:25-28).
#pragma once

#include "tpt_bsdf.h"
#include "tpt_kernels.h"
#include "tpt_trace32.h"

namespace tpt {

typedef vec3<float> f3;

__device__ __forceinline__ f3 f3of(float4 a) { return mk(a.x, a.y, a.z); }

// ---------------------------------------------------------------- environment (env.rs), FP32
__device__ __forceinline__ f3 sky32(f3 d) {  // integrator.rs:25-28
  float t = 0.5f * (d.y + 1.0f);
  return (1.0f - t) * mk(0.9f, 0.95f, 1.0f) + t * mk(0.6f, 0.75f, 1.0f);
}
__device__ __forceinline__ void dir_uv(f3 dir, float* u, float* v, float* theta) {
  f3 d = vnorm(dir);
  float th = acosf(fminf(fmaxf(d.y, -1.f), 1.f));
  float phi = atan2f(d.z, d.x);
  if (phi < 0.f) phi += (float)TPT_TAU;
  *u = phi / (float)TPT_TAU;
  *v = th / (float)TPT_PI;
  *theta = th;
}
// EnvMap::sample (env.rs:92-121): bilinear, x = u W (wraps), y = v (H - 1).
__device__ __forceinline__ f3 env_lookup32(const EnvDev& E, f3 dir) {
  float u, v, th;
  dir_uv(dir, &u, &v, &th);
  float x = u * (float)E.w;
  float y = v * ((float)E.h - 1.f);
  float fxf = floorf(x);
  int x0 = ((int)fxf) % E.w;
  if (x0 < 0) x0 += E.w;
  int x1 = (x0 + 1) % E.w;
  int y0 = (int)floorf(fminf(fmaxf(y, 0.f), (float)(E.h - 1)));
  int y1 = min(y0 + 1, E.h - 1);
  float fx = x - fxf, fy = y - floorf(y);
  float4 c00 = E.rgb32[y0 * E.w + x0], c10 = E.rgb32[y0 * E.w + x1];
  float4 c01 = E.rgb32[y1 * E.w + x0], c11 = E.rgb32[y1 * E.w + x1];
  f3 c0 = f3of(c00) * (1.f - fx) + f3of(c10) * fx;
  f3 c1 = f3of(c01) * (1.f - fx) + f3of(c11) * fx;
  return c0 * (1.f - fy) + c1 * fy;
}
// EnvMap::pdf (env.rs:156-182) with the per-cell numerator precomputed in FP64.
__device__ __forceinline__ float env_pdf32(const EnvDev& E, f3 dir) {
  if (!(E.total > 0.0)) return 1.0f / (4.0f * (float)TPT_PI);
  float u, v, th;
  dir_uv(dir, &u, &v, &th);
  int x = (int)fminf(fmaxf(floorf(u * (float)E.w), 0.f), (float)(E.w - 1));
  int y = (int)fminf(fmaxf(floorf(v * (float)E.h), 0.f), (float)(E.h - 1));
  return E.cell_pdf[y * E.w + x] / fmaxf(sinf(th), 1e-6f);
}
// EnvMap::sample_dir (env.rs:124-152): the cell is drawn from the alias table, which realises
// the same cell probabilities weight/total as the reference's row/column CDFs.
__device__ __forceinline__ void env_sample32(const EnvDev& E, RngF32& rng, f3* dir, f3* li, float* pdf) {
  if (!(E.total > 0.0)) {  // sample_uniform_env (env.rs:200-210)
    float u = rng.uni(), v = rng.uni();
    float z = 1.f - 2.f * u;
    float r = sqrtf(fmaxf(1.f - z * z, 0.f));
    float phi = (float)TPT_TAU * v;
    *dir = mk(r * sample_cos(phi), z, r * sample_sin(phi));
    *li = env_lookup32(E, *dir);
    *pdf = 1.0f / (4.0f * (float)TPT_PI);
    return;
  }
  const int N = E.w * E.h;
  int cell = min((int)(rng.uni() * (float)N), N - 1);
  if (rng.uni() >= E.alias_prob[cell]) cell = E.alias_idx[cell];
  int x = cell % E.w, y = cell / E.w;
  float u = ((float)x + rng.uni()) / (float)E.w;
  float v = ((float)y + rng.uni()) / (float)E.h;
  float theta = (float)TPT_PI * v, phi = (float)TPT_TAU * u;
  float st = sinf(theta);
  *dir = mk(cosf(phi) * st, cosf(theta), sinf(phi) * st);
  *li = env_lookup32(E, *dir);
  *pdf = env_pdf32(E, *dir);
}

// ---------------------------------------------------------------- lights (world.rs:171-272), FP32
struct LightSample32 {
  f3 pos, emit;
  float pdf;
  int mat;
  int idx;        // light drawn
  double u, v;    // its surface uniforms (53-bit draws; the FP32 position uses them rounded)
};
// The light a sample draws (world.rs:173-176): its index from one uniform and the CDF.
__device__ __forceinline__ int light_pick32(const LightsDev& Ls, RngF32& rng) {
  float r = rng.uni() * Ls.total32;
  return min(cdf_search(Ls.cdf32, Ls.n, r), Ls.n - 1);
}
// Light sample of world.rs:171-209.  The surface uniforms are 53-bit draws, like the
// reference's: the FP64 self-occlusion decision (tpt_selfocc64.h) depends on their low bits,
// and regenerates them from the stream state before this call (light_pick32, then two uni53).
__device__ __forceinline__ bool sample_light32(const LightsDev& Ls, RngF32& rng, f3 p, LightSample32* out) {
  if (!(Ls.total32 > 0.f) || Ls.n == 0) return false;
  const int idx = light_pick32(Ls, rng);
  const LightF32& L = Ls.l32[idx];
  float pdf_select = L.emit[3] / Ls.total32;
  f3 pos, nrm;
  float area;
  const double u53 = rng.uni53(), v53 = rng.uni53();
  const float u = (float)u53, v = (float)v53;
  if (L.kind == 0) {
    float z = 1.f - 2.f * u;
    float rr = sqrtf(fmaxf(1.f - z * z, 0.f));
    float phi = (float)TPT_TAU * v;
    f3 dir = mk(rr * sample_cos(phi), z, rr * sample_sin(phi));
    pos = ld3(L.p0) + dir * L.p0[3];
    nrm = dir;
    area = 4.f * (float)TPT_PI * L.p0[3] * L.p0[3];
  } else {
    f3 a = ld3(L.p0), b = ld3(L.p1), c = ld3(L.p2);
    float su = sqrtf(u);
    float b0 = 1.f - su, b1 = v * su, b2 = 1.f - b0 - b1;
    pos = a * b0 + b * b1 + c * b2;
    f3 n = cross(b - a, c - a);
    area = 0.5f * vlen(n);
    nrm = area > 0.f ? n / (2.f * area) : mk(0.f, 1.f, 0.f);
  }
  if (area <= 0.f) return false;
  f3 to_light = pos - p;
  float dist2 = dot(to_light, to_light);
  if (dist2 <= 1e-12f) return false;
  float dist = sqrtf(dist2);
  f3 wi = to_light / dist;
  float cos_light = fmaxf(dot(nrm, -wi), 0.f);
  if (cos_light <= 0.f) return false;
  float pdf = pdf_select * ((1.f / area) * dist2 / cos_light);
  if (!(pdf > 0.f)) return false;
  out->pos = pos;
  out->emit = mk(L.emit[0], L.emit[1], L.emit[2]);
  out->pdf = pdf;
  out->mat = L.mat;
  out->idx = idx;
  out->u = u53;
  out->v = v53;
  return true;
}

// ---------------------------------------------------------------- camera (ray.rs:91-120), FP32
// Pixel jitter (render.rs:153-157) then Camera::ray.  The shutter time only moves motion-blurred
// triangles, and the scene format has none, so it is not drawn.
__device__ __forceinline__ void camera_ray32(const CamF32& C, uint32_t x, uint32_t y, float inv_w, float inv_h, RngF32& rng,
                                             f3* o, f3* d) {
  float jx = rng.uni(), jy = rng.uni();
  float sx = ((float)x + jx) * inv_w * 2.f - 1.f;
  float sy = 1.f - ((float)y + jy) * inv_h * 2.f;
  f3 dir_base = vnorm(-ld3(C.w) + (sx * C.half_w) * ld3(C.u) + (sy * C.half_h) * ld3(C.v));
  f3 origin = ld3(C.o);
  f3 dir = dir_base;
  if (C.lens_radius > 0.f) {
    float dx = 0.f, dy = 0.f;
    for (int tries = 0; tries < 256; tries++) {  // sample_unit_disk (ray.rs:112-120)
      float a = 2.f * rng.uni() - 1.f, b = 2.f * rng.uni() - 1.f;
      if (a * a + b * b < 1.f) { dx = a; dy = b; break; }
    }
    origin = origin + ld3(C.u) * (dx * C.lens_radius) + ld3(C.v) * (dy * C.lens_radius);
    dir = vnorm((ld3(C.o) + dir_base * C.focus_dist) - origin);
  }
  *o = origin;
  *d = dir;
}

// Hit point, geometric normal (world.rs:99-100 / geometry.rs:331) and effective material
// (world.rs:106) of an FP32 hit record.
__device__ __forceinline__ void hit_geometry32(const SceneF32& S, int32_t inst, int32_t prim, float t, f3 o, f3 d, f3* p,
                                               f3* ng, int* mat) {
  *p = o + d * t;
  if ((uint32_t)inst & TPT_ITEM_SPHERE) {
    int s = (int)((uint32_t)inst & ~TPT_ITEM_SPHERE);
    const double* sp = S.sph + 4 * s;
    *ng = (*p - mk((float)sp[0], (float)sp[1], (float)sp[2])) / (float)sp[3];
    *mat = S.sph_mat[s];
  } else if ((uint32_t)inst & TPT_HIT_FLAT) {
    // world-space triangle: e1 x e2 (the direction of the record's row 2), negated where the
    // instance's A reverses orientation
    const float4* tp = reinterpret_cast<const float4*>(S.tris + prim);
    f3 no = vnorm(f3of(tp[2]));
    *ng = (S.tri_inst[prim] & (int32_t)TPT_TRI_FLIP) ? -no : no;
    *mat = S.tri_mat[prim];
  } else {
    const InstF32& I = S.inst[inst];
    const float4* tp = reinterpret_cast<const float4*>(S.tris + prim);
    f3 no = f3of(tp[2]);  // along e1 x e2 (the record's row 2)
    const float4* nm = reinterpret_cast<const float4*>(I.nrm);
    float4 n0 = nm[0], n1 = nm[1], n2 = nm[2];
    *ng = vnorm(mk(n0.x * no.x + n0.y * no.y + n0.z * no.z, n1.x * no.x + n1.y * no.y + n1.z * no.z,
                   n2.x * no.x + n2.y * no.y + n2.z * no.z));
    *mat = I.mat_override >= 0 ? I.mat_override : S.tri_mat[prim];
  }
}

}  // namespace tpt
