// tpt_selfocc64.h -- the reference's shadow test of a sampled area-light point against the
// light's own surface, restated in FP64 for the FP32 kernels.
//
// nee_area_light (integrator.rs:178-213) shoots the shadow ray from hit_p + 1e-4 wi with
// tmax = dist - 1e-4, where dist = |pos - hit_p|: the light surface itself lies at t = tmax up to
// rounding.  World::hit accepts an instance hit when t_world < tmax (world.rs:104) and a sphere
// hit when t <= tmax (geometry.rs:324-329), so whether the light shadows its own sample is
// decided by the last bits of the f64 computation -- about 30% of samples are self-occluded on
// a rectangle light.  FP32 rounding gives a very different rate.  The FP32 kernels therefore
// trace their shadow ray with the light's own surface excluded (tmax pulled in by 2e-4
// relative, light_shadow_tmax32) and take the self-occlusion decision from
// light_self_occluded64, which repeats the reference's f64 operations in the reference's order.
//
// The decision's rate depends on the low bits of its inputs, so they must have the reference's
// provenance: the surface uniforms are 53-bit draws (RngF32::uni53), and the shading point is
// the one the reference's World::hit computes in FP64 for the primitive the FP32 traversal
// found, along the FP32 ray (hit_point64).  Fed the FP32 hit point o + t d and 24-bit uniforms
// instead, the decision's rate moved by +0.6 points on the Cornell box and +3 points under a
// rotated rectangle light (a CPU comparison with the FP64 oracle on the same samples), which
// darkened the direct light by 1-4%.  Contraction is off in every function here so each
// operation rounds as the reference's does.
#pragma once

#include "tpt_kernels.h"
#include "tpt_math.h"

namespace tpt {

struct D3 {
  double x, y, z;
};

namespace s64 {
TPT_HD D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
TPT_HD D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
TPT_HD D3 mul(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
TPT_HD double dot(D3 a, D3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
TPT_HD D3 cross(D3 a, D3 b) {
#pragma clang fp contract(off)
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
TPT_HD D3 mv(const double* m, D3 v) {  // Mat3::mul_vec, row-major
#pragma clang fp contract(off)
  return D3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
TPT_HD D3 ld(const double* p) { return D3{p[0], p[1], p[2]}; }
// v / l per component, each quotient correctly rounded as by a division, from one correctly
// rounded reciprocal y = 1 / l: q = v.x * y lies within an ulp of v.x / l, the residual
// v.x - q l is exact under an FMA, and q + residual * y rounded once is v.x / l rounded to
// nearest (Markstein's theorem: P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Develop. 34(1), 1990).  Valid while no quotient or
// residual underflows, which holds for these direction components (|v.x| <= l, and l far above
// the subnormal range); one division instead of three.  The residual is formed as q l - x and
// subtracted so that a zero component keeps its sign, as the division's quotient does.
TPT_HD double div_by(double x, double l, double y) {
  const double q = x * y;
  return fma(-fma(q, l, -x), y, q);
}
TPT_HD D3 div3(D3 v, double l) {
  const double y = 1.0 / l;
  return D3{div_by(v.x, l, y), div_by(v.y, l, y), div_by(v.z, l, y)};
}
}  // namespace s64

// World::hit on triangle `tri` (mesh-local, original order) of instance `inst`
// (world.rs:92-108, Triangle::hit geometry.rs:236-273 on the static vertices): the world
// distance along d and the world hit point, or false when the triangle test rejects the ray.
TPT_HD bool ref_instance_tri64(const SceneF64& S, int inst, int tri, D3 o, D3 d, double tmin, double* t_world,
                               D3* p_world) {
#pragma clang fp contract(off)
  using namespace s64;
  const InstF64& I = S.inst[inst];
  const double* tv = S.tris + 9 * (size_t)(S.meshes[I.mesh].tri_off + tri);
  const D3 v0 = ld(tv), v1 = ld(tv + 3), v2 = ld(tv + 6);
  const D3 t = ld(I.t);
  D3 o_obj = mv(I.a_inv, sub(o, t));
  D3 d_obj = mv(I.a_inv, d);
  double l = fmax(sqrt(dot(d_obj, d_obj)), 1e-30);
  d_obj = div3(d_obj, l);
  D3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  D3 pvec = cross(d_obj, e2);
  double det = dot(e1, pvec);
  if (fabs(det) < 1e-10) return false;
  double inv_det = 1.0 / det;
  D3 tvec = sub(o_obj, v0);
  double uu = dot(tvec, pvec) * inv_det;
  if (uu < 0.0 || uu > 1.0) return false;
  D3 qvec = cross(tvec, e1);
  double vv = dot(d_obj, qvec) * inv_det;
  if (vv < 0.0 || uu + vv > 1.0) return false;
  double th = dot(e2, qvec) * inv_det;
  if (th < tmin || th > 1e30) return false;
  D3 p_obj = add(o_obj, mul(d_obj, th));
  *p_world = add(mv(I.a, p_obj), t);
  *t_world = dot(sub(*p_world, o), d);
  return true;
}

// Sphere::hit (geometry.rs:174-195) on the world ray over [tmin, tmax]: t and the point o + d t.
TPT_HD bool ref_sphere64(const double* s4, D3 o, D3 d, double tmin, double tmax, double* t_out, D3* p) {
#pragma clang fp contract(off)
  using namespace s64;
  const D3 c = ld(s4);
  const double rs = s4[3];
  D3 oc = sub(o, c);
  double a = dot(d, d);
  double b = dot(oc, d);
  double cc = dot(oc, oc) - rs * rs;
  double disc = b * b - a * cc;
  if (disc <= 0.0) return false;
  double sd = sqrt(disc);
  double t = (-b - sd) / a;
  if (t < tmin || t > tmax) {
    t = (-b + sd) / a;
    if (t < tmin || t > tmax) return false;
  }
  *t_out = t;
  *p = add(o, mul(d, t));
  return true;
}

// The shading point as the reference's World::hit computes it for the primitive the FP32
// traversal found (instance code and triangle slot of the FP32 hit record), along the FP32 ray
// (o, d) promoted to FP64.  Falls back to the FP32 point p32 in the rare case that the FP64
// test rejects what the FP32 one accepted (an edge).
TPT_HD D3 hit_point64(const SceneF64& S, const int32_t* tri_orig, int32_t inst_code, int32_t slot, vec3<float> o,
                      vec3<float> d, vec3<float> p32) {
  const D3 o64{(double)o.x, (double)o.y, (double)o.z}, d64{(double)d.x, (double)d.y, (double)d.z};
  const double eps = 1e-4;
  D3 p{(double)p32.x, (double)p32.y, (double)p32.z};
  double t;
  if ((uint32_t)inst_code & TPT_ITEM_SPHERE) {
    const int s = (int)((uint32_t)inst_code & ~TPT_ITEM_SPHERE);
    D3 q;
    if (ref_sphere64(S.sph + 4 * s, o64, d64, eps, 1e30, &t, &q)) p = q;
  } else {
    const int inst = (int)((uint32_t)inst_code & ~TPT_HIT_FLAT);
    D3 q;
    if (ref_instance_tri64(S, inst, tri_orig[slot], o64, d64, eps, &t, &q)) p = q;
  }
  return p;
}

// Light index `idx`, shading point hit_p, light-sample uniforms (u, v): true when the light's
// own surface blocks the reference's shadow ray to the sample.
TPT_HD bool light_self_occluded64(const SceneF64& S, const LightsDev& Ls, int idx, D3 hit_p, double u, double v) {
#pragma clang fp contract(off)
  using namespace s64;
  const double eps = 1e-4;
  const LightF64& L = Ls.l64[idx];
  if (L.kind == 0) {
    // Light::sample_surface for a sphere (world.rs:244-255), then Sphere::hit (geometry.rs:174-195)
    const double* s = S.sph + 4 * L.idx;
    const D3 c = ld(s);
    const double rs = s[3];
    double z = 1.0 - 2.0 * u;
    double r = sqrt(fmax(1.0 - z * z, 0.0));
    double phi = 6.28318530717958647692 * v;
    D3 dir{r * cos(phi), z, r * sin(phi)};
    D3 pos = add(c, mul(dir, rs));
    D3 tl = sub(pos, hit_p);
    double dist2 = dot(tl, tl);
    double dist = sqrt(dist2);
    D3 wi = div3(tl, dist);
    D3 o = add(hit_p, mul(wi, eps));
    double tmax = fmax(dist - eps, eps);
    double t;
    D3 q;
    return ref_sphere64(s, o, wi, eps, tmax, &t, &q);
  }
  // triangle light: tri_world_verts (world.rs:276-292), sample point (world.rs:256-269)
  const InstF64& I = S.inst[L.inst];
  const double* tv = S.tris + 9 * (size_t)(S.meshes[L.mesh].tri_off + L.tri);
  const D3 v0 = ld(tv), v1 = ld(tv + 3), v2 = ld(tv + 6);
  const D3 t = ld(I.t);
  const D3 w0 = add(mv(I.a, v0), t), w1 = add(mv(I.a, v1), t), w2 = add(mv(I.a, v2), t);
  double su = sqrt(u);
  double b0 = 1.0 - su, b1 = v * su, b2 = 1.0 - b0 - b1;
  D3 pos = add(add(mul(w0, b0), mul(w1, b1)), mul(w2, b2));
  D3 tl = sub(pos, hit_p);
  double dist2 = dot(tl, tl);
  double dist = sqrt(dist2);
  D3 wi = div3(tl, dist);
  D3 o = add(hit_p, mul(wi, eps));
  double tmax = fmax(dist - eps, eps);
  // World::hit for this instance (world.rs:92-108), Triangle::hit on the sampled triangle
  double t_world;
  D3 p_world;
  if (!ref_instance_tri64(S, L.inst, L.tri, o, wi, eps, &t_world, &p_world)) return false;
  return t_world > eps && t_world < tmax;
}

// Far end of the FP32 shadow ray towards a sampled light point at distance `dist`: the
// reference's tmax = max(dist - 1e-4, 1e-4) pulled in by 2e-4 relative (at least 4e-6), which
// keeps the FP32 ray off the light's own surface for all but grazing (cos < ~1e-3) arrivals.
__device__ __forceinline__ float light_shadow_tmax32(float dist) {
  const float eps = 1e-4f;
  return fmaxf(dist - eps, eps) - fmaxf(2e-4f * dist, 4e-6f);
}

}  // namespace tpt
