// tpt_oracle.cpp -- CPU FP64 oracle for the tinypt path-tracing hot path.
//
// TEST INFRASTRUCTURE, NOT PRODUCT.  Only the test suite (tests/), __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker / CPU baseline.
// Nothing in tinypt_amd/ links, loads or calls it.
//
// It restates the reference renderer's CPU path in IEEE binary64, one function per reference
// function, in the reference's operation order (no FMA contraction: build with
// -ffp-contract=off; glibc libm for sin/cos/tan/exp/log/acos/atan2/pow, as Rust's std uses on
// Linux).  Rust semantics kept: f64::max/min ignore NaN (fmax/fmin), `as usize` saturates
// (NaN -> 0), sort_by_key is stable (std::stable_sort), Aabb::empty is +-inf.
// Additions that do not change results: primitive IDs on hits (kind / instance / primitive),
// per-thread work counters, optional per-pixel sums of squares, and an optional per-sample
// seeding mode.
//
// Files followed: rng.rs, math.rs, ray.rs, geometry.rs, bvh.rs, transform.rs, world.rs,
// material.rs, env.rs, integrator.rs, render.rs:42-299, output.rs:96-102, constants.rs,
// checkpoint.rs:48-138.
//
// How each part is pinned without the reference binary (tests/test_oracle_*.py,
// tests/test_loader.py, tests/test_output.py): closed forms, exact rational arithmetic, brute
// force, the reference's own unit tests restated, and independent restatements of the integer
// specifications.  Every function exported below has such a check; parity unpinned: only the
// bit equality of the whole with the Rust binary, which no build here can run.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../include/tinypt_gpu.h"

namespace orc {

static const double PI = 3.14159265358979323846;
static const double TAU = 6.28318530717958647692;
static const double INF = std::numeric_limits<double>::infinity();

// constants.rs:4-24
static const double RAY_EPSILON = 1e-4;
static const double RAY_T_MAX = 1e30;
static const double FIREFLY_CLAMP = 50.0;
static const size_t SAH_BINS = 8;
static const size_t LEAF_SIZE = 4;

// Rust `f64 as usize`: saturating, NaN -> 0.
static inline size_t as_usize(double v) {
  if (!(v > 0.0)) return 0;  // NaN, negative, zero
  if (v >= 18446744073709551615.0) return SIZE_MAX;
  return (size_t)v;
}
// Rust f64::max / f64::min (NaN-ignoring)
static inline double fmx(double a, double b) { return std::fmax(a, b); }
static inline double fmn(double a, double b) { return std::fmin(a, b); }
// math.rs:251 clamp(x, a, b) = x.max(a).min(b)
static inline double clampd(double x, double a, double b) { return fmn(fmx(x, a), b); }
// std f64::clamp (env.rs:393-394): NaN stays NaN
static inline double std_clamp(double x, double a, double b) {
  if (x < a) x = a;
  if (x > b) x = b;
  return x;
}

// ---------------------------------------------------------------- math.rs
struct V3 {
  double x, y, z;
};
static inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator*(double s, V3 a) { return a * s; }  // math.rs:165 (v * self)
static inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // math.rs:57
static inline V3 cross(V3 a, V3 b) {                                                 // math.rs:59-65
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len(V3 a) { return std::sqrt(dot(a, a)); }  // math.rs:67
static inline V3 norm(V3 a) {                                      // math.rs:69-72
  double l = fmx(len(a), 1e-30);
  return a / l;
}
static inline V3 had(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double lum(V3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }  // math.rs:40
static inline V3 clamp_lum(V3 c, double m) {                                             // math.rs:42-49
  double l = lum(c);
  if (l > m && l > 0.0) return c * (m / l);
  return c;
}
double srgb_to_linear(double c) {  // math.rs:176-182
  if (c <= 0.04045) return c / 12.92;
  return std::pow((c + 0.055) / 1.055, 2.4);
}
double linear_to_srgb(double c) {  // math.rs:185-191
  if (c <= 0.0031308) return 12.92 * c;
  return 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
}
static inline V3 reflect(V3 v, V3 n) { return v - (2.0 * dot(v, n)) * n; }  // math.rs:193
static inline bool refract(V3 v, V3 n, double eta, V3* out) {                // math.rs:199-207
  double cosi = fmn(fmx(dot(-v, n), -1.0), 1.0);
  double sin2t = eta * eta * (1.0 - cosi * cosi);
  if (sin2t >= 1.0) return false;
  double cost = std::sqrt(1.0 - sin2t);
  *out = eta * v + (eta * cosi - cost) * n;
  return true;
}

struct Mat3 {  // math.rs:100-170
  double m[3][3];
  static Mat3 identity() {
    Mat3 r{};
    r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0;
    return r;
  }
  Mat3 mul(const Mat3& b) const {
    Mat3 r{};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = m[i][0] * b.m[0][j] + m[i][1] * b.m[1][j] + m[i][2] * b.m[2][j];
    return r;
  }
  V3 mul_vec(V3 v) const {
    return v3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
  }
  Mat3 transpose() const {
    Mat3 r{};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = m[j][i];
    return r;
  }
  Mat3 invert() const {  // cofactor inverse, math.rs:134-150
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double inv_det = 1.0 / det;
    Mat3 r{};
    r.m[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv_det;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv_det;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv_det;
    r.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv_det;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv_det;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv_det;
    r.m[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv_det;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv_det;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv_det;
    return r;
  }
};

// ---------------------------------------------------------------- transform.rs:11-93
struct Transform {
  Mat3 a, a_inv, normal_mat;
  V3 t;
  static Transform from_affine(const Mat3& a, V3 t) {
    Transform x;
    x.a = a;
    x.t = t;
    x.a_inv = a.invert();
    x.normal_mat = x.a_inv.transpose();
    return x;
  }
  V3 apply_point(V3 p) const { return a.mul_vec(p) + t; }
  V3 apply_point_inv(V3 p) const { return a_inv.mul_vec(p - t); }
  V3 apply_vec_inv(V3 v) const { return a_inv.mul_vec(v); }
  V3 apply_normal(V3 n) const { return norm(normal_mat.mul_vec(n)); }
};

// ---------------------------------------------------------------- work counters (additions)
struct Counters {
  uint64_t aabb = 0, tri = 0, sph = 0, inst = 0, rays_ext = 0, rays_shadow = 0, paths = 0, ext_bounces = 0;
  // the share of aabb / tri / sph / inst spent on extension (closest-hit) rays
  uint64_t aabb_ext = 0, tri_ext = 0, sph_ext = 0, inst_ext = 0;
  // area-light samples with a non-zero contribution that reach the shadow test, and those of
  // them whose shadow ray the sampled light's own surface blocks (other geometry ignored)
  uint64_t light_samples = 0, light_self = 0;
  void add(const Counters& o) {
    aabb += o.aabb; tri += o.tri; sph += o.sph; inst += o.inst;
    rays_ext += o.rays_ext; rays_shadow += o.rays_shadow; paths += o.paths; ext_bounces += o.ext_bounces;
    aabb_ext += o.aabb_ext; tri_ext += o.tri_ext; sph_ext += o.sph_ext; inst_ext += o.inst_ext;
    light_samples += o.light_samples; light_self += o.light_self;
  }
};
static thread_local Counters* tl_cnt = nullptr;
#define CNT(field) \
  do {             \
    if (tl_cnt) tl_cnt->field++; \
  } while (0)

// ---------------------------------------------------------------- rng.rs:9-47
struct Rng {
  uint64_t state;
  explicit Rng(uint64_t seed) { state = seed == 0 ? 0x9E3779B97F4A7C15ull : seed; }
  uint32_t next_u32() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  double next_f64() {
    uint64_t hi = ((uint64_t)next_u32()) << 21;
    uint64_t lo = ((uint64_t)next_u32()) & ((1ull << 21) - 1);
    uint64_t u = hi | lo;
    return (double)u * (1.0 / (double)(1ull << 53));
  }
};
static inline uint64_t seed_for(uint32_t px, uint32_t py, uint32_t s) {
  uint64_t h = 0xD1B54A32D192ED03ull;
  h ^= (uint64_t)px * 0x9E3779B185EBCA87ull;
  h ^= (uint64_t)py * 0xC2B2AE3D27D4EB4Full;
  h ^= (uint64_t)s * 0x165667B19E3779F9ull;
  return h ^ (h >> 32);
}

// ---------------------------------------------------------------- ray.rs
struct Ray {
  V3 o, d;
  double time;
  V3 at(double t) const { return o + d * t; }
};
struct Camera {  // ray.rs:22-33
  V3 o, u, v, w;
  double half_h, half_w, lens_radius, focus_dist, shutter_open, shutter_close;
  // look_at_dof (ray.rs:59-86); look_at (ray.rs:36-56) equals it with focus |eye-target|, aperture 0
  static Camera look_at_dof(V3 eye, V3 target, V3 up, double fov_deg, double aspect, double focus, double aperture) {
    Camera c;
    double theta = fov_deg * (PI / 180.0);
    c.half_h = std::tan(theta * 0.5);
    c.half_w = aspect * c.half_h;
    c.w = norm(eye - target);
    c.u = norm(cross(up, c.w));
    c.v = cross(c.w, c.u);
    c.o = eye;
    c.lens_radius = 0.5 * fmx(aperture, 0.0);
    c.focus_dist = fmx(focus, 1e-6);
    c.shutter_open = 0.0;
    c.shutter_close = 1.0;
    return c;
  }
  Ray ray(double sx, double sy, Rng& rng) const {  // ray.rs:91-108
    V3 dir_base = norm(((-w) + (sx * half_w) * u) + (sy * half_h) * v);
    V3 lens_offset = v3(0, 0, 0);
    if (lens_radius > 0.0) {
      double dx, dy;
      for (;;) {  // sample_unit_disk ray.rs:112-120
        double x = 2.0 * rng.next_f64() - 1.0;
        double y = 2.0 * rng.next_f64() - 1.0;
        if (x * x + y * y < 1.0) { dx = x; dy = y; break; }
      }
      lens_offset = u * (dx * lens_radius) + v * (dy * lens_radius);
    }
    V3 origin = o + lens_offset;
    V3 focus_point = o + dir_base * focus_dist;
    V3 dir = norm(focus_point - origin);
    double time = shutter_open + (shutter_close - shutter_open) * rng.next_f64();
    return Ray{origin, dir, time};
  }
};

// ---------------------------------------------------------------- geometry.rs
enum HitKind { HIT_NONE = 0, HIT_TRI = 1, HIT_SPHERE = 2 };
struct Hit {  // geometry.rs:14-23 + primitive IDs
  double t;
  V3 p, n;
  int mat_id;
  int kind, inst, prim;
};
struct Sphere {
  V3 c;
  double r;
  int mat_id;
  bool hit(const Ray& r_, double tmin, double tmax, Hit* h) const {  // geometry.rs:174-195
    CNT(sph);
    V3 oc = r_.o - c;
    double a = dot(r_.d, r_.d);
    double b = dot(oc, r_.d);
    double cc = dot(oc, oc) - r * r;
    double d = b * b - a * cc;
    if (d <= 0.0) return false;
    double sd = std::sqrt(d);
    double t = (-b - sd) / a;
    if (t < tmin || t > tmax) {
      t = (-b + sd) / a;
      if (t < tmin || t > tmax) return false;
    }
    V3 p = r_.at(t);
    h->t = t;
    h->p = p;
    h->n = (p - c) / r;
    h->mat_id = mat_id;
    return true;
  }
};
struct Aabb {
  V3 mn, mx;
  static Aabb empty() { return Aabb{v3(INF, INF, INF), v3(-INF, -INF, -INF)}; }
  Aabb grow(V3 p) const {
    return Aabb{v3(fmn(mn.x, p.x), fmn(mn.y, p.y), fmn(mn.z, p.z)), v3(fmx(mx.x, p.x), fmx(mx.y, p.y), fmx(mx.z, p.z))};
  }
  Aabb unite(const Aabb& b) const {
    return Aabb{v3(fmn(mn.x, b.mn.x), fmn(mn.y, b.mn.y), fmn(mn.z, b.mn.z)),
                v3(fmx(mx.x, b.mx.x), fmx(mx.y, b.mx.y), fmx(mx.z, b.mx.z))};
  }
  static Aabb from_points(V3 a, V3 b, V3 c) {  // geometry.rs:70-80
    Aabb bb = empty().grow(a).grow(b).grow(c);
    const double eps = 1e-9;
    bb.mn = v3(bb.mn.x - eps, bb.mn.y - eps, bb.mn.z - eps);
    bb.mx = v3(bb.mx.x + eps, bb.mx.y + eps, bb.mx.z + eps);
    return bb;
  }
  V3 centroid() const { return (mn + mx) * 0.5; }
  V3 extent() const { return mx - mn; }
  // geometry.rs:117-136 and 146-165: one slab routine; hit_inv == range_inv(...) != None
  bool range_inv(const Ray& r, V3 inv, double tmin, double tmax, double* t0out) const {
    CNT(aabb);
    double t0x = (mn.x - r.o.x) * inv.x, t1x = (mx.x - r.o.x) * inv.x;
    tmin = fmx(tmin, fmn(t0x, t1x));
    tmax = fmn(tmax, fmx(t0x, t1x));
    if (tmax <= tmin) return false;
    double t0y = (mn.y - r.o.y) * inv.y, t1y = (mx.y - r.o.y) * inv.y;
    tmin = fmx(tmin, fmn(t0y, t1y));
    tmax = fmn(tmax, fmx(t0y, t1y));
    if (tmax <= tmin) return false;
    double t0z = (mn.z - r.o.z) * inv.z, t1z = (mx.z - r.o.z) * inv.z;
    tmin = fmx(tmin, fmn(t0z, t1z));
    tmax = fmn(tmax, fmx(t0z, t1z));
    if (tmax <= tmin) return false;
    if (t0out) *t0out = tmin;
    return true;
  }
};
struct Triangle {  // geometry.rs:41-47 (static: open == close)
  V3 v0_0, v1_0, v2_0, v0_1, v1_1, v2_1;
  int mat_id;
  static Triangle new_static(V3 a, V3 b, V3 c, int mat) { return Triangle{a, b, c, a, b, c, mat}; }
  void vertices_at(double t, V3* a, V3* b, V3* c) const {  // geometry.rs:212-218
    *a = v0_0 * (1.0 - t) + v0_1 * t;
    *b = v1_0 * (1.0 - t) + v1_1 * t;
    *c = v2_0 * (1.0 - t) + v2_1 * t;
  }
  Aabb bounds() const {  // geometry.rs:221-225
    return Aabb::from_points(v0_0, v1_0, v2_0).unite(Aabb::from_points(v0_1, v1_1, v2_1));
  }
  bool hit(const Ray& r, double tmin, double tmax, Hit* h) const {  // geometry.rs:236-273
    CNT(tri);
    V3 v0, v1, v2;
    vertices_at(r.time, &v0, &v1, &v2);
    V3 e1 = v1 - v0, e2 = v2 - v0;
    V3 pvec = cross(r.d, e2);
    double det = dot(e1, pvec);
    const double EPS = 1e-10;
    if (std::fabs(det) < EPS) return false;
    double inv_det = 1.0 / det;
    V3 tvec = r.o - v0;
    double u = dot(tvec, pvec) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    V3 qvec = cross(tvec, e1);
    double v = dot(r.d, qvec) * inv_det;
    if (v < 0.0 || u + v > 1.0) return false;
    double thit = dot(e2, qvec) * inv_det;
    if (thit < tmin || thit > tmax) return false;
    h->t = thit;
    h->p = r.at(thit);
    h->n = norm(cross(e1, e2));
    h->mat_id = mat_id;
    return true;
  }
};

// ---------------------------------------------------------------- bvh.rs
struct BvhNode {  // bvh.rs:24-35
  Aabb bbox;
  int32_t left, right;
  uint32_t start, count;
};
struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<size_t> indices;

  static Bvh build(const std::vector<Triangle>& tris) {  // bvh.rs:51-242
    Bvh b;
    b.indices.resize(tris.size());
    for (size_t i = 0; i < tris.size(); i++) b.indices[i] = i;
    std::vector<Aabb> bounds(tris.size());
    std::vector<V3> cents(tris.size());
    for (size_t i = 0; i < tris.size(); i++) {
      bounds[i] = tris[i].bounds();
      cents[i] = bounds[i].centroid();
    }
    if (!b.indices.empty()) b.build_node(bounds, cents, 0, b.indices.size());
    return b;
  }

  static double axis_of(V3 c, int axis) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); }

  int32_t build_node(const std::vector<Aabb>& bounds, const std::vector<V3>& cents, size_t start, size_t end) {
    size_t n = end - start;
    Aabb bbox = Aabb::empty(), cbox = Aabb::empty();  // bounds_for_range bvh.rs:58-67
    for (size_t k = start; k < end; k++) {
      size_t i = indices[k];
      bbox = bbox.unite(bounds[i]);
      cbox = cbox.grow(cents[i]);
    }
    int32_t node_index = (int32_t)nodes.size();
    nodes.push_back(BvhNode{bbox, -1, -1, 0, 0});
    if (n <= LEAF_SIZE) {
      nodes[node_index].start = (uint32_t)start;
      nodes[node_index].count = (uint32_t)n;
      return node_index;
    }
    V3 ext = cbox.extent();  // bvh.rs:101-108
    int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    size_t mid;
    if (n >= SAH_BINS * 2) {  // bvh.rs:110-198
      double minc = axis_of(cbox.mn, axis), maxc = axis_of(cbox.mx, axis);
      double extent = maxc - minc;
      if (extent > 1e-12) {
        double inv_extent = 1.0 / extent;
        Aabb bin_box[SAH_BINS];
        size_t bin_cnt[SAH_BINS];
        for (size_t i = 0; i < SAH_BINS; i++) { bin_box[i] = Aabb::empty(); bin_cnt[i] = 0; }
        auto bin_of = [&](size_t idx) {
          double cv = axis_of(cents[idx], axis);
          size_t bi = as_usize((cv - minc) * inv_extent * (double)SAH_BINS);
          if (bi >= SAH_BINS) bi = SAH_BINS - 1;
          return bi;
        };
        for (size_t k = start; k < end; k++) {
          size_t idx = indices[k];
          size_t bi = bin_of(idx);
          bin_box[bi] = bin_box[bi].unite(bounds[idx]);
          bin_cnt[bi] += 1;
        }
        Aabb left_bbox[SAH_BINS], right_bbox[SAH_BINS];
        size_t left_count[SAH_BINS], right_count[SAH_BINS];
        Aabb acc = Aabb::empty();
        size_t accn = 0;
        for (size_t i = 0; i < SAH_BINS; i++) {
          if (bin_cnt[i] > 0) { acc = acc.unite(bin_box[i]); accn += bin_cnt[i]; }
          left_bbox[i] = acc;
          left_count[i] = accn;
        }
        acc = Aabb::empty();
        accn = 0;
        for (size_t ii = SAH_BINS; ii-- > 0;) {
          if (bin_cnt[ii] > 0) { acc = acc.unite(bin_box[ii]); accn += bin_cnt[ii]; }
          right_bbox[ii] = acc;
          right_count[ii] = accn;
        }
        double best_cost = INF;
        size_t best_split = 0;
        for (size_t i = 0; i < SAH_BINS - 1; i++) {  // bvh.rs:159-174
          size_t lc = left_count[i], rc = right_count[i + 1];
          if (lc == 0 || rc == 0) continue;
          V3 la = left_bbox[i].extent(), ra = right_bbox[i + 1].extent();
          double left_area = 2.0 * (la.x * la.y + la.y * la.z + la.z * la.x);
          double right_area = 2.0 * (ra.x * ra.y + ra.y * ra.z + ra.z * ra.x);
          double cost = left_area * (double)lc + right_area * (double)rc;
          if (cost < best_cost) { best_cost = cost; best_split = i + 1; }
        }
        if (std::isfinite(best_cost)) {
          std::stable_sort(indices.begin() + start, indices.begin() + end,
                           [&](size_t a, size_t b) { return bin_of(a) < bin_of(b); });
          size_t left_total = 0;
          for (size_t i = 0; i < best_split; i++) left_total += bin_cnt[i];
          mid = (left_total > 0 && left_total < n) ? start + left_total : start + n / 2;
        } else {
          mid = start + n / 2;
        }
      } else {
        mid = start + n / 2;
      }
    } else {
      mid = start + n / 2;
    }
    // bvh.rs:203-222 (mid == start || mid == end) cannot occur: n > LEAF_SIZE gives n/2 >= 2
    // and the SAH split is taken only when 0 < left_total < n.
    int32_t l = build_node(bounds, cents, start, mid);
    int32_t r = build_node(bounds, cents, mid, end);
    nodes[node_index].left = l;
    nodes[node_index].right = r;
    nodes[node_index].start = 0;
    nodes[node_index].count = 0;
    return node_index;
  }

  bool hit(const std::vector<Triangle>& tris, const Ray& r, double tmin, double tmax, Hit* best, int* best_ti) const {
    // bvh.rs:248-351
    if (nodes.empty()) return false;
    V3 inv = v3(1.0 / r.d.x, 1.0 / r.d.y, 1.0 / r.d.z);
    // bvh.rs:255-291: a 64-entry array plus a heap vector.  Once the array is full the heap
    // takes a copy of it and all later pushes/pops go to the heap; `sp` is not reset, so after
    // the heap drains the array's entries are popped a second time.  Kept as is.
    int32_t stack_buf[64];
    size_t sp = 0;
    stack_buf[sp++] = 0;
    std::vector<int32_t> heap_stack;
    auto push = [&](int32_t id) {
      if (heap_stack.empty()) {
        if (sp < 64) stack_buf[sp++] = id;
        else { heap_stack.assign(stack_buf, stack_buf + sp); heap_stack.push_back(id); }
      } else {
        heap_stack.push_back(id);
      }
    };
    bool found = false;
    for (;;) {
      int32_t nid;
      if (heap_stack.empty()) {
        if (sp == 0) break;
        nid = stack_buf[--sp];
      } else {
        nid = heap_stack.back();
        heap_stack.pop_back();
      }
      const BvhNode& n = nodes[(size_t)nid];
      if (!n.bbox.range_inv(r, inv, tmin, tmax, nullptr)) continue;
      if (n.left == -1 && n.right == -1) {
        for (size_t k = n.start; k < (size_t)n.start + n.count; k++) {
          size_t ti = indices[k];
          Hit h;
          if (tris[ti].hit(r, tmin, tmax, &h)) {
            tmax = h.t;
            *best = h;
            *best_ti = (int)ti;
            found = true;
          }
        }
      } else {
        int32_t a_id = n.left, b_id = n.right;
        if (a_id == -1) {
          if (b_id != -1) push(b_id);
          continue;
        }
        if (b_id == -1) {
          push(a_id);
          continue;
        }
        double a_t0 = 0, b_t0 = 0;
        bool ah = nodes[(size_t)a_id].bbox.range_inv(r, inv, tmin, tmax, &a_t0);
        bool bh = nodes[(size_t)b_id].bbox.range_inv(r, inv, tmin, tmax, &b_t0);
        if (ah && bh) {
          if (a_t0 <= b_t0) { push(b_id); push(a_id); }
          else { push(a_id); push(b_id); }
        } else if (ah) {
          push(a_id);
        } else if (bh) {
          push(b_id);
        }
      }
    }
    return found;
  }
};

// ---------------------------------------------------------------- material.rs
enum MatKind { M_LAMBERT = 0, M_METAL = 1, M_DIELECTRIC = 2, M_GGX = 3, M_SUBSURFACE = 4, M_LIGHT = 5 };
struct Material {
  int kind;
  V3 c;      // albedo / absorption / emit
  double s;  // ior / alpha / scatter_dist
  bool is_delta() const { return kind == M_METAL || kind == M_DIELECTRIC; }
  bool emitted(V3* e) const {
    if (kind == M_LIGHT) { *e = c; return true; }
    return false;
  }
};
struct BsdfSample {
  Ray scattered;
  V3 weight;
  double pdf;
  bool is_delta;
};

static inline double pow5(double x) { double x2 = x * x; return x2 * x2 * x; }  // material.rs:259-262
static inline V3 fresnel_schlick(double cos_theta, V3 f0) {                      // material.rs:265-269
  V3 one = v3(1, 1, 1);
  double x = 1.0 - fmx(cos_theta, 0.0);
  return f0 + (one - f0) * pow5(x);
}
static inline double ggx_distribution(double alpha, double cos_h) {  // material.rs:272-279
  if (cos_h <= 0.0) return 0.0;
  double a2 = alpha * alpha;
  double denom = cos_h * cos_h * (a2 - 1.0) + 1.0;
  return a2 / (PI * denom * denom);
}
static inline double ggx_smith_g1(double alpha, double c) {  // material.rs:282-288
  if (c <= 0.0) return 0.0;
  double a = alpha * std::sqrt(fmx(1.0 - c * c, 0.0)) / fmx(c, 1e-6);
  return 2.0 / (1.0 + std::sqrt(1.0 + a * a));
}
static inline double ggx_smith(double alpha, double ci, double co) { return ggx_smith_g1(alpha, ci) * ggx_smith_g1(alpha, co); }
static V3 ggx_eval(V3 albedo, double alpha, V3 n, V3 wo, V3 wi) {  // material.rs:227-240
  double cos_i = dot(n, wo), cos_o = dot(n, wi);
  if (cos_i <= 0.0 || cos_o <= 0.0) return v3(0, 0, 0);
  V3 m = norm(wo + wi);
  double cos_h = fmx(dot(n, m), 0.0);
  double d = ggx_distribution(alpha, cos_h);
  double g = ggx_smith(alpha, cos_i, cos_o);
  V3 f = fresnel_schlick(std::fabs(dot(wi, m)), albedo);
  double denom = 4.0 * cos_i * cos_o + 1e-6;
  return f * (d * g / denom);
}
static double ggx_pdf(double alpha, V3 n, V3 wo, V3 wi) {  // material.rs:243-256
  double cos_i = dot(n, wo), cos_o = dot(n, wi);
  if (cos_i <= 0.0 || cos_o <= 0.0) return 0.0;
  V3 m = norm(wo + wi);
  double cos_h = fmx(dot(n, m), 0.0);
  double d = ggx_distribution(alpha, cos_h);
  double g1 = ggx_smith_g1(alpha, cos_i);
  double pdf_m = d * g1 * cos_h / fmx(cos_i, 1e-6);
  double denom = 4.0 * fmx(std::fabs(dot(wi, m)), 1e-6);
  return fmx(pdf_m / denom, 0.0);
}
static void tangent_frame(V3 n, V3* t, V3* b) {  // material.rs:313-318
  V3 a = std::fabs(n.x) > 0.9 ? v3(0, 1, 0) : v3(1, 0, 0);
  *t = norm(cross(n, a));
  *b = cross(*t, n);
}
static V3 sample_cosine_hemisphere(V3 n, Rng& rng) {  // material.rs:299-310
  double u = rng.next_f64();
  double v = rng.next_f64();
  double r = std::sqrt(u);
  double phi = 2.0 * PI * v;
  double x = r * std::cos(phi);
  double y = r * std::sin(phi);
  double z = std::sqrt(1.0 - u);
  V3 t, b;
  tangent_frame(n, &t, &b);
  return t * x + b * y + n * z;
}
static V3 sample_ggx_vndf(V3 wo, double alpha, double u1, double u2) {  // material.rs:327-347
  V3 v = norm(v3(alpha * wo.x, alpha * wo.y, wo.z));
  V3 t1, t2;
  if (v.z < 0.9999) {
    t1 = norm(v3(-v.y, v.x, 0.0));
    t2 = cross(v, t1);
  } else {
    t1 = v3(1, 0, 0);
    t2 = v3(0, 1, 0);
  }
  double r = std::sqrt(u1);
  double phi = TAU * u2;
  double t1p = r * std::cos(phi);
  double t2p_init = r * std::sin(phi);
  double s = 0.5 * (1.0 + v.z);
  double t2p = std::sqrt(1.0 - s) * t2p_init + s * std::sqrt(fmx(1.0 - t1p * t1p, 0.0));
  V3 nh = norm(t1 * t1p + t2 * t2p + v * std::sqrt(fmx(1.0 - t1p * t1p - t2p * t2p, 0.0)));
  return norm(v3(alpha * nh.x, alpha * nh.y, nh.z));
}

static bool mat_sample(const Material& m, const Ray& ray_in, const Hit& hit, Rng& rng, BsdfSample* out) {
  // material.rs:76-201
  bool entering = dot(hit.n, ray_in.d) < 0.0;
  V3 n = entering ? hit.n : -hit.n;
  switch (m.kind) {
    case M_LAMBERT: {
      V3 d = sample_cosine_hemisphere(n, rng);
      *out = BsdfSample{Ray{hit.p + RAY_EPSILON * d, d, ray_in.time}, m.c, fmx(dot(n, d), 0.0) / PI, false};
      return true;
    }
    case M_METAL: {
      V3 d = reflect(ray_in.d, n);
      *out = BsdfSample{Ray{hit.p + RAY_EPSILON * d, d, ray_in.time}, m.c, 0.0, true};
      return true;
    }
    case M_DIELECTRIC: {
      double ior = m.s;
      double eta = entering ? 1.0 / ior : ior;
      double cos_i = fmx(dot(-ray_in.d, n), 0.0);
      double f0 = ((ior - 1.0) / (ior + 1.0)) * ((ior - 1.0) / (ior + 1.0));
      double fresnel = f0 + (1.0 - f0) * pow5(1.0 - cos_i);
      V3 beta = v3(1, 1, 1);
      if (!entering) {
        double dist = fmx(hit.t, 0.0);
        V3 att = v3(std::exp(-m.c.x * dist), std::exp(-m.c.y * dist), std::exp(-m.c.z * dist));
        beta = had(beta, att);
      }
      V3 refl = reflect(ray_in.d, n);
      V3 refr;
      bool can_refr = refract(ray_in.d, n, eta, &refr);
      bool choose_refl = !can_refr || rng.next_f64() < fresnel;
      V3 d;
      if (choose_refl) {
        beta = beta * (fresnel / fmx(fresnel, 1e-6));
        d = refl;
      } else {
        double scale = (1.0 - fresnel) / fmx(1.0 - fresnel, 1e-6);
        beta = beta * scale * (eta * eta);
        d = refr;
      }
      *out = BsdfSample{Ray{hit.p + RAY_EPSILON * d, d, ray_in.time}, beta, 0.0, true};
      return true;
    }
    case M_GGX: {
      double alpha = fmx(m.s, 1e-3);
      V3 t, b;
      tangent_frame(n, &t, &b);
      V3 wo = norm(-ray_in.d);
      V3 wo_local = v3(dot(wo, t), dot(wo, b), dot(wo, n));
      if (wo_local.z <= 0.0) return false;
      double u1 = rng.next_f64();
      double u2 = rng.next_f64();
      V3 ml = sample_ggx_vndf(wo_local, alpha, u1, u2);
      V3 mm = t * ml.x + b * ml.y + n * ml.z;
      V3 d = reflect(ray_in.d, mm);
      double cos_o = dot(d, n);
      double cos_i = dot(-ray_in.d, n);
      if (cos_o <= 1e-6 || cos_i <= 1e-6) return false;
      double cos_h = fmx(dot(n, mm), 0.0);
      double dg = ggx_distribution(alpha, cos_h);
      double g = ggx_smith(alpha, cos_i, cos_o);
      V3 f = fresnel_schlick(std::fabs(dot(d, mm)), m.c);
      double denom = 4.0 * cos_i * cos_o + 1e-6;
      V3 spec = f * (dg * g / denom);
      double pdf = ggx_pdf(alpha, n, wo, d);
      *out = BsdfSample{Ray{hit.p + RAY_EPSILON * d, d, ray_in.time}, spec * (cos_o / fmx(pdf, 1e-6)), pdf, false};
      return true;
    }
    case M_SUBSURFACE: {
      double scale = fmx(m.s, 1e-4);
      double u = fmx(rng.next_f64(), 1e-12);
      double dist = -std::log(u) * scale;
      V3 normal_out = hit.n;
      V3 d = sample_cosine_hemisphere(normal_out, rng);
      V3 att = m.c * std::exp(-dist / scale);
      V3 origin = hit.p - normal_out * dist + RAY_EPSILON * d;
      *out = BsdfSample{Ray{origin, d, ray_in.time}, att, fmx(dot(n, d), 0.0) / PI, false};
      return true;
    }
    default:
      return false;
  }
}
static void mat_eval(const Material& m, V3 wo, V3 wi, V3 n, V3* f, double* pdf) {  // material.rs:207-223
  switch (m.kind) {
    case M_LAMBERT:
    case M_SUBSURFACE: {
      double c = fmx(dot(n, wi), 0.0);
      if (c <= 0.0) { *f = v3(0, 0, 0); *pdf = 0.0; }
      else { *f = m.c * (1.0 / PI); *pdf = c / PI; }
      return;
    }
    case M_GGX: {
      double a = fmx(m.s, 1e-3);
      *f = ggx_eval(m.c, a, n, wo, wi);
      *pdf = ggx_pdf(a, n, wo, wi);
      return;
    }
    default:
      *f = v3(0, 0, 0);
      *pdf = 0.0;
  }
}

// ---------------------------------------------------------------- env.rs
static size_t cdf_search(const double* cdf, size_t len, double x) {  // env.rs:185-197 == world.rs:324-336
  size_t lo = 0;
  size_t hi = len == 0 ? 0 : len - 1;
  while (lo + 1 < hi) {
    size_t mid = (lo + hi) / 2;
    if (cdf[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}
struct EnvMap {
  size_t width = 0, height = 0;
  std::vector<V3> data;
  std::vector<double> row_cdf, col_cdf;
  double total_weight = 0;

  static EnvMap from_pixels(size_t w, size_t h, std::vector<V3> data) {  // env.rs:61-89
    EnvMap e;
    e.width = w;
    e.height = h;
    e.row_cdf.assign(h + 1, 0.0);
    e.col_cdf.assign(h * (w + 1), 0.0);
    double total = 0.0;
    for (size_t y = 0; y < h; y++) {
      double theta0 = PI * (double)y / (double)h;
      double theta1 = PI * (double)(y + 1) / (double)h;
      double row_weight = fmx(std::cos(theta0) - std::cos(theta1), 0.0);
      size_t row_offset = y * (w + 1);
      double row_sum = 0.0;
      e.col_cdf[row_offset] = 0.0;
      for (size_t x = 0; x < w; x++) {
        double l = fmx(lum(data[y * w + x]), 0.0);
        row_sum += l * row_weight;
        e.col_cdf[row_offset + x + 1] = row_sum;
      }
      total += row_sum;
      e.row_cdf[y + 1] = total;
    }
    e.data = std::move(data);
    e.total_weight = total;
    return e;
  }
  V3 sample(V3 dir) const {  // env.rs:92-121
    V3 d = norm(dir);
    double theta = std::acos(clampd(d.y, -1.0, 1.0));
    double phi = std::atan2(d.z, d.x);
    if (phi < 0.0) phi += TAU;
    double u = phi / TAU;
    double v = theta / PI;
    double x = u * (double)width;
    double y = v * ((double)height - 1.0);
    size_t x0 = as_usize(std::floor(x)) % width;
    size_t x1 = (x0 + 1) % width;
    size_t y0 = as_usize(std::floor(clampd(y, 0.0, (double)(height - 1))));
    size_t y1 = std::min(y0 + 1, height - 1);
    double fx = x - std::floor(x);
    double fy = y - std::floor(y);
    V3 c00 = data[y0 * width + x0], c10 = data[y0 * width + x1];
    V3 c01 = data[y1 * width + x0], c11 = data[y1 * width + x1];
    V3 c0 = c00 * (1.0 - fx) + c10 * fx;
    V3 c1 = c01 * (1.0 - fx) + c11 * fx;
    return c0 * (1.0 - fy) + c1 * fy;
  }
  void sample_uniform(Rng& rng, V3* dir, V3* rad, double* pdf) const {  // env.rs:200-210
    double u = rng.next_f64();
    double v = rng.next_f64();
    double z = 1.0 - 2.0 * u;
    double r = std::sqrt(fmx(1.0 - z * z, 0.0));
    double phi = TAU * v;
    *dir = v3(r * std::cos(phi), z, r * std::sin(phi));
    *rad = sample(*dir);
    *pdf = 1.0 / (4.0 * PI);
  }
  void sample_dir(Rng& rng, V3* dir, V3* rad, double* pdf_out) const {  // env.rs:124-152
    if (total_weight <= 0.0) { sample_uniform(rng, dir, rad, pdf_out); return; }
    double r0 = rng.next_f64() * total_weight;
    size_t y = std::min(cdf_search(row_cdf.data(), row_cdf.size(), r0), height - 1);
    double row_start = row_cdf[y], row_end = row_cdf[y + 1];
    double row_sum = fmx(row_end - row_start, 0.0);
    if (row_sum <= 0.0) { sample_uniform(rng, dir, rad, pdf_out); return; }
    double r1 = rng.next_f64() * row_sum;
    size_t row_offset = y * (width + 1);
    size_t x = std::min(cdf_search(col_cdf.data() + row_offset, width + 1, r1), width - 1);
    double u = ((double)x + rng.next_f64()) / (double)width;
    double v = ((double)y + rng.next_f64()) / (double)height;
    double theta = PI * v;
    double phi = TAU * u;
    double sin_theta = std::sin(theta);
    *dir = v3(std::cos(phi) * sin_theta, std::cos(theta), std::sin(phi) * sin_theta);
    *rad = sample(*dir);
    *pdf_out = pdf(*dir);
  }
  double pdf(V3 dir) const {  // env.rs:156-182
    if (total_weight <= 0.0) return 1.0 / (4.0 * PI);
    V3 d = norm(dir);
    double theta = std::acos(clampd(d.y, -1.0, 1.0));
    double phi = std::atan2(d.z, d.x);
    if (phi < 0.0) phi += TAU;
    double u = phi / TAU;
    double v = theta / PI;
    size_t x = as_usize(std_clamp(std::floor(u * (double)width), 0.0, (double)(width - 1)));
    size_t y = as_usize(std_clamp(std::floor(v * (double)height), 0.0, (double)(height - 1)));
    double theta0 = PI * (double)y / (double)height;
    double theta1 = PI * (double)(y + 1) / (double)height;
    double sin_theta = fmx(std::sin(theta), 1e-6);
    double row_weight = fmx(std::cos(theta0) - std::cos(theta1), 0.0);
    double l = fmx(lum(data[y * width + x]), 0.0);
    double weight = l * row_weight;
    if (weight <= 0.0) return 0.0;
    double pdf_uv = (weight / total_weight) * (double)width * (double)height;
    return pdf_uv / (2.0 * PI * PI * sin_theta);
  }
};

// ---------------------------------------------------------------- world.rs
struct Mesh {
  std::vector<Triangle> tris;
  Bvh bvh;
};
struct Instance {
  int mesh_id;
  Transform xform;
  int mat_override;  // -1 = None
};
struct LightInfo {
  int kind;  // 0 sphere, 1 triangle
  int idx, mesh_id, tri_id, inst_id;
  V3 emit;
  double weight;
};
struct LightSample {
  V3 position, normal, emit;
  double pdf;
  size_t idx;  // the light drawn (for the work counters)
};

struct Scene {
  Camera cam;
  std::vector<Sphere> spheres;
  std::vector<Mesh> meshes;
  std::vector<Instance> instances;
  std::vector<LightInfo> lights;
  std::vector<double> light_cdf;
  double light_total = 0;
  std::vector<Material> mats;
  bool has_env = false;
  EnvMap env;

  bool tri_world_verts(int mesh_id, int tri_id, int inst_id, double time, V3* a, V3* b, V3* c) const {
    // world.rs:276-292
    if (mesh_id < 0 || (size_t)mesh_id >= meshes.size()) return false;
    const Mesh& m = meshes[(size_t)mesh_id];
    if (tri_id < 0 || (size_t)tri_id >= m.tris.size()) return false;
    if (inst_id < 0 || (size_t)inst_id >= instances.size()) return false;
    const Instance& in = instances[(size_t)inst_id];
    V3 v0, v1, v2;
    m.tris[(size_t)tri_id].vertices_at(time, &v0, &v1, &v2);
    *a = in.xform.apply_point(v0);
    *b = in.xform.apply_point(v1);
    *c = in.xform.apply_point(v2);
    return true;
  }
  double light_area(const LightInfo& L, double time) const {  // world.rs:226-238
    if (L.kind == 0) {
      if (L.idx < 0 || (size_t)L.idx >= spheres.size()) return 0.0;
      double r = spheres[(size_t)L.idx].r;
      return 4.0 * PI * r * r;
    }
    V3 a, b, c;
    if (!tri_world_verts(L.mesh_id, L.tri_id, L.inst_id, time, &a, &b, &c)) return 0.0;
    return 0.5 * len(cross(b - a, c - a));
  }
  void build_lights() {  // world.rs:127-167
    lights.clear();
    light_cdf.clear();
    double total = 0.0;
    auto add = [&](LightInfo L, V3 emit) {
      double area = light_area(L, 0.5);
      double weight = area * lum(emit);
      if (weight > 0.0) {
        total += weight;
        L.emit = emit;
        L.weight = weight;
        lights.push_back(L);
        light_cdf.push_back(total);
      }
    };
    V3 e;
    for (size_t i = 0; i < spheres.size(); i++) {
      int mid = spheres[i].mat_id;
      if (mid >= 0 && (size_t)mid < mats.size() && mats[(size_t)mid].emitted(&e)) add(LightInfo{0, (int)i, 0, 0, 0, {}, 0}, e);
    }
    for (size_t ii = 0; ii < instances.size(); ii++) {
      const Instance& in = instances[ii];
      if (in.mesh_id < 0 || (size_t)in.mesh_id >= meshes.size()) continue;
      const Mesh& m = meshes[(size_t)in.mesh_id];
      for (size_t ti = 0; ti < m.tris.size(); ti++) {
        int mid = in.mat_override >= 0 ? in.mat_override : m.tris[ti].mat_id;
        if (mid >= 0 && (size_t)mid < mats.size() && mats[(size_t)mid].emitted(&e))
          add(LightInfo{1, 0, in.mesh_id, (int)ti, (int)ii, {}, 0}, e);
      }
    }
    light_total = total;
  }
  bool sample_surface(const LightInfo& L, double time, Rng& rng, V3* pos, V3* nrm, double* area) const {
    // world.rs:242-272: a sphere light draws (u, v) first; a triangle light only once its
    // world vertices exist
    if (L.kind == 0) {
      if (L.idx < 0 || (size_t)L.idx >= spheres.size()) return false;
    } else {
      V3 a, b, c;
      if (!tri_world_verts(L.mesh_id, L.tri_id, L.inst_id, time, &a, &b, &c)) return false;
    }
    double u = rng.next_f64();
    double v = rng.next_f64();
    return surface_at(L, time, u, v, pos, nrm, area);
  }
  // the point of light L at surface uniforms (u, v) (world.rs:244-269)
  bool surface_at(const LightInfo& L, double time, double u, double v, V3* pos, V3* nrm, double* area) const {
    if (L.kind == 0) {
      if (L.idx < 0 || (size_t)L.idx >= spheres.size()) return false;
      const Sphere& s = spheres[(size_t)L.idx];
      double z = 1.0 - 2.0 * u;
      double r = std::sqrt(fmx(1.0 - z * z, 0.0));
      double phi = TAU * v;
      V3 dir = v3(r * std::cos(phi), z, r * std::sin(phi));
      *pos = s.c + dir * s.r;
      *nrm = dir;
      *area = 4.0 * PI * s.r * s.r;
      return true;
    }
    V3 a, b, c;
    if (!tri_world_verts(L.mesh_id, L.tri_id, L.inst_id, time, &a, &b, &c)) return false;
    double su = std::sqrt(u);
    double b0 = 1.0 - su;
    double b1 = v * su;
    double b2 = 1.0 - b0 - b1;
    *pos = a * b0 + b * b1 + c * b2;
    V3 n = cross(b - a, c - a);
    double ar = 0.5 * len(n);
    *nrm = ar > 0.0 ? n / (2.0 * ar) : v3(0, 1, 0);
    *area = ar;
    return true;
  }
  bool sample_light(Rng& rng, double time, V3 p, LightSample* out) const {  // world.rs:171-209
    if (light_total <= 0.0 || lights.empty()) return false;
    double r = rng.next_f64() * light_total;
    size_t idx = std::min(cdf_search(light_cdf.data(), light_cdf.size(), r), lights.size() - 1);
    const LightInfo& info = lights[idx];
    double pdf_select = info.weight / light_total;
    V3 pos, nrm;
    double area;
    if (!sample_surface(info, time, rng, &pos, &nrm, &area)) return false;
    if (area <= 0.0) return false;
    V3 to_light = pos - p;
    double dist2 = dot(to_light, to_light);
    if (dist2 <= 1e-12) return false;
    double dist = std::sqrt(dist2);
    V3 wi = to_light / dist;
    double cos_light = fmx(dot(nrm, -wi), 0.0);
    if (cos_light <= 0.0) return false;
    double pdf_area = 1.0 / area;
    double pdf_omega = pdf_area * dist2 / cos_light;
    double pdf = pdf_select * pdf_omega;
    if (pdf <= 0.0) return false;
    *out = LightSample{pos, nrm, info.emit, pdf, idx};
    return true;
  }

  // The part of World::hit that tests only light L's own primitive: its sphere, or its
  // instance's mesh (world.rs:86-118), over [tmin, tmax).
  bool light_blocks_itself(const LightInfo& L, const Ray& r, double tmin, double tmax) const {
    Hit h;
    if (L.kind == 0) return spheres[(size_t)L.idx].hit(r, tmin, tmax, &h);
    const Instance& in = instances[(size_t)L.inst_id];
    const Mesh& mesh = meshes[(size_t)in.mesh_id];
    V3 o_obj = in.xform.apply_point_inv(r.o);
    V3 d_obj = norm(in.xform.apply_vec_inv(r.d));
    Ray r_obj{o_obj, d_obj, r.time};
    int ti = -1;
    if (!mesh.bvh.hit(mesh.tris, r_obj, tmin, 1e30, &h, &ti)) return false;
    double t_world = dot(in.xform.apply_point(h.p) - r.o, r.d);
    return t_world > tmin && t_world < tmax;
  }
  bool hit(const Ray& r, double tmin, double tmax, Hit* best) const {  // world.rs:81-121
    double closest = tmax;
    bool found = false;
    for (size_t ii = 0; ii < instances.size(); ii++) {
      const Instance& in = instances[ii];
      if (in.mesh_id < 0 || (size_t)in.mesh_id >= meshes.size()) continue;
      const Mesh& mesh = meshes[(size_t)in.mesh_id];
      CNT(inst);
      V3 o_obj = in.xform.apply_point_inv(r.o);
      V3 d_obj = norm(in.xform.apply_vec_inv(r.d));
      Ray r_obj{o_obj, d_obj, r.time};
      Hit h_obj;
      int ti = -1;
      if (mesh.bvh.hit(mesh.tris, r_obj, tmin, 1e30, &h_obj, &ti)) {
        V3 p_world = in.xform.apply_point(h_obj.p);
        V3 n_world = in.xform.apply_normal(h_obj.n);
        double t_world = dot(p_world - r.o, r.d);
        if (t_world > tmin && t_world < closest) {
          closest = t_world;
          int mat_id = in.mat_override >= 0 ? in.mat_override : h_obj.mat_id;
          *best = Hit{t_world, p_world, n_world, mat_id, HIT_TRI, (int)ii, ti};
          found = true;
        }
      }
    }
    for (size_t si = 0; si < spheres.size(); si++) {
      Hit h;
      if (spheres[si].hit(r, tmin, closest, &h)) {
        closest = h.t;
        h.kind = HIT_SPHERE;
        h.inst = (int)si;
        h.prim = -1;
        *best = h;
        found = true;
      }
    }
    return found;
  }
};

// ---------------------------------------------------------------- integrator.rs
static V3 sky(V3 d) {  // integrator.rs:25-28
  double t = 0.5 * (d.y + 1.0);
  return (1.0 - t) * v3(0.9, 0.95, 1.0) + t * v3(0.6, 0.75, 1.0);
}
static inline double mis_weight(double a, double b) {  // integrator.rs:221-225
  double a2 = a * a, b2 = b * b;
  return a2 + b2 > 0.0 ? a2 / (a2 + b2) : 0.0;
}
static inline V3 oriented_normal(V3 n, V3 d) { return dot(n, d) < 0.0 ? n : -n; }  // integrator.rs:138-140

static V3 nee_environment(const Scene& sc, const Material& mat, V3 beta, V3 hit_p, V3 n, const Ray& ray, Rng& rng) {
  // integrator.rs:148-172
  V3 wi, li;
  double pdf_env;
  sc.env.sample_dir(rng, &wi, &li, &pdf_env);
  double c = fmx(dot(n, wi), 0.0);
  if (c <= 0.0 || pdf_env <= 0.0) return v3(0, 0, 0);
  Ray shadow{hit_p + RAY_EPSILON * wi, wi, ray.time};
  Hit tmp;
  if (tl_cnt) tl_cnt->rays_shadow++;
  if (sc.hit(shadow, RAY_EPSILON, RAY_T_MAX, &tmp)) return v3(0, 0, 0);
  V3 f;
  double pdf_bsdf;
  mat_eval(mat, norm(-ray.d), wi, n, &f, &pdf_bsdf);
  double w = mis_weight(pdf_env, pdf_bsdf);
  return had(had(beta, f), li) * (c * w / pdf_env);
}
static V3 nee_area_light(const Scene& sc, const Material& mat, V3 beta, V3 hit_p, V3 n, const Ray& ray, const LightSample& ls) {
  // integrator.rs:178-213
  V3 to_light = ls.position - hit_p;
  double dist2 = dot(to_light, to_light);
  double dist = std::sqrt(dist2);
  if (dist <= 1e-6) return v3(0, 0, 0);
  V3 wi = to_light / dist;
  double c = fmx(dot(n, wi), 0.0);
  if (c <= 0.0) return v3(0, 0, 0);
  Ray shadow{hit_p + RAY_EPSILON * wi, wi, ray.time};
  double tmax = fmx(dist - RAY_EPSILON, RAY_EPSILON);
  Hit tmp;
  if (tl_cnt) {
    tl_cnt->rays_shadow++;
    // counters only (no effect on the result): samples with a non-zero contribution, and the
    // share of them that the light's own surface shadows
    V3 f0;
    double pb0;
    mat_eval(mat, norm(-ray.d), wi, n, &f0, &pb0);
    V3 c0 = had(had(beta, f0), ls.emit) * (c * mis_weight(ls.pdf, pb0) / ls.pdf);
    if (ls.pdf > 0.0 && (c0.x != 0.0 || c0.y != 0.0 || c0.z != 0.0)) {
      Counters* keep = tl_cnt;
      tl_cnt = nullptr;  // the extra intersection test is not traversal work of the reference
      keep->light_samples++;
      if (sc.light_blocks_itself(sc.lights[ls.idx], shadow, RAY_EPSILON, tmax)) keep->light_self++;
      tl_cnt = keep;
    }
  }
  if (sc.hit(shadow, RAY_EPSILON, tmax, &tmp)) return v3(0, 0, 0);
  V3 f;
  double pdf_bsdf;
  mat_eval(mat, norm(-ray.d), wi, n, &f, &pdf_bsdf);
  if (ls.pdf <= 0.0) return v3(0, 0, 0);
  double w = mis_weight(ls.pdf, pdf_bsdf);
  return had(had(beta, f), ls.emit) * (c * w / ls.pdf);
}

static V3 radiance(const Scene& sc, Ray ray, Rng& rng, int max_bounces, int rr_start) {  // integrator.rs:53-135
  V3 acc = v3(0, 0, 0);
  V3 beta = v3(1, 1, 1);
  double last_bsdf_pdf = 0.0;
  bool last_non_delta = false;
  for (int bounce = 0; bounce < max_bounces; bounce++) {
    Hit hit;
    Counters before;
    if (tl_cnt) { tl_cnt->rays_ext++; tl_cnt->ext_bounces++; before = *tl_cnt; }
    const bool found = sc.hit(ray, RAY_EPSILON, RAY_T_MAX, &hit);
    if (tl_cnt) {
      tl_cnt->aabb_ext += tl_cnt->aabb - before.aabb;
      tl_cnt->tri_ext += tl_cnt->tri - before.tri;
      tl_cnt->sph_ext += tl_cnt->sph - before.sph;
      tl_cnt->inst_ext += tl_cnt->inst - before.inst;
    }
    if (!found) {
      V3 bg = sc.has_env ? sc.env.sample(ray.d) : sky(ray.d);
      V3 contrib = had(beta, bg);
      if (sc.has_env && last_non_delta) {
        double pdf_env = sc.env.pdf(ray.d);
        contrib = contrib * mis_weight(last_bsdf_pdf, pdf_env);
      }
      acc = acc + clamp_lum(contrib, FIREFLY_CLAMP);
      break;
    }
    if (bounce >= rr_start) {
      double p = fmx(fmn(fmx(fmx(beta.x, beta.y), beta.z), 0.95), 0.05);
      if (rng.next_f64() > p) break;
      beta = beta / p;
    }
    const Material& mat = sc.mats[(size_t)hit.mat_id];
    V3 emit;
    if (mat.emitted(&emit)) {
      acc = acc + clamp_lum(had(beta, emit), FIREFLY_CLAMP);
      break;
    }
    V3 n = oriented_normal(hit.n, ray.d);
    if (!mat.is_delta()) {
      if (sc.has_env) acc = acc + nee_environment(sc, mat, beta, hit.p, n, ray, rng);
      LightSample ls;
      if (sc.sample_light(rng, ray.time, hit.p, &ls)) acc = acc + nee_area_light(sc, mat, beta, hit.p, n, ray, ls);
    }
    BsdfSample bs;
    if (!mat_sample(mat, ray, hit, rng, &bs)) break;
    last_bsdf_pdf = bs.pdf;
    last_non_delta = !bs.is_delta;
    beta = had(beta, bs.weight);
    ray = bs.scattered;
  }
  return acc;
}

// ---------------------------------------------------------------- render.rs
static uint64_t interleave_bits(uint64_t v) {  // render.rs:291-299
  uint64_t x = v & 0x00000000ffffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
static uint64_t morton2(size_t tx, size_t ty, size_t tiles_x, size_t tiles_y) {  // render.rs:284-288
  uint64_t x = (uint64_t)std::min(tx, tiles_x == 0 ? 0 : tiles_x - 1);
  uint64_t y = (uint64_t)std::min(ty, tiles_y == 0 ? 0 : tiles_y - 1);
  return interleave_bits(x) | (interleave_bits(y) << 1);
}

struct Task {
  size_t x0, y0, x1, y1;
};
// The tile tasks of a w x h image in task-id order (render.rs:76-106): row-major tiles, then a
// stable sort by the Morton code of the tile position when enabled.
static std::vector<Task> make_tasks(size_t w, size_t h, size_t tile, bool morton) {
  std::vector<Task> tasks;
  for (size_t y = 0; y < h; y += tile)
    for (size_t x = 0; x < w; x += tile) tasks.push_back(Task{x, y, std::min(x + tile, w), std::min(y + tile, h)});
  if (morton) {
    size_t tiles_x = (w + tile - 1) / tile, tiles_y = (h + tile - 1) / tile;
    std::stable_sort(tasks.begin(), tasks.end(), [&](const Task& a, const Task& b) {
      return morton2(a.x0 / tile, a.y0 / tile, tiles_x, tiles_y) < morton2(b.x0 / tile, b.y0 / tile, tiles_x, tiles_y);
    });
  }
  return tasks;
}

struct RenderOpts {
  int width, height, spp, max_bounces, rr_start, tile;
  bool adaptive;
  int adaptive_min_spp;
  double adaptive_threshold;
  uint64_t seed;
  bool morton;
  int threads;
  bool per_sample_seed;  // addition: seed_for(x, y, s) ^ seed per sample (FP32-wavefront convention)
  int x0, y0, x1, y1;    // addition: pixel window (crop) to render; others stay 0
};

static void render_pixel(const Scene& sc, const RenderOpts& o, size_t x, size_t y, double inv_w, double inv_h, V3* c_out,
                         double* w_out, V3* sq_out) {
  // render.rs:137-191
  V3 c = v3(0, 0, 0), sq = v3(0, 0, 0);
  Rng rng(seed_for((uint32_t)x, (uint32_t)y, 0) ^ o.seed);
  size_t spp = (size_t)std::max(o.spp, 0);
  auto one = [&](size_t s) {
    if (o.per_sample_seed) rng = Rng(seed_for((uint32_t)x, (uint32_t)y, (uint32_t)s) ^ o.seed);
    double jx = rng.next_f64();
    double jy = rng.next_f64();
    double sx = ((double)x + jx) * inv_w * 2.0 - 1.0;
    double sy = 1.0 - ((double)y + jy) * inv_h * 2.0;
    Ray ray = sc.cam.ray(sx, sy, rng);
    V3 L = radiance(sc, ray, rng, o.max_bounces, o.rr_start);
    if (tl_cnt) tl_cnt->paths++;
    c = c + L;
    sq = sq + had(L, L);
    return L;
  };
  if (o.adaptive) {
    size_t max_spp = std::max(spp, (size_t)1);
    size_t min_spp = std::min(std::max((size_t)std::max(o.adaptive_min_spp, 0), (size_t)1), max_spp);
    size_t n = 0;
    double mean = 0.0, m2 = 0.0;
    for (size_t s = 0; s < spp; s++) {
      V3 L = one(s);
      n += 1;
      double l = lum(L);
      double delta = l - mean;
      mean += delta / (double)n;
      double delta2 = l - mean;
      m2 += delta * delta2;
      if (n >= min_spp) {
        double var = n > 1 ? m2 / (double)(n - 1) : INF;
        double denom = fmx(std::fabs(mean), 1e-4);
        double rel_std = std::sqrt(var) / denom;
        if (rel_std < o.adaptive_threshold) break;
      }
    }
    *w_out = (double)n;
  } else {
    for (size_t s = 0; s < spp; s++) one(s);
    *w_out = (double)spp;
  }
  *c_out = c;
  *sq_out = sq;
}

static void render(const Scene& sc, const RenderOpts& o, double* acc, double* acc_w, double* acc_sq, Counters* total) {
  // render.rs:42-278: tiles, Morton order, dynamic queue over worker threads.  Each pixel
  // belongs to exactly one task and render() merges into zeroed buffers, so the merge order
  // does not change any value; results are written straight to the output buffers.
  size_t w = (size_t)o.width, h = (size_t)o.height;
  double inv_w = 1.0 / (double)w, inv_h = 1.0 / (double)h;
  std::vector<Task> tasks = make_tasks(w, h, (size_t)std::max(o.tile, 1), o.morton);
  int nthreads = o.threads > 0 ? o.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<size_t> next{0};
  std::vector<Counters> cnts((size_t)nthreads);
  bool crop = o.x1 > o.x0 && o.y1 > o.y0;
  auto worker = [&](int tid) {
    tl_cnt = &cnts[(size_t)tid];
    for (;;) {
      size_t ti = next.fetch_add(1);
      if (ti >= tasks.size()) break;
      const Task& t = tasks[ti];
      for (size_t y = t.y0; y < t.y1; y++)
        for (size_t x = t.x0; x < t.x1; x++) {
          size_t g = y * w + x;
          if (crop && ((int)x < o.x0 || (int)x >= o.x1 || (int)y < o.y0 || (int)y >= o.y1)) {
            acc[3 * g] = acc[3 * g + 1] = acc[3 * g + 2] = 0.0;
            acc_w[g] = 0.0;
            if (acc_sq) acc_sq[3 * g] = acc_sq[3 * g + 1] = acc_sq[3 * g + 2] = 0.0;
            continue;
          }
          V3 c, sq;
          double ww;
          render_pixel(sc, o, x, y, inv_w, inv_h, &c, &ww, &sq);
          acc[3 * g] = c.x;
          acc[3 * g + 1] = c.y;
          acc[3 * g + 2] = c.z;
          acc_w[g] = ww;
          if (acc_sq) {
            acc_sq[3 * g] = sq.x;
            acc_sq[3 * g + 1] = sq.y;
            acc_sq[3 * g + 2] = sq.z;
          }
        }
    }
    tl_cnt = nullptr;
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nthreads; i++) th.emplace_back(worker, i);
  worker(0);
  for (auto& t : th) t.join();
  if (total)
    for (auto& c : cnts) total->add(c);
}

// ---------------------------------------------------------------- scene from descriptor
static Scene* scene_from_desc(const tpt_scene_desc* d) {
  std::unique_ptr<Scene> sc(new Scene());
  const tpt_camera_desc& c = d->camera;
  sc->cam = Camera::look_at_dof(v3(c.eye[0], c.eye[1], c.eye[2]), v3(c.target[0], c.target[1], c.target[2]),
                                v3(c.up[0], c.up[1], c.up[2]), c.vfov_deg, c.aspect, c.focus_dist, c.aperture);
  for (int i = 0; i < d->n_materials; i++) {
    const double* p = d->mat_param + 4 * i;
    sc->mats.push_back(Material{d->mat_kind[i], v3(p[0], p[1], p[2]), p[3]});
  }
  for (int i = 0; i < d->n_spheres; i++)
    sc->spheres.push_back(Sphere{v3(d->sphere_center[3 * i], d->sphere_center[3 * i + 1], d->sphere_center[3 * i + 2]),
                                 d->sphere_radius[i], d->sphere_mat[i]});
  for (int m = 0; m < d->n_meshes; m++) {
    Mesh mesh;
    for (int64_t t = d->mesh_tri_begin[m]; t < d->mesh_tri_begin[m + 1]; t++) {
      const double* v = d->tri_vertices + 9 * t;
      mesh.tris.push_back(Triangle::new_static(v3(v[0], v[1], v[2]), v3(v[3], v[4], v[5]), v3(v[6], v[7], v[8]), d->tri_mat[t]));
    }
    mesh.bvh = Bvh::build(mesh.tris);
    sc->meshes.push_back(std::move(mesh));
  }
  for (int i = 0; i < d->n_instances; i++) {
    const double* x = d->inst_xform + 12 * i;
    Mat3 a{};
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) a.m[r][k] = x[3 * r + k];
    sc->instances.push_back(Instance{d->inst_mesh[i], Transform::from_affine(a, v3(x[9], x[10], x[11])), d->inst_mat_override[i]});
  }
  if (d->env_kind == TPT_ENV_IMAGE) {
    size_t w = (size_t)d->env_width, h = (size_t)d->env_height;
    std::vector<V3> px(w * h);
    for (size_t i = 0; i < w * h; i++) px[i] = v3(d->env_rgb[3 * i], d->env_rgb[3 * i + 1], d->env_rgb[3 * i + 2]);
    sc->env = EnvMap::from_pixels(w, h, std::move(px));
    sc->has_env = true;
  }
  sc->build_lights();
  return sc.release();
}

}  // namespace orc

// ================================================================ C API (ctypes)
using namespace orc;

extern "C" {

typedef struct orc_counters {
  uint64_t aabb, tri, sph, inst, rays_ext, rays_shadow, paths, ext_bounces;
  uint64_t aabb_ext, tri_ext, sph_ext, inst_ext;  // part of the above spent on extension rays
  uint64_t light_samples, light_self;             // area-light samples; those the light itself shadows
} orc_counters;

typedef struct orc_render_opts {
  int32_t width, height, spp, max_bounces, rr_start, tile;
  int32_t adaptive, adaptive_min_spp;
  double adaptive_threshold;
  uint64_t seed;
  int32_t morton, threads, per_sample_seed;
  int32_t crop_x0, crop_y0, crop_x1, crop_y1;
} orc_render_opts;

// ---- RNG (rng.rs)
void orc_rng_u32(uint64_t seed, int64_t n, uint32_t* out) {
  Rng r(seed);
  for (int64_t i = 0; i < n; i++) out[i] = r.next_u32();
}
void orc_rng_f64(uint64_t seed, int64_t n, double* out) {
  Rng r(seed);
  for (int64_t i = 0; i < n; i++) out[i] = r.next_f64();
}
uint64_t orc_seed_for(uint32_t px, uint32_t py, uint32_t s) { return seed_for(px, py, s); }
uint64_t orc_morton2(uint64_t tx, uint64_t ty, uint64_t tiles_x, uint64_t tiles_y) {
  return morton2((size_t)tx, (size_t)ty, (size_t)tiles_x, (size_t)tiles_y);
}
double orc_srgb_to_linear(double c) { return srgb_to_linear(c); }
double orc_linear_to_srgb(double c) { return linear_to_srgb(c); }

// ---- primitives (geometry.rs)
// ray[7] = o, d, time; out[7] = t, p, n.  Returns 1 on hit.
int orc_sphere_hit(const double* c, double r, const double* ray, double tmin, double tmax, double* out) {
  Sphere s{v3(c[0], c[1], c[2]), r, 0};
  Ray rr{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), ray[6]};
  Hit h;
  if (!s.hit(rr, tmin, tmax, &h)) return 0;
  out[0] = h.t; out[1] = h.p.x; out[2] = h.p.y; out[3] = h.p.z; out[4] = h.n.x; out[5] = h.n.y; out[6] = h.n.z;
  return 1;
}
int orc_triangle_hit(const double* v, const double* ray, double tmin, double tmax, double* out) {
  Triangle t = Triangle::new_static(v3(v[0], v[1], v[2]), v3(v[3], v[4], v[5]), v3(v[6], v[7], v[8]), 0);
  Ray rr{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), ray[6]};
  Hit h;
  if (!t.hit(rr, tmin, tmax, &h)) return 0;
  out[0] = h.t; out[1] = h.p.x; out[2] = h.p.y; out[3] = h.p.z; out[4] = h.n.x; out[5] = h.n.y; out[6] = h.n.z;
  return 1;
}
// box[6] = min, max.  Returns 1 if hit; *t0 = entry.
int orc_aabb_hit(const double* box, const double* ray, double tmin, double tmax, double* t0) {
  Aabb b{v3(box[0], box[1], box[2]), v3(box[3], box[4], box[5])};
  Ray rr{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), ray[6]};
  V3 inv = v3(1.0 / rr.d.x, 1.0 / rr.d.y, 1.0 / rr.d.z);
  return b.range_inv(rr, inv, tmin, tmax, t0) ? 1 : 0;
}

// ---- the area-light shadow test of nee_area_light (integrator.rs:178-213) for light `light[i]`
// sampled at uniforms uv[2i..2i+1] from shading point p[3i..3i+2] at ray time time[i]
// (world.rs:242-272): out[i] = 0 unoccluded, 1 blocked by the light's own primitive, 2 blocked by
// other geometry, 3 blocked by another triangle of the light's instance, -1 no shadow ray
// (dist <= 1e-6 or no surface point)
void orc_light_shadow(void* s, int64_t n, const int32_t* light, const double* p, const double* uv, const double* time,
                      int32_t* out) {
  const Scene& sc = *(Scene*)s;
  for (int64_t i = 0; i < n; i++) {
    out[i] = -1;
    if (light[i] < 0 || (size_t)light[i] >= sc.lights.size()) continue;
    const LightInfo& L = sc.lights[(size_t)light[i]];
    V3 pos, nrm;
    double area;
    if (!sc.surface_at(L, time[i], uv[2 * i], uv[2 * i + 1], &pos, &nrm, &area)) continue;
    const V3 hp = v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    V3 to_light = pos - hp;
    double dist = std::sqrt(dot(to_light, to_light));
    if (dist <= 1e-6) continue;
    V3 wi = to_light / dist;
    Ray shadow{hp + RAY_EPSILON * wi, wi, time[i]};
    double tmax = fmx(dist - RAY_EPSILON, RAY_EPSILON);
    Hit h;
    if (!sc.hit(shadow, RAY_EPSILON, tmax, &h)) {
      out[i] = 0;
      continue;
    }
    if (L.kind == 0)
      out[i] = h.kind == HIT_SPHERE && h.inst == L.idx ? 1 : 2;
    else if (h.kind == HIT_TRI && h.inst == L.inst_id)
      out[i] = h.prim == L.tri_id ? 1 : 3;
    else
      out[i] = 2;
  }
}

// ---- distance to one given primitive along each ray (World::hit restricted to it, world.rs:86-118):
// kind 1 = triangle prim[i] (mesh-local) of instance inst[i], kind 2 = sphere inst[i]; rays[8i..]
// = o, d, tmin, tmax; t_out[i] = the world distance, or +inf when the primitive is not hit;
// margin[i] = for a triangle, the smallest barycentric coordinate of the ray's crossing of its
// plane (negative outside the triangle), for a sphere 1
void orc_prim_distance(void* s, int64_t n, const double* rays, const int32_t* kind, const int32_t* inst,
                       const int32_t* prim, double* t_out, double* margin) {
  const Scene& sc = *(Scene*)s;
  for (int64_t i = 0; i < n; i++) {
    const double* q = rays + 8 * i;
    Ray r{v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]), 0.0};
    const double tmin = q[6], tmax = q[7];
    t_out[i] = INFINITY;
    margin[i] = -INFINITY;
    Hit h;
    if (kind[i] == 2) {
      margin[i] = 1.0;
      if (inst[i] >= 0 && (size_t)inst[i] < sc.spheres.size() && sc.spheres[(size_t)inst[i]].hit(r, tmin, tmax, &h))
        t_out[i] = h.t;
    } else if (kind[i] == 1 && inst[i] >= 0 && (size_t)inst[i] < sc.instances.size()) {
      const Instance& in = sc.instances[(size_t)inst[i]];
      const Mesh& mesh = sc.meshes[(size_t)in.mesh_id];
      if (prim[i] < 0 || (size_t)prim[i] >= mesh.tris.size()) continue;
      Ray r_obj{in.xform.apply_point_inv(r.o), norm(in.xform.apply_vec_inv(r.d)), 0.0};
      const Triangle& tr = mesh.tris[(size_t)prim[i]];
      V3 v0, v1, v2;
      tr.vertices_at(0.0, &v0, &v1, &v2);
      V3 e1 = v1 - v0, e2 = v2 - v0, pv = cross(r_obj.d, e2);
      double det = dot(e1, pv);
      if (det != 0.0) {
        V3 tv = r_obj.o - v0, qv = cross(tv, e1);
        double u = dot(tv, pv) / det, v = dot(r_obj.d, qv) / det;
        margin[i] = std::min(std::min(u, v), 1.0 - u - v);
      }
      if (!tr.hit(r_obj, tmin, 1e30, &h)) continue;
      double t_world = dot(in.xform.apply_point(h.p) - r.o, r.d);
      if (t_world > tmin && t_world < tmax) t_out[i] = t_world;
    }
  }
}

// ---- tile tasks (render.rs:76-106): n tasks, out[4i..] = x0, y0, x1, y1 in task-id order
int64_t orc_task_tiles(int32_t w, int32_t h, int32_t tile, int32_t morton, int64_t* out) {
  std::vector<Task> t = make_tasks((size_t)w, (size_t)h, (size_t)std::max(tile, 1), morton != 0);
  if (out)
    for (size_t i = 0; i < t.size(); i++) {
      out[4 * i] = (int64_t)t[i].x0; out[4 * i + 1] = (int64_t)t[i].y0;
      out[4 * i + 2] = (int64_t)t[i].x1; out[4 * i + 3] = (int64_t)t[i].y1;
    }
  return (int64_t)t.size();
}

// ---- checkpoint files (checkpoint.rs:48-138): "HYPCKPT\0", u32 version 1, u64 scene hash,
// u32 w, u32 h, u64 next task id, w*h*3 f64 acc, w*h f64 acc_w; little endian (this host is);
// written to path.tmp and renamed
int32_t orc_checkpoint_save(const char* path, uint64_t hash, uint32_t w, uint32_t h, uint64_t next_id,
                            const double* acc, const double* acc_w) {
  std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return -1;
  const uint32_t ver = 1;
  bool ok = std::fwrite("HYPCKPT\0", 1, 8, f) == 8 && std::fwrite(&ver, 4, 1, f) == 1 && std::fwrite(&hash, 8, 1, f) == 1 &&
            std::fwrite(&w, 4, 1, f) == 1 && std::fwrite(&h, 4, 1, f) == 1 && std::fwrite(&next_id, 8, 1, f) == 1;
  const size_t n = (size_t)w * h;
  ok = ok && std::fwrite(acc, 8, 3 * n, f) == 3 * n && std::fwrite(acc_w, 8, n, f) == n;
  ok = std::fclose(f) == 0 && ok;
  if (!ok) return -1;
  return std::rename(tmp.c_str(), path) == 0 ? 0 : -1;
}
// returns 1 and fills the outputs when the file exists with this magic, version, hash and size;
// 0 when it does not (load_checkpoint's None); -1 on a read error
int32_t orc_checkpoint_load(const char* path, uint64_t hash, uint32_t w, uint32_t h, uint64_t* next_id, double* acc,
                            double* acc_w) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 0;
  char magic[8];
  uint32_t ver = 0, fw = 0, fh = 0;
  uint64_t fhash = 0, nid = 0;
  int rc = 0;
  if (std::fread(magic, 1, 8, f) != 8) rc = -1;
  else if (std::memcmp(magic, "HYPCKPT\0", 8) != 0) rc = 0;
  else if (std::fread(&ver, 4, 1, f) != 1) rc = -1;
  else if (ver != 1) rc = 0;
  else if (std::fread(&fhash, 8, 1, f) != 1) rc = -1;
  else if (fhash != hash) rc = 0;
  else if (std::fread(&fw, 4, 1, f) != 1 || std::fread(&fh, 4, 1, f) != 1) rc = -1;
  else if (fw != w || fh != h) rc = 0;
  else if (std::fread(&nid, 8, 1, f) != 1) rc = -1;
  else {
    const size_t n = (size_t)w * h;
    rc = std::fread(acc, 8, 3 * n, f) == 3 * n && std::fread(acc_w, 8, n, f) == n ? 1 : -1;
    *next_id = nid;
  }
  std::fclose(f);
  return rc;
}

// ---- transform (transform.rs:24-93).  xf[12] = A row-major, t; op 0 apply_point, 1 apply_point_inv,
// 2 apply_vec_inv, 3 apply_normal; n vectors in[3n] -> out[3n]
void orc_transform_apply(const double* xf, int32_t op, int64_t n, const double* in, double* out) {
  Mat3 a;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) a.m[r][c] = xf[3 * r + c];
  Transform x = Transform::from_affine(a, v3(xf[9], xf[10], xf[11]));
  for (int64_t i = 0; i < n; i++) {
    V3 v = v3(in[3 * i], in[3 * i + 1], in[3 * i + 2]), r;
    switch (op) {
      case 0: r = x.apply_point(v); break;
      case 1: r = x.apply_point_inv(v); break;
      case 2: r = x.apply_vec_inv(v); break;
      default: r = x.apply_normal(v); break;
    }
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
}

// ---- camera (ray.rs)
void orc_camera_rays(const tpt_camera_desc* c, uint64_t seed, int64_t n, const double* sxy, double* out /*n*7*/) {
  Camera cam = Camera::look_at_dof(v3(c->eye[0], c->eye[1], c->eye[2]), v3(c->target[0], c->target[1], c->target[2]),
                                   v3(c->up[0], c->up[1], c->up[2]), c->vfov_deg, c->aspect, c->focus_dist, c->aperture);
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) {
    Ray r = cam.ray(sxy[2 * i], sxy[2 * i + 1], rng);
    double* o = out + 7 * i;
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z; o[6] = r.time;
  }
}

// ---- materials (material.rs).  mat: kind + param[4].  ray[7], hit[8] = t, p, n, (unused).
// Draws n samples from one stream; out per sample: valid, d(3), weight(3), pdf, is_delta, o(3) = 12.
void orc_material_sample(int kind, const double* param, const double* ray, const double* hit, uint64_t seed, int64_t n,
                         double* out) {
  Material m{kind, v3(param[0], param[1], param[2]), param[3]};
  Ray r{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5]), ray[6]};
  Hit h{hit[0], v3(hit[1], hit[2], hit[3]), v3(hit[4], hit[5], hit[6]), 0, 0, 0, 0};
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) {
    BsdfSample s;
    double* o = out + 12 * i;
    if (!mat_sample(m, r, h, rng, &s)) {
      for (int k = 0; k < 12; k++) o[k] = 0.0;
      continue;
    }
    o[0] = 1.0;
    o[1] = s.scattered.d.x; o[2] = s.scattered.d.y; o[3] = s.scattered.d.z;
    o[4] = s.weight.x; o[5] = s.weight.y; o[6] = s.weight.z;
    o[7] = s.pdf;
    o[8] = s.is_delta ? 1.0 : 0.0;
    o[9] = s.scattered.o.x; o[10] = s.scattered.o.y; o[11] = s.scattered.o.z;
  }
}
// out[4] = f rgb, pdf
void orc_material_eval(int kind, const double* param, const double* wo, const double* wi, const double* n, double* out) {
  Material m{kind, v3(param[0], param[1], param[2]), param[3]};
  V3 f;
  double pdf;
  mat_eval(m, v3(wo[0], wo[1], wo[2]), v3(wi[0], wi[1], wi[2]), v3(n[0], n[1], n[2]), &f, &pdf);
  out[0] = f.x; out[1] = f.y; out[2] = f.z; out[3] = pdf;
}
void orc_ggx_vndf(const double* wo, double alpha, double u1, double u2, double* out) {
  V3 r = sample_ggx_vndf(v3(wo[0], wo[1], wo[2]), alpha, u1, u2);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// ---- scenes
void* orc_scene_create(const tpt_scene_desc* d) { return (void*)scene_from_desc(d); }
void orc_scene_destroy(void* s) { delete (Scene*)s; }
int64_t orc_scene_n_lights(void* s) { return (int64_t)((Scene*)s)->lights.size(); }
double orc_scene_light_total(void* s) { return ((Scene*)s)->light_total; }
// out per light: kind, idx, mesh, tri, inst, emit(3), weight, cdf = 10
void orc_scene_lights(void* s, double* out) {
  Scene* sc = (Scene*)s;
  for (size_t i = 0; i < sc->lights.size(); i++) {
    const LightInfo& L = sc->lights[i];
    double* o = out + 10 * i;
    o[0] = L.kind; o[1] = L.idx; o[2] = L.mesh_id; o[3] = L.tri_id; o[4] = L.inst_id;
    o[5] = L.emit.x; o[6] = L.emit.y; o[7] = L.emit.z; o[8] = L.weight; o[9] = sc->light_cdf[i];
  }
}
// light samples from one stream at point p: out per draw: valid, pos(3), normal(3), emit(3), pdf = 11
void orc_scene_sample_light(void* s, const double* p, double time, uint64_t seed, int64_t n, double* out) {
  Scene* sc = (Scene*)s;
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) {
    LightSample ls;
    double* o = out + 11 * i;
    if (!sc->sample_light(rng, time, v3(p[0], p[1], p[2]), &ls)) {
      for (int k = 0; k < 11; k++) o[k] = 0.0;
      continue;
    }
    o[0] = 1; o[1] = ls.position.x; o[2] = ls.position.y; o[3] = ls.position.z;
    o[4] = ls.normal.x; o[5] = ls.normal.y; o[6] = ls.normal.z;
    o[7] = ls.emit.x; o[8] = ls.emit.y; o[9] = ls.emit.z; o[10] = ls.pdf;
  }
}
int64_t orc_scene_bvh_nodes(void* s, int32_t mesh) { return (int64_t)((Scene*)s)->meshes[(size_t)mesh].bvh.nodes.size(); }
void orc_scene_bvh(void* s, int32_t mesh, double* bbox, int32_t* left, int32_t* right, uint32_t* start, uint32_t* count,
                   int64_t* indices) {
  const Bvh& b = ((Scene*)s)->meshes[(size_t)mesh].bvh;
  for (size_t i = 0; i < b.nodes.size(); i++) {
    const BvhNode& n = b.nodes[i];
    double* x = bbox + 6 * i;
    x[0] = n.bbox.mn.x; x[1] = n.bbox.mn.y; x[2] = n.bbox.mn.z; x[3] = n.bbox.mx.x; x[4] = n.bbox.mx.y; x[5] = n.bbox.mx.z;
    left[i] = n.left; right[i] = n.right; start[i] = n.start; count[i] = n.count;
  }
  for (size_t i = 0; i < b.indices.size(); i++) indices[i] = (int64_t)b.indices[i];
}
// Closest hit of n rays (rays[n*8]: o, d, tmin, tmax; time = 0.5).  out per ray: t, kind, inst,
// prim, mat, n(3), p(3) = 11.  counters optional.
void orc_scene_trace(void* s, int64_t n, const double* rays, double* out, orc_counters* cnt, int32_t threads) {
  Scene* sc = (Scene*)s;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  std::vector<Counters> cs((size_t)nt);
  std::atomic<int64_t> next{0};
  auto work = [&](int tid) {
    tl_cnt = &cs[(size_t)tid];
    for (;;) {
      int64_t b0 = next.fetch_add(1024);
      if (b0 >= n) break;
      int64_t b1 = std::min(n, b0 + 1024);
      for (int64_t i = b0; i < b1; i++) {
        const double* r = rays + 8 * i;
        Ray ray{v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), 0.5};
        Hit h;
        double* o = out + 11 * i;
        tl_cnt->rays_ext++;
        if (sc->hit(ray, r[6], r[7], &h)) {
          o[0] = h.t; o[1] = h.kind; o[2] = h.inst; o[3] = h.prim; o[4] = h.mat_id;
          o[5] = h.n.x; o[6] = h.n.y; o[7] = h.n.z; o[8] = h.p.x; o[9] = h.p.y; o[10] = h.p.z;
        } else {
          o[0] = INF; o[1] = 0; o[2] = -1; o[3] = -1; o[4] = -1;
          for (int k = 5; k < 11; k++) o[k] = 0;
        }
      }
    }
    tl_cnt = nullptr;
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nt; i++) th.emplace_back(work, i);
  work(0);
  for (auto& t : th) t.join();
  if (cnt) {
    Counters tot;
    for (auto& c : cs) tot.add(c);
    // every query here is a closest-hit query
    *cnt = orc_counters{tot.aabb, tot.tri, tot.sph, tot.inst, tot.rays_ext, tot.rays_shadow, tot.paths, tot.ext_bounces,
                        tot.aabb, tot.tri, tot.sph, tot.inst, 0, 0};
  }
}
// env (env.rs): only for TPT_ENV_IMAGE scenes
double orc_env_total(void* s) { return ((Scene*)s)->env.total_weight; }
void orc_env_lookup(void* s, int64_t n, const double* dirs, double* rgb, double* pdf) {
  Scene* sc = (Scene*)s;
  for (int64_t i = 0; i < n; i++) {
    V3 d = v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    V3 c = sc->env.sample(d);
    rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
    pdf[i] = sc->env.pdf(d);
  }
}
// n draws of sample_dir from one stream: out per draw: dir(3), rgb(3), pdf = 7
void orc_env_sample(void* s, uint64_t seed, int64_t n, double* out) {
  Scene* sc = (Scene*)s;
  Rng rng(seed);
  for (int64_t i = 0; i < n; i++) {
    V3 d, c;
    double p;
    sc->env.sample_dir(rng, &d, &c, &p);
    double* o = out + 7 * i;
    o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = c.x; o[4] = c.y; o[5] = c.z; o[6] = p;
  }
}
// radiance() of given camera-space samples: per path seed, jitter-free ray given: used for
// small targeted checks.  rays[n*7]; out[n*3].
void orc_scene_radiance(void* s, int64_t n, const double* rays, const uint64_t* seeds, int32_t max_bounces, int32_t rr_start,
                        double* out) {
  Scene* sc = (Scene*)s;
  for (int64_t i = 0; i < n; i++) {
    const double* r = rays + 7 * i;
    Rng rng(seeds[i]);
    V3 L = radiance(*sc, Ray{v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), r[6]}, rng, max_bounces, rr_start);
    out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
  }
}
// render() (render.rs:42-278) into caller buffers; returns elapsed seconds of the pixel loop.
double orc_render(void* s, const orc_render_opts* ro, double* acc, double* acc_w, double* acc_sq, orc_counters* cnt) {
  Scene* sc = (Scene*)s;
  RenderOpts o;
  o.width = ro->width; o.height = ro->height; o.spp = ro->spp; o.max_bounces = ro->max_bounces; o.rr_start = ro->rr_start;
  o.tile = ro->tile > 0 ? ro->tile : 16;
  o.adaptive = ro->adaptive != 0; o.adaptive_min_spp = ro->adaptive_min_spp; o.adaptive_threshold = ro->adaptive_threshold;
  o.seed = ro->seed; o.morton = ro->morton != 0; o.threads = ro->threads; o.per_sample_seed = ro->per_sample_seed != 0;
  o.x0 = ro->crop_x0; o.y0 = ro->crop_y0; o.x1 = ro->crop_x1; o.y1 = ro->crop_y1;
  Counters tot;
  auto t0 = std::chrono::steady_clock::now();
  render(*sc, o, acc, acc_w, acc_sq, &tot);
  auto t1 = std::chrono::steady_clock::now();
  if (cnt) *cnt = orc_counters{tot.aabb, tot.tri, tot.sph, tot.inst, tot.rays_ext, tot.rays_shadow, tot.paths, tot.ext_bounces,
                               tot.aabb_ext, tot.tri_ext, tot.sph_ext, tot.inst_ext, tot.light_samples, tot.light_self};
  return std::chrono::duration<double>(t1 - t0).count();
}
// output.rs:96-102
void orc_resolve_pixels(int64_t n, const double* acc, const double* acc_w, double* out) {
  for (int64_t i = 0; i < n; i++) {
    double w = fmx(acc_w[i], 1.0);
    out[3 * i] = acc[3 * i] / w; out[3 * i + 1] = acc[3 * i + 1] / w; out[3 * i + 2] = acc[3 * i + 2] / w;
  }
}

}  // extern "C"
