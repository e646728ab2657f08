// k_plain64.hip -- the plain FP64 GPU form: one thread per pixel, the reference's per-pixel
// random stream across all samples (render.rs:142), reference data layout, f64 arithmetic
// without FMA contraction (this file is compiled with -ffp-contract=off).  It restates
// render.rs:137-191, integrator.rs, world.rs, bvh.rs, geometry.rs, env.rs and material.rs on
// the device and is the baseline the FP32 wavefront is measured against.
//
// Samples can be split over launches: the per-pixel RNG state, sums and Welford state persist
// in device buffers between chunks, so a chunked render equals an unchunked one.
#include "plain64.h"
#include "tpt_bsdf.h"

namespace tpt {

typedef vec3<double> d3;

struct Hit64 {
  double t;
  d3 p, n;
  int mat;
};

// ---------------------------------------------------------------- geometry.rs
__device__ __forceinline__ bool sphere_hit(const double* s, d3 o, d3 d, double tmin, double tmax, Hit64* h) {
  d3 c = ld3(s);
  double r = s[3];
  d3 oc = o - c;
  double a = dot(d, d);
  double b = dot(oc, d);
  double cc = dot(oc, oc) - r * r;
  double disc = b * b - a * cc;
  if (disc <= 0.0) return false;
  double sd = sqrt(disc);
  double t = (-b - sd) / a;
  if (t < tmin || t > tmax) {
    t = (-b + sd) / a;
    if (t < tmin || t > tmax) return false;
  }
  h->t = t;
  h->p = o + d * t;
  h->n = (h->p - c) / r;
  return true;
}
__device__ __forceinline__ bool tri_hit(const double* v, d3 o, d3 d, double time, double tmin, double tmax, Hit64* h) {
  // vertices_at(time) of a static triangle (geometry.rs:212-218), then Moller-Trumbore
  d3 v0 = ld3(v) * (1.0 - time) + ld3(v) * time;
  d3 v1 = ld3(v + 3) * (1.0 - time) + ld3(v + 3) * time;
  d3 v2 = ld3(v + 6) * (1.0 - time) + ld3(v + 6) * time;
  d3 e1 = v1 - v0, e2 = v2 - v0;
  d3 pvec = cross(d, e2);
  double det = dot(e1, pvec);
  if (fabs(det) < 1e-10) return false;
  double inv_det = 1.0 / det;
  d3 tvec = o - v0;
  double u = dot(tvec, pvec) * inv_det;
  if (u < 0.0 || u > 1.0) return false;
  d3 qvec = cross(tvec, e1);
  double vv = dot(d, qvec) * inv_det;
  if (vv < 0.0 || u + vv > 1.0) return false;
  double th = dot(e2, qvec) * inv_det;
  if (th < tmin || th > tmax) return false;
  h->t = th;
  h->p = o + d * th;
  h->n = vnorm(cross(e1, e2));
  return true;
}
__device__ __forceinline__ bool aabb_range(const double* b, d3 o, d3 inv, double tmin, double tmax, double* t0) {
  double t0x = (b[0] - o.x) * inv.x, t1x = (b[3] - o.x) * inv.x;
  tmin = fmax(tmin, fmin(t0x, t1x));
  tmax = fmin(tmax, fmax(t0x, t1x));
  if (tmax <= tmin) return false;
  double t0y = (b[1] - o.y) * inv.y, t1y = (b[4] - o.y) * inv.y;
  tmin = fmax(tmin, fmin(t0y, t1y));
  tmax = fmin(tmax, fmax(t0y, t1y));
  if (tmax <= tmin) return false;
  double t0z = (b[2] - o.z) * inv.z, t1z = (b[5] - o.z) * inv.z;
  tmin = fmax(tmin, fmin(t0z, t1z));
  tmax = fmin(tmax, fmax(t0z, t1z));
  if (tmax <= tmin) return false;
  *t0 = tmin;
  return true;
}

// ---------------------------------------------------------------- bvh.rs:248-351
// The reference's 64-entry stack; the host rejects trees deeper than 63 levels for this form,
// so the heap spill path of the reference is never needed.
__device__ bool bvh_hit(const SceneF64& S, const MeshDev& M, d3 o, d3 d, double time, double tmin, double tmax, Hit64* best,
                        uint32_t* n_box) {
  if (M.n_nodes == 0) return false;
  const RefNodeDev* nodes = S.nodes + M.node_off;
  const int32_t* idx = S.indices + M.idx_off;
  d3 inv = mk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  int32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  bool found = false;
  double t0;
  while (sp > 0) {
    const RefNodeDev& n = nodes[stack[--sp]];
    (*n_box)++;
    if (!aabb_range(n.bbox, o, inv, tmin, tmax, &t0)) continue;
    if (n.left == -1 && n.right == -1) {
      for (uint32_t k = n.start; k < n.start + n.count; k++) {
        int32_t ti = idx[k];
        Hit64 h;
        if (tri_hit(S.tris + 9 * (size_t)(M.tri_off + ti), o, d, time, tmin, tmax, &h)) {
          h.mat = S.tri_mat[M.tri_off + ti];
          tmax = h.t;
          *best = h;
          found = true;
        }
      }
    } else {
      int32_t a_id = n.left, b_id = n.right;
      if (a_id == -1) {
        if (b_id != -1) stack[sp++] = b_id;
        continue;
      }
      if (b_id == -1) {
        stack[sp++] = a_id;
        continue;
      }
      double a0 = 0, b0 = 0;
      *n_box += 2;
      bool ah = aabb_range(nodes[a_id].bbox, o, inv, tmin, tmax, &a0);
      bool bh = aabb_range(nodes[b_id].bbox, o, inv, tmin, tmax, &b0);
      if (ah && bh) {
        if (a0 <= b0) { stack[sp++] = b_id; stack[sp++] = a_id; }
        else { stack[sp++] = a_id; stack[sp++] = b_id; }
      } else if (ah) {
        stack[sp++] = a_id;
      } else if (bh) {
        stack[sp++] = b_id;
      }
    }
  }
  return found;
}

__device__ __forceinline__ d3 mat_vec(const double* m, d3 v) {
  return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// ---------------------------------------------------------------- world.rs:81-121
__device__ bool world_hit(const SceneF64& S, d3 o, d3 d, double time, double tmin, double tmax, Hit64* best, uint32_t* n_box) {
  double closest = tmax;
  bool found = false;
  for (int ii = 0; ii < S.n_inst; ii++) {
    const InstF64& I = S.inst[ii];
    if (I.mesh < 0 || I.mesh >= S.n_meshes) continue;
    d3 o_obj = mat_vec(I.a_inv, o - ld3(I.t));
    d3 d_obj = vnorm(mat_vec(I.a_inv, d));
    Hit64 h;
    if (bvh_hit(S, S.meshes[I.mesh], o_obj, d_obj, time, tmin, 1e30, &h, n_box)) {
      d3 p_world = mat_vec(I.a, h.p) + ld3(I.t);
      d3 n_world = vnorm(mat_vec(I.nrm, h.n));
      double t_world = dot(p_world - o, d);
      if (t_world > tmin && t_world < closest) {
        closest = t_world;
        best->t = t_world;
        best->p = p_world;
        best->n = n_world;
        best->mat = I.mat_override >= 0 ? I.mat_override : h.mat;
        found = true;
      }
    }
  }
  for (int s = 0; s < S.n_spheres; s++) {
    Hit64 h;
    if (sphere_hit(S.sph + 4 * s, o, d, tmin, closest, &h)) {
      closest = h.t;
      h.mat = S.sph_mat[s];
      *best = h;
      found = true;
    }
  }
  return found;
}

// ---------------------------------------------------------------- env.rs
__device__ __forceinline__ void dir_uv64(d3 dir, double* u, double* v, double* theta) {
  d3 d = vnorm(dir);
  double th = acos(fmin(fmax(d.y, -1.0), 1.0));
  double phi = atan2(d.z, d.x);
  if (phi < 0.0) phi += TPT_TAU;
  *u = phi / TPT_TAU;
  *v = th / TPT_PI;
  *theta = th;
}
__device__ __forceinline__ size_t as_usize(double v) { return v > 0.0 ? (size_t)v : 0; }
__device__ d3 env_lookup64(const EnvDev& E, d3 dir) {  // env.rs:92-121
  double u, v, th;
  dir_uv64(dir, &u, &v, &th);
  double x = u * (double)E.w;
  double y = v * ((double)E.h - 1.0);
  size_t x0 = as_usize(floor(x)) % (size_t)E.w;
  size_t x1 = (x0 + 1) % (size_t)E.w;
  size_t y0 = as_usize(floor(fmin(fmax(y, 0.0), (double)(E.h - 1))));
  size_t y1 = y0 + 1 < (size_t)(E.h - 1) ? y0 + 1 : (size_t)(E.h - 1);
  double fx = x - floor(x), fy = y - floor(y);
  d3 c00 = ld3(E.rgb64 + 3 * (y0 * E.w + x0)), c10 = ld3(E.rgb64 + 3 * (y0 * E.w + x1));
  d3 c01 = ld3(E.rgb64 + 3 * (y1 * E.w + x0)), c11 = ld3(E.rgb64 + 3 * (y1 * E.w + x1));
  d3 c0 = c00 * (1.0 - fx) + c10 * fx;
  d3 c1 = c01 * (1.0 - fx) + c11 * fx;
  return c0 * (1.0 - fy) + c1 * fy;
}
__device__ double env_pdf64(const EnvDev& E, d3 dir) {  // env.rs:156-182
  if (E.total <= 0.0) return 1.0 / (4.0 * TPT_PI);
  double u, v, theta;
  dir_uv64(dir, &u, &v, &theta);
  double fxv = floor(u * (double)E.w), fyv = floor(v * (double)E.h);
  if (fxv < 0.0) fxv = 0.0;
  if (fxv > (double)(E.w - 1)) fxv = (double)(E.w - 1);
  if (fyv < 0.0) fyv = 0.0;
  if (fyv > (double)(E.h - 1)) fyv = (double)(E.h - 1);
  size_t x = as_usize(fxv), y = as_usize(fyv);
  double theta0 = TPT_PI * (double)y / (double)E.h;
  double theta1 = TPT_PI * (double)(y + 1) / (double)E.h;
  double sin_theta = fmax(sin(theta), 1e-6);
  double row_weight = fmax(cos(theta0) - cos(theta1), 0.0);
  double l = fmax(luminance(ld3(E.rgb64 + 3 * (y * E.w + x))), 0.0);
  double weight = l * row_weight;
  if (weight <= 0.0) return 0.0;
  double pdf_uv = (weight / E.total) * (double)E.w * (double)E.h;
  return pdf_uv / (2.0 * TPT_PI * TPT_PI * sin_theta);
}
__device__ void env_sample64(const EnvDev& E, RngF64& rng, d3* dir, d3* li, double* pdf) {  // env.rs:124-152, 200-210
  bool uniform = E.total <= 0.0;
  int y = 0;
  double row_sum = 0.0;
  if (!uniform) {
    double r0 = rng.uni() * E.total;
    y = min(cdf_search(E.row_cdf, E.h + 1, r0), E.h - 1);
    row_sum = fmax(E.row_cdf[y + 1] - E.row_cdf[y], 0.0);
    if (row_sum <= 0.0) uniform = true;
  }
  if (uniform) {
    double u = rng.uni(), v = rng.uni();
    double z = 1.0 - 2.0 * u;
    double r = sqrt(fmax(1.0 - z * z, 0.0));
    double phi = TPT_TAU * v;
    *dir = mk(r * cos(phi), z, r * sin(phi));
    *li = env_lookup64(E, *dir);
    *pdf = 1.0 / (4.0 * TPT_PI);
    return;
  }
  double r1 = rng.uni() * row_sum;
  int x = min(cdf_search(E.col_cdf + (size_t)y * (E.w + 1), E.w + 1, r1), E.w - 1);
  double u = ((double)x + rng.uni()) / (double)E.w;
  double v = ((double)y + rng.uni()) / (double)E.h;
  double theta = TPT_PI * v, phi = TPT_TAU * u;
  double st = sin(theta);
  *dir = mk(cos(phi) * st, cos(theta), sin(phi) * st);
  *li = env_lookup64(E, *dir);
  *pdf = env_pdf64(E, *dir);
}

// ---------------------------------------------------------------- lights (world.rs:171-272)
__device__ __forceinline__ void tri_world_verts(const SceneF64& S, const LightF64& L, double time, d3* a, d3* b, d3* c) {
  const InstF64& I = S.inst[L.inst];
  const double* v = S.tris + 9 * (size_t)(S.meshes[L.mesh].tri_off + L.tri);
  d3 v0 = ld3(v) * (1.0 - time) + ld3(v) * time;
  d3 v1 = ld3(v + 3) * (1.0 - time) + ld3(v + 3) * time;
  d3 v2 = ld3(v + 6) * (1.0 - time) + ld3(v + 6) * time;
  *a = mat_vec(I.a, v0) + ld3(I.t);
  *b = mat_vec(I.a, v1) + ld3(I.t);
  *c = mat_vec(I.a, v2) + ld3(I.t);
}
struct LightSample64 {
  d3 pos, nrm, emit;
  double pdf;
};
__device__ bool sample_light64(const SceneF64& S, const LightsDev& Ls, RngF64& rng, double time, d3 p, LightSample64* out) {
  if (Ls.total <= 0.0 || Ls.n == 0) return false;
  double r = rng.uni() * Ls.total;
  int idx = min(cdf_search(Ls.cdf64, Ls.n, r), Ls.n - 1);
  const LightF64& L = Ls.l64[idx];
  double pdf_select = L.weight / Ls.total;
  d3 pos, nrm;
  double area;
  if (L.kind == 0) {
    const double* s = S.sph + 4 * L.idx;
    double u = rng.uni(), v = rng.uni();
    double z = 1.0 - 2.0 * u;
    double rr = sqrt(fmax(1.0 - z * z, 0.0));
    double phi = TPT_TAU * v;
    d3 dir = mk(rr * cos(phi), z, rr * sin(phi));
    pos = ld3(s) + dir * s[3];
    nrm = dir;
    area = 4.0 * TPT_PI * s[3] * s[3];
  } else {
    d3 a, b, c;
    tri_world_verts(S, L, time, &a, &b, &c);
    double u = rng.uni(), v = rng.uni();
    double su = sqrt(u);
    double b0 = 1.0 - su, b1 = v * su, b2 = 1.0 - b0 - b1;
    pos = a * b0 + b * b1 + c * b2;
    d3 n = cross(b - a, c - a);
    area = 0.5 * vlen(n);
    nrm = area > 0.0 ? n / (2.0 * area) : mk(0.0, 1.0, 0.0);
  }
  if (area <= 0.0) return false;
  d3 to_light = pos - p;
  double dist2 = dot(to_light, to_light);
  if (dist2 <= 1e-12) return false;
  double dist = sqrt(dist2);
  d3 wi = to_light / dist;
  double cos_light = fmax(dot(nrm, -wi), 0.0);
  if (cos_light <= 0.0) return false;
  double pdf_area = 1.0 / area;
  double pdf_omega = pdf_area * dist2 / cos_light;
  double pdf = pdf_select * pdf_omega;
  if (pdf <= 0.0) return false;
  out->pos = pos;
  out->nrm = nrm;
  out->emit = mk(L.emit[0], L.emit[1], L.emit[2]);
  out->pdf = pdf;
  return true;
}

// ---------------------------------------------------------------- integrator.rs:53-213
struct Count64 {
  uint32_t ext, shadow, box;
};
__device__ d3 radiance64(const Plain64Params& P, d3 ro, d3 rd, double time, RngF64& rng, Count64* cnt) {
  const SceneF64& S = P.S;
  d3 acc = mk(0.0, 0.0, 0.0), beta = mk(1.0, 1.0, 1.0);
  double last_bsdf_pdf = 0.0;
  bool last_non_delta = false;
  const bool has_env = P.env.kind == TPT_ENV_IMAGE;
  for (int bounce = 0; bounce < P.max_bounces; bounce++) {
    Hit64 hit;
    cnt->ext++;
    if (!world_hit(S, ro, rd, time, TPT_RAY_EPSILON, 1e30, &hit, &cnt->box)) {
      d3 bg;
      if (has_env) bg = env_lookup64(P.env, rd);
      else {
        double t = 0.5 * (rd.y + 1.0);
        bg = (1.0 - t) * mk(0.9, 0.95, 1.0) + t * mk(0.6, 0.75, 1.0);
      }
      d3 contrib = had(beta, bg);
      if (has_env && last_non_delta) contrib = contrib * mis_weight(last_bsdf_pdf, env_pdf64(P.env, rd));
      acc = acc + clamp_luminance(contrib, TPT_FIREFLY_CLAMP);
      break;
    }
    if (bounce >= P.rr_start) {
      double p = fmax(fmin(fmax(fmax(beta.x, beta.y), beta.z), 0.95), 0.05);
      if (rng.uni() > p) break;
      beta = beta / p;
    }
    const int kind = P.mats.kind[hit.mat];
    const double* mp = P.mats.p64 + 4 * hit.mat;
    const d3 mc = ld3(mp);
    if (kind == TPT_MAT_DIFFUSE_LIGHT) {
      acc = acc + clamp_luminance(had(beta, mc), TPT_FIREFLY_CLAMP);
      break;
    }
    d3 n = dot(hit.n, rd) < 0.0 ? hit.n : -hit.n;
    if (!mat_is_delta(kind)) {
      if (has_env) {  // nee_environment
        d3 wi, li;
        double pdf_env;
        env_sample64(P.env, rng, &wi, &li, &pdf_env);
        double cs = fmax(dot(n, wi), 0.0);
        if (cs > 0.0 && pdf_env > 0.0) {
          Hit64 tmp;
          cnt->shadow++;
          if (!world_hit(S, hit.p + TPT_RAY_EPSILON * wi, wi, time, TPT_RAY_EPSILON, 1e30, &tmp, &cnt->box)) {
            d3 f;
            double pdf_b;
            bsdf_eval(kind, mc, mp[3], vnorm(-rd), wi, n, &f, &pdf_b);
            double w = mis_weight(pdf_env, pdf_b);
            acc = acc + had(had(beta, f), li) * (cs * w / pdf_env);
          }
        }
      }
      LightSample64 ls;  // nee_area_light
      if (sample_light64(S, P.lights, rng, time, hit.p, &ls)) {
        d3 to_light = ls.pos - hit.p;
        double dist2 = dot(to_light, to_light);
        double dist = sqrt(dist2);
        if (dist > 1e-6) {
          d3 wi = to_light / dist;
          double cs = fmax(dot(n, wi), 0.0);
          if (cs > 0.0) {
            double tmax = fmax(dist - TPT_RAY_EPSILON, TPT_RAY_EPSILON);
            Hit64 tmp;
            cnt->shadow++;
            if (!world_hit(S, hit.p + TPT_RAY_EPSILON * wi, wi, time, TPT_RAY_EPSILON, tmax, &tmp, &cnt->box)) {
              d3 f;
              double pdf_b;
              bsdf_eval(kind, mc, mp[3], vnorm(-rd), wi, n, &f, &pdf_b);
              if (ls.pdf > 0.0) acc = acc + had(had(beta, f), ls.emit) * (cs * mis_weight(ls.pdf, pdf_b) / ls.pdf);
            }
          }
        }
      }
    }
    BsdfSample<double> bs;
    if (!bsdf_sample(kind, mc, mp[3], rd, hit.p, hit.n, hit.t, rng, &bs)) break;
    last_bsdf_pdf = bs.pdf;
    last_non_delta = !bs.is_delta;
    beta = had(beta, bs.weight);
    ro = bs.o;
    rd = bs.d;
  }
  return acc;
}

// ---------------------------------------------------------------- render.rs:137-191
__global__ void __launch_bounds__(TPT_PLAIN_BLOCK) k_plain64(Plain64Params P) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;  // domain pixel: indexes the sums and state
  Count64 cnt{0, 0, 0};
  uint32_t paths = 0;
  if (pix < P.n_pix && !P.done[pix]) {
    const int img = P.pix_map ? P.pix_map[pix] : pix;
    const int x = img % P.width, y = img / P.width;
    RngF64 rng(1);
    rng.p.state = P.s_begin == 0 ? Pcg32(seed_for((uint32_t)x, (uint32_t)y, 0) ^ P.seed).state : P.rng[pix];
    d3 c = P.s_begin == 0 ? mk(0.0, 0.0, 0.0) : ld3(P.acc + 3 * pix);
    d3 sq = P.s_begin == 0 || !P.acc_sq ? mk(0.0, 0.0, 0.0) : ld3(P.acc_sq + 3 * pix);
    uint32_t n = P.s_begin == 0 ? 0u : P.nsamp[pix];
    double mean = P.s_begin == 0 ? 0.0 : P.welford[2 * pix], m2 = P.s_begin == 0 ? 0.0 : P.welford[2 * pix + 1];
    bool done = false;
    const CameraF64& C = P.cam;
    const uint32_t max_spp = (uint32_t)max(P.spp, 1);
    const uint32_t min_spp = min((uint32_t)max(P.min_spp, 1), max_spp);
    for (int s = P.s_begin; s < P.s_end && !done; s++) {
      double jx = rng.uni(), jy = rng.uni();
      double sx = ((double)x + jx) * P.inv_w * 2.0 - 1.0;
      double sy = 1.0 - ((double)y + jy) * P.inv_h * 2.0;
      // Camera::ray (ray.rs:91-108)
      d3 dir_base = vnorm(-ld3(C.w) + (sx * C.half_w) * ld3(C.u) + (sy * C.half_h) * ld3(C.v));
      d3 lens = mk(0.0, 0.0, 0.0);
      if (C.lens_radius > 0.0) {
        double dx = 0.0, dy = 0.0;
        for (int tries = 0; tries < 1024; tries++) {
          double a = 2.0 * rng.uni() - 1.0, b = 2.0 * rng.uni() - 1.0;
          if (a * a + b * b < 1.0) { dx = a; dy = b; break; }
        }
        lens = ld3(C.u) * (dx * C.lens_radius) + ld3(C.v) * (dy * C.lens_radius);
      }
      d3 origin = ld3(C.o) + lens;
      d3 focus_point = ld3(C.o) + dir_base * C.focus_dist;
      d3 dir = vnorm(focus_point - origin);
      double time = C.shutter_open + (C.shutter_close - C.shutter_open) * rng.uni();
      d3 L = radiance64(P, origin, dir, time, rng, &cnt);
      paths++;
      c = c + L;
      sq = sq + had(L, L);
      n += 1;
      if (P.adaptive) {
        double lum = luminance(L);
        double delta = lum - mean;
        mean += delta / (double)n;
        double delta2 = lum - mean;
        m2 += delta * delta2;
        if (n >= min_spp) {
          double var = n > 1 ? m2 / (double)(n - 1) : __longlong_as_double(0x7ff0000000000000ll);
          double rel = sqrt(var) / fmax(fabs(mean), 1e-4);
          if (rel < P.threshold) done = true;
        }
      }
    }
    P.rng[pix] = rng.p.state;
    P.acc[3 * pix] = c.x; P.acc[3 * pix + 1] = c.y; P.acc[3 * pix + 2] = c.z;
    if (P.acc_sq) { P.acc_sq[3 * pix] = sq.x; P.acc_sq[3 * pix + 1] = sq.y; P.acc_sq[3 * pix + 2] = sq.z; }
    P.nsamp[pix] = n;
    P.welford[2 * pix] = mean;
    P.welford[2 * pix + 1] = m2;
    P.acc_w[pix] = (double)n;
    if (done) P.done[pix] = 1;
  }
  // one atomic per wave per counter
  unsigned long long e = cnt.ext, sh = cnt.shadow, pa = paths;
  for (int off = 32; off > 0; off >>= 1) {
    e += (unsigned long long)__shfl_down((long long)e, off);
    sh += (unsigned long long)__shfl_down((long long)sh, off);
    pa += (unsigned long long)__shfl_down((long long)pa, off);
  }
  if (__lane_id() == 0) {
    if (e) atomicAdd((unsigned long long*)&P.stats[0], e);
    if (sh) atomicAdd((unsigned long long*)&P.stats[1], sh);
    if (pa) atomicAdd((unsigned long long*)&P.stats[2], pa);
  }
}

hipError_t launch_plain64(const Plain64Params& P, hipStream_t s) {
  if (P.n_pix <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plain64, dim3((P.n_pix + TPT_PLAIN_BLOCK - 1) / TPT_PLAIN_BLOCK), dim3(TPT_PLAIN_BLOCK), 0, s, P);
  return hipGetLastError();
}

}  // namespace tpt
