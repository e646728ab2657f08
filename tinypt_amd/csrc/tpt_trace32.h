// tpt_trace32.h -- FP32 closest-hit / any-hit traversal of the two-level scene (device).
//
// Semantics follow World::hit (world.rs:81-121) with Bvh::hit (bvh.rs:248-351), Triangle::hit
// (geometry.rs:236-273) and Sphere::hit (geometry.rs:174-195):
//   * instance rays are taken to object space (A^-1 (o - t), A^-1 d) but d is NOT normalised,
//     so object-space t equals world t and the running closest distance culls every level;
//     the reference's object-space bound t_obj >= tmin on the normalised ray becomes
//     t >= tmin / |A^-1 d|, combined with the world bound t > tmin;
//   * an entering ray is first re-based, in FP64, to its point nearest the instance's centre
//     (t shifted by s, the instance's distances taken relative to it): an FP32 object-space
//     origin far from the object (a camera 50 units away, an instance scaled by 0.2) would
//     carry absolute errors of 1e-5 -- a few thousandths of a triangle of a 1.3M-triangle mesh
//     -- and let rays through the cracks between neighbouring triangles;
//   * Moller-Trumbore in FP32, two-sided, |det| < 1e-10 * |A^-1 d| rejected (the reference's
//     1e-10 applies to the normalised object ray);
//   * spheres use the Ray Tracing Gems (Haines et al. 2019, ch. 7) quadratic evaluated in FP64
//     on the FP32 ray, so large spheres (r = 1000 ground) keep sub-1e-9 hit accuracy;
//   * equal-t ties resolve as the reference's loop order does: a later triangle of the same
//     instance wins, an earlier instance wins over a later one, a sphere wins over instances
//     and over earlier spheres.  Candidates carry a key; (t, key) is compared
//     lexicographically with larger keys winning ties.
// The traversal keeps the current node in a register and pushes only far children onto a
// per-lane stack in LDS (stride = block size, so lanes hit distinct banks), overflowing to a
// per-thread global area when a tree is deeper than the LDS part.
// Small scenes are flattened at build time: every instance's triangles in world space under one
// tree, held by an identity pseudo-instance (SceneF32::flat_inst).  A hit there reports the
// triangle's own instance (TPT_HIT_FLAT | instance) and takes that instance's tie key.
#pragma once

#include "tpt_math.h"

namespace tpt {

struct SceneF32 {
  const Node4* __restrict__ nodes;
  const TriF32* __restrict__ tris;
  const int32_t* __restrict__ tri_mat;
  const int32_t* __restrict__ tri_orig;
  const int32_t* __restrict__ tri_inst;  // flattened slots: original instance | TPT_TRI_FLIP
  const InstF32* __restrict__ inst;
  const InstF64* __restrict__ inst64;    // A^-1 and t in FP64, to re-base entering rays
  const double* __restrict__ inst_c64;   // per instance: the world point rays are re-based to
  const int32_t* __restrict__ tlas_items;
  const double* __restrict__ sph;  // c.xyz, r (FP64)
  const int32_t* __restrict__ sph_mat;
  int32_t tlas_root;
  int32_t n_spheres;
  int32_t flat_inst;  // the pseudo-instance holding the flattened world-space tree, or -1
  int32_t flat_root;  // that tree's root when the top level holds nothing else, or TPT_REF_EMPTY
};

struct HitF32 {
  float t;
  int32_t inst;  // instance index (| TPT_HIT_FLAT for a flattened triangle), TPT_ITEM_SPHERE | sphere, or -1
  int32_t prim;  // reordered triangle slot, -1 for spheres / miss
};

// OVF = false: the tree's stack need fits the LDS part, so there is no overflow path.
template <bool OVF>
struct LaneStackT {
  int32_t* lds;     // this lane's first LDS entry
  int32_t* glob;    // this lane's global overflow area (may be null if never needed)
  int stride;       // LDS stride in entries (block size)
  int cap;          // LDS entries per lane
  TPT_HD void push(int& sp, int32_t v) const {
    if (!OVF || sp < cap) lds[sp * stride] = v;
    else glob[sp - cap] = v;
    ++sp;
  }
  TPT_HD int32_t pop(int& sp) const {
    --sp;
    return !OVF || sp < cap ? lds[sp * stride] : glob[sp - cap];
  }
};
using LaneStack = LaneStackT<true>;
// Whether a traversal stack type carries the overflow path (deep trees); true unless known.
template <class Stack> struct StackOverflows { static constexpr bool value = true; };
template <> struct StackOverflows<LaneStackT<false>> { static constexpr bool value = false; };

TPT_HD float safe_rcp(float d) { return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d)); }

// 1 / x to about 2^-22 relative: the hardware reciprocal on the device (a correctly rounded FP64
// division costs a dozen instructions and, in the traversal kernels, registers that spill);
// the host build of the traversal divides.
TPT_HD double approx_rcp64(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcp(x);
#else
  return 1.0 / x;
#endif
}

#define TPT_KEY_SPHERE(s) (0x40000000 + (int32_t)(s))
#define TPT_KEY_INST(i) (0x3fffffff - (int32_t)(i))
#define TPT_STACK_SENTINEL TPT_REF_EMPTY

// An empty asm statement that makes v opaque to the optimiser at this point: FP64 terms that
// depend only on the ray (its direction's squared length and reciprocal) would otherwise be
// hoisted out of the traversal loop from the sphere test and the instance entry, and hold FP64
// registers -- spilled to scratch -- on every ray, whether it meets a sphere or instance or not.
// No instruction is emitted.  The sphere test always keeps its terms; the instance entry keeps
// them in the kernels for trees within the LDS stack (scenes of few instances or one flattened
// tree, entered by few rays or none), while the deep-tree kernels, whose rays enter instances
// at most steps, are faster with them hoisted.
TPT_HD void opaque(float& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v));
#endif
}

// FP64 sphere root in [tmin, tmax] (near root first, then far), geometry.rs:174-195.
TPT_HD bool sphere_hit64(const double* s, vec3<float> of, vec3<float> df, double tmin, double tmax,
                                             double* t_out) {
  opaque(df.x);  // the direction's FP64 terms stay in this test (see opaque)
  opaque(df.y);
  opaque(df.z);
  vec3<double> o = mk((double)of.x, (double)of.y, (double)of.z), d = mk((double)df.x, (double)df.y, (double)df.z);
  vec3<double> c = mk(s[0], s[1], s[2]);
  double r = s[3];
  vec3<double> f = o - c;
  double a = dot(d, d);
  double ia = 1.0 / a;  // d is a unit FP32 vector: a is 1 within a few ulp, and x * ia is x / a
                        // within an FP64 ulp, far below the FP32 rounding of the root
  double bp = -dot(f, d);
  vec3<double> l = f + (bp * ia) * d;
  double disc = a * (r * r - dot(l, l));  // = b^2 - a c without cancellation
  if (disc <= 0.0) return false;
  double cc = dot(f, f) - r * r;
  double q = bp + copysign(sqrt(disc), bp);
  double ta = cc / q, tb = q * ia;
  double tn = fmin(ta, tb), tf = fmax(ta, tb);
  double t = tn;
  if (t < tmin || t > tmax) {
    t = tf;
    if (t < tmin || t > tmax) return false;
  }
  *t_out = t;
  return true;
}

// Byte offsets of the near planes of a Node4 for a ray direction: lo_x / hi_x at 0 / 16,
// lo_y / hi_y at 32 / 48, lo_z / hi_z at 64 / 80; the far plane is at near ^ 16.  Along an axis
// with a positive inverse direction the lower plane gives the smaller distance (the FMA is
// monotone in the plane coordinate), so loading near and far planes by the ray's signs gives
// the slab entry and exit distances without a min / max per axis -- bit for bit the same values.
TPT_HD uint32_t near_off(float inv, uint32_t axis) { return axis * 32u + (inv < 0.f ? 16u : 0u); }
static_assert(offsetof(Node4, hi_x) == 16 && offsetof(Node4, lo_y) == 32 && offsetof(Node4, hi_y) == 48 &&
                  offsetof(Node4, lo_z) == 64 && offsetof(Node4, hi_z) == 80 && offsetof(Node4, child) == 96,
              "near_off assumes the Node4 plane layout");

// One ray's traversal, resumable between steps so that a wave can swap finished rays for new
// ones (k_extend).  ANY: stop at the first hit in range (shadow rays).
// A step is one pass of the loop: an inner loop descends through inner nodes (near child kept
// in `ref`, far children pushed); the step then handles the leaf, top-level item or instance
// exit that the descent stopped at, and pops the next entry.
template <bool ANY>
struct TraceF32 {
  vec3<float> o, d, inv_w, oi_w;  // world ray
  vec3<float> co, cd, inv, oi;    // ray of the current level (object space inside an instance)
  uint32_t nx_w, ny_w, nz_w, nx, ny, nz;
  float tmin;
  float shift;    // world t of the current level's origin (0 at the top level)
  float lo;       // lower t bound at the current level
  float det_eps;
  float best_t;
  int32_t best_key;
  int32_t cur_inst;  // instance being traversed (-1: top level)
  int32_t cur_key;
  int32_t ref;
  int sp;
  bool found;
  HitF32 hit;

  // False when there is nothing to traverse (the ray misses; hit is set to "no hit").
  TPT_HD bool init(const SceneF32& S, vec3<float> ro, vec3<float> rd, float rtmin, float rtmax) {
    hit.t = rtmax;
    hit.inst = -1;
    hit.prim = -1;
    found = false;
    ref = S.tlas_root;
    if (ref == TPT_REF_EMPTY) return false;
    cur_inst = -1;
    if (S.flat_root != TPT_REF_EMPTY) {  // start inside the flattened tree (world space)
      ref = S.flat_root;
      cur_inst = S.flat_inst;
    }
    o = ro;
    d = rd;
    tmin = rtmin;
    inv_w = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    oi_w = mk(o.x * inv_w.x, o.y * inv_w.y, o.z * inv_w.z);
    co = o; cd = d; inv = inv_w; oi = oi_w;
    nx_w = near_off(inv_w.x, 0); ny_w = near_off(inv_w.y, 1); nz_w = near_off(inv_w.z, 2);
    nx = nx_w; ny = ny_w; nz = nz_w;
    lo = tmin;
    shift = 0.f;
    det_eps = 1e-10f;
    best_t = rtmax;
    best_key = -1;
    cur_key = 0;
    sp = 0;
    return true;
  }

  // One step; false when the ray is done (hit and found hold the result).
  template <class Stack>
  TPT_HD bool step(const SceneF32& S, const Stack& st) {
    // ---- descend inner nodes: four child boxes from one 128 B node; the nearest hit child
    // is entered, the others are pushed farthest first
    bool have = true;
    while (ref >= 0 && ref != TPT_STACK_SENTINEL) {
      // 32-bit byte offsets from the node array (the host keeps it under 4 GiB), so each load
      // is the array base plus one offset register
      const char* nb = reinterpret_cast<const char*>(S.nodes);
      const uint32_t pn = (uint32_t)ref * (uint32_t)sizeof(Node4);
      const float4 nx4 = *reinterpret_cast<const float4*>(nb + (pn + nx));
      const float4 fx4 = *reinterpret_cast<const float4*>(nb + (pn + (nx ^ 16u)));
      const float4 ny4 = *reinterpret_cast<const float4*>(nb + (pn + ny));
      const float4 fy4 = *reinterpret_cast<const float4*>(nb + (pn + (ny ^ 16u)));
      const float4 nz4 = *reinterpret_cast<const float4*>(nb + (pn + nz));
      const float4 fz4 = *reinterpret_cast<const float4*>(nb + (pn + (nz ^ 16u)));
      const int4 ch = *reinterpret_cast<const int4*>(nb + (pn + 96u));
      float dn[4];
      int32_t cr[4] = {ch.x, ch.y, ch.z, ch.w};
      const float bnx[4] = {nx4.x, nx4.y, nx4.z, nx4.w}, bfx[4] = {fx4.x, fx4.y, fx4.z, fx4.w};
      const float bny[4] = {ny4.x, ny4.y, ny4.z, ny4.w}, bfy[4] = {fy4.x, fy4.y, fy4.z, fy4.w};
      const float bnz[4] = {nz4.x, nz4.y, nz4.z, nz4.w}, bfz[4] = {fz4.x, fz4.y, fz4.z, fz4.w};
      int nhit = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float xn = fmaf(bnx[k], inv.x, -oi.x), xf = fmaf(bfx[k], inv.x, -oi.x);
        const float yn = fmaf(bny[k], inv.y, -oi.y), yf = fmaf(bfy[k], inv.y, -oi.y);
        const float zn = fmaf(bnz[k], inv.z, -oi.z), zf = fmaf(bfz[k], inv.z, -oi.z);
        const float tn = fmaxf(fmaxf(xn, yn), fmaxf(zn, lo));
        const float tf = fminf(fminf(xf, yf), fminf(zf, best_t)) * 1.0000003f;
        const bool h = tn <= tf && cr[k] != TPT_REF_EMPTY;
        dn[k] = h ? tn : INFINITY;
        nhit += h ? 1 : 0;
      }
      if (nhit == 0) {
        have = false;
        break;
      }
      if (ANY) {
        // any hit ends the ray, so the visiting order does not matter: enter the first hit
        // child, push the others unsorted
        int32_t nref = TPT_REF_EMPTY;
        bool got = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (dn[k] != INFINITY) {
            if (got) st.push(sp, cr[k]);
            else nref = cr[k];
            got = true;
          }
        }
        ref = nref;
        continue;
      }
      // sort (distance, child) ascending: 5-comparator network; misses sort last
#define TPT_CSWAP(a, b)                                              \
  if (dn[b] < dn[a]) {                                               \
    const float td = dn[a]; dn[a] = dn[b]; dn[b] = td;                \
    const int32_t tr = cr[a]; cr[a] = cr[b]; cr[b] = tr;              \
  }
      TPT_CSWAP(0, 1) TPT_CSWAP(2, 3) TPT_CSWAP(0, 2) TPT_CSWAP(1, 3) TPT_CSWAP(1, 2)
#undef TPT_CSWAP
      if (nhit > 3) st.push(sp, cr[3]);
      if (nhit > 2) st.push(sp, cr[2]);
      if (nhit > 1) st.push(sp, cr[1]);
      ref = cr[0];
    }
    if (have) {
      if (ref == TPT_STACK_SENTINEL) {
        // ---- leaving an instance: back to the world ray (best_t back in world units: it is the
        // best hit's distance, or the ray's tmax while nothing is hit)
        co = o; cd = d; inv = inv_w; oi = oi_w; lo = tmin; det_eps = 1e-10f; cur_inst = -1;
        nx = nx_w; ny = ny_w; nz = nz_w;
        best_t = hit.t;
        shift = 0.f;
      } else if (cur_inst < 0) {
        // ---- top-level leaf: one item
        const int32_t item = S.tlas_items[(uint32_t)ref & TPT_LEAF_START_MASK];
        if ((uint32_t)item & TPT_ITEM_SPHERE) {
          const int32_t s = (int32_t)((uint32_t)item & ~TPT_ITEM_SPHERE);
          double th;
          if (sphere_hit64(S.sph + 4 * s, o, d, (double)tmin, (double)best_t, &th)) {
            const float tf = (float)th;
            const int32_t key = TPT_KEY_SPHERE(s);
            if (tf < best_t || key >= best_key) {
              best_t = tf; best_key = key; found = true;
              hit.t = tf; hit.inst = (int32_t)((uint32_t)s | TPT_ITEM_SPHERE); hit.prim = -1;
              if (ANY) return false;
            }
          }
        } else {
          const InstF32& I = S.inst[item];
          const int32_t root = I.root;
          if (root != TPT_REF_EMPTY && item == S.flat_inst) {
            // the flattened world-space tree: the ray stays in world space
            cur_inst = item;
            st.push(sp, TPT_STACK_SENTINEL);
            ref = root;
            return true;
          }
          if (root != TPT_REF_EMPTY) {
            const float4* w = reinterpret_cast<const float4*>(I.w2o);
            const float4 r0 = w[0], r1 = w[1], r2 = w[2];
            // re-base the origin in FP64 to o + s d, the ray's point nearest the instance's
            // centre c (s is an FP32 number, so that the shift of t is exactly s; any s near
            // the nearest point serves, so |d|^2 is divided by with an approximate reciprocal)
            {
              const double* c = S.inst_c64 + 3 * (size_t)item;
              const InstF64& J = S.inst64[item];
              float fdx = d.x, fdy = d.y, fdz = d.z, fox = o.x, foy = o.y, foz = o.z;
              if constexpr (!StackOverflows<Stack>::value) {
                opaque(fdx); opaque(fdy); opaque(fdz); opaque(fox); opaque(foy); opaque(foz);
              }
              const double dx = fdx, dy = fdy, dz = fdz, ox = fox, oy = foy, oz = foz;
              shift = (float)(((c[0] - ox) * dx + (c[1] - oy) * dy + (c[2] - oz) * dz) *
                              approx_rcp64(dx * dx + dy * dy + dz * dz));
              const double s = shift;
              const double px = ox + s * dx - J.t[0], py = oy + s * dy - J.t[1], pz = oz + s * dz - J.t[2];
              const double* m = J.a_inv;
              co = mk((float)(m[0] * px + m[1] * py + m[2] * pz), (float)(m[3] * px + m[4] * py + m[5] * pz),
                      (float)(m[6] * px + m[7] * py + m[8] * pz));
            }
            cd = mk(fmaf(r0.x, d.x, fmaf(r0.y, d.y, r0.z * d.z)), fmaf(r1.x, d.x, fmaf(r1.y, d.y, r1.z * d.z)),
                    fmaf(r2.x, d.x, fmaf(r2.y, d.y, r2.z * d.z)));
            const float len = sqrtf(dot(cd, cd));
            lo = fmaxf(tmin, tmin / len) - shift;
            best_t -= shift;
            det_eps = 1e-10f * len;
            inv = mk(safe_rcp(cd.x), safe_rcp(cd.y), safe_rcp(cd.z));
            oi = mk(co.x * inv.x, co.y * inv.y, co.z * inv.z);
            nx = near_off(inv.x, 0); ny = near_off(inv.y, 1); nz = near_off(inv.z, 2);
            cur_inst = item;
            cur_key = TPT_KEY_INST(item);
            st.push(sp, TPT_STACK_SENTINEL);
            ref = root;
            return true;
          }
        }
      } else {
        // ---- bottom-level leaf: up to 8 triangles in consecutive slots
        const uint32_t start = (uint32_t)ref & TPT_LEAF_START_MASK;
        const uint32_t cnt = (((uint32_t)ref >> 28) & 7u) + 1u;
        for (uint32_t k = 0; k < cnt; k++) {
          const float4* tp = reinterpret_cast<const float4*>(S.tris + start + k);
          const float4 a = tp[0], b = tp[1], c = tp[2];
          // the ray in the triangle's frame (TriF32): the plane at w = 0, the triangle where
          // u, v >= 0 and u + v <= 1 -- the hits of Triangle::hit (geometry.rs:236-273) up to
          // rounding; comparisons written so that NaN (degenerate triangles, rays in the
          // plane) never hits
          const vec3<float> tv = mk(co.x - a.w, co.y - b.w, co.z - c.w);
          const float dz = c.x * cd.x + c.y * cd.y + c.z * cd.z;
          // the reference's |det| < 1e-10 test: det = -(n . d) = -dz |n|^2, |n|^2 = 1 / |row 2|^2
          if (!(fabsf(dz) >= det_eps * (c.x * c.x + c.y * c.y + c.z * c.z))) continue;
          const float th = -(c.x * tv.x + c.y * tv.y + c.z * tv.z) / dz;
          if (!(th >= lo && th <= best_t)) continue;
          const float u = fmaf(th, a.x * cd.x + a.y * cd.y + a.z * cd.z, a.x * tv.x + a.y * tv.y + a.z * tv.z);
          if (!(u >= 0.f && u <= 1.f)) continue;
          const float v = fmaf(th, b.x * cd.x + b.y * cd.y + b.z * cd.z, b.x * tv.x + b.y * tv.y + b.z * tv.z);
          if (!(v >= 0.f && u + v <= 1.f)) continue;
          int32_t key = cur_key, hinst = cur_inst;
          if (cur_inst == S.flat_inst) {  // world-space triangle: its own instance decides ties
            const int32_t i0 = S.tri_inst[start + k] & (int32_t)~TPT_TRI_FLIP;
            key = TPT_KEY_INST(i0);
            hinst = (int32_t)((uint32_t)i0 | TPT_HIT_FLAT);
          }
          if (th < best_t || key >= best_key) {
            best_t = th; best_key = key; found = true;
            hit.t = th + shift; hit.inst = hinst; hit.prim = (int32_t)(start + k);
            if (ANY) return false;
          }
        }
      }
    }
    if (sp == 0) return false;
    ref = st.pop(sp);
    return true;
  }
};

// Returns true on a hit.  ANY: stop at the first hit in range (shadow rays).
template <bool ANY, class Stack>
TPT_HD bool trace_f32(const SceneF32& S, vec3<float> o, vec3<float> d, float tmin, float tmax, HitF32* hit,
                      const Stack& st) {
  TraceF32<ANY> tr;
  if (tr.init(S, o, d, tmin, tmax)) {
    while (tr.step(S, st)) {
    }
  }
  *hit = tr.hit;
  return tr.found;
}

}  // namespace tpt
