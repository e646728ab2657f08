// host_trace.hip -- test hook: the FP32 traversal (tpt_trace32.h) compiled for the host and run
// over the product's host scene build, so CPU tests can compare it with the oracle's World::hit
// without a GPU.  Not part of the product library; built by tests/native/build_host_trace.py.
#include <cmath>
#include <vector>

#include "scene_host.h"
#include "tpt_selfocc64.h"
#include "tpt_trace32.h"

using namespace tpt;

// The kernels' stack (a few "LDS" entries, the rest in the overflow area) with the deepest
// entry count recorded, so tests can hold it against HostScene::stack4.
struct CheckedStack {
  LaneStack in;
  int* max_sp;
  void push(int& sp, int32_t v) const {
    in.push(sp, v);
    if (sp > *max_sp) *max_sp = sp;
  }
  int32_t pop(int& sp) const { return in.pop(sp); }
};

extern "C" int tpt_test_host_trace(const tpt_scene_desc* d, int64_t n, const float* rays, float* t, int32_t* kind,
                                   int32_t* inst, int32_t* prim, int32_t* flat, int32_t* stack_bound,
                                   int32_t* stack_max) {
  HostScene H;
  if (!H.build(d)) return -1;
  SceneF32 S{};
  S.nodes = H.nodes4.data();
  S.tris = H.tris32.data();
  S.tri_mat = H.tri_mat32.data();
  S.tri_orig = H.tri_orig.data();
  S.tri_inst = H.flat ? H.tri_inst32.data() : nullptr;
  S.inst = H.inst32.data();
  S.inst64 = H.inst64.data();
  S.inst_c64 = H.inst_c64.data();
  S.tlas_items = H.tlas_items.data();
  S.sph = H.sph.data();
  S.sph_mat = H.sph_mat.data();
  S.tlas_root = H.tlas_root;
  S.n_spheres = (int32_t)H.sph_mat.size();
  S.flat_inst = H.flat ? H.flat_inst : -1;
  S.flat_root = H.flat_only_root();
  *flat = H.flat ? 1 : 0;
  // room far beyond the bound, so that a bound that is too small shows up in stack_max rather
  // than as a write past the buffer
  std::vector<int32_t> stack((size_t)H.stack4 + 4096);
  int max_sp = 0;
  CheckedStack st{LaneStack{stack.data(), stack.data() + 4, 1, 4}, &max_sp};  // 4 "LDS" entries, the rest overflow
  *stack_bound = H.stack4;
  for (int64_t i = 0; i < n; i++) {
    const float* r = rays + 8 * i;
    HitF32 h;
    const bool f = trace_f32<false>(S, mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), r[6], r[7], &h, st);
    if (!f) {
      t[i] = INFINITY; kind[i] = TPT_HIT_NONE; inst[i] = -1; prim[i] = -1;
    } else if ((uint32_t)h.inst & TPT_ITEM_SPHERE) {
      t[i] = h.t; kind[i] = TPT_HIT_SPHERE; inst[i] = (int32_t)((uint32_t)h.inst & ~TPT_ITEM_SPHERE); prim[i] = -1;
    } else {
      t[i] = h.t; kind[i] = TPT_HIT_TRI; inst[i] = (int32_t)((uint32_t)h.inst & ~TPT_HIT_FLAT);
      prim[i] = H.tri_orig[(size_t)h.prim];
    }
    if (max_sp > H.stack4 + 4000) return -2;
  }
  *stack_max = max_sp;
  return 0;
}

// World box of an instance of a triangle mesh as the scene build makes it (the cut of the mesh's
// FP32 tree through A, t), and, for comparison, the box of the root box's eight corners.
extern "C" int tpt_test_instance_box(int64_t n_tris, const double* tri_v, const double* a, const double* t,
                                     double* box_cut, double* box_root) {
  std::vector<DAabb> bounds((size_t)n_tris);
  for (int64_t i = 0; i < n_tris; i++) bounds[(size_t)i] = triangle_bounds(tri_v + 9 * i);
  const RefBvh tree = build_sah_bvh(bounds, 4);
  if (tree.nodes.empty()) return -1;
  const DAabb c = instance_world_box(tree_cut_boxes(tree, kInstanceCutDepth), a, t);
  const DAabb r = instance_world_box(std::vector<DAabb>{tree.nodes[0].bbox}, a, t);
  for (int k = 0; k < 3; k++) {
    box_cut[k] = c.mn[k]; box_cut[3 + k] = c.mx[k];
    box_root[k] = r.mn[k]; box_root[3 + k] = r.mx[k];
  }
  return 0;
}

// The FP32 kernels' light self-occlusion decision for n light samples, with the inputs the
// kernels give it: each FP32 ray rays[8i..] (o, d, tmin, tmax) is traced by the FP32 traversal,
// its shading point rebuilt in FP64 (hit_point64), and light light[i] tested at the 53-bit
// uniforms uv[2i..].  out[i] = -1 when the ray misses, else 1 when the sampled light shadows its
// own sample, 0 when not.
extern "C" int tpt_test_selfocc(const tpt_scene_desc* d, int64_t n, const float* rays, const int32_t* light,
                                const double* uv, int32_t* out) {
  HostScene H;
  if (!H.build(d)) return -1;
  SceneF32 S{};
  S.nodes = H.nodes4.data();
  S.tris = H.tris32.data();
  S.tri_mat = H.tri_mat32.data();
  S.tri_orig = H.tri_orig.data();
  S.tri_inst = H.flat ? H.tri_inst32.data() : nullptr;
  S.inst = H.inst32.data();
  S.inst64 = H.inst64.data();
  S.inst_c64 = H.inst_c64.data();
  S.tlas_items = H.tlas_items.data();
  S.sph = H.sph.data();
  S.sph_mat = H.sph_mat.data();
  S.tlas_root = H.tlas_root;
  S.n_spheres = (int32_t)H.sph_mat.size();
  S.flat_inst = H.flat ? H.flat_inst : -1;
  S.flat_root = H.flat_only_root();
  SceneF64 S64{};
  S64.nodes = H.ref_nodes.data();
  S64.indices = H.ref_indices.data();
  S64.tris = H.tri_v.data();
  S64.tri_mat = H.tri_mat.data();
  S64.meshes = H.meshes.data();
  S64.inst = H.inst64.data();
  S64.sph = H.sph.data();
  S64.sph_mat = H.sph_mat.data();
  S64.n_inst = (int32_t)H.inst64.size();
  S64.n_meshes = (int32_t)H.meshes.size();
  S64.n_spheres = (int32_t)H.sph_mat.size();
  LightsDev Ls{};
  Ls.n = (int32_t)H.lights.size();
  Ls.l64 = H.lights.data();
  Ls.cdf64 = H.light_cdf.data();
  std::vector<int32_t> stack((size_t)H.stack4 + 8);
  LaneStack st{stack.data(), stack.data() + 4, 1, 4};
  for (int64_t i = 0; i < n; i++) {
    if (light[i] < 0 || light[i] >= Ls.n) return -2;
    const float* r = rays + 8 * i;
    const vec3<float> o = mk(r[0], r[1], r[2]), dd = mk(r[3], r[4], r[5]);
    HitF32 h;
    if (!trace_f32<false>(S, o, dd, r[6], r[7], &h, st)) {
      out[i] = -1;
      continue;
    }
    const vec3<float> p32 = o + dd * h.t;
    const D3 p = hit_point64(S64, S.tri_orig, h.inst, h.prim, o, dd, p32);
    out[i] = light_self_occluded64(S64, Ls, light[i], p, uv[2 * i], uv[2 * i + 1]) ? 1 : 0;
  }
  return 0;
}

// The self-occlusion test's per-component quotient (s64::div3: one reciprocal, a residual
// correction per component) for n vectors v[3i..] and divisors l[i], written to out[3i..], to be
// compared with plain division.
extern "C" int tpt_test_div3(int64_t n, const double* v, const double* l, double* out) {
  for (int64_t i = 0; i < n; i++) {
    const D3 q = s64::div3(D3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}, l[i]);
    out[3 * i] = q.x;
    out[3 * i + 1] = q.y;
    out[3 * i + 2] = q.z;
  }
  return 0;
}
