// scene_host.h -- host-side scene build for the MI355X backend.
//
// Turns a tpt_scene_desc into the arrays the device kernels read:
//   * per mesh, the bottom-level BVH exactly as Bvh::build builds it (bvh.rs:51-242), kept in
//     the reference's FP64 layout for the plain FP64 kernel and converted to compressed
//     two-child FP32 nodes (Node2) for the FP32 kernels;
//   * a top-level tree over instances and spheres (the reference loops over them linearly,
//     world.rs:86-118); its leaves hold one item each;
//   * the light list and CDF of World::build_lights (world.rs:127-167);
//   * the environment's FP64 row/column CDFs (env.rs:61-89) and, over the same cell weights,
//     a Walker/Vose alias table for the FP32 kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tinypt_gpu.h"
#include "tpt_layout.h"

namespace tpt {

struct DAabb {
  double mn[3], mx[3];
};

// Reference BVH node (bvh.rs:24-35).
struct RefNode {
  DAabb bbox;
  int32_t left, right;
  uint32_t start, count;
};

struct RefBvh {
  std::vector<RefNode> nodes;
  std::vector<int64_t> indices;
  int depth = 0;  // levels, root = 1
};

// Bvh::build over arbitrary primitive boxes (bvh.rs:51-242); `leaf_size` is LEAF_SIZE (4) for
// meshes.  Centroids are box centres, as Triangle::centroid (geometry.rs:241-243).
RefBvh build_ref_bvh(const std::vector<DAabb>& prim_bounds, size_t leaf_size);

// The FP32 kernels' trees: binned SAH over three axes with cost-based leaves of at most
// `max_leaf` primitives (own design; the node form of RefBvh).
RefBvh build_sah_bvh(const std::vector<DAabb>& prim_bounds, size_t max_leaf);

// Triangle::bounds (geometry.rs:221-225) of a static triangle.
DAabb triangle_bounds(const double* v9);

// Boxes of the frontier of a tree at depth `max_depth` (with the leaves above it): together they
// hold every primitive.  Through an instance's rotation their corners bound the instance much
// more tightly than the root box's eight corners, whose world box grows by up to sqrt(3) per
// axis for a rotated round mesh -- rays would then enter instances they cannot hit.
constexpr int kInstanceCutDepth = 10;  // up to 1024 boxes per mesh
std::vector<DAabb> tree_cut_boxes(const RefBvh& t, int max_depth);

// World box of an instance (A, t) of a mesh with cut boxes `cut`: the box of their corners
// through the transform, padded by 1e-6 relative.
DAabb instance_world_box(const std::vector<DAabb>& cut, const double* a, const double* t);

// Scenes with at most this many instanced triangles are flattened into world space.
constexpr int64_t kFlatMaxTris = 1 << 20;

struct HostScene {
  std::string error;
  // ---- copies of the descriptor
  CameraF64 cam;
  std::vector<int32_t> mat_kind;
  std::vector<double> mat_param;  // 4 per material
  std::vector<double> sph;        // 4 per sphere: c, r
  std::vector<int32_t> sph_mat;
  std::vector<int64_t> mesh_tri_begin;
  std::vector<double> tri_v;  // 9 per triangle (original order)
  std::vector<int32_t> tri_mat;
  std::vector<int32_t> inst_mesh;
  std::vector<double> inst_xform;  // 12 per instance (A row-major, t)
  std::vector<int32_t> inst_override;
  int env_kind = TPT_ENV_SKY;
  int env_w = 0, env_h = 0;
  std::vector<double> env_rgb;

  // ---- derived, FP64 (plain kernel, reference layout)
  std::vector<RefBvh> blas;
  std::vector<InstF64> inst64;       // A, t, A^-1, A^-T
  std::vector<RefNodeDev> ref_nodes;  // all meshes' nodes, concatenated
  std::vector<int32_t> ref_indices;   // all meshes' indices (mesh-local values), concatenated
  std::vector<MeshDev> meshes;        // offsets into the concatenated arrays
  std::vector<LightF64> lights;
  std::vector<double> light_cdf;
  double light_total = 0.0;
  std::vector<double> env_row_cdf, env_col_cdf;
  double env_total = 0.0;

  // ---- derived, FP32 (wavefront kernels)
  std::vector<Node2> nodes2;      // BLAS nodes of all meshes, then TLAS nodes (binary form)
  std::vector<Node4> nodes4;      // the same trees collapsed four-wide; roots below index these
  std::vector<int32_t> mesh_root2;  // per mesh: child ref of its root
  int32_t tlas_root = TPT_REF_EMPTY;
  std::vector<int32_t> tlas_items;  // per TLAS leaf slot: item code (kind bit | index)
  std::vector<TriF32> tris32;       // reordered in leaf order
  std::vector<int32_t> tri_orig;    // reordered -> mesh-local original index
  std::vector<int32_t> tri_mat32;   // reordered -> material (instance override applied for flattened slots)
  std::vector<int32_t> tri_inst32;  // flattened slots: instance | TPT_TRI_FLIP (empty when not flattened)
  bool flat = false;                // instances flattened into one world-space tree
  int32_t flat_inst = -1;           // index of the identity pseudo-instance that holds that tree
  std::vector<InstF32> inst32;
  std::vector<double> inst_c64;   // 3 per inst32 entry: the world point an entering ray is re-based
                                  // to (the centre of the instance's world box), FP64
  std::vector<float> sph32;  // 4 per sphere (FP32 copy, TLAS boxes only)
  std::vector<LightF32> lights32;
  std::vector<float> light_cdf32;
  std::vector<float> env_rgb32;      // 4 per texel (rgb, 0)
  std::vector<float> alias_prob;     // per cell
  std::vector<int32_t> alias_idx;    // per cell
  std::vector<float> cell_pdf;       // per cell: (w/total) * W * H / (2 pi^2), the pdf numerator
  int max_blas_depth = 0, max_tlas_depth = 0;
  int flat_depth = 0;  // levels of the flattened tree
  int stack4 = 0;      // traversal stack entries the FP32 (four-wide) kernels can need

  bool build(const tpt_scene_desc* d);
  void rebuild_lights();
  size_t n_tris() const { return tri_mat.size(); }
  // Root of the flattened world-space tree when the top level holds nothing else (no spheres):
  // rays then start there, without the top-level node and the instance entry.  Else
  // TPT_REF_EMPTY.
  int32_t flat_only_root() const {
    return flat && sph_mat.empty() ? inst32[(size_t)flat_inst].root : TPT_REF_EMPTY;
  }
};

}  // namespace tpt
