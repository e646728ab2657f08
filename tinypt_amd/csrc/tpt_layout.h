// tpt_layout.h -- device memory layout shared by the host scene build and the HIP kernels.
//
// FP32 kernels read four-wide nodes (Node4, 128 B: up to four child boxes and refs in one node,
// so a traversal step is one 128 B fetch and each child box is tested once), triangles as
// v0/e1/e2 (48 B), and per-instance world->object rows.  The host builds binary Node2 trees
// first and collapses them into Node4.  The plain FP64 kernel reads the reference's own layout
// (bvh.rs:24-35 nodes, geometry.rs triangles, transform.rs matrices).
#pragma once

#include <cstdint>

// Child reference encoding (Node2::child, mesh roots, TLAS root):
//   ref >= 0                : index of an inner Node2
//   ref & TPT_LEAF_BIT      : leaf; (ref >> 28) & 7 = count - 1; ref & 0x0fffffff = first slot
//   TPT_REF_EMPTY           : empty tree
#define TPT_LEAF_BIT 0x80000000u
#define TPT_REF_EMPTY 0x7fffffff
#define TPT_LEAF_START_MASK 0x0fffffffu

// TLAS item codes: bit 31 set = sphere (index in low bits), clear = instance index.
#define TPT_ITEM_SPHERE 0x80000000u

// FP32 hit record: `inst` of a hit on a flattened (world-space) triangle carries this bit over
// the original instance index.
#define TPT_HIT_FLAT 0x40000000u
// Flattened triangle's instance word: the instance's A has det < 0, so the world-space
// e1 x e2 points against the reference's normal.
#define TPT_TRI_FLIP 0x40000000u

// Hit kind in the FP32 hit record (high bits of `inst`).
#define TPT_HIT_NONE 0
#define TPT_HIT_TRI 1
#define TPT_HIT_SPHERE 2

struct alignas(16) Node2 {
  // b[0..5] = left child min.xyz max.xyz, b[6..11] = right child min.xyz max.xyz
  float b[12];
  int32_t child[2];
  int32_t pad[2];
};
static_assert(sizeof(Node2) == 64, "Node2 is one 64 B fetch");

// Four-wide node, the form the FP32 kernels traverse: each binary Node2 subtree is collapsed so
// that one 128 B fetch tests up to four child boxes (x, y, z bounds as float4 over the children).
struct alignas(16) Node4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int32_t child[4];  // Node2 reference encoding; TPT_REF_EMPTY for an unused slot
  int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is one 128 B fetch");

// A triangle as the rows of the inverse of [e1 e2 n] (e1 = v1 - v0, e2 = v2 - v0, n = e1 x e2;
// computed in FP64, rounded), each row's w one coordinate of v0: for a point q, the rows give
// (u, v, w) of q - v0 in the triangle's own frame, so a ray hits the plane at w = 0 and the
// triangle where u, v >= 0 and u + v <= 1.  Row 2 is n / |n|^2.  Degenerate triangles have zero
// rows (never hit).
struct alignas(16) TriF32 {
  float r0[4];  // (e2 x n) / |n|^2, v0.x
  float r1[4];  // (n x e1) / |n|^2, v0.y
  float r2[4];  // n / |n|^2, v0.z
};

struct alignas(16) InstF32 {
  float w2o[12];   // rows of [A^-1 | -A^-1 t]
  float nrm[12];   // rows of A^-T (w unused)
  int32_t root;    // child ref of the mesh's BLAS root
  int32_t mat_override;
  int32_t mesh;
  int32_t pad;
};

struct alignas(16) LightF32 {
  float p0[4];  // sphere: c.xyz, r   | triangle: v0 world
  float p1[4];  // triangle: v1 world
  float p2[4];  // triangle: v2 world
  float emit[4];  // rgb, selection weight
  int32_t kind;   // 0 sphere, 1 triangle
  int32_t mat;    // emitting material
  int32_t pad[2];
};

// ---- FP64, reference layout
struct RefNodeDev {  // bvh.rs:24-35
  double bbox[6];
  int32_t left, right;
  uint32_t start, count;
};
static_assert(sizeof(RefNodeDev) == 64, "reference node is 64 B");

struct InstF64 {  // transform.rs:11-20 + world.rs:39-46
  double a[9], t[3], a_inv[9], nrm[9];
  int32_t mesh, mat_override;
};

struct MeshDev {
  int32_t node_off;  // into the reference node array
  int32_t idx_off;   // into the reference index array
  int32_t tri_off;   // into the original-order FP64 triangles
  int32_t n_nodes;
  int32_t n_tris;
  int32_t pad;
};

struct LightF64 {  // world.rs:218-221, 296-300
  int32_t kind;  // 0 sphere, 1 triangle
  int32_t idx;   // sphere index
  int32_t mesh, tri, inst;
  int32_t pad;
  double emit[3];
  double weight;
};

struct CameraF64 {  // ray.rs:22-33
  double o[3], u[3], v[3], w[3];
  double half_h, half_w, lens_radius, focus_dist, shutter_open, shutter_close;
};
