// api.hip -- the C ABI (include/tinypt_gpu.h): contexts, scene upload, render dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tinypt_gpu.h"
#include "adjoint.h"
#include "pass_plan.h"
#include "plain64.h"
#include "scene_host.h"
#include "wavefront.h"

using namespace tpt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

namespace tpt {
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace tpt

namespace {

#define HIPC(x)                                                                                     \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) return fail(TPT_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);  // release path: nothing to report to
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t n) {  // grow-only
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, n ? n : 16);
    if (e == hipSuccess) bytes = n ? n : 16;
    return e;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// HIP events kept by a context between calls: creating and destroying an event per timed
// launch cost tens of microseconds per small render.
struct EventPool {
  std::vector<hipEvent_t> ev;
  EventPool() = default;
  EventPool(const EventPool&) = delete;
  EventPool& operator=(const EventPool&) = delete;
  ~EventPool() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);  // teardown: nothing to report to
  }
};

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  hipError_t e = b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s);
}

}  // namespace

// One set of wavefront path state (grow-only) with the stream its passes run on.  A context has
// two: passes of a non-adaptive render alternate between them, so one set's kernels fill the
// ramp-down of the other's.
struct PathSet {
  hipStream_t stream = nullptr;
  DevBuf ps_o[2], ps_d[2], ps_beta[2], ps_rng[2], ps_hit, ps_L, sh_o[2], sh_d[2], sh_c[2], sh_l, sh_lv, shl_keep, counters,
      stats, gstack;
  size_t ps_cap = 0;
  EventPool ev_trace;  // timing events of the extension launches (one call at a time)
  hipEvent_t done = nullptr;  // recorded after the set's last accumulation of a render
};

struct tpt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // the context's stream (set[0]'s)
  PathSet set[2];
  DevBuf stats;  // plain FP64 form counters
  // adaptive / plain per-pixel state
  DevBuf act0, act1, nsamp, welford, rngpix, done;
  // adjoint (grow-only, kept between calls)
  DevBuf adj[ADJ_SCRATCH_N];
  // device sums of renders into host buffers (grow-only)
  DevBuf out_acc, out_w, out_sq;
  EventPool ev_total;  // timing events of the render calls (one call at a time)
};

struct tpt_scene {
  tpt_ctx* ctx = nullptr;  // null for a host-only scene (tpt_scene_build_host)
  HostScene host;
  DevBuf nodes4, tris32, tri_mat32, tri_orig, tri_inst, inst32, inst_c64, tlas_items, sph64, sph_mat;
  DevBuf ref_nodes, ref_indices, tri_v, tri_mat, meshes, inst64;
  DevBuf mat_kind, mat_p32, mat_p64;
  DevBuf l32, cdf32, l64, cdf64;
  DevBuf env_rgb32, alias_prob, alias_idx, cell_pdf, env_rgb64, row_cdf, col_cdf;
  SceneF32 s32{};
  SceneF64 s64{};
  EnvDev env{};
  LightsDev lights{};
  MatsDev mats{};
  CamF32 cam32{};
  size_t device_bytes = 0;
};

// Fixed scratch of tpt_trace_closest_noalloc: the traversal kernel's per-thread overflow stacks
// (the entries beyond the 24 per lane it keeps in LDS), for callers that cannot allocate.  One
// copy per device, held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address is looked up once per device in
// tpt_ctx_create (hipGetSymbolAddress, with that device current) and only that address is
// passed to kernels.
#define TPT_NOALLOC_SCRATCH_BYTES (16u << 20)
#define TPT_MAX_DEVICES 256
static __device__ __align__(256) int32_t scratch[TPT_NOALLOC_SCRATCH_BYTES / 4];
static int32_t* g_scratch_dev[TPT_MAX_DEVICES];  // device address of `scratch` per device
static std::mutex g_scratch_mu;

static int upload_materials(tpt_scene* sc) {
  hipStream_t s = sc->ctx->stream;
  const HostScene& H = sc->host;
  std::vector<float4> p32(H.mat_kind.size());
  for (size_t i = 0; i < p32.size(); i++)
    p32[i] = make_float4((float)H.mat_param[4 * i], (float)H.mat_param[4 * i + 1], (float)H.mat_param[4 * i + 2],
                         (float)H.mat_param[4 * i + 3]);
  HIPC(upload(sc->mat_kind, H.mat_kind, s));
  HIPC(upload(sc->mat_p32, p32, s));
  HIPC(upload(sc->mat_p64, H.mat_param, s));
  sc->mats.n = (int32_t)H.mat_kind.size();
  sc->mats.kind = sc->mat_kind.as<int32_t>();
  sc->mats.p32 = sc->mat_p32.as<float4>();
  sc->mats.p64 = sc->mat_p64.as<double>();
  return TPT_OK;
}

static int upload_lights(tpt_scene* sc) {
  hipStream_t s = sc->ctx->stream;
  const HostScene& H = sc->host;
  HIPC(upload(sc->l32, H.lights32, s));
  HIPC(upload(sc->cdf32, H.light_cdf32, s));
  HIPC(upload(sc->l64, H.lights, s));
  HIPC(upload(sc->cdf64, H.light_cdf, s));
  sc->lights.n = (int32_t)H.lights.size();
  sc->lights.total = H.light_total;
  sc->lights.total32 = (float)H.light_total;
  sc->lights.l32 = sc->l32.as<LightF32>();
  sc->lights.cdf32 = sc->cdf32.as<float>();
  sc->lights.l64 = sc->l64.as<LightF64>();
  sc->lights.cdf64 = sc->cdf64.as<double>();
  return TPT_OK;
}

extern "C" {

int tpt_abi_version(void) { return TPT_ABI_VERSION; }

const char* tpt_last_error(void) { return g_err.c_str(); }

int tpt_ctx_create(int device, tpt_ctx** out) {
  if (!out) return fail(TPT_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  int n = 0;
  HIPC(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(TPT_ERR_INVALID_ARGUMENT, "no such HIP device");
  if (device >= TPT_MAX_DEVICES) return fail(TPT_ERR_UNSUPPORTED, "device index above the library's device table");
  HIPC(hipSetDevice(device));
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
      void* p = nullptr;
      HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return fail(TPT_ERR_DEVICE, "no device address for the fixed scratch");
      g_scratch_dev[device] = static_cast<int32_t*>(p);
    }
  }
  std::unique_ptr<tpt_ctx> c(new tpt_ctx());
  c->device = device;
  HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->set[0].stream = c->stream;
  HIPC(hipStreamCreateWithFlags(&c->set[1].stream, hipStreamNonBlocking));
  for (PathSet& ps : c->set) HIPC(hipEventCreateWithFlags(&ps.done, hipEventDisableTiming));
  *out = c.release();
  return TPT_OK;
}

int tpt_ctx_destroy(tpt_ctx* ctx) {
  if (!ctx) return TPT_OK;
  (void)hipSetDevice(ctx->device);  // teardown: errors cannot be acted on
  for (PathSet& ps : ctx->set) {
    if (ps.stream) (void)hipStreamSynchronize(ps.stream);
    if (ps.done) (void)hipEventDestroy(ps.done);
  }
  if (ctx->set[1].stream) (void)hipStreamDestroy(ctx->set[1].stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TPT_OK;
}

int tpt_scene_upload(tpt_ctx* ctx, const tpt_scene_desc* desc, tpt_scene** out) {
  if (!ctx || !desc || !out) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  HIPC(hipSetDevice(ctx->device));
  std::unique_ptr<tpt_scene> sc(new tpt_scene());
  sc->ctx = ctx;
  HostScene& H = sc->host;
  try {
    if (!H.build(desc)) return fail(TPT_ERR_INVALID_ARGUMENT, H.error);
  } catch (const std::bad_alloc&) {
    return fail(TPT_ERR_OUT_OF_MEMORY, "host allocation failed while building the scene");
  }
  hipStream_t s = ctx->stream;
  std::vector<double> sph64 = H.sph;
  HIPC(upload(sc->nodes4, H.nodes4, s));
  HIPC(upload(sc->tris32, H.tris32, s));
  HIPC(upload(sc->tri_mat32, H.tri_mat32, s));
  HIPC(upload(sc->tri_orig, H.tri_orig, s));
  if (H.flat) HIPC(upload(sc->tri_inst, H.tri_inst32, s));
  HIPC(upload(sc->inst32, H.inst32, s));
  HIPC(upload(sc->inst_c64, H.inst_c64, s));
  HIPC(upload(sc->tlas_items, H.tlas_items, s));
  HIPC(upload(sc->sph64, sph64, s));
  HIPC(upload(sc->sph_mat, H.sph_mat, s));
  HIPC(upload(sc->ref_nodes, H.ref_nodes, s));
  HIPC(upload(sc->ref_indices, H.ref_indices, s));
  HIPC(upload(sc->tri_v, H.tri_v, s));
  HIPC(upload(sc->tri_mat, H.tri_mat, s));
  HIPC(upload(sc->meshes, H.meshes, s));
  HIPC(upload(sc->inst64, H.inst64, s));
  int rc = upload_materials(sc.get());
  if (rc) return rc;
  rc = upload_lights(sc.get());
  if (rc) return rc;
  // environment
  EnvDev& E = sc->env;
  E.kind = H.env_kind;
  E.w = H.env_w;
  E.h = H.env_h;
  E.total = H.env_total;
  E.black = H.env_kind == TPT_ENV_IMAGE &&
            std::all_of(H.env_rgb.begin(), H.env_rgb.end(), [](double v) { return v == 0.0; });
  if (H.env_kind == TPT_ENV_IMAGE) {
    std::vector<float4> rgb(H.env_rgb32.size() / 4);
    std::memcpy(rgb.data(), H.env_rgb32.data(), sizeof(float) * H.env_rgb32.size());
    HIPC(upload(sc->env_rgb32, rgb, s));
    HIPC(upload(sc->alias_prob, H.alias_prob, s));
    HIPC(upload(sc->alias_idx, H.alias_idx, s));
    HIPC(upload(sc->cell_pdf, H.cell_pdf, s));
    HIPC(upload(sc->env_rgb64, H.env_rgb, s));
    HIPC(upload(sc->row_cdf, H.env_row_cdf, s));
    HIPC(upload(sc->col_cdf, H.env_col_cdf, s));
    E.rgb32 = sc->env_rgb32.as<float4>();
    E.alias_prob = sc->alias_prob.as<float>();
    E.alias_idx = sc->alias_idx.as<int32_t>();
    E.cell_pdf = sc->cell_pdf.as<float>();
    E.rgb64 = sc->env_rgb64.as<double>();
    E.row_cdf = sc->row_cdf.as<double>();
    E.col_cdf = sc->col_cdf.as<double>();
  }
  // FP32 view
  SceneF32& S = sc->s32;
  S.nodes = sc->nodes4.as<Node4>();
  S.tris = sc->tris32.as<TriF32>();
  S.tri_mat = sc->tri_mat32.as<int32_t>();
  S.tri_orig = sc->tri_orig.as<int32_t>();
  S.tri_inst = H.flat ? sc->tri_inst.as<int32_t>() : nullptr;
  S.flat_inst = H.flat ? H.flat_inst : -1;
  S.flat_root = H.flat_only_root();
  S.inst = sc->inst32.as<InstF32>();
  S.inst64 = sc->inst64.as<InstF64>();
  S.inst_c64 = sc->inst_c64.as<double>();
  S.tlas_items = sc->tlas_items.as<int32_t>();
  S.sph = sc->sph64.as<double>();
  S.sph_mat = sc->sph_mat.as<int32_t>();
  S.tlas_root = H.tlas_root;
  S.n_spheres = (int32_t)H.sph_mat.size();
  // FP64 view
  SceneF64& D = sc->s64;
  D.nodes = sc->ref_nodes.as<RefNodeDev>();
  D.indices = sc->ref_indices.as<int32_t>();
  D.tris = sc->tri_v.as<double>();
  D.tri_mat = sc->tri_mat.as<int32_t>();
  D.meshes = sc->meshes.as<MeshDev>();
  D.inst = sc->inst64.as<InstF64>();
  D.sph = sc->sph64.as<double>();
  D.sph_mat = sc->sph_mat.as<int32_t>();
  D.n_inst = (int32_t)H.inst64.size();
  D.n_meshes = (int32_t)H.meshes.size();
  D.n_spheres = (int32_t)H.sph_mat.size();
  // camera
  CamF32& C = sc->cam32;
  for (int k = 0; k < 3; k++) {
    C.o[k] = (float)H.cam.o[k];
    C.u[k] = (float)H.cam.u[k];
    C.v[k] = (float)H.cam.v[k];
    C.w[k] = (float)H.cam.w[k];
  }
  C.half_h = (float)H.cam.half_h;
  C.half_w = (float)H.cam.half_w;
  C.lens_radius = (float)H.cam.lens_radius;
  C.focus_dist = (float)H.cam.focus_dist;
  HIPC(hipStreamSynchronize(s));
  for (DevBuf* b : {&sc->nodes4, &sc->tris32, &sc->tri_mat32, &sc->tri_orig, &sc->tri_inst, &sc->inst32, &sc->inst_c64, &sc->tlas_items, &sc->sph64,
                    &sc->sph_mat, &sc->ref_nodes, &sc->ref_indices, &sc->tri_v, &sc->tri_mat, &sc->meshes, &sc->inst64,
                    &sc->mat_kind, &sc->mat_p32, &sc->mat_p64, &sc->l32, &sc->cdf32, &sc->l64, &sc->cdf64, &sc->env_rgb32,
                    &sc->alias_prob, &sc->alias_idx, &sc->cell_pdf, &sc->env_rgb64, &sc->row_cdf, &sc->col_cdf})
    sc->device_bytes += b->bytes;
  *out = sc.release();
  return TPT_OK;
}

int tpt_scene_destroy(tpt_scene* scene) {
  if (!scene) return TPT_OK;
  if (scene->ctx) {
    (void)hipSetDevice(scene->ctx->device);
    (void)hipStreamSynchronize(scene->ctx->stream);
  }
  delete scene;
  return TPT_OK;
}

int tpt_scene_build_host(const tpt_scene_desc* desc, tpt_scene** out) {
  if (!desc || !out) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::unique_ptr<tpt_scene> sc(new tpt_scene());
  try {
    if (!sc->host.build(desc)) return fail(TPT_ERR_INVALID_ARGUMENT, sc->host.error);
  } catch (const std::bad_alloc&) {
    return fail(TPT_ERR_OUT_OF_MEMORY, "host allocation failed while building the scene");
  }
  *out = sc.release();
  return TPT_OK;
}

int tpt_scene_get_lights(const tpt_scene* sc, int32_t* n_lights, double* out) {
  if (!sc || !n_lights) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  const HostScene& H = sc->host;
  *n_lights = (int32_t)H.lights.size();
  if (!out) return TPT_OK;
  for (size_t i = 0; i < H.lights.size(); i++) {
    const LightF64& L = H.lights[i];
    double* o = out + 10 * i;
    o[0] = L.kind; o[1] = L.idx; o[2] = L.mesh; o[3] = L.tri; o[4] = L.inst;
    o[5] = L.emit[0]; o[6] = L.emit[1]; o[7] = L.emit[2]; o[8] = L.weight; o[9] = H.light_cdf[i];
  }
  return TPT_OK;
}

int tpt_scene_get_env_tables(const tpt_scene* sc, double* total_weight, float* alias_prob, int32_t* alias_idx,
                             float* cell_pdf) {
  if (!sc) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  const HostScene& H = sc->host;
  if (H.env_kind != TPT_ENV_IMAGE) return fail(TPT_ERR_INVALID_ARGUMENT, "scene has no image environment");
  if (total_weight) *total_weight = H.env_total;
  size_t n = H.alias_prob.size();
  if (alias_prob) std::memcpy(alias_prob, H.alias_prob.data(), n * sizeof(float));
  if (alias_idx) std::memcpy(alias_idx, H.alias_idx.data(), n * sizeof(int32_t));
  if (cell_pdf) std::memcpy(cell_pdf, H.cell_pdf.data(), n * sizeof(float));
  return TPT_OK;
}

int tpt_scene_get_info(const tpt_scene* sc, tpt_scene_info* info) {
  if (!sc || !info) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  const HostScene& H = sc->host;
  info->n_tris = (int64_t)H.tri_mat.size();
  int64_t blas_nodes = 0;
  for (const RefBvh& b : H.blas) blas_nodes += (int64_t)b.nodes.size();
  info->n_blas_nodes = blas_nodes;
  info->n_tlas_nodes = (int64_t)H.tlas_items.size();
  info->max_blas_depth = H.max_blas_depth;
  info->max_tlas_depth = H.max_tlas_depth;
  info->n_lights = (int32_t)H.lights.size();
  info->traversal_stack = H.stack4;
  info->device_bytes = (int64_t)sc->device_bytes;
  return TPT_OK;
}

int tpt_scene_get_bvh(const tpt_scene* sc, int32_t mesh, int64_t* n_nodes, double* node_bbox, int32_t* node_left,
                      int32_t* node_right, uint32_t* node_start, uint32_t* node_count, int64_t* indices) {
  if (!sc || !n_nodes) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (mesh < 0 || mesh >= (int32_t)sc->host.blas.size()) return fail(TPT_ERR_INVALID_ARGUMENT, "mesh out of range");
  const RefBvh& b = sc->host.blas[(size_t)mesh];
  *n_nodes = (int64_t)b.nodes.size();
  if (!node_bbox) return TPT_OK;
  for (size_t i = 0; i < b.nodes.size(); i++) {
    const RefNode& n = b.nodes[i];
    for (int k = 0; k < 3; k++) {
      node_bbox[6 * i + (size_t)k] = n.bbox.mn[k];
      node_bbox[6 * i + 3 + (size_t)k] = n.bbox.mx[k];
    }
    node_left[i] = n.left;
    node_right[i] = n.right;
    node_start[i] = n.start;
    node_count[i] = n.count;
  }
  if (indices)
    for (size_t i = 0; i < b.indices.size(); i++) indices[i] = b.indices[i];
  return TPT_OK;
}

int tpt_scene_update_materials(tpt_ctx* ctx, tpt_scene* sc, const double* mat_param) {
  if (!ctx || !sc || !mat_param) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (sc->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  HIPC(hipSetDevice(ctx->device));
  HostScene& H = sc->host;
  std::memcpy(H.mat_param.data(), mat_param, sizeof(double) * H.mat_param.size());
  H.rebuild_lights();
  int rc = upload_materials(sc);
  if (rc) return rc;
  rc = upload_lights(sc);
  if (rc) return rc;
  HIPC(hipStreamSynchronize(ctx->stream));
  return TPT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- render
namespace {

int check_cfg(const tpt_render_config* c) {
  if (!c) return fail(TPT_ERR_INVALID_ARGUMENT, "config is null");
  if (c->width <= 0 || c->height <= 0) return fail(TPT_ERR_INVALID_ARGUMENT, "width and height must be positive");
  if ((int64_t)c->width * c->height > (int64_t)1 << 28) return fail(TPT_ERR_INVALID_ARGUMENT, "image too large");
  if (c->spp < 0 || c->max_bounces < 0 || c->max_bounces > 255 || c->rr_start < 0)
    return fail(TPT_ERR_INVALID_ARGUMENT, "spp, max_bounces (<= 255) and rr_start must be non-negative");
  if (c->precision != TPT_FP32_WAVEFRONT && c->precision != TPT_FP64_PLAIN)
    return fail(TPT_ERR_INVALID_ARGUMENT, "unknown precision");
  return TPT_OK;
}

// Deepest traversal stack any ray can need in the four-wide trees (HostScene::stack4).
int stack_need(const HostScene& H) { return H.stack4; }

// Global overflow entries per lane beyond the TPT_WF_STACK entries kept in LDS.
int needed_gstack(const HostScene& H) {
  int need = stack_need(H);
  return need > TPT_WF_STACK ? need - TPT_WF_STACK : 0;
}

// Lanes below which a wave of the persistent traversal kernels takes new rays (0: one ray per
// lane, the plain kernels), for the first bounce's extension, the later extensions and the
// shadow rays.  Refilling pays where the rays of a wave need very different amounts of work and
// costs a global cursor atomic per refill.  Measured on MI355X (FP32 form, images unchanged): a
// two-level scene of 4096 instances (top-level depth 16, 3.4M BLAS nodes) at 48 -- extension
// -17%, shadow rays -9%, adaptive render 26.3 -> 19.0 s (its adaptive passes' camera rays, of
// scattered active pixels, gain too); a single tree of 1.31M triangles at 16 -- later bounces
// -5%, shadow rays -17%, but its first bounce, coherent camera rays, 75% slower, so that one
// keeps one ray per lane; the 36-triangle Cornell box at 16 -- extension 2.1x slower (its rays
// finish within a few steps, so waves refill almost every step).
void refill_thresholds(const HostScene& H, int32_t requested, WfParams* W) {
  int32_t first = 0, later = 0;
  if (requested > 0) {
    first = later = std::min(requested, 63);
  } else if (requested == 0) {
    if (!H.flat && H.inst32.size() >= 64) first = later = 48;
    else if (H.stack4 > TPT_WF_STACK) later = 16;
  }
  W->refill_first = first;
  W->refill = W->refill_sh = later;
}

// Queue arrays hold TPT_WF_SEGS segments of wf_seg_cap(P) entries; hit and L are dense.
int ensure_paths(PathSet* c, size_t P, int gcap) {
  if (P > c->ps_cap) {
    const size_t Q = (size_t)TPT_WF_SEGS * wf_seg_cap((uint32_t)P);
    for (int b = 0; b < 2; b++) {
      HIPC(c->ps_o[b].ensure(Q * 16));
      HIPC(c->ps_d[b].ensure(Q * 16));
      HIPC(c->ps_beta[b].ensure(Q * 16));
      HIPC(c->ps_rng[b].ensure(Q * 8));
    }
    HIPC(c->ps_hit.ensure(P * sizeof(HitRec)));
    HIPC(c->ps_L.ensure(P * 16));
    HIPC(c->sh_o[0].ensure(Q * 16));  // the light queue keeps its origin in sh_l
    for (int q = 0; q < 2; q++) {
      HIPC(c->sh_d[q].ensure(Q * 16));
      HIPC(c->sh_c[q].ensure(Q * 16));
    }
    HIPC(c->sh_l.ensure(Q * 16));
    HIPC(c->sh_lv.ensure(Q * 8));
    HIPC(c->shl_keep.ensure(Q * 4));
    c->ps_cap = P;
  }
  HIPC(c->counters.ensure(TPT_NUM_CNT * 4));
  HIPC(c->stats.ensure(TPT_STAT_WORDS * 8));
  if (gcap > 0) HIPC(c->gstack.ensure((size_t)wf_grid_threads((uint32_t)P) * (size_t)gcap * 4));
  return TPT_OK;
}

WfParams make_wf(PathSet* c, const tpt_scene* sc, const tpt_render_config* cfg, int gcap) {
  WfParams W{};
  W.S = sc->s32;
  W.S64 = sc->s64;
  W.env = sc->env;
  W.lights = sc->lights;
  W.mats = sc->mats;
  W.cam = sc->cam32;
  PathState& ps = W.ps;
  for (int b = 0; b < 2; b++) {
    ps.o[b] = c->ps_o[b].as<float4>();
    ps.d[b] = c->ps_d[b].as<float4>();
    ps.beta[b] = c->ps_beta[b].as<float4>();
    ps.rng[b] = c->ps_rng[b].as<uint64_t>();
  }
  ps.hit = c->ps_hit.as<HitRec>();
  ps.L = c->ps_L.as<float4>();
  for (int q = 0; q < 2; q++) {
    ps.sh_o[q] = q == 0 ? c->sh_o[0].as<float4>() : nullptr;
    ps.sh_d[q] = c->sh_d[q].as<float4>();
    ps.sh_c[q] = c->sh_c[q].as<float4>();
  }
  ps.sh_l = c->sh_l.as<float4>();
  ps.sh_lv = c->sh_lv.as<uint64_t>();
  ps.shl_keep = c->shl_keep.as<uint32_t>();
  ps.counters = c->counters.as<uint32_t>();
  ps.stats = c->stats.as<uint64_t>();
  ps.seg_cap = wf_seg_cap((uint32_t)c->ps_cap);
  W.gstack = gcap > 0 ? c->gstack.as<int32_t>() : nullptr;
  W.gstack_cap = gcap;
  refill_thresholds(sc->host, cfg->traversal_refill, &W);
  W.width = cfg->width;
  W.height = cfg->height;
  W.max_bounces = cfg->max_bounces;
  W.rr_start = cfg->rr_start;
  W.inv_w = (float)(1.0 / (double)cfg->width);
  W.inv_h = (float)(1.0 / (double)cfg->height);
  W.seed = cfg->seed;
  return W;
}

// Records events on a stream from a pool (growing it as needed); pairs_ms sums the intervals
// between marks 2i and 2i + 1.  A pool serves one timer at a time.
struct EventTimer {
  hipStream_t s;
  bool on;
  EventPool* pool;
  size_t used = 0;
  EventTimer(hipStream_t st, bool enable, EventPool* p) : s(st), on(enable), pool(p) {}
  hipError_t mark() {
    if (!on) return hipSuccess;
    if (used == pool->ev.size()) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      pool->ev.push_back(e);
    }
    return hipEventRecord(pool->ev[used++], s);
  }
  // Two events for a dispatch to record itself (its start and end), or nulls when timing is off.
  hipError_t reserve_pair(hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!on) return hipSuccess;
    while (pool->ev.size() < used + 2) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      pool->ev.push_back(e);
    }
    *a = pool->ev[used];
    *b = pool->ev[used + 1];
    used += 2;
    return hipSuccess;
  }
  double pairs_ms() {
    double t = 0;
    for (size_t i = 0; i + 1 < used; i += 2) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, pool->ev[i], pool->ev[i + 1]) == hipSuccess) t += ms;
    }
    return t;
  }
};

// Bounce loop of one pass over the A.n_paths paths that A describes; the first bounce's kernels
// generate the camera rays themselves.
int wf_bounces(PathSet* c, const WfParams& W, const PassArgs& A, EventTimer& tt, uint64_t* launches) {
  hipStream_t s = c->stream;
  const uint32_t P = A.n_paths;
  if ((size_t)P > c->ps_cap) return fail(TPT_ERR_INTERNAL, "a pass holds more paths than were allocated");
  if (W.max_bounces == 0) HIPC(wf_launch_clear_radiance(W, P, s));
  int cur = 0;
  for (int b = 0; b < W.max_bounces; b++) {
    const PassArgs* gen = b == 0 ? &A : nullptr;
    hipEvent_t e0, e1;
    HIPC(tt.reserve_pair(&e0, &e1));
    HIPC(wf_launch_extend(W, cur, gen, P, s, e0, e1));
    (*launches)++;
    HIPC(wf_launch_shade(W, cur, gen, P, s));
    if (W.lights.n > 0) {
      HIPC(wf_launch_light_selfocc(W, cur, gen, P, s));
      HIPC(wf_launch_shadow(W, 1, P, s));
    }
    if (W.env.kind == TPT_ENV_IMAGE && !W.env.black) HIPC(wf_launch_shadow(W, 0, P, s));
    cur ^= 1;
  }
  return TPT_OK;
}

// Renders the n_pix pixels of the domain: pixel q of the domain is image pixel pix_map[q]
// (q itself when pix_map is null), and its sums go to acc[3q..], acc_w[q], acc_sq[3q..].
int render_wavefront(tpt_ctx* c, const tpt_scene* sc, const tpt_render_config* cfg, uint32_t npix,
                     const int32_t* pix_map, double* acc, double* acc_w, double* acc_sq, tpt_stats* st) {
  hipStream_t s = c->stream;
  const uint32_t spp = (uint32_t)cfg->spp;
  const bool adaptive = cfg->adaptive_enabled != 0;
  // Paths in flight: up to 2^26 (about 18 GB of path state and queues), so that a 512^2 x 256
  // spp render is one drain of the bounce loop per path set; less when the device has less free.
  uint32_t target = 1u << 26;
  {
    size_t free_b = 0, total_b = 0;
    const size_t have = c->set[0].ps_cap + c->set[1].ps_cap;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && have < target) {
      const size_t per_path = 280;  // both path sets, hit, L, both shadow queues, segment slack
      while (target > (1u << 20) && (size_t)target * per_path > free_b / 2) target >>= 1;
    }
  }
  // A non-adaptive render of two or more passes alternates them between the two path sets,
  // each holding half the paths: a set's kernels then fill the other's ramp-downs (measured on
  // the Cornell box: two 128-spp renders on two streams 30.4 ms, one after the other 32.0 ms).
  // Accumulation stays in pass order, so the sums do not depend on the overlap.  An adaptive
  // pass's chunks alternate between the sets too; they hold disjoint pixels.
  const RenderPlan rp = plan_render_sets(npix, spp, cfg->samples_per_pass, adaptive, target);
  const PassPlan plan = rp.pass;
  const int nsets = rp.sets;
  const uint32_t k = plan.k, pix_chunk = plan.pix_chunk;
  const int gcap = needed_gstack(sc->host);
  const size_t P = (size_t)plan.paths_alloc;
  WfParams W[2];
  for (int j = 0; j < nsets; j++) {
    PathSet* S = &c->set[j];
    int rc = ensure_paths(S, P, gcap);
    if (rc) return rc;
    W[j] = make_wf(S, sc, cfg, gcap);
    HIPC(hipMemsetAsync(W[j].ps.stats, 0, TPT_STAT_WORDS * 8, S->stream));
    HIPC(hipMemsetAsync(W[j].ps.counters, 0, TPT_NUM_CNT * 4, S->stream));
  }
  EventTimer total(s, st != nullptr, &c->ev_total);
  EventTimer tt0(c->set[0].stream, st != nullptr, &c->set[0].ev_trace);
  EventTimer tt1(c->set[1].stream, st != nullptr, &c->set[1].ev_trace);
  EventTimer* tt[2] = {&tt0, &tt1};
  uint64_t launches = 0;
  int rc = TPT_OK;
  HIPC(total.mark());
  if (nsets == 2 && !adaptive) {  // set 1 starts after the work already queued on the context stream
    HIPC(hipEventRecord(c->set[0].done, s));
    HIPC(hipStreamWaitEvent(c->set[1].stream, c->set[0].done, 0));
  }
  if (!adaptive) {
    uint64_t pass = 0;
    for (uint32_t p0 = 0; p0 < npix; p0 += pix_chunk) {
      uint32_t np = std::min(pix_chunk, npix - p0);
      for (uint32_t s0 = 0; s0 < spp; s0 += k) {
        const int j = nsets == 2 ? (int)(pass & 1) : 0;
        PathSet* S = &c->set[j];
        uint32_t kk = std::min(k, spp - s0);
        PassArgs A{np * kk, np, p0, s0, nullptr, nullptr, pix_map};
        rc = wf_bounces(S, W[j], A, *tt[j], &launches);
        if (rc) return rc;
        AccumArgs B{np, p0, kk, acc, acc_w, acc_sq};
        // after the previous pass's accumulation (queued on the other set's stream)
        if (nsets == 2 && pass > 0) HIPC(hipStreamWaitEvent(S->stream, c->set[j ^ 1].done, 0));
        HIPC(wf_launch_accumulate(W[j], B, S->stream));
        if (nsets == 2) HIPC(hipEventRecord(S->done, S->stream));
        pass++;
      }
    }
    if (nsets == 2) HIPC(hipStreamWaitEvent(s, c->set[1].done, 0));
  } else {
    // render.rs:143-179 per pixel; each pass gives every active pixel k more samples
    HIPC(c->act0.ensure((size_t)npix * 4));
    HIPC(c->act1.ensure((size_t)npix * 4));
    HIPC(c->nsamp.ensure((size_t)npix * 4));
    HIPC(c->welford.ensure((size_t)npix * 16));
    HIPC(hipMemsetAsync(c->nsamp.p, 0, (size_t)npix * 4, s));
    HIPC(hipMemsetAsync(c->welford.p, 0, (size_t)npix * 16, s));
    std::vector<int32_t> all(npix);
    for (uint32_t i = 0; i < npix; i++) all[i] = (int32_t)i;
    HIPC(hipMemcpyAsync(c->act0.p, all.data(), (size_t)npix * 4, hipMemcpyHostToDevice, s));
    uint32_t n_active = spp > 0 ? npix : 0;
    uint32_t taken = 0;
    const uint32_t max_spp = std::max(spp, 1u);
    const uint32_t min_spp = std::min((uint32_t)std::max(cfg->adaptive_min_spp, 1), max_spp);
    int32_t* act_in = c->act0.as<int32_t>();
    int32_t* act_out = c->act1.as<int32_t>();
    // the survivors of every chunk, on either set, are appended to act_out through one counter
    uint32_t* n_out = W[0].ps.counters + TPT_CNT_ACTIVE;
    while (n_active > 0 && taken < spp) {
      uint32_t kk = 1, chunk = 1;
      plan_adaptive_pass(n_active, taken, min_spp, spp, rp.pass_cap, P, &kk, &chunk);
      HIPC(hipMemsetAsync(n_out, 0, 4, s));
      // set 1 starts after this pass's set-up on the context stream
      HIPC(hipEventRecord(c->set[0].done, s));
      HIPC(hipStreamWaitEvent(c->set[1].stream, c->set[0].done, 0));
      // the active list in chunks of at most P / kk pixels, alternating between the sets (a
      // pixel's samples do not depend on its chunk or its position in the list)
      int j = 0;
      for (uint32_t c0 = 0; c0 < n_active; c0 += chunk, j ^= 1) {
        PathSet* S = &c->set[j];
        const uint32_t nc = std::min(chunk, n_active - c0);
        PassArgs A{nc * kk, nc, 0, 0, act_in + c0, c->nsamp.as<uint32_t>(), pix_map};
        rc = wf_bounces(S, W[j], A, *tt[j], &launches);
        if (rc) return rc;
        AdaptiveArgs B{nc, kk, min_spp, spp, cfg->adaptive_threshold, act_in + c0, act_out, n_out,
                       c->nsamp.as<uint32_t>(), c->welford.as<double>(), acc, acc_w, acc_sq};
        HIPC(wf_launch_accumulate_adaptive(W[j], B, S->stream));
      }
      HIPC(hipEventRecord(c->set[1].done, c->set[1].stream));
      HIPC(hipStreamWaitEvent(s, c->set[1].done, 0));
      HIPC(hipMemcpyAsync(&n_active, n_out, 4, hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
      taken += kk;
      std::swap(act_in, act_out);
    }
  }
  HIPC(total.mark());
  HIPC(hipStreamSynchronize(s));
  uint64_t h[5] = {0, 0, 0, 0, 0};
  for (int j = 0; j < nsets; j++) {
    uint32_t overflow = 0;
    HIPC(hipMemcpy(&overflow, W[j].ps.counters + TPT_CNT_OVERFLOW, 4, hipMemcpyDeviceToHost));
    if (overflow) return fail(TPT_ERR_INTERNAL, "a path queue segment overflowed");
    if (st) {
      std::vector<uint64_t> seg(TPT_STAT_WORDS);
      HIPC(hipMemcpy(seg.data(), W[j].ps.stats, TPT_STAT_WORDS * 8, hipMemcpyDeviceToHost));
      for (int g = 0; g < TPT_WF_SEGS; g++)
        for (int q = 0; q < 5; q++) h[q] += seg[(size_t)g * TPT_STAT_STRIDE + q];
    }
  }
  if (st) {
    st->paths = h[2];
    st->rays_extension = h[0];
    st->rays_shadow = h[1];
    st->elapsed_ms = total.pairs_ms();
    st->trace_kernel_ms = tt0.pairs_ms() + tt1.pairs_ms();
    st->trace_launches = launches;
    st->trace_rays = h[0];
    st->light_samples = h[3];
    st->light_self_occluded = h[4];
  }
  return TPT_OK;
}

int render_plain(tpt_ctx* c, const tpt_scene* sc, const tpt_render_config* cfg, uint32_t npix, const int32_t* pix_map,
                 double* acc, double* acc_w, double* acc_sq, tpt_stats* st) {
  if (sc->host.max_blas_depth > 63)
    return fail(TPT_ERR_UNSUPPORTED, "the plain FP64 form keeps the reference's 64-entry BVH stack; tree too deep");
  hipStream_t s = c->stream;
  HIPC(c->rngpix.ensure((size_t)npix * 8));
  HIPC(c->nsamp.ensure((size_t)npix * 4));
  HIPC(c->welford.ensure((size_t)npix * 16));
  HIPC(c->done.ensure(npix));
  HIPC(c->stats.ensure(64));
  HIPC(hipMemsetAsync(c->done.p, 0, npix, s));
  HIPC(hipMemsetAsync(c->stats.p, 0, 64, s));
  Plain64Params P{};
  P.S = sc->s64;
  P.env = sc->env;
  P.lights = sc->lights;
  P.mats = sc->mats;
  P.cam = sc->host.cam;
  P.width = cfg->width;
  P.height = cfg->height;
  P.n_pix = (int32_t)npix;
  P.pix_map = pix_map;
  P.spp = cfg->spp;
  P.max_bounces = cfg->max_bounces;
  P.rr_start = cfg->rr_start;
  P.adaptive = cfg->adaptive_enabled;
  P.min_spp = cfg->adaptive_min_spp;
  P.threshold = cfg->adaptive_threshold;
  P.inv_w = 1.0 / (double)cfg->width;
  P.inv_h = 1.0 / (double)cfg->height;
  P.seed = cfg->seed;
  P.rng = c->rngpix.as<uint64_t>();
  P.acc = acc;
  P.acc_sq = acc_sq;
  P.acc_w = acc_w;
  P.nsamp = c->nsamp.as<uint32_t>();
  P.welford = c->welford.as<double>();
  P.done = c->done.as<uint8_t>();
  P.stats = c->stats.as<uint64_t>();
  EventTimer total(s, st != nullptr, &c->ev_total);
  HIPC(total.mark());
  const int chunk = 8;
  if (cfg->spp == 0) {
    P.s_begin = 0;
    P.s_end = 0;
    HIPC(launch_plain64(P, s));
  }
  for (int s0 = 0; s0 < cfg->spp; s0 += chunk) {
    P.s_begin = s0;
    P.s_end = std::min(cfg->spp, s0 + chunk);
    HIPC(launch_plain64(P, s));
  }
  HIPC(total.mark());
  HIPC(hipStreamSynchronize(s));
  if (st) {
    uint64_t h[8] = {0};
    HIPC(hipMemcpy(h, c->stats.p, 64, hipMemcpyDeviceToHost));
    st->paths = h[2];
    st->rays_extension = h[0];
    st->rays_shadow = h[1];
    st->elapsed_ms = total.pairs_ms();
    st->trace_kernel_ms = st->elapsed_ms;
    st->trace_launches = (uint64_t)((cfg->spp + chunk - 1) / chunk);
    st->trace_rays = h[0] + h[1];
    st->light_samples = 0;  // the plain form traces every shadow ray in full (no separate decision)
    st->light_self_occluded = 0;
  }
  return TPT_OK;
}

// Zeroes the npix-entry sums, then renders the domain (see render_wavefront).
int render_device(tpt_ctx* c, const tpt_scene* sc, const tpt_render_config* cfg, uint32_t npix, const int32_t* pix_map,
                  double* acc, double* acc_w, double* acc_sq, tpt_stats* st) {
  hipStream_t s = c->stream;
  HIPC(hipMemsetAsync(acc, 0, (size_t)npix * 24, s));
  HIPC(hipMemsetAsync(acc_w, 0, (size_t)npix * 8, s));
  if (acc_sq) HIPC(hipMemsetAsync(acc_sq, 0, (size_t)npix * 24, s));
  if (st) std::memset(st, 0, sizeof(*st));
  if (cfg->precision == TPT_FP64_PLAIN) return render_plain(c, sc, cfg, npix, pix_map, acc, acc_w, acc_sq, st);
  return render_wavefront(c, sc, cfg, npix, pix_map, acc, acc_w, acc_sq, st);
}

// Host-side sums for a domain of npix pixels: device sums, render, copy back.
int render_to_host(tpt_ctx* c, const tpt_scene* sc, const tpt_render_config* cfg, uint32_t npix, const int32_t* pix_map,
                   double* acc_rgb, double* acc_w, double* acc_sq, tpt_stats* st) {
  DevBuf &a = c->out_acc, &w = c->out_w, &q = c->out_sq;
  HIPC(a.ensure((size_t)npix * 24));
  HIPC(w.ensure((size_t)npix * 8));
  if (acc_sq) HIPC(q.ensure((size_t)npix * 24));
  int rc = render_device(c, sc, cfg, npix, pix_map, a.as<double>(), w.as<double>(), acc_sq ? q.as<double>() : nullptr, st);
  if (rc) return rc;
  HIPC(hipMemcpy(acc_rgb, a.p, (size_t)npix * 24, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(acc_w, w.p, (size_t)npix * 8, hipMemcpyDeviceToHost));
  if (acc_sq) HIPC(hipMemcpy(acc_sq, q.p, (size_t)npix * 24, hipMemcpyDeviceToHost));
  return TPT_OK;
}

}  // namespace

extern "C" {

int tpt_render_device(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg, double* d_acc_rgb, double* d_acc_w,
                      double* d_acc_sq, tpt_stats* stats) {
  if (!ctx || !scene || !d_acc_rgb || !d_acc_w) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  int rc = check_cfg(cfg);
  if (rc) return rc;
  HIPC(hipSetDevice(ctx->device));
  return render_device(ctx, scene, cfg, (uint32_t)cfg->width * (uint32_t)cfg->height, nullptr, d_acc_rgb, d_acc_w,
                       d_acc_sq, stats);
}

int tpt_render_pixels(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg, int64_t n,
                      const int32_t* pixels, double* sum_rgb, double* sum_w, double* sum_sq, tpt_stats* stats) {
  if (!ctx || !scene || n < 0 || (n > 0 && (!pixels || !sum_rgb || !sum_w)))
    return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  int rc = check_cfg(cfg);
  if (rc) return rc;
  const int64_t npix = (int64_t)cfg->width * cfg->height;
  if (n > npix) return fail(TPT_ERR_INVALID_ARGUMENT, "more pixels listed than the image holds");
  for (int64_t i = 0; i < n; i++)
    if (pixels[i] < 0 || pixels[i] >= npix) return fail(TPT_ERR_INVALID_ARGUMENT, "pixel index outside the image");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return TPT_OK;
  HIPC(hipSetDevice(ctx->device));
  DevBuf map;
  HIPC(map.ensure((size_t)n * 4));
  HIPC(hipMemcpy(map.p, pixels, (size_t)n * 4, hipMemcpyHostToDevice));
  return render_to_host(ctx, scene, cfg, (uint32_t)n, map.as<int32_t>(), sum_rgb, sum_w, sum_sq, stats);
}

int tpt_render(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg, double* acc_rgb, double* acc_w,
               double* acc_sq, tpt_stats* stats) {
  if (!ctx || !scene || !acc_rgb || !acc_w) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  int rc = check_cfg(cfg);
  if (rc) return rc;
  HIPC(hipSetDevice(ctx->device));
  return render_to_host(ctx, scene, cfg, (uint32_t)cfg->width * (uint32_t)cfg->height, nullptr, acc_rgb, acc_w, acc_sq,
                        stats);
}

int tpt_render_plan(const tpt_render_config* cfg, int64_t npix, uint32_t path_budget, uint32_t n_active, uint32_t taken,
                    tpt_pass_plan* out) {
  if (!out) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (npix <= 0 || npix > (int64_t)1 << 28) return fail(TPT_ERR_INVALID_ARGUMENT, "pixel count out of range");
  const uint32_t spp = (uint32_t)cfg->spp;
  const RenderPlan rp = plan_render_sets((uint32_t)npix, spp, cfg->samples_per_pass, cfg->adaptive_enabled != 0,
                                         path_budget);
  const PassPlan plan = rp.pass;
  out->paths_alloc = plan.paths_alloc;
  out->path_sets = (uint32_t)rp.sets;
  out->samples_per_pass = plan.k;
  out->pixels_per_pass = plan.pix_chunk;
  if (cfg->adaptive_enabled) {
    if (n_active == 0 || n_active > npix) return fail(TPT_ERR_INVALID_ARGUMENT, "n_active must be in 1..npix");
    const uint32_t max_spp = std::max(spp, 1u);
    const uint32_t min_spp = std::min((uint32_t)std::max(cfg->adaptive_min_spp, 1), max_spp);
    plan_adaptive_pass(n_active, taken, min_spp, spp, rp.pass_cap, plan.paths_alloc, &out->samples_per_pass,
                       &out->pixels_per_pass);
  }
  return TPT_OK;
}

int tpt_trace_closest(tpt_ctx* ctx, const tpt_scene* sc, int64_t n, const float* rays, float* hit_t, int32_t* hit_kind,
                      int32_t* hit_inst, int32_t* hit_prim) {
  if (!ctx || !sc || n < 0 || (n > 0 && (!rays || !hit_t || !hit_kind || !hit_inst || !hit_prim)))
    return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (sc->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  if (n == 0) return TPT_OK;
  HIPC(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf r, t, k, i, p, g;
  HIPC(r.ensure((size_t)n * 32));
  HIPC(t.ensure((size_t)n * 4));
  HIPC(k.ensure((size_t)n * 4));
  HIPC(i.ensure((size_t)n * 4));
  HIPC(p.ensure((size_t)n * 4));
  int gcap = needed_gstack(sc->host);
  if (gcap > 0) HIPC(g.ensure((size_t)trace_grid_threads(n) * (size_t)gcap * 4));
  HIPC(hipMemcpyAsync(r.p, rays, (size_t)n * 32, hipMemcpyHostToDevice, s));
  HIPC(launch_trace_closest(sc->s32, n, r.as<float>(), t.as<float>(), k.as<int32_t>(), i.as<int32_t>(), p.as<int32_t>(),
                            gcap > 0 ? g.as<int32_t>() : nullptr, gcap, s));
  HIPC(hipMemcpyAsync(hit_t, t.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(hit_kind, k.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(hit_inst, i.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(hit_prim, p.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return TPT_OK;
}

int tpt_trace_closest_noalloc(tpt_ctx* ctx, const tpt_scene* sc, int64_t n, const float* d_rays, float* d_hit_t,
                              int32_t* d_hit_kind, int32_t* d_hit_inst, int32_t* d_hit_prim) {
  if (!ctx || !sc || n < 0 || (n > 0 && (!d_rays || !d_hit_t || !d_hit_kind || !d_hit_inst || !d_hit_prim)))
    return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (sc->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  if (n == 0) return TPT_OK;
  int32_t* scr = g_scratch_dev[ctx->device];
  if (!scr) return fail(TPT_ERR_INTERNAL, "the fixed scratch of this device was not set up");
  const int gcap = needed_gstack(sc->host);
  uint32_t max_blocks = 0;  // default grid when the stack fits in LDS (no scratch used)
  if (gcap > 0) {
    // a grid whose overflow stacks fit the scratch; the kernel loops over the rays
    const uint64_t fit = (uint64_t)(TPT_NOALLOC_SCRATCH_BYTES / 4) / (uint64_t)gcap / trace_block_threads();
    if (fit == 0)
      return fail(TPT_ERR_UNSUPPORTED, "the scene's traversal stack (" + std::to_string(sc->host.stack4) +
                                           " entries) needs more overflow space per block than the " +
                                           std::to_string(TPT_NOALLOC_SCRATCH_BYTES >> 20) +
                                           " MiB fixed scratch holds; use tpt_trace_closest");
    max_blocks = (uint32_t)std::min<uint64_t>(fit, trace_grid_threads(n) / trace_block_threads());
  }
  HIPC(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  HIPC(launch_trace_closest(sc->s32, n, d_rays, d_hit_t, d_hit_kind, d_hit_inst, d_hit_prim, gcap > 0 ? scr : nullptr,
                            gcap, s, max_blocks));
  HIPC(hipStreamSynchronize(s));
  return TPT_OK;
}

int tpt_render_adjoint(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg, const float* d_loss_d_pixel,
                       double* grad_params, tpt_stats* stats) {
  if (!ctx || !scene || !d_loss_d_pixel || !grad_params) return fail(TPT_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->ctx != ctx) return fail(TPT_ERR_INVALID_ARGUMENT, "scene was not uploaded with this context");
  int rc = check_cfg(cfg);
  if (rc) return rc;
  HIPC(hipSetDevice(ctx->device));
  size_t bytes[ADJ_SCRATCH_N];
  adjoint_scratch_bytes(cfg, scene->mats.n, stack_need(scene->host), bytes);
  void* p[ADJ_SCRATCH_N];
  for (int k = 0; k < ADJ_SCRATCH_N; k++) {
    p[k] = nullptr;
    if (bytes[k] > 0) {
      HIPC(ctx->adj[k].ensure(bytes[k]));
      p[k] = ctx->adj[k].p;
    }
  }
  const AdjScratch scratch{p[ADJ_SCRATCH_DLDP], p[ADJ_SCRATCH_GRAD], p[ADJ_SCRATCH_STATS], p[ADJ_SCRATCH_APM],
                           p[ADJ_SCRATCH_GSTACK]};
  return adjoint_render(ctx->stream, scene->s32, scene->s64, scene->env, scene->lights, scene->mats, scene->cam32,
                        scene->host.mat_param.data(), stack_need(scene->host), cfg, d_loss_d_pixel, grad_params, stats,
                        scratch, &g_err);
}

}  // extern "C"
