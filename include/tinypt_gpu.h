/*
 * tinypt_gpu.h -- C ABI of the MI355X (gfx950) path-tracing backend.
 *
 * The reference renderer (a Rust crate) has no FFI; its boundary is the crate's public API
 * (lib.rs:38-43):
 *
 *   pub fn render(scene: &Scene, config: &RenderConfig, ckpt_file: &str)
 *       -> io::Result<RenderOutput>                                    (render.rs:42)
 *   RenderOutput { acc: Vec<Color>, acc_w: Vec<f64> }                  (render.rs:34-39)
 *   Scene { cam: Camera, world: World, mats: Vec<Material>, env: Option<EnvMap> }  (scene.rs:14-23)
 *
 * This header replaces that API with plain C:
 *   - tpt_scene_desc   : a flat, host-owned description of Scene/World/Material/EnvMap/Camera.
 *                        The library builds the per-mesh BVH itself, with the same build as
 *                        Bvh::build (bvh.rs:51-242), and the light CDF as World::build_lights
 *                        (world.rs:127-167).
 *   - tpt_render_config: the RenderConfig fields the path reads (config.rs:5-46).
 *   - tpt_render       : render() -> acc / acc_w (render.rs:42-278), plus optional per-pixel
 *                        sums of squares for Monte Carlo error estimates.
 *
 * Conventions: every entry point returns 0 on success and a negative tpt_status on failure;
 * tpt_last_error() gives the message (thread-local).  Nothing throws across the ABI.  Output
 * buffers are caller-owned.  Calls are synchronous unless stated otherwise.  One context per
 * host thread.
 */
#ifndef TINYPT_GPU_H
#define TINYPT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPT_ABI_VERSION 3

/* ---- status codes ------------------------------------------------------------------------ */
typedef enum tpt_status {
  TPT_OK = 0,
  TPT_ERR_INVALID_ARGUMENT = -1,
  TPT_ERR_DEVICE = -2,         /* a HIP runtime call failed */
  TPT_ERR_OUT_OF_MEMORY = -3,
  TPT_ERR_UNSUPPORTED = -4,    /* a scene or configuration this backend does not handle */
  TPT_ERR_INTERNAL = -5
} tpt_status;

/* ---- materials (material.rs:30-43) --------------------------------------------------------- */
typedef enum tpt_material_kind {
  TPT_MAT_LAMBERT = 0,       /* param: albedo rgb                                 */
  TPT_MAT_METAL = 1,         /* param: albedo rgb                                 */
  TPT_MAT_DIELECTRIC = 2,    /* param: absorption rgb, [3] = ior                  */
  TPT_MAT_GGX = 3,           /* param: albedo (F0) rgb, [3] = alpha               */
  TPT_MAT_SUBSURFACE = 4,    /* param: albedo rgb, [3] = scatter_dist             */
  TPT_MAT_DIFFUSE_LIGHT = 5  /* param: emit rgb                                   */
} tpt_material_kind;

/* ---- camera (ray.rs:22-86): look_at_dof(eye, target, up, vfov, aspect, focus, aperture) ------ */
typedef struct tpt_camera_desc {
  double eye[3];
  double target[3];
  double up[3];
  double vfov_deg;
  double aspect;
  double focus_dist;  /* Camera::look_at uses |eye - target|; look_at_dof takes it explicitly */
  double aperture;    /* lens_radius = 0.5 * max(aperture, 0) (ray.rs:81)                    */
} tpt_camera_desc;

/* ---- environment (env.rs:20-89; integrator.rs:25-37) --------------------------------------- */
typedef enum tpt_env_kind {
  TPT_ENV_SKY = 0,    /* Scene.env == None: procedural sky() gradient (integrator.rs:25-28)     */
  TPT_ENV_IMAGE = 1   /* Scene.env == Some(EnvMap::from_pixels(w, h, rgb)) (env.rs:61-89)       */
} tpt_env_kind;

/* ---- flat scene description ---------------------------------------------------------------- */
typedef struct tpt_scene_desc {
  tpt_camera_desc camera;

  int32_t n_materials;
  const int32_t* mat_kind;       /* [n_materials] tpt_material_kind                            */
  const double* mat_param;       /* [n_materials * 4]                                          */

  int32_t n_spheres;             /* World.spheres (world.rs:50-51)                             */
  const double* sphere_center;   /* [n_spheres * 3]                                            */
  const double* sphere_radius;   /* [n_spheres]                                                */
  const int32_t* sphere_mat;     /* [n_spheres]                                                */

  int32_t n_meshes;              /* World.meshes (world.rs:52-53); static triangles only        */
  const int64_t* mesh_tri_begin; /* [n_meshes + 1] prefix offsets into the triangle arrays     */
  const double* tri_vertices;    /* [n_tris * 9] v0, v1, v2 in object space                    */
  const int32_t* tri_mat;        /* [n_tris]                                                   */

  int32_t n_instances;           /* World.instances (world.rs:39-46)                           */
  const int32_t* inst_mesh;      /* [n_instances]                                              */
  const double* inst_xform;      /* [n_instances * 12]: A (3x3 row-major), then t (3)          */
  const int32_t* inst_mat_override; /* [n_instances], -1 = None                                */

  int32_t env_kind;              /* tpt_env_kind                                               */
  int32_t env_width;
  int32_t env_height;
  const double* env_rgb;         /* [env_width * env_height * 3], linear RGB, row-major        */
} tpt_scene_desc;

/* ---- render configuration (config.rs:5-46, the fields render() and radiance() read) -------- */
typedef enum tpt_precision {
  TPT_FP32_WAVEFRONT = 0,  /* FP32 wavefront kernels, FP64 per-pixel accumulation; one RNG
                              stream per (pixel, sample): seed_for(x, y, s) ^ seed            */
  TPT_FP64_PLAIN = 1       /* FP64 megakernel, one thread per pixel, one RNG stream per pixel
                              across all samples exactly as render.rs:142                     */
} tpt_precision;

typedef struct tpt_render_config {
  int32_t width;
  int32_t height;
  int32_t spp;
  int32_t max_bounces;          /* RenderConfig.max_bounces                                   */
  int32_t rr_start;             /* RenderConfig.rr_start                                      */
  int32_t adaptive_enabled;     /* RenderConfig.adaptive_enabled                              */
  int32_t adaptive_min_spp;     /* RenderConfig.adaptive_min_spp                              */
  int32_t precision;            /* tpt_precision                                              */
  double adaptive_threshold;    /* RenderConfig.adaptive_threshold                            */
  uint64_t seed;                /* RenderConfig.seed                                          */
  int32_t samples_per_pass;     /* wavefront batching: samples of every pixel in flight per
                                   pass (0 = library default); no effect on results           */
  int32_t traversal_refill;     /* FP32 wavefront traversal: 0 = chosen by the scene's shape;
                                   -1 = one ray per lane; n > 0 = persistent waves that take
                                   new rays once fewer than min(n, 63) lanes still trace (no
                                   effect on results)                                         */
} tpt_render_config;

/* ---- counters ------------------------------------------------------------------------------ */
typedef struct tpt_stats {
  uint64_t paths;            /* (pixel, sample) camera paths completed                          */
  uint64_t rays_extension;   /* closest-hit queries                                             */
  uint64_t rays_shadow;      /* any-hit (shadow) queries                                        */
  double elapsed_ms;         /* device time from first launch to last completion                */
  double trace_kernel_ms;    /* summed device time of the extension traversal kernel            */
  uint64_t trace_launches;   /* launches of the extension traversal kernel                      */
  uint64_t trace_rays;       /* rays processed by those launches                                */
  uint64_t light_samples;    /* area-light samples with a non-zero contribution (FP32 form)     */
  uint64_t light_self_occluded; /* of those, the samples whose light shadows its own surface
                                   point (integrator.rs:178-213, the reference's f64 decision)  */
} tpt_stats;

/* ---- opaque handles ------------------------------------------------------------------------ */
typedef struct tpt_ctx tpt_ctx;
typedef struct tpt_scene tpt_scene;

/* Library ABI version (TPT_ABI_VERSION). */
int tpt_abi_version(void);

/* Last error message of the calling thread ("" if none). */
const char* tpt_last_error(void);

/* Create / destroy a context bound to HIP device `device`.  Replaces nothing in the reference
 * (the CPU renderer has no device); it owns the stream every call below runs on. */
int tpt_ctx_create(int device, tpt_ctx** out);
int tpt_ctx_destroy(tpt_ctx* ctx);

/* Build device-side scene data from a descriptor: per-mesh BVH as built by Bvh::build
 * (bvh.rs:51-242), a top-level tree over instances and spheres, the light CDF of
 * World::build_lights (world.rs:127-167), and for an image environment the cell weights of
 * EnvMap::from_pixels (env.rs:61-89) as FP64 and an alias table over the same cells.
 * The descriptor's arrays are copied; the caller may free them on return. */
int tpt_scene_upload(tpt_ctx* ctx, const tpt_scene_desc* desc, tpt_scene** out);
int tpt_scene_destroy(tpt_scene* scene);

/* Build the same host-side tables as tpt_scene_upload without touching a device (no context
 * needed): lets a caller validate a scene and inspect its BVHs, lights and environment tables
 * on a machine without a GPU.  Such a scene can be queried and destroyed but not rendered
 * (render / trace entry points return TPT_ERR_INVALID_ARGUMENT).  No reference counterpart:
 * the reference builds the same tables inside load_scene / build_scene. */
int tpt_scene_build_host(const tpt_scene_desc* desc, tpt_scene** out);

/* Scene facts: node counts, BVH depths, light count, triangle count. */
typedef struct tpt_scene_info {
  int64_t n_tris;
  int64_t n_blas_nodes;
  int64_t n_tlas_nodes;
  int32_t max_blas_depth;
  int32_t max_tlas_depth;
  int32_t n_lights;
  int32_t traversal_stack;  /* deepest traversal stack of the FP32 kernels (entries per ray); the
                               kernels keep 24 per lane in LDS and the rest in a global area */
  int64_t device_bytes;
} tpt_scene_info;
int tpt_scene_get_info(const tpt_scene* scene, tpt_scene_info* info);

/* Copy out the bottom-level BVH of mesh `mesh` as built (bvh.rs:24-43 layout, FP64 boxes):
 * node_bbox[n*6] (min xyz, max xyz), node_left[n], node_right[n], node_start[n],
 * node_count[n], indices[n_tris_of_mesh].  Pass NULL pointers to query *n_nodes only. */
int tpt_scene_get_bvh(const tpt_scene* scene, int32_t mesh, int64_t* n_nodes, double* node_bbox,
                      int32_t* node_left, int32_t* node_right, uint32_t* node_start,
                      uint32_t* node_count, int64_t* indices);

/* Light list as World::build_lights builds it (world.rs:127-167): per light kind (0 sphere,
 * 1 triangle), sphere index, mesh, triangle, instance, emit rgb, weight, cumulative CDF value
 * (10 doubles per light, light_out[n_lights*10]).  NULL light_out queries *n_lights only. */
int tpt_scene_get_lights(const tpt_scene* scene, int32_t* n_lights, double* light_out);

/* Environment sampling tables of an image environment: the FP64 total weight (env.rs:32), and
 * per cell (w*h) the alias-table probability and alias index and the pdf numerator
 * (weight/total) W H / (2 pi^2).  Any output pointer may be NULL. */
int tpt_scene_get_env_tables(const tpt_scene* scene, double* total_weight, float* alias_prob, int32_t* alias_idx,
                             float* cell_pdf);

/* render() (render.rs:42-278): fills acc_rgb[w*h*3] (sum of path radiance), acc_w[w*h]
 * (samples taken), and, if acc_sq is non-NULL, acc_sq[w*h*3] (sum of squared radiance per
 * channel).  Host pointers.  Checkpointing is handled by the caller. */
int tpt_render(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg,
               double* acc_rgb, double* acc_w, double* acc_sq, tpt_stats* stats);

/* Same as tpt_render with device pointers (e.g. torch tensors on the context's device);
 * the call is synchronous with respect to the host. */
int tpt_render_device(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg,
                      double* d_acc_rgb, double* d_acc_w, double* d_acc_sq, tpt_stats* stats);

/* A batch of tile tasks (render.rs:119-198, task.rs:9-40): every listed image pixel (index
 * y * width + x) is rendered with all cfg->spp samples -- adaptive and either precision as cfg
 * says -- and its sums are written, in list order, to sum_rgb[3i..3i+2], sum_w[i] and
 * sum_sq[3i..3i+2] (host pointers; sum_sq may be null).  A pixel's samples do not depend on
 * which other pixels share the call, so merging the sums in task order gives the full render
 * (bit for bit with the plain FP64 form, and with the FP32 form when cfg->samples_per_pass is
 * fixed, since its FP64 sums are then grouped alike).  The tiled, checkpointed render loop of
 * render.rs is built on this; indices outside the image are rejected. */
int tpt_render_pixels(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg, int64_t n,
                      const int32_t* pixels, double* sum_rgb, double* sum_w, double* sum_sq, tpt_stats* stats);

/* Pass sizing of an FP32 wavefront render (host arithmetic; no device, no context).  The
 * wavefront form keeps pixels x samples paths in flight per pass.  For npix pixels and at most
 * path_budget paths in flight (the renders use 2^26, halved while device memory is short):
 * paths_alloc = path slots of one path set; path_sets = the sets the render allocates (2 for
 * a fixed-spp render of two or more passes, which alternate between the sets, each then at
 * most half the budget, and for an adaptive render, whose chunks alternate between them; else
 * 1); samples_per_pass / pixels_per_pass = the fixed-spp passes, or, when
 * cfg->adaptive_enabled, the adaptive pass that starts after `taken` samples with n_active
 * pixels still sampling (its active list is launched in chunks of pixels_per_pass pixels).  No
 * pass or chunk holds more than paths_alloc paths.  No reference
 * counterpart: the reference's worker threads take one sample of one pixel at a time
 * (render.rs:131-191). */
typedef struct tpt_pass_plan {
  uint64_t paths_alloc;
  uint32_t samples_per_pass;
  uint32_t pixels_per_pass;
  uint32_t path_sets;
  uint32_t reserved;
} tpt_pass_plan;
int tpt_render_plan(const tpt_render_config* cfg, int64_t npix, uint32_t path_budget, uint32_t n_active,
                    uint32_t taken, tpt_pass_plan* out);

/* Closest hit of n rays (World::hit, world.rs:81-121) in FP32.  rays[n*8]:
 * ox oy oz dx dy dz tmin tmax (d need not be unit length; it is used as given).  Outputs per
 * ray: hit_t (world t, +inf on miss), hit_kind (0 miss, 1 triangle, 2 sphere), hit_inst
 * (instance index, or sphere index for spheres), hit_prim (triangle index within its mesh in
 * the mesh's original order, -1 for spheres).  Host pointers. */
int tpt_trace_closest(tpt_ctx* ctx, const tpt_scene* scene, int64_t n, const float* rays,
                      float* hit_t, int32_t* hit_kind, int32_t* hit_inst, int32_t* hit_prim);

/* tpt_trace_closest for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid allocation): rays and outputs are device
 * pointers of the context's device, laid out as for tpt_trace_closest, and the call allocates
 * nothing -- the traversal's overflow stacks live in a fixed 16 MiB scratch per device, held
 * until the process ends (its grid is capped so that they fit).  Fails with
 * TPT_ERR_UNSUPPORTED for a scene whose traversal stack is too deep for that scratch.  Calls
 * that overlap on one device race on the scratch: issue them from one host thread at a time.
 * Synchronous.  No counterpart in the reference. */
int tpt_trace_closest_noalloc(tpt_ctx* ctx, const tpt_scene* scene, int64_t n, const float* d_rays, float* d_hit_t,
                              int32_t* d_hit_kind, int32_t* d_hit_inst, int32_t* d_hit_prim);

/* Differentiable parameters of the adjoint: per material, d/d(param[0..3]) as laid out in
 * mat_param (albedo rgb + alpha for GGX, albedo rgb for Lambert, emit rgb for DiffuseLight). */
#define TPT_GRAD_PER_MATERIAL 4

/* Path-replay adjoint of the FP32 render: given dLoss/dPixel (d_loss_d_pixel[w*h*3], the
 * derivative w.r.t. the resolved pixel acc/acc_w), regenerates every path from its seed and
 * accumulates grad_params[n_materials*4] += sum over paths of dLoss/dparam.  Albedo / F0 /
 * absorption / emission are differentiated in reverse mode along the path (sampling pdfs and
 * Russian-roulette probabilities held fixed).  GGX roughness (param[3]) takes the central
 * difference (h = 1e-3) of each path's loss term, replayed with the same random numbers,
 * because the reference's VNDF sampler (material.rs:327-347) draws from a density whose
 * support moves with alpha.  No reference counterpart.  Host pointers. */
int tpt_render_adjoint(tpt_ctx* ctx, const tpt_scene* scene, const tpt_render_config* cfg,
                       const float* d_loss_d_pixel, double* grad_params, tpt_stats* stats);

/* Update the material table of an uploaded scene in place (mat_param[n_materials*4]); the
 * light CDF is rebuilt when emitter radiance changes.  Used by inverse rendering loops. */
int tpt_scene_update_materials(tpt_ctx* ctx, tpt_scene* scene, const double* mat_param);

/* ---- output and checkpoints --------------------------------------------------------------- */
typedef enum tpt_tonemap {
  TPT_TONEMAP_NONE = 0,  /* Tonemap::None (config.rs:50-56) */
  TPT_TONEMAP_ACES = 1   /* Tonemap::Aces: Narkowicz fit (output.rs:137-148) */
} tpt_tonemap;

/* resolve_pixels (output.rs:96-102): rgb_out[i] = acc_rgb[i] / max(acc_w[i], 1), n pixels. */
int tpt_resolve_pixels(int64_t n, const double* acc_rgb, const double* acc_w, double* rgb_out);

/* OutputFormat::from_path + write (output.rs:45-94) on the accumulated buffers: ".exr" (any
 * case) writes linear ACEScg float32 R, G, B with ACEScg chromaticities (an uncompressed
 * OpenEXR file; the reference's encoder compresses), ".hdr" writes Radiance RGBE (hdr.rs:65-88),
 * anything else an ASCII PPM with exposure (EV), the tonemap and the sRGB encoding. */
int tpt_write_image(const char* path, int32_t width, int32_t height, const double* acc_rgb, const double* acc_w,
                    double exposure, int32_t tonemap);

/* save_checkpoint / load_checkpoint (checkpoint.rs:48-138): the HYPCKPT v1 file (magic,
 * version, scene hash, width, height, next id, acc, acc_w; little endian), written through a
 * temporary file and a rename.  load sets *found = 0 (and returns TPT_OK) when the file is
 * absent or its magic, version, scene hash or size differ, as the reference returns None. */
int tpt_checkpoint_save(const char* path, uint64_t scene_hash, int32_t width, int32_t height, uint64_t next_id,
                        const double* acc_rgb, const double* acc_w);
int tpt_checkpoint_load(const char* path, uint64_t scene_hash, int32_t width, int32_t height, uint64_t* next_id,
                        double* acc_rgb, double* acc_w, int32_t* found);

/* ---- scene files ------------------------------------------------------------------------- */
typedef struct tpt_loaded_scene tpt_loaded_scene;

/* load_scene (mitsuba.rs:137-223): the Mitsuba-XML subset the reference reads -- perspective
 * sensor, sphere / rectangle / cube / disk / obj shapes with to_world transforms, diffuse /
 * conductor / roughconductor / dielectric / twosided bsdfs, area emitters, constant and
 * envmap (Radiance .hdr) scene emitters, and film / sampler / integrator settings, which
 * update cfg's width, height, spp, max_bounces and rr_start.  Unsupported elements are skipped
 * with a warning on stderr ("[mitsuba] warning: ..."), as the reference does; the warnings
 * are also kept with the loaded scene.  Fails on unreadable files, malformed XML or a root
 * that is not <scene>. */
int tpt_scene_load_xml(const char* path, tpt_render_config* cfg, tpt_loaded_scene** out);

/* The loaded scene as a descriptor for tpt_scene_upload / tpt_scene_build_host; its arrays
 * live until tpt_loaded_scene_destroy.  No reference counterpart (the reference returns a
 * Scene value). */
int tpt_loaded_scene_desc(const tpt_loaded_scene* scene, const tpt_scene_desc** desc);

/* Newline-separated warnings of the load ("" if none); valid until destroy. */
const char* tpt_loaded_scene_warnings(const tpt_loaded_scene* scene);

int tpt_loaded_scene_destroy(tpt_loaded_scene* scene);

/* load_obj_triangles (obj_loader.rs:101-109): triangles of a Wavefront OBJ file (v and f
 * records; polygons fan-triangulated; negative indices count from the end; other records
 * ignored).  tri_vertices[n*9] (v0, v1, v2); pass NULL to query *n_tris only.  With a buffer,
 * *n_tris on entry is its capacity in triangles (TPT_ERR_INVALID_ARGUMENT, nothing written, if
 * the file holds more) and on return the number of triangles. */
int tpt_obj_load(const char* path, int64_t* n_tris, double* tri_vertices);

#ifdef __cplusplus
}
#endif

#endif /* TINYPT_GPU_H */
