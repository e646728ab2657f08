/* c_abi_render.c -- a plain C client of include/tinypt_gpu.h: what a non-Python caller (the
 * Rust crate through its extern "C" block) does.  Loads a Mitsuba-XML scene, uploads it,
 * renders it at the given size and spp with the FP32 wavefront form and the plain FP64 form,
 * writes each image's sums as raw doubles (acc_rgb then acc_w) and a PPM of the FP32 render.
 *
 *   c_abi_render scene.xml width height spp out_prefix
 *
 * Exit status 0 on success; on failure the library's message goes to stderr. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tinypt_gpu.h"

static int fail(const char* what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, tpt_last_error());
  return 1;
}

static int write_raw(const char* path, const double* a, size_t na, const double* b, size_t nb) {
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  int bad = fwrite(a, sizeof(double), na, f) != na || fwrite(b, sizeof(double), nb, f) != nb;
  return fclose(f) != 0 || bad;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s scene.xml width height spp out_prefix\n", argv[0]);
    return 2;
  }
  if (tpt_abi_version() != TPT_ABI_VERSION) {
    fprintf(stderr, "ABI version %d, header %d\n", tpt_abi_version(), TPT_ABI_VERSION);
    return 1;
  }
  tpt_render_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.width = atoi(argv[2]);  /* the film size before loading sets the camera's aspect ratio */
  cfg.height = atoi(argv[3]);
  cfg.spp = atoi(argv[4]);
  cfg.max_bounces = 8;
  cfg.rr_start = 3;
  cfg.seed = 7;
  tpt_loaded_scene* loaded = NULL;
  int rc = tpt_scene_load_xml(argv[1], &cfg, &loaded);
  if (rc) return fail("tpt_scene_load_xml", rc);
  cfg.width = atoi(argv[2]);
  cfg.height = atoi(argv[3]);
  cfg.spp = atoi(argv[4]);
  const tpt_scene_desc* desc = NULL;
  if ((rc = tpt_loaded_scene_desc(loaded, &desc))) return fail("tpt_loaded_scene_desc", rc);

  tpt_ctx* ctx = NULL;
  if ((rc = tpt_ctx_create(0, &ctx))) return fail("tpt_ctx_create", rc);
  tpt_scene* scene = NULL;
  if ((rc = tpt_scene_upload(ctx, desc, &scene))) return fail("tpt_scene_upload", rc);

  const size_t npix = (size_t)cfg.width * (size_t)cfg.height;
  double* acc = calloc(npix * 3, sizeof(double));
  double* acc_w = calloc(npix, sizeof(double));
  if (!acc || !acc_w) return 1;
  const int32_t precisions[2] = {TPT_FP32_WAVEFRONT, TPT_FP64_PLAIN};
  const char* names[2] = {"fp32", "fp64"};
  for (int k = 0; k < 2; k++) {
    cfg.precision = precisions[k];
    tpt_stats st;
    if ((rc = tpt_render(ctx, scene, &cfg, acc, acc_w, NULL, &st))) return fail("tpt_render", rc);
    char path[4096];
    snprintf(path, sizeof path, "%s_%s.bin", argv[5], names[k]);
    if (write_raw(path, acc, npix * 3, acc_w, npix)) {
      fprintf(stderr, "cannot write %s\n", path);
      return 1;
    }
    printf("%s: %llu paths, %llu extension rays\n", names[k], (unsigned long long)st.paths,
           (unsigned long long)st.rays_extension);
    if (k == 0) {
      snprintf(path, sizeof path, "%s_fp32.ppm", argv[5]);
      if ((rc = tpt_write_image(path, cfg.width, cfg.height, acc, acc_w, 0.0, TPT_TONEMAP_ACES)))
        return fail("tpt_write_image", rc);
    }
  }
  free(acc);
  free(acc_w);
  if ((rc = tpt_scene_destroy(scene))) return fail("tpt_scene_destroy", rc);
  if ((rc = tpt_ctx_destroy(ctx))) return fail("tpt_ctx_destroy", rc);
  if ((rc = tpt_loaded_scene_destroy(loaded))) return fail("tpt_loaded_scene_destroy", rc);
  return 0;
}
