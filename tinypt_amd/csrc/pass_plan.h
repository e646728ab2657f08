// pass_plan.h -- how an FP32 wavefront render splits its paths into passes (host arithmetic).
//
// All paths of a pass are in flight at once: a pass of `pixels` pixels with `k` samples each
// holds pixels * k path slots.  The plan fixes the slot count the render allocates, and every
// pass -- fixed-spp or adaptive -- is sized so that it never holds more.
#pragma once

#include <stdint.h>

#include <algorithm>

namespace tpt {

struct PassPlan {
  uint64_t paths_alloc;  // path slots the render allocates; no pass holds more
  uint32_t k;            // fixed-spp passes: samples of each pixel per pass
  uint32_t pix_chunk;    // fixed-spp passes: pixels per pass
};

// npix pixels at spp samples, at most `target` paths in flight by default (the caller's
// samples_per_pass, if positive, fixes k instead).  Images too large for one sample of every
// pixel in 2 * target slots are split into pixel chunks of `target` pixels at one sample.
inline PassPlan plan_render(uint32_t npix, uint32_t spp, int32_t samples_per_pass, bool adaptive, uint32_t target) {
  target = std::max(target, 1u);
  uint32_t k = samples_per_pass > 0 ? (uint32_t)samples_per_pass : std::max(1u, target / std::max(npix, 1u));
  k = std::max(1u, std::min(k, std::max(spp, 1u)));
  uint32_t pix_chunk = npix;
  if ((uint64_t)npix * k > (uint64_t)target * 2) {
    k = 1;
    pix_chunk = std::min(npix, target);
  }
  uint64_t P = (uint64_t)pix_chunk * k;
  // adaptive passes give each active pixel up to 64 samples, within the same in-flight budget
  if (adaptive) P = std::max<uint64_t>(P, std::min<uint64_t>((uint64_t)npix * 64, target));
  return PassPlan{std::max<uint64_t>(P, 1), k, std::max(pix_chunk, 1u)};
}

// The render's plan.  A fixed-spp render of two or more passes alternates them between two path
// sets, each of at most half the budget (one set's kernels then fill the other's ramp-downs); a
// one-pass render uses one set of the whole budget.  An adaptive render keeps the whole budget
// for the size of its passes (the samples per pixel per pass) and launches each pass's active
// list in chunks that alternate between two sets of half the slots.
struct RenderPlan {
  PassPlan pass;      // pass.paths_alloc: slots of one set
  int sets;           // path sets of pass.paths_alloc slots each
  uint64_t pass_cap;  // adaptive: paths of one pass over both sets
};
inline RenderPlan plan_render_sets(uint32_t npix, uint32_t spp, int32_t samples_per_pass, bool adaptive,
                                   uint32_t target) {
  if (adaptive) {
    PassPlan p = plan_render(npix, spp, samples_per_pass, true, target);
    const uint64_t cap = p.paths_alloc;
    p.paths_alloc = std::max<uint64_t>(1, (cap + 1) / 2);
    return RenderPlan{p, 2, cap};
  }
  const PassPlan half = plan_render(npix, spp, samples_per_pass, false, std::max(target / 2, 1u));
  const uint64_t passes = (uint64_t)((npix + half.pix_chunk - 1) / half.pix_chunk) * ((spp + half.k - 1) / half.k);
  if (passes >= 2) return RenderPlan{half, 2, half.paths_alloc};
  const PassPlan one = plan_render(npix, spp, samples_per_pass, false, target);
  return RenderPlan{one, 1, one.paths_alloc};
}

// One adaptive pass: n_active pixels are still sampling after `taken` samples each.  Every
// active pixel gets *kk more samples (at least the rest of the min_spp phase, else up to 64 as
// the pass's cap allows); the active list is launched in chunks of *chunk pixels, with
// chunk * kk <= set_cap (the slots of one path set).  A pass of at least 2^21 paths is split into
// at least two chunks, so that both path sets work.
inline void plan_adaptive_pass(uint32_t n_active, uint32_t taken, uint32_t min_spp, uint32_t spp, uint64_t pass_cap,
                               uint64_t set_cap, uint32_t* kk_out, uint32_t* chunk_out) {
  n_active = std::max(n_active, 1u);
  const uint64_t fill = std::max<uint64_t>(1, pass_cap / n_active);
  uint32_t kk = taken < min_spp ? min_spp - taken : 1;
  kk = std::max<uint32_t>(kk, (uint32_t)std::min<uint64_t>(fill, 64));
  kk = std::min(kk, spp > taken ? spp - taken : 1u);
  kk = (uint32_t)std::min<uint64_t>(kk, fill);
  kk = (uint32_t)std::min<uint64_t>(kk, std::max<uint64_t>(1, set_cap));
  kk = std::max(kk, 1u);
  uint64_t chunk = std::min<uint64_t>(n_active, std::max<uint64_t>(1, set_cap / kk));
  if (chunk == n_active && n_active >= 2 && (uint64_t)n_active * kk >= ((uint64_t)1 << 21)) chunk = (n_active + 1) / 2;
  *kk_out = kk;
  *chunk_out = (uint32_t)chunk;
}

}  // namespace tpt
