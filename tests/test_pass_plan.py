"""Pass sizing of the FP32 wavefront render (tpt_render_plan, host arithmetic; no device).

The wavefront form keeps pixels x samples paths in flight per pass.  These tests pin that no
pass -- fixed-spp or adaptive, including images with more pixels than the path budget -- holds
more paths than the render allocates, and that the passes cover every pixel and sample."""
import ctypes as C

import pytest

import tinypt_amd as T
from tinypt_amd import _abi


def plan_full(cfg, npix, budget, n_active=0, taken=0):
    out = _abi.PassPlan()
    c = cfg.to_c()
    rc = T.load_library().tpt_render_plan(C.byref(c), npix, budget, n_active, taken, C.byref(out))
    if rc:
        raise T.TptError(rc, T.load_library().tpt_last_error().decode())
    return out


def plan(cfg, npix, budget, n_active=0, taken=0):
    out = plan_full(cfg, npix, budget, n_active, taken)
    return out.paths_alloc, out.samples_per_pass, out.pixels_per_pass


def cfg_for(w, h, spp, adaptive=False, min_spp=64, spp_pass=0):
    return T.RenderConfig(width=w, height=h, spp=spp, adaptive_enabled=adaptive, adaptive_min_spp=min_spp,
                          samples_per_pass=spp_pass)


@pytest.mark.parametrize("w,h,spp,spp_pass,budget", [
    (512, 512, 256, 0, 1 << 26), (1920, 1080, 1024, 0, 1 << 26), (16384, 8193, 4, 0, 1 << 26),
    (16384, 8193, 4, 0, 1 << 20), (256, 256, 64, 0, 1 << 20), (7, 3, 1000, 0, 1 << 20), (512, 512, 256, 3, 1 << 26),
    (16384, 16384, 2, 64, 1 << 26), (1, 1, 0, 0, 1 << 26)])
def test_fixed_spp_passes_fit_and_cover(w, h, spp, spp_pass, budget):
    npix = w * h
    full = plan_full(cfg_for(w, h, spp, spp_pass=spp_pass), npix, budget)
    alloc, k, chunk, sets = full.paths_alloc, full.samples_per_pass, full.pixels_per_pass, full.path_sets
    assert k >= 1 and chunk >= 1
    assert chunk * k <= alloc
    assert sets in (1, 2) and sets * alloc <= 2 * budget
    n_passes = -(-npix // chunk) * -(-spp // k)
    assert (sets == 2) == (n_passes >= 2), (sets, n_passes)
    # the loops of the render: pixel chunks x sample groups cover every (pixel, sample) once
    covered = 0
    for p0 in range(0, npix, chunk):
        np_ = min(chunk, npix - p0)
        for s0 in range(0, spp, k):
            kk = min(k, spp - s0)
            assert np_ * kk <= alloc
            covered += np_ * kk
    assert covered == npix * spp


def simulate_adaptive(w, h, spp, min_spp, budget, stop_rule):
    """Runs the render's adaptive loop on the host: every pass's chunks must fit the allocation
    and each pixel must get exactly the samples the stop rule lets it take."""
    npix = w * h
    cfg = cfg_for(w, h, spp, adaptive=True, min_spp=min_spp)
    alloc, _, _ = plan(cfg, npix, budget, n_active=npix, taken=0)
    taken, n_active, passes = 0, npix, 0
    while n_active > 0 and taken < spp:
        _, kk, chunk = plan(cfg, npix, budget, n_active=n_active, taken=taken)
        assert kk >= 1 and chunk >= 1 and chunk * kk <= alloc, (taken, n_active, kk, chunk, alloc)
        assert taken + kk <= spp
        if taken < min(min_spp, spp) and n_active * (min(min_spp, spp) - taken) <= alloc:
            assert taken + kk >= min(min_spp, spp)  # the min-spp phase is one pass when it fits
        launched = sum(min(chunk, n_active - c0) for c0 in range(0, n_active, chunk))
        assert launched == n_active
        taken += kk
        n_active = stop_rule(taken, n_active)
        passes += 1
    return passes, alloc


def test_adaptive_passes_fit_when_pixels_exceed_the_budget():
    # 16384 x 8193 pixels (accepted: below 2^28) with 2^26 paths in flight: before the chunked
    # active list, a pass launched one path per pixel into 2^26 slots
    # (two path sets of half the slots each; a pass's chunks alternate between them)
    passes, alloc = simulate_adaptive(16384, 8193, 8, 2, 1 << 26, lambda t, n: n - n // 3)
    assert alloc == 1 << 25 and passes >= 2
    passes, alloc = simulate_adaptive(4096, 4096, 16, 4, 1 << 20, lambda t, n: n // 2)
    assert alloc == 1 << 19


@pytest.mark.parametrize("w,h,spp,min_spp", [(1920, 1080, 4096, 64), (512, 512, 256, 64), (32, 32, 256, 64),
                                             (64, 64, 10, 64), (3, 5, 100, 1)])
def test_adaptive_passes_fit_and_reach_max_spp(w, h, spp, min_spp):
    passes, alloc = simulate_adaptive(w, h, spp, min_spp, 1 << 26, lambda t, n: max(n - n // 10, 1))
    assert passes >= 1


def test_plan_rejects_bad_arguments():
    with pytest.raises(T.TptError):
        plan(cfg_for(4, 4, 8, adaptive=True), 16, 1 << 20, n_active=0)
    with pytest.raises(T.TptError):
        plan(cfg_for(4, 4, 8, adaptive=True), 16, 1 << 20, n_active=17)
    with pytest.raises(T.TptError):
        plan(cfg_for(4, 4, 8), 0, 1 << 20)


def test_two_path_sets_for_renders_of_several_passes():
    # the Cornell benchmark (512^2 x 256 spp, 2^26 paths in flight): two passes of 128 samples,
    # one per path set of 2^25 paths
    full = plan_full(cfg_for(512, 512, 256), 512 * 512, 1 << 26)
    assert (full.path_sets, full.samples_per_pass, full.pixels_per_pass, full.paths_alloc) == (2, 128, 512 * 512, 1 << 25)
    # a render that fits one pass keeps one set of the whole budget
    full = plan_full(cfg_for(256, 256, 64), 256 * 256, 1 << 26)
    assert (full.path_sets, full.samples_per_pass, full.paths_alloc) == (1, 64, 256 * 256 * 64)
    # adaptive renders: two sets of half the slots; a pass of 2^21 paths or more has at least two
    # chunks, so both sets work (its samples per pixel are sized by the whole budget)
    full = plan_full(cfg_for(512, 512, 256, adaptive=True), 512 * 512, 1 << 26, n_active=512 * 512)
    assert full.path_sets == 2 and full.paths_alloc == 1 << 25
    assert full.samples_per_pass == 64 and full.pixels_per_pass == 512 * 512 // 2
    full = plan_full(cfg_for(512, 512, 256, adaptive=True), 512 * 512, 1 << 26, n_active=1000, taken=64)
    assert full.samples_per_pass == 64 and full.pixels_per_pass == 1000
