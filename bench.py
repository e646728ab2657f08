#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Workload: the Cornell box built from rectangles and two cubes (white .73 / red .65,.05,.05 /
green .12,.45,.15, ceiling light 17,12,4), 512x512, 256 samples per pixel, max_depth 12,
rr_depth 4, seed 2, FP32 wavefront kernels with FP64 per-pixel accumulation.  One step = one
full render of that image (67.1M camera paths).

Printed (one JSON line):  paths/s of the FP32 wavefront render (value), Mrays/s, the plain
FP64 GPU form's paths/s on the same workload, RMSE of both against a 2048-spp FP64 plain
render, the roofline of the dominant kernel (extension traversal) and the FP64 CPU oracle
timed on a bounded sample on this host.  Under "workloads", the other configurations, each
timed once or a few times: the default scene at 256x256, 64 spp (FP32 and plain FP64), the
1.31M-triangle mesh at 1920x1080, 1024 spp, and the 4096-instance scene at 1920x1080, adaptive
64..4096 spp, threshold 0.01.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-heavy]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 512
SPP = 256
MAX_BOUNCES = 12
RR_START = 4
SEED = 2
HBM_PEAK_GBS = 8000.0
METRIC = "paths/s & Mrays/s at fixed RMSE vs FP64 oracle; FP32 vs FP64 plain; grad s/iter"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reference_bytes_per_ray(counters):
    """Bytes the reference's traversal would read and write per extension (closest-hit) query on
    the same scene, from the oracle's work counts: 24 per AABB slab test (FP32 box), 36 per
    triangle test, 16 per sphere test, 48 per instance visit (FP32 affine inverse), 32 ray in +
    16 hit out.  A diagnostic of the work the four-wide trees and the caches absorb, not the
    kernel's traffic."""
    c = counters
    return (24.0 * c["aabb_ext"] + 36.0 * c["tri_ext"] + 16.0 * c["sph_ext"] + 48.0 * c["inst_ext"]) / c["rays_ext"] + 48.0


# Compulsory HBM bytes of one extension ray: the first bounce generates its camera rays and
# writes a 12-byte hit record; later bounces read the 32-byte ray (origin + pdf, direction +
# flags) and write the hit.  The scene's nodes and triangles are read through the caches.
EXTEND_BYTES_FIRST = 12.0
EXTEND_BYTES_LATER = 44.0


def cpu_baseline(scene, seconds_target=15.0):
    """The FP64 oracle (a C++ restatement of the reference CPU renderer with its scheduler:
    16x16 tiles, Morton order, one thread per core) on a bounded sample of the workload."""
    import oracle
    o = oracle.OracleScene(scene)
    threads = len(os.sched_getaffinity(0)) or 1
    # calibrate on 1 spp, then size the sample to ~seconds_target
    r = o.render(W, H, 1, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED, threads=threads)
    rate = W * H / max(r["seconds"], 1e-6)
    spp = int(max(1, min(64, seconds_target * rate / (W * H))))
    r = o.render(W, H, spp, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED, threads=threads)
    paths = W * H * spp
    c = r["counters"]
    return {"value": paths / r["seconds"], "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"512x512 at {spp} spp of the same scene ({paths} paths, {r['seconds']:.2f} s)",
            "mrays_per_s": (c["rays_ext"] + c["rays_shadow"]) / r["seconds"] / 1e6, "cpu_model": cpu_model()}, c


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


TRAFFIC_PROFILE = "profiles/wavefront_cornell_512x512_256spp_hbm_counters.md"
VALU_PROFILE = "profiles/wavefront_cornell_512x512_256spp_valu_l2_counters.md"


def measured_traffic():
    """HBM bytes per k_extend launch on this workload: rocprofv3 FETCH_SIZE + WRITE_SIZE, each
    scaled by the counter's reading for a 1 GiB copy (the two counter runs and the calibration
    are recorded in TRAFFIC_PROFILE, whose last line holds the figure)."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("k_extend_traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def measured_valu():
    """VALU busy and lane utilisation (percent) of the later-bounce extension kernel, from the
    counter table in VALU_PROFILE."""
    path = os.path.join(ROOT, VALU_PROFILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| `void tpt::k_extend<false, false>"):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            return {"valu_busy_pct": float(cols[3]), "valu_lane_utilisation_pct": float(cols[4]),
                    "occupancy_pct": float(cols[5]), "source": VALU_PROFILE}
    return None


def hip_copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` through the HIP runtime's copy kernel: a kernel of
    known traffic, against which the FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def _timed(ds, cfg, reps, warm=True):
    """Mean wall time of `reps` renders into device buffers (after one warm-up render when
    `warm`), with the stats of the last."""
    import tinypt_amd as T
    import torch
    dev = torch.device("cuda", ds.ctx.device)
    acc = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float64, device=dev)
    acc_w = torch.empty((cfg.height, cfg.width), dtype=torch.float64, device=dev)
    if warm:
        T.render_device(ds, cfg, acc, acc_w)
    t0 = time.perf_counter()
    for _ in range(reps):
        st = T.render_device(ds, cfg, acc, acc_w)
    secs = (time.perf_counter() - t0) / reps
    return secs, st, acc_w


def other_workloads(ctx, heavy=True, only=None):
    """The other configurations: the default scene at 256x256, 64 spp, depth 8, rr 3, seed 1 in
    both precisions; with `heavy`, the 1.31M-triangle mesh (1920x1080, 1024 spp, depth 16, rr 3)
    and the 4096-instance scene (1920x1080, adaptive 64..4096 spp, threshold 0.01, depth 8).
    `only` names one of them ("default", "mesh", "instanced") to run it alone."""
    import tinypt_amd as T
    from tinypt_amd import scenes
    out = {}
    if only not in (None, "default"):
        return _heavy_workloads(ctx, out, only)
    ds = T.DeviceScene(scenes.default_scene(256, 256, "black"), ctx)
    kw = dict(width=256, height=256, spp=64, max_bounces=8, rr_start=3, seed=1)
    s32, st32, _ = _timed(ds, T.RenderConfig(precision=T.FP32_WAVEFRONT, **kw), 20)
    s64, _, _ = _timed(ds, T.RenderConfig(precision=T.FP64_PLAIN, **kw), 5)
    paths = 256 * 256 * 64
    out["default_256x256_64spp"] = {"fp32_ms": s32 * 1e3, "fp32_paths_per_s": paths / s32,
                                    "fp32_mrays_per_s": (st32["rays_extension"] + st32["rays_shadow"]) / s32 / 1e6,
                                    "fp64_plain_ms": s64 * 1e3, "fp64_plain_paths_per_s": paths / s64,
                                    "speedup_fp32_vs_fp64_plain": s64 / s32}
    log(f"default 256x256 64 spp: fp32 {s32 * 1e3:.3f} ms, plain fp64 {s64 * 1e3:.3f} ms ({s64 / s32:.1f}x)")
    ds.close()
    if not heavy or only == "default":
        return out
    return _heavy_workloads(ctx, out, None)


def _heavy_workloads(ctx, out, only):
    import tinypt_amd as T
    from tinypt_amd import scenes
    if only == "instanced":
        return _instanced_workload(ctx, out)
    t0 = time.perf_counter()
    ds = T.DeviceScene(scenes.mesh_scene(1920, 1080), ctx)
    build = time.perf_counter() - t0
    cfg = T.RenderConfig(width=1920, height=1080, spp=1024, max_bounces=16, rr_start=3, seed=3)
    secs, st, _ = _timed(ds, cfg, 1, warm=False)
    paths = 1920 * 1080 * 1024
    out["mesh_1920x1080_1024spp"] = {"ms": secs * 1e3, "paths": paths, "paths_per_s": paths / secs,
                                     "mrays_per_s": (st["rays_extension"] + st["rays_shadow"]) / secs / 1e6,
                                     "extension_kernel_ms": st["trace_kernel_ms"], "scene_build_s": build,
                                     "triangles": 1310722}
    log(f"mesh 1080p 1024 spp: {secs:.3f} s, {paths / secs / 1e6:.1f} M paths/s")
    ds.close()
    if only == "mesh":
        return out
    return _instanced_workload(ctx, out)


def _instanced_workload(ctx, out):
    import tinypt_amd as T
    from tinypt_amd import scenes
    t0 = time.perf_counter()
    ds = T.DeviceScene(scenes.instanced_scene(1920, 1080), ctx)
    build = time.perf_counter() - t0
    cfg = T.RenderConfig(width=1920, height=1080, spp=4096, max_bounces=8, rr_start=3, seed=3, adaptive_enabled=True,
                         adaptive_min_spp=64, adaptive_threshold=0.01)
    secs, st, acc_w = _timed(ds, cfg, 1, warm=False)
    w = acc_w.cpu().numpy()
    out["instanced_4096_1920x1080_adaptive_64_4096"] = {
        "ms": secs * 1e3, "paths": st["paths"], "paths_per_s": st["paths"] / secs,
        "mrays_per_s": (st["rays_extension"] + st["rays_shadow"]) / secs / 1e6, "mean_spp": float(w.mean()),
        "min_spp": float(w.min()), "max_spp": float(w.max()), "extension_kernel_ms": st["trace_kernel_ms"],
        "scene_build_s": build}
    log(f"instanced 4096 adaptive: {secs:.2f} s, {st['paths'] / secs / 1e6:.1f} M paths/s, mean spp {w.mean():.0f}")
    ds.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-rmse", action="store_true", help="skip the 2048-spp FP64 reference image")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU render")
    ap.add_argument("--no-inverse", action="store_true", help="skip the inverse-rendering gradient loop")
    ap.add_argument("--inverse-steps", type=int, default=200)
    ap.add_argument("--no-heavy", action="store_true", help="skip the mesh and 4096-instance workloads")
    ap.add_argument("--no-workloads", action="store_true", help="skip every workload but the Cornell box")
    ap.add_argument("--only", choices=["default", "mesh", "instanced"],
                    help="time one of the other workloads alone and print its figures (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (a known-size kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the renderer has no multi-GPU path; run it with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import tinypt_amd as T
    from tinypt_amd import scenes

    if args.only:
        print(json.dumps({"workloads": other_workloads(T.default_context(), only=args.only)}), flush=True)
        return

    scene = scenes.cornell_box(W, H)
    ds = T.DeviceScene(scene)
    cfg = T.RenderConfig(width=W, height=H, spp=SPP, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED,
                         precision=T.FP32_WAVEFRONT)
    # The timed steps render into device-resident FP64 buffers (tpt_render_device): inputs and
    # outputs stay in HBM.  The same steps with host buffers (tpt_render, which adds the
    # device-to-host copy of the accumulators) are timed after them and reported beside.
    import torch
    dev = torch.device("cuda", ds.ctx.device)
    acc = torch.empty((H, W, 3), dtype=torch.float64, device=dev)
    acc_w = torch.empty((H, W), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        T.render_device(ds, cfg, acc, acc_w)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(T.render_device(ds, cfg, acc, acc_w))  # synchronous: returns after the device finishes
    t1 = time.perf_counter()
    secs = (t1 - t0) / args.steps
    out_acc, out_w = acc.cpu().numpy(), acc_w.cpu().numpy()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        T.render(ds, cfg, want_sq=False)
    host_secs = (time.perf_counter() - t0) / args.steps
    paths = W * H * SPP
    rays = np.mean([s["rays_extension"] + s["rays_shadow"] for s in stats])
    value = paths / secs
    mrays = rays / secs / 1e6
    trace_ms = np.mean([s["trace_kernel_ms"] for s in stats])
    trace_launches = np.mean([s["trace_launches"] for s in stats])
    trace_rays = np.mean([s["trace_rays"] for s in stats])
    img32 = T.resolve_pixels(out_acc, out_w)
    log(f"fp32 wavefront: {secs * 1e3:.1f} ms/step, {value / 1e6:.1f} M paths/s, {mrays:.1f} Mrays/s "
        f"(host buffers: {host_secs * 1e3:.1f} ms/step)")

    if args.calibrate_copy:
        hip_copy_calibration(1 << 30)
        log("calibration copy: 1 GiB device-to-device (1 GiB read + 1 GiB written)")

    # plain FP64 GPU form on the same workload
    plain_rate = None
    img64 = None
    if not args.no_plain:
        cfg64 = T.RenderConfig(width=W, height=H, spp=SPP, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED,
                               precision=T.FP64_PLAIN)
        t0 = time.perf_counter()
        T.render_device(ds, cfg64, acc, acc_w)
        secs64 = time.perf_counter() - t0
        img64 = T.resolve_pixels(acc.cpu().numpy(), acc_w.cpu().numpy())
        plain_rate = paths / secs64
        log(f"fp64 plain: {secs64 * 1e3:.1f} ms, {plain_rate / 1e6:.1f} M paths/s")

    rmse = None
    if not args.no_rmse:
        ref_cfg = T.RenderConfig(width=W, height=H, spp=2048, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED + 1000,
                                 precision=T.FP64_PLAIN)
        ref = T.render(ds, ref_cfg, want_sq=False)
        ref_img = T.resolve_pixels(ref.acc, ref.acc_w)
        rmse = {"fp32_wavefront_256spp": float(np.sqrt(np.mean((img32 - ref_img) ** 2))),
                "fp64_plain_256spp": float(np.sqrt(np.mean((img64 - ref_img) ** 2))) if img64 is not None else None,
                "reference": "FP64 plain GPU render, 2048 spp, independent seed"}
        log(f"rmse: {rmse}")

    cpu = None
    bytes_per_ray = None
    if not args.no_cpu:
        cpu, counters = cpu_baseline(scene)
        bytes_per_ray = reference_bytes_per_ray(counters)
        log(f"cpu oracle: {cpu}")
    else:
        import oracle
        r = oracle.OracleScene(scene).render(64, 64, 16, max_bounces=MAX_BOUNCES, rr_start=RR_START, seed=SEED)
        bytes_per_ray = reference_bytes_per_ray(r["counters"])

    inverse = None
    if not args.no_inverse:
        # Cornell box + 16 GGX spheres (64 parameters), 256x256, 16 spp forward, depth 8, target at
        # 1024 spp; one step = forward render + path-replay adjoint + Adam update
        from tinypt_amd.inverse import InverseProblem, optimize
        prob = InverseProblem(ctx=ds.ctx)
        hist = optimize(prob, steps=args.inverse_steps, lr=0.02, eval_spp=512)
        inverse = {"s_per_iter": hist["s_per_iter"], "steps": args.inverse_steps, "loss_first": hist["loss"][0],
                   "loss_last10_mean": float(np.mean(hist["loss"][-10:])),
                   "param_error_init": hist["init_error"], "param_error_final": hist["param_error"][-1],
                   "f0_error_init": hist["init_f0_error"], "f0_error_final": hist["f0_error"][-1],
                   "alpha_error_init": hist["init_alpha_error"], "alpha_error_final": hist["alpha_error"][-1],
                   "image_error_init": hist["image_error_init"], "image_error_final": hist["image_error_final"],
                   "image_error_truth": hist["image_error_truth"],
                   "image_error_note": "MSE vs the target of a 512-spp render; the truth value is the noise floor",
                   "config": "cornell + 16 GGX spheres, 256x256, 16 spp, depth 8, target 1024 spp, Adam lr 0.02"}
        log(f"inverse: {inverse}")

    workloads = None if args.no_workloads else other_workloads(ds.ctx, heavy=not args.no_heavy)

    # roofline of the extension kernel: its compulsory bytes per launch over its measured launch
    # time (HIP events around every launch, on the stream of the path set that launches it; the
    # render's two path sets run side by side, so a launch shares the GPU with the other set's)
    avg_launch_ms = trace_ms / max(trace_launches, 1)
    rays_per_launch = trace_rays / max(trace_launches, 1)
    first_rays = float(paths)  # one camera ray per path on the first bounce
    bytes_per_render = EXTEND_BYTES_FIRST * first_rays + EXTEND_BYTES_LATER * max(trace_rays - first_rays, 0.0)
    bytes_per_launch = bytes_per_render / max(trace_launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic, "kernel": "k_extend (FP32 closest-hit traversal)",
                "algorithmic_bytes": "per ray: 12 B hit written on the first bounce (camera rays generated), "
                                     "32 B ray read + 12 B hit written on later bounces; nodes and triangles cached",
                "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_launch_ms, "rays_per_launch": rays_per_launch,
                "traffic_source": TRAFFIC_PROFILE if traffic else None,
                "hbm_gbs_from_traffic": traffic / (avg_launch_ms * 1e-3) / 1e9 if traffic else None,
                "limiter": {"kind": "VALU issue (divergent per-ray traversal work)", **(measured_valu() or {})},
                "concurrency": "two path sets on two streams: each launch's duration includes the other set's "
                               "kernels running beside it",
                "reference_traversal_bytes_per_ray": bytes_per_ray}

    line = {"metric": METRIC, "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": secs * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (scene generated in code)",
            "config": {"workload": "cornell_box_512x512_256spp_maxdepth12_rrdepth4_seed2", "width": W, "height": H,
                       "spp": SPP, "max_depth": MAX_BOUNCES, "rr_depth": RR_START, "paths_per_step": paths,
                       "accumulation": "fp64"},
            "mrays_per_s": mrays, "host_buffers_ms_per_step": host_secs * 1e3,
            "host_buffers_paths_per_s": paths / host_secs, "fp64_plain_paths_per_s": plain_rate,
            "speedup_vs_fp64_plain": value / plain_rate if plain_rate else None,
            "rmse": rmse, "grad_s_per_iter": inverse["s_per_iter"] if inverse else None, "inverse": inverse,
            "roofline": roofline, "cpu_baseline": cpu, "workloads": workloads}
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
