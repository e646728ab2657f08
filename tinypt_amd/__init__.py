"""tinypt_amd -- MI355X (gfx950) backend for the tinypt path-tracing hot path.

Host mirror of the reference crate's render entry (lib.rs:38-43, render.rs:42):

    scene = tinypt_amd.scenes.cornell_box(512, 512)
    cfg = tinypt_amd.RenderConfig(width=512, height=512, spp=256, max_bounces=12, rr_start=4, seed=2)
    out = tinypt_amd.render(scene, cfg)          # RenderOutput(acc (h,w,3), acc_w (h,w))
    img = tinypt_amd.resolve_pixels(out.acc, out.acc_w)

Everything runs through the C ABI of ``lib/libtinypt_gpu.so`` (include/tinypt_gpu.h).  There is
no CPU fallback: without the built library or a HIP device every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi
from ._abi import FP32_WAVEFRONT, FP64_PLAIN
from .scene import (Camera, DiffuseLight, Dielectric, EnvMap, Ggx, Instance, Lambert, Material, Mesh, Metal, RenderConfig,
                    Scene, Sphere, Subsurface, Transform, World, color_from_srgb, luminance)

__all__ = ["Camera", "DiffuseLight", "Dielectric", "EnvMap", "Ggx", "Instance", "Lambert", "Material", "Mesh", "Metal",
           "RenderConfig", "Scene", "Sphere", "Subsurface", "Transform", "World", "Context", "DeviceScene", "HostScene",
           "RenderOutput",
           "render", "render_pixels", "render_device", "render_adjoint", "resolve_pixels", "lib_path", "load_library", "TptError", "FP32_WAVEFRONT", "FP64_PLAIN"]

import atexit
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libtinypt_gpu.so")
_lib = None


class TptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tinypt_gpu error {code}: {msg}")
        self.code = code


def lib_path() -> str:
    return _LIB_PATH


def load_library():
    """Load the built C-ABI library (never builds; run tinypt_amd.build first)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (soname libamdhip64.so.7).
        # Loaded first, it is the runtime this library binds to; loaded after this library it
        # would be a second runtime, and PyTorch would then find no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(_LIB_PATH):
            raise TptError(_abi.TPT_ERR_INTERNAL, f"{_LIB_PATH} is not built (python -m tinypt_amd.build)")
        L = C.CDLL(_LIB_PATH)
        for name, (res, args) in _abi.ENTRY_POINTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.tpt_abi_version() != _abi.ABI_VERSION:
            raise TptError(_abi.TPT_ERR_INTERNAL, "ABI version mismatch")
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise TptError(rc, load_library().tpt_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# Live handles, released in dependency order before the HIP runtime tears itself down at exit.
_live_scenes: "weakref.WeakSet" = weakref.WeakSet()
_live_ctxs: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    for s in list(_live_scenes):
        s.close()
    for c in list(_live_ctxs):
        c.close()


class Context:
    """A HIP device context (one stream).  ``tpt_ctx_create``."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.tpt_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        _live_ctxs.add(self)

    def close(self):
        if getattr(self, "h", None):
            self._lib.tpt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class _SceneTables:
    """Queries shared by host-only and uploaded scenes."""

    def info(self) -> dict:
        inf = _abi.SceneInfo()
        _check(self._lib.tpt_scene_get_info(self.h, C.byref(inf)))
        return inf.as_dict()

    def bvh(self, mesh: int) -> dict:
        """The mesh's BVH as built (bvh.rs:24-43 layout)."""
        n = C.c_int64()
        _check(self._lib.tpt_scene_get_bvh(self.h, mesh, C.byref(n), None, None, None, None, None, None))
        n = n.value
        ntri = len(self.scene.world.meshes[mesh].tris)
        bbox = np.zeros((max(n, 1), 6))
        left = np.zeros(max(n, 1), np.int32)
        right = np.zeros(max(n, 1), np.int32)
        start = np.zeros(max(n, 1), np.uint32)
        count = np.zeros(max(n, 1), np.uint32)
        idx = np.zeros(max(ntri, 1), np.int64)
        nn = C.c_int64()
        _check(self._lib.tpt_scene_get_bvh(self.h, mesh, C.byref(nn), _dp(bbox), left.ctypes.data_as(C.POINTER(C.c_int32)),
                                           right.ctypes.data_as(C.POINTER(C.c_int32)),
                                           start.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           count.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           idx.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(bbox=bbox[:n], left=left[:n], right=right[:n], start=start[:n], count=count[:n], indices=idx[:ntri])

    def lights(self) -> np.ndarray:
        """(n, 10): kind, sphere, mesh, tri, instance, emit rgb, weight, cdf (world.rs:127-167)."""
        n = C.c_int32()
        _check(self._lib.tpt_scene_get_lights(self.h, C.byref(n), None))
        out = np.zeros((max(n.value, 1), 10))
        _check(self._lib.tpt_scene_get_lights(self.h, C.byref(n), _dp(out)))
        return out[:n.value]

    def env_tables(self) -> dict:
        """Alias table and pdf numerators of an image environment."""
        env = self.scene.env
        N = env.width * env.height
        total = C.c_double()
        prob = np.zeros(N, np.float32)
        alias = np.zeros(N, np.int32)
        pdf = np.zeros(N, np.float32)
        _check(self._lib.tpt_scene_get_env_tables(self.h, C.byref(total), prob.ctypes.data_as(C.POINTER(C.c_float)),
                                                  alias.ctypes.data_as(C.POINTER(C.c_int32)),
                                                  pdf.ctypes.data_as(C.POINTER(C.c_float))))
        return dict(total=total.value, prob=prob, alias=alias, cell_pdf=pdf)


class HostScene(_SceneTables):
    """Host-side scene tables only (``tpt_scene_build_host``): no device needed, cannot render."""

    def __init__(self, scene: Scene):
        self._lib = load_library()
        self.scene = scene
        desc, keep = scene.to_desc()
        h = C.c_void_p()
        _check(self._lib.tpt_scene_build_host(C.byref(desc), C.byref(h)))
        self.h = h
        _live_scenes.add(self)

    def close(self):
        if getattr(self, "h", None):
            self._lib.tpt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceScene(_SceneTables):
    """A scene uploaded to the device (``tpt_scene_upload``): BVHs built as Bvh::build does,
    top-level tree, lights, environment tables."""

    def __init__(self, scene: Scene, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        self.scene = scene
        desc, keep = scene.to_desc()
        h = C.c_void_p()
        _check(self._lib.tpt_scene_upload(self.ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        self.n_materials = len(scene.mats)
        _live_scenes.add(self)

    def close(self):
        if getattr(self, "h", None):
            self._lib.tpt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def update_materials(self, mat_param: np.ndarray):
        """Replace every material's 4 parameters (``tpt_scene_update_materials``); a table equal to
        the one uploaded last is not uploaded again."""
        p = np.ascontiguousarray(np.asarray(mat_param, dtype=np.float64).reshape(self.n_materials, 4))
        last = getattr(self, "_last_materials", None)
        if last is not None and np.array_equal(last, p):
            return
        _check(self._lib.tpt_scene_update_materials(self.ctx.h, self.h, _dp(p)))
        self._last_materials = p.copy()

    def trace_closest(self, rays: np.ndarray) -> dict:
        """FP32 closest hit (World::hit semantics).  rays (n, 8): o, d, tmin, tmax."""
        r = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(-1, 8))
        n = len(r)
        t = np.zeros(n, np.float32)
        kind = np.zeros(n, np.int32)
        inst = np.zeros(n, np.int32)
        prim = np.zeros(n, np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _check(self._lib.tpt_trace_closest(self.ctx.h, self.h, n, r.ctypes.data_as(C.POINTER(C.c_float)),
                                           t.ctypes.data_as(C.POINTER(C.c_float)), ip(kind), ip(inst), ip(prim)))
        return dict(t=t, kind=kind, inst=inst, prim=prim)


@dataclass
class RenderOutput:
    """render.rs:34-39 plus per-pixel sums of squares and run counters."""
    acc: np.ndarray     # (h, w, 3) sum of path radiance
    acc_w: np.ndarray   # (h, w) samples per pixel
    acc_sq: Optional[np.ndarray]
    stats: dict


def render(scene, config: RenderConfig, want_sq: bool = True, ctx: Optional[Context] = None,
           ckpt_file: Optional[str] = None, progress: bool = False) -> RenderOutput:
    """render() (render.rs:42-278) on the GPU.  ``scene`` is a Scene or an uploaded
    DeviceScene.

    With ``ckpt_file`` set, ``config.checkpoint_enabled`` and ``checkpoint_every_tasks`` > 0 the
    render follows the reference's tile tasks: it resumes from a matching checkpoint, renders
    the remaining tiles in task order (batches of whole checkpoint intervals, one GPU call per
    batch), merges them in task-id order and saves the HYPCKPT file every
    ``checkpoint_every_tasks`` tasks and at the end.  ``acc_sq`` is then None when the render
    resumed (the file holds no squares).  ``progress`` prints the reference's stderr lines
    (``Render: ...``, ``Progress: n/N tiles (p%)`` at most every 250 ms, ``Done in ...``) and
    also renders tile batches.  Otherwise the whole image is one GPU call."""
    dscene = scene if isinstance(scene, DeviceScene) else DeviceScene(scene, ctx)
    ckpt_on = ckpt_file is not None and config.checkpoint_enabled and config.checkpoint_every_tasks > 0
    if ckpt_on or progress:
        return _render_tiled(dscene, config, ckpt_file if ckpt_on else None, want_sq, progress)
    lib = dscene._lib
    h, w = int(config.height), int(config.width)
    acc = np.zeros((h, w, 3))
    acc_w = np.zeros((h, w))
    acc_sq = np.zeros((h, w, 3)) if want_sq else None
    st = _abi.Stats()
    cfg = config.to_c()
    _check(lib.tpt_render(dscene.ctx.h, dscene.h, C.byref(cfg), _dp(acc), _dp(acc_w),
                          _dp(acc_sq) if want_sq else None, C.byref(st)))
    return RenderOutput(acc, acc_w, acc_sq, st.as_dict())


def render_pixels(scene: "DeviceScene", config: RenderConfig, pixels: np.ndarray, want_sq: bool = False) -> tuple:
    """``tpt_render_pixels``: every listed image pixel (y * width + x) rendered with all its
    samples; returns (sum (n, 3), sum_w (n,), sum_sq (n, 3) or None, counters) in list order
    (the TileResult sums of task.rs:28-40)."""
    pix = np.ascontiguousarray(pixels, dtype=np.int32).reshape(-1)
    n = pix.shape[0]
    s = np.zeros((n, 3))
    sw = np.zeros(n)
    sq = np.zeros((n, 3)) if want_sq else None
    st = _abi.Stats()
    cfg = config.to_c()
    _check(scene._lib.tpt_render_pixels(scene.ctx.h, scene.h, C.byref(cfg), n, pix.ctypes.data_as(C.POINTER(C.c_int32)),
                                        _dp(s), _dp(sw), _dp(sq) if want_sq else None, C.byref(st)))
    return s, sw, sq, st.as_dict()


_CKPT_BATCH_PATHS = 1 << 22  # samples per GPU call below which more checkpoint intervals join a batch
_TILE_BATCH_PATHS = 1 << 24  # samples per GPU call of a tiled render without checkpoints
_PROGRESS_INTERVAL_S = 0.25  # constants.rs:30


def _render_tiled(dscene: "DeviceScene", config: RenderConfig, ckpt_file: Optional[str], want_sq: bool,
                  progress: bool) -> RenderOutput:
    """render.rs:48-278 over tile tasks: tiles in task-id order (Morton order when enabled),
    rendered in batches by tpt_render_pixels, merged in task-id order; with ``ckpt_file`` the
    render resumes from and saves HYPCKPT checkpoints, with ``progress`` it reports progress."""
    import sys
    import time
    from .output import load_checkpoint, save_checkpoint, task_pixels, tile_tasks
    h, w = int(config.height), int(config.width)
    acc = np.zeros((h, w, 3))
    acc_w = np.zeros((h, w))
    resume = 0
    if ckpt_file is not None:
        try:
            got = load_checkpoint(ckpt_file, config.scene_hash, w, h)
        except TptError:
            got = None  # render.rs:57 takes only Ok(Some(..))
        if got is not None:
            resume, acc, acc_w = got
            print(f"Resumed from checkpoint: {ckpt_file} (next task id: {resume})", file=sys.stderr)
    acc_sq = np.zeros((h, w, 3)) if want_sq and resume == 0 else None
    tasks = tile_tasks(w, h, int(config.tile), bool(config.morton_enabled))
    tid = len(tasks)
    every = int(config.checkpoint_every_tasks) if ckpt_file is not None else 1
    batch_paths = _CKPT_BATCH_PATHS if ckpt_file is not None else _TILE_BATCH_PATHS
    flat_acc, flat_w = acc.reshape(-1, 3), acc_w.reshape(-1)
    flat_sq = acc_sq.reshape(-1, 3) if acc_sq is not None else None
    stats = {}
    next_id = min(resume, tid)
    sizes = [(x1 - x0) * (y1 - y0) for x0, y0, x1, y1 in tasks]
    if progress:  # render.rs:101-105, 216-222
        print(f"Render: {w}x{h}, spp={config.spp}, tile={config.tile}, device=hip:{dscene.ctx.device}, "
              f"morton={str(bool(config.morton_enabled)).lower()}, seed={config.seed}", file=sys.stderr)
        if tid > 0:
            print(f"\rProgress: {next_id}/{tid} tiles ({next_id * 100.0 / tid:.1f}%)", end="", file=sys.stderr,
                  flush=True)
    start = last_print = time.monotonic()
    while next_id < tid:
        end, n_pix = next_id, 0
        while end < tid:
            stop = min(tid, (end // every + 1) * every)
            n_pix += sum(sizes[end:stop])
            end = stop
            if n_pix * max(int(config.spp), 1) >= batch_paths:
                break
        pix = task_pixels(tasks[next_id:end], w)
        s, sw, sq, st = render_pixels(dscene, config, pix, want_sq=flat_sq is not None)
        for k, v in st.items():
            stats[k] = stats.get(k, 0) + v
        off = 0
        for t in range(next_id, end):
            sl = slice(off, off + sizes[t])
            p = pix[sl]
            flat_acc[p] = flat_acc[p] + s[sl]
            flat_w[p] += sw[sl]
            if flat_sq is not None:
                flat_sq[p] += sq[sl]
            off += sizes[t]
            now = time.monotonic()
            if progress and (now - last_print >= _PROGRESS_INTERVAL_S or t + 1 == tid):  # render.rs:244-251
                print(f"\rProgress: {t + 1}/{tid} tiles ({(t + 1) * 100.0 / tid:.1f}%)", end="", file=sys.stderr,
                      flush=True)
                last_print = now
            if ckpt_file is not None and (t + 1) % every == 0:
                save_checkpoint(ckpt_file, config.scene_hash, t + 1, acc, acc_w)
        next_id = end
    if progress:
        print(file=sys.stderr)
        print(f"Done in {time.monotonic() - start:.2f}s", file=sys.stderr)
    if ckpt_file is not None:
        save_checkpoint(ckpt_file, config.scene_hash, tid, acc, acc_w)  # render.rs:267-272
    return RenderOutput(acc, acc_w, acc_sq, stats)


def render_device(scene: "DeviceScene", config: RenderConfig, acc, acc_w, acc_sq=None) -> dict:
    """``tpt_render_device``: render into caller-owned device buffers (e.g. float64 torch
    tensors on the context's device, acc (h, w, 3), acc_w (h, w), acc_sq (h, w, 3) or None).
    Synchronous with respect to the host.  Returns the run counters."""
    h, w = int(config.height), int(config.width)
    for name, t, n in (("acc", acc, h * w * 3), ("acc_w", acc_w, h * w), ("acc_sq", acc_sq, h * w * 3)):
        if t is None:
            continue
        if getattr(t, "dtype", None) is not None and str(t.dtype) not in ("torch.float64", "float64"):
            raise ValueError(f"{name} must be float64")
        if hasattr(t, "numel") and (t.numel() != n or not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous buffer of {n} doubles")
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    st = _abi.Stats()
    cfg = config.to_c()
    _check(scene._lib.tpt_render_device(scene.ctx.h, scene.h, C.byref(cfg), ptr(acc), ptr(acc_w), ptr(acc_sq),
                                        C.byref(st)))
    return st.as_dict()


def render_adjoint(scene: "DeviceScene", config: RenderConfig, d_loss_d_pixel: np.ndarray,
                   grad: Optional[np.ndarray] = None) -> tuple:
    """Gradient of a loss with respect to every material's 4 parameters (``tpt_render_adjoint``).

    ``d_loss_d_pixel`` (h, w, 3) is dLoss/dpixel of the resolved image (acc / spp).  The render
    is the FP32 wavefront estimator with the same seeds, replayed path by path; the result is
    added into ``grad`` (n_materials, 4), which is returned together with the run counters."""
    lib = scene._lib
    h, w = int(config.height), int(config.width)
    dl = np.ascontiguousarray(np.asarray(d_loss_d_pixel, dtype=np.float32).reshape(h, w, 3))
    g = np.zeros((scene.n_materials, 4)) if grad is None else grad
    if g.dtype != np.float64 or not g.flags.c_contiguous or g.shape != (scene.n_materials, 4):
        raise ValueError("grad must be a C-contiguous float64 array of shape (n_materials, 4)")
    st = _abi.Stats()
    cfg = config.to_c()
    _check(lib.tpt_render_adjoint(scene.ctx.h, scene.h, C.byref(cfg), dl.ctypes.data_as(C.POINTER(C.c_float)), _dp(g),
                                  C.byref(st)))
    return g, st.as_dict()


def resolve_pixels(acc: np.ndarray, acc_w: np.ndarray) -> np.ndarray:
    """output.rs:96-102: acc / max(acc_w, 1)."""
    return acc / np.maximum(acc_w, 1.0)[..., None]
