"""Image files and checkpoints: the steps after the path, on the FP64 accumulation buffers.

Thin wrappers over the host entry points of the C ABI (include/tinypt_gpu.h):

* ``write_image``      -- OutputFormat::from_path + write (output.rs:45-94): PPM, Radiance HDR
                          (hdr.rs:63-86) or OpenEXR in ACEScg (exr.rs:44-75, aces.rs:66-88)
* ``save_checkpoint`` / ``load_checkpoint`` -- the HYPCKPT v1 file (checkpoint.rs:48-138)
* ``ckpt_path``        -- checkpoint.rs:15-18
* ``morton2`` / ``tile_tasks`` -- the tile list and its Morton order (render.rs:72-98, 282-298)

None of these touch the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Tuple

import numpy as np

from . import _abi

TONEMAPS = {"none": _abi.TONEMAP_NONE, "aces": _abi.TONEMAP_ACES}


def tonemap_from_str(s: str) -> Optional[int]:
    """Tonemap::from_str (config.rs:58-66): case-insensitive "none" / "aces", else None."""
    return TONEMAPS.get(s.lower())


def _lib():
    from . import load_library
    return load_library()


def _check(rc: int):
    from . import _check as check
    check(rc)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _buffers(acc: np.ndarray, acc_w: np.ndarray):
    a = np.ascontiguousarray(acc, dtype=np.float64)
    w = np.ascontiguousarray(acc_w, dtype=np.float64)
    if a.ndim != 3 or a.shape[2] != 3 or w.shape != a.shape[:2]:
        raise ValueError("acc must be (h, w, 3) and acc_w (h, w)")
    return a, w


def write_image(path: str, acc: np.ndarray, acc_w: np.ndarray, exposure: float = 0.0, tonemap="aces") -> None:
    """Resolve acc / max(acc_w, 1) and write it in the format the extension names: ``.exr``
    (linear ACEScg float32), ``.hdr`` (linear RGBE), anything else PPM (exposure in EV, the
    tonemap, sRGB, 8 bit).  ``tonemap`` is "none", "aces" or a TONEMAP_* code."""
    a, w = _buffers(acc, acc_w)
    tm = tonemap_from_str(tonemap) if isinstance(tonemap, str) else int(tonemap)
    if tm is None:
        raise ValueError(f"unknown tonemap {tonemap!r}")
    h, wd = a.shape[:2]
    _check(_lib().tpt_write_image(path.encode(), wd, h, _dp(a), _dp(w), float(exposure), tm))


def resolve_pixels_native(acc: np.ndarray, acc_w: np.ndarray) -> np.ndarray:
    """tpt_resolve_pixels: the C ABI form of resolve_pixels (output.rs:96-102)."""
    a, w = _buffers(acc, acc_w)
    out = np.empty_like(a)
    _check(_lib().tpt_resolve_pixels(a.shape[0] * a.shape[1], _dp(a), _dp(w), _dp(out)))
    return out


def ckpt_path(scene_hash: int) -> str:
    """checkpoint.rs:15-18."""
    return f"checkpoint_{scene_hash & 0xFFFFFFFFFFFFFFFF:016x}.bin"


def save_checkpoint(path: str, scene_hash: int, next_id: int, acc: np.ndarray, acc_w: np.ndarray) -> None:
    a, w = _buffers(acc, acc_w)
    h, wd = a.shape[:2]
    _check(_lib().tpt_checkpoint_save(path.encode(), scene_hash & 0xFFFFFFFFFFFFFFFF, wd, h, int(next_id), _dp(a),
                                      _dp(w)))


def load_checkpoint(path: str, scene_hash: int, width: int, height: int
                    ) -> Optional[Tuple[int, np.ndarray, np.ndarray]]:
    """(next_id, acc (h, w, 3), acc_w (h, w)), or None when the file is absent or was written
    for another scene hash, size or format version.  A truncated file raises."""
    a = np.zeros((height, width, 3))
    w = np.zeros((height, width))
    nid = C.c_uint64(0)
    found = C.c_int32(0)
    _check(_lib().tpt_checkpoint_load(path.encode(), scene_hash & 0xFFFFFFFFFFFFFFFF, width, height, C.byref(nid),
                                      _dp(a), _dp(w), C.byref(found)))
    if not found.value:
        return None
    return int(nid.value), a, w


def _interleave_bits(v: int) -> int:
    x = v & 0x00000000FFFFFFFF
    x = (x | (x << 16)) & 0x0000FFFF0000FFFF
    x = (x | (x << 8)) & 0x00FF00FF00FF00FF
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0F
    x = (x | (x << 2)) & 0x3333333333333333
    x = (x | (x << 1)) & 0x5555555555555555
    return x


def morton2(tx: int, ty: int, tiles_x: int, tiles_y: int) -> int:
    """Z-order key of a tile (render.rs:282-298)."""
    x = min(tx, max(tiles_x - 1, 0))
    y = min(ty, max(tiles_y - 1, 0))
    return _interleave_bits(x) | (_interleave_bits(y) << 1)


def tile_tasks(width: int, height: int, tile: int, morton: bool = True) -> List[Tuple[int, int, int, int]]:
    """The render's tile tasks (x0, y0, x1, y1) in task-id order (render.rs:72-98): row-major
    tiles, stably sorted by Morton key when ``morton``."""
    if tile <= 0:
        raise ValueError("tile must be positive")
    tasks = [(x, y, min(x + tile, width), min(y + tile, height))
             for y in range(0, height, tile) for x in range(0, width, tile)]
    if morton:
        tx = (width + tile - 1) // tile
        ty = (height + tile - 1) // tile
        tasks.sort(key=lambda t: morton2(t[0] // tile, t[1] // tile, tx, ty))
    return tasks


def task_pixels(tasks, width: int) -> np.ndarray:
    """Image pixel indices (y * width + x) of the tasks, tile after tile, rows inside a tile."""
    parts = []
    for x0, y0, x1, y1 in tasks:
        ys, xs = np.meshgrid(np.arange(y0, y1, dtype=np.int64), np.arange(x0, x1, dtype=np.int64), indexing="ij")
        parts.append((ys * width + xs).reshape(-1))
    return np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)
