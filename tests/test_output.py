"""Image output and checkpoint files (host entry points of the C ABI; no GPU).

Pinned against the reference's own output tests (output.rs:150-199: from_path, resolve,
sRGB encode, HDR round trip), the published sRGB -> ACEScg matrix (Bradford D65 -> D60, as
OCIO and colour-science print it), closed forms of the sRGB curve and the Narkowicz fit, the
OpenEXR and Radiance file layouts read by decoders written here from the format
specifications, and the byte layout of the HYPCKPT file (checkpoint.rs:48-84).
"""
import math
import os
import struct

import numpy as np
import pytest

import tinypt_amd as T
from tinypt_amd import output as O


def _img(vals):
    """(h, w, 3) radiance with acc_w = 1."""
    a = np.asarray(vals, dtype=np.float64)
    return a, np.ones(a.shape[:2])


# ---------------------------------------------------------------- format choice (output.rs:45-56)
@pytest.mark.parametrize("name,magic", [("out.exr", b"\x76\x2f\x31\x01"), ("OUT.EXR", b"\x76\x2f\x31\x01"),
                                        ("out.hdr", b"#?RADIANCE"), ("out.HDR", b"#?RADIANCE"),
                                        ("out.ppm", b"P3\n"), ("out.txt", b"P3\n"), ("out", b"P3\n")])
def test_from_path_picks_format(tmp_path, name, magic):
    a, w = _img([[[0.2, 0.4, 0.6]]])
    p = str(tmp_path / name)
    O.write_image(p, a, w, tonemap="none")
    assert open(p, "rb").read(len(magic)) == magic


def test_resolve_divides_by_weight():
    acc = np.array([[[4.0, 2.0, 1.0], [0.0, 0.0, 0.0]]])
    acc_w = np.array([[2.0, 0.0]])
    for px in (O.resolve_pixels_native(acc, acc_w), T.resolve_pixels(acc, acc_w)):
        assert np.array_equal(px, np.array([[[2.0, 1.0, 0.5], [0.0, 0.0, 0.0]]]))


def _ppm_values(path):
    tok = open(path).read().split()
    assert tok[0] == "P3"
    w, h, mx = int(tok[1]), int(tok[2]), int(tok[3])
    assert mx == 255
    return np.array([int(t) for t in tok[4:]]).reshape(h, w, 3)


def test_ppm_srgb_encode(tmp_path):
    # linear 0.5 encodes to 0.735357 (the sRGB curve), i.e. 187.99 -> 188; clamping at 0 and 1
    a, w = _img([[[0.5, 0.0, 1.0], [2.0, -1.0, 0.0031308]]])
    p = str(tmp_path / "a.ppm")
    O.write_image(p, a, w, tonemap="none")
    v = _ppm_values(p)
    assert v[0, 0].tolist() == [188, 0, 255]
    # 0.0031308 is the linear-segment knee: 12.92 * 0.0031308 * 255 + 0.5 = 10.81 -> 10
    assert v[0, 1].tolist() == [255, 0, 10]
    assert open(p).read().startswith("P3\n2 1\n255\n")


def test_ppm_exposure_and_aces(tmp_path):
    a, w = _img([[[0.25, 1.0, 100.0]]])
    p = str(tmp_path / "b.ppm")
    O.write_image(p, a, w, exposure=1.0, tonemap="none")  # 2^1 scale: 0.25 -> 0.5
    assert _ppm_values(p)[0, 0].tolist() == [188, 255, 255]
    O.write_image(p, a, w, exposure=0.0, tonemap="ACES")
    v = _ppm_values(p)[0, 0]
    # Narkowicz fit at x = 1: (2.51 + 0.03) / (2.43 + 0.59 + 0.14) = 2.54 / 3.16; the fit tends
    # to 2.51 / 2.43 > 1 and is clamped to 1 for large x
    y = 2.54 / 3.16
    srgb = 1.055 * y ** (1 / 2.4) - 0.055
    assert v[1] == int(srgb * 255 + 0.5) == 232
    assert v[2] == 255
    # x = 0.25: 0.25 * (0.6275 + 0.03) / (0.25 * (0.6075 + 0.59) + 0.14) = 0.164375 / 0.439375
    y = 0.164375 / 0.439375
    assert v[0] == int((1.055 * y ** (1 / 2.4) - 0.055) * 255 + 0.5)
    with pytest.raises(ValueError):
        O.write_image(p, a, w, tonemap="filmic")


# ---------------------------------------------------------------- Radiance HDR
def _read_rgbe(path):
    """Radiance RGBE reader from the format description (flat and new-style RLE scanlines)."""
    data = open(path, "rb").read()
    lines = []
    pos = 0
    while True:
        e = data.index(b"\n", pos)
        lines.append(data[pos:e].decode())
        pos = e + 1
        if lines[-1].startswith("-Y"):
            break
    _, h, _, w = lines[-1].split()
    h, w = int(h), int(w)
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w <= 0x7FFF:
            assert data[pos:pos + 2] == b"\x02\x02" and (data[pos + 2] << 8 | data[pos + 3]) == w
            pos += 4
            for ch in range(4):
                x = 0
                while x < w:
                    c = data[pos]
                    pos += 1
                    if c > 128:
                        out[y, x:x + c - 128, ch] = data[pos]
                        pos += 1
                        x += c - 128
                    else:
                        assert c > 0
                        out[y, x:x + c, ch] = np.frombuffer(data[pos:pos + c], np.uint8)
                        pos += c
                        x += c
        else:
            out[y] = np.frombuffer(data[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    assert pos == len(data)
    e = out[..., 3].astype(np.int32)
    f = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136))
    return lines, out[..., :3] * f[..., None]


def test_hdr_roundtrip_preserves_linear(tmp_path):
    px = np.array([[[0.25, 0.5, 1.0], [2.0, 0.1, 0.0]], [[0.0, 0.0, 0.0], [8.0, 4.0, 2.0]]])
    p = str(tmp_path / "a.hdr")
    O.write_image(p, px, np.ones((2, 2)))
    lines, back = _read_rgbe(p)
    assert lines == ["#?RADIANCE", "FORMAT=32-bit_rle_rgbe", "", "-Y 2 +X 2"]
    pmax = px.max(axis=2, keepdims=True)
    assert np.all(np.abs(back - px) <= pmax / 128 + 1e-6)


def test_hdr_run_length_scanlines(tmp_path):
    # 40 wide: new-style scanlines with runs (constant stretches) and literals (a ramp)
    rng = np.random.default_rng(3)
    h, w = 3, 40
    px = np.zeros((h, w, 3))
    px[0, :] = 0.75
    px[1, :] = rng.uniform(0.0, 50.0, (w, 3))
    px[2, :13] = 3.0
    px[2, 13:] = np.linspace(0.01, 1.0, w - 13)[:, None]
    p = str(tmp_path / "b.hdr")
    O.write_image(p, px, np.ones((h, w)))
    _, back = _read_rgbe(p)
    pmax = px.max(axis=2, keepdims=True)
    assert np.all(np.abs(back - px) <= pmax / 128 + 1e-6)
    # RGBE truncates the mantissa: values never round up
    assert np.all(back <= px + 1e-12)
    # the constant row is one 4-byte header plus one run per channel
    raw = open(p, "rb").read()
    header = raw.index(b"+X 40\n") + len(b"+X 40\n")
    assert raw[header:header + 4] == b"\x02\x02\x00\x28"
    assert raw[header + 4:header + 12] == bytes([128 + 40, 192, 128 + 40, 192, 128 + 40, 192, 128 + 40, 128])


# ---------------------------------------------------------------- OpenEXR (ACEScg)
def _read_exr(path):
    """Single-part uncompressed scanline OpenEXR reader (from the file-layout specification)."""
    d = open(path, "rb").read()
    magic, ver = struct.unpack_from("<II", d, 0)
    assert magic == 20000630 and ver == 2
    pos = 8
    attrs = {}
    while d[pos] != 0:
        e = d.index(b"\0", pos)
        name = d[pos:e].decode()
        e2 = d.index(b"\0", e + 1)
        typ = d[e + 1:e2].decode()
        (size,) = struct.unpack_from("<i", d, e2 + 1)
        attrs[name] = (typ, d[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    chl = attrs["channels"][1]
    chans = []
    q = 0
    while chl[q] != 0:
        e = chl.index(b"\0", q)
        nm = chl[q:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iiii", chl, e + 1)
        chans.append((nm, ptype, xs, ys))
        q = e + 17
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", d, pos)
    img = {}
    for y, off in enumerate(offsets):
        yy, size = struct.unpack_from("<ii", d, off)
        assert yy == y and size == 4 * w * len(chans)
        for c, (nm, _, _, _) in enumerate(chans):
            img.setdefault(nm, np.zeros((h, w), np.float32))[y] = np.frombuffer(d, np.float32, w, off + 8 + 4 * w * c)
    return attrs, chans, img


# sRGB (D65) -> ACEScg (D60), Bradford adaptation, as published by OCIO / colour-science
ACESCG_FROM_SRGB = np.array([[0.6130974, 0.3395231, 0.0473795],
                             [0.0701937, 0.9163539, 0.0134524],
                             [0.0206156, 0.1095698, 0.8698147]])


def test_exr_layout_and_acescg(tmp_path):
    px = np.array([[[1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
                   [[0.0, 0.0, 1.0], [0.3, 2.0, 7.5], [0.0, 0.0, 0.0]]])
    p = str(tmp_path / "a.exr")
    O.write_image(p, px * 2.0, np.full((2, 3), 2.0), exposure=3.0)  # HDR formats ignore exposure and tonemap
    attrs, chans, img = _read_exr(p)
    assert [c[0] for c in chans] == ["B", "G", "R"] and all(c[1:] == (2, 1, 1) for c in chans)  # FLOAT
    assert attrs["compression"][1] == b"\0" and attrs["lineOrder"][1] == b"\0"
    chroma = np.frombuffer(attrs["chromaticities"][1], np.float32)
    assert np.array_equal(chroma, np.array([0.713, 0.293, 0.165, 0.830, 0.128, 0.044, 0.32168, 0.33767], np.float32))
    assert attrs["name"][1] == b"ACEScg"
    got = np.stack([img["R"], img["G"], img["B"]], axis=-1).astype(np.float64)
    # white stays white (both white points map to RGB (1, 1, 1))
    assert np.allclose(got[0, 0], 1.0, atol=2e-5)
    # the primaries are the matrix's columns
    for k, (y, x) in enumerate([(0, 1), (0, 2), (1, 0)]):
        assert np.allclose(got[y, x], ACESCG_FROM_SRGB[:, k], atol=2e-6)
    # linear: any colour is the matrix applied to it
    assert np.allclose(got[1, 1], ACESCG_FROM_SRGB @ px[1, 1], rtol=2e-6, atol=1e-6)
    assert np.array_equal(got[1, 2], np.zeros(3))


# ---------------------------------------------------------------- checkpoints (checkpoint.rs)
def test_ckpt_path():
    assert O.ckpt_path(0x4859503000000001) == "checkpoint_4859503000000001.bin"
    assert O.ckpt_path(0xAB) == "checkpoint_00000000000000ab.bin"


def test_checkpoint_roundtrip_and_layout(tmp_path):
    rng = np.random.default_rng(9)
    h, w = 3, 5
    acc = rng.normal(size=(h, w, 3))
    acc_w = rng.integers(0, 9, size=(h, w)).astype(np.float64)
    p = str(tmp_path / "c.bin")
    O.save_checkpoint(p, 0x4859503000000001, 7, acc, acc_w)
    assert not os.path.exists(p + ".tmp")
    raw = open(p, "rb").read()
    # magic, u32 version, u64 scene hash, u32 w, u32 h, u64 next id, f64 acc (x, y, z), f64 acc_w
    assert raw[:8] == b"HYPCKPT\0"
    assert struct.unpack_from("<IQIIQ", raw, 8) == (1, 0x4859503000000001, w, h, 7)
    assert len(raw) == 36 + 8 * 4 * h * w
    assert np.array_equal(np.frombuffer(raw, "<f8", 3 * h * w, 36), acc.reshape(-1))
    assert np.array_equal(np.frombuffer(raw, "<f8", h * w, 36 + 24 * h * w), acc_w.reshape(-1))
    nid, a2, w2 = O.load_checkpoint(p, 0x4859503000000001, w, h)
    assert nid == 7 and np.array_equal(a2, acc) and np.array_equal(w2, acc_w)


def test_checkpoint_mismatch_is_none(tmp_path):
    p = str(tmp_path / "c.bin")
    acc, acc_w = np.ones((2, 2, 3)), np.ones((2, 2))
    assert O.load_checkpoint(p, 1, 2, 2) is None  # absent
    O.save_checkpoint(p, 1, 3, acc, acc_w)
    assert O.load_checkpoint(p, 2, 2, 2) is None  # scene hash
    assert O.load_checkpoint(p, 1, 3, 2) is None  # width
    assert O.load_checkpoint(p, 1, 2, 1) is None  # height
    raw = bytearray(open(p, "rb").read())
    raw[8] = 2  # version
    open(p, "wb").write(bytes(raw))
    assert O.load_checkpoint(p, 1, 2, 2) is None
    raw[0] = ord("X")  # magic
    open(p, "wb").write(bytes(raw))
    assert O.load_checkpoint(p, 1, 2, 2) is None


def test_checkpoint_truncated_raises(tmp_path):
    p = str(tmp_path / "c.bin")
    O.save_checkpoint(p, 1, 3, np.ones((2, 2, 3)), np.ones((2, 2)))
    raw = open(p, "rb").read()
    for cut in (4, 20, 40, len(raw) - 1):
        open(p, "wb").write(raw[:cut])
        with pytest.raises(T.TptError):
            O.load_checkpoint(p, 1, 2, 2)


# ---------------------------------------------------------------- tiles (render.rs:72-98, 282-298)
def _interleave_loop(v):
    r = 0
    for b in range(32):
        r |= ((v >> b) & 1) << (2 * b)
    return r


def test_morton_matches_bit_interleave():
    for tx, ty in [(0, 0), (1, 0), (0, 1), (3, 5), (1023, 77), (65535, 65535)]:
        assert O.morton2(tx, ty, 1 << 20, 1 << 20) == _interleave_loop(tx) | (_interleave_loop(ty) << 1)
    assert O.morton2(9, 9, 4, 3) == O.morton2(3, 2, 4, 3)  # clamped to the tile grid


def test_tile_tasks_cover_image_once():
    tasks = O.tile_tasks(40, 24, 16)
    assert tasks[:4] == [(0, 0, 16, 16), (16, 0, 32, 16), (0, 16, 16, 24), (16, 16, 32, 24)]
    assert tasks[4:] == [(32, 0, 40, 16), (32, 16, 40, 24)]
    pix = O.task_pixels(tasks, 40)
    assert np.array_equal(np.sort(pix), np.arange(40 * 24))
    assert O.tile_tasks(40, 24, 16, morton=False)[:3] == [(0, 0, 16, 16), (16, 0, 32, 16), (32, 0, 40, 16)]
    assert pix[:17].tolist() == list(range(16)) + [40]  # rows inside a tile


# ---------------------------------------------------------------- the checkpointed loop (host logic)
class _Stop(RuntimeError):
    pass


def _fake_render_pixels(calls, stop_at=None):
    """Stands in for the GPU call: pixel p's sum is (p, 2p, 3p) over spp samples."""
    def f(dscene, config, pixels, want_sq=False):
        calls.append(pixels.copy())
        if stop_at is not None and len(calls) == stop_at:
            raise _Stop()
        p = pixels.astype(np.float64)
        s = np.stack([p, 2 * p, 3 * p], axis=1)
        sw = np.full(p.size, float(config.spp))
        return s, sw, (s * s if want_sq else None), {"paths": int(p.size * config.spp)}
    return f


def test_checkpointed_loop_merges_and_saves(tmp_path, monkeypatch):
    w, h = 40, 24
    cfg = T.RenderConfig(width=w, height=h, spp=2, tile=8, checkpoint_enabled=True, checkpoint_every_tasks=4)
    ck = str(tmp_path / "ck.bin")
    calls = []
    monkeypatch.setattr(T, "_CKPT_BATCH_PATHS", 1)
    monkeypatch.setattr(T, "render_pixels", _fake_render_pixels(calls, stop_at=3))
    with pytest.raises(_Stop):
        T.render(object.__new__(T.DeviceScene), cfg, ckpt_file=ck)
    tasks = O.tile_tasks(w, h, 8)
    assert len(tasks) == 15 and len(calls) == 3
    assert np.array_equal(calls[0], O.task_pixels(tasks[0:4], w))
    assert np.array_equal(calls[1], O.task_pixels(tasks[4:8], w))
    nid, a, aw = O.load_checkpoint(ck, cfg.scene_hash, w, h)
    assert nid == 8
    done = O.task_pixels(tasks[:8], w)
    assert aw.reshape(-1)[done].tolist() == [2.0] * done.size and aw.sum() == 2.0 * done.size
    assert np.array_equal(a.reshape(-1, 3)[done, 1], 2.0 * done)
    # resume: only the remaining tasks are rendered, in task order, with a final checkpoint
    calls.clear()
    monkeypatch.setattr(T, "render_pixels", _fake_render_pixels(calls))
    out = T.render(object.__new__(T.DeviceScene), cfg, ckpt_file=ck)
    assert np.array_equal(np.concatenate(calls), O.task_pixels(tasks[8:], w))
    idx = np.arange(w * h, dtype=np.float64).reshape(h, w)
    assert np.array_equal(out.acc, np.stack([idx, 2 * idx, 3 * idx], axis=-1))
    assert np.array_equal(out.acc_w, np.full((h, w), 2.0)) and out.acc_sq is None
    assert O.load_checkpoint(ck, cfg.scene_hash, w, h)[0] == 15
    # batches of several intervals when the interval is small: one call for everything
    calls.clear()
    monkeypatch.setattr(T, "_CKPT_BATCH_PATHS", 1 << 22)
    os.remove(ck)
    out = T.render(object.__new__(T.DeviceScene), cfg, ckpt_file=ck)
    assert len(calls) == 1 and out.acc_sq is not None


def test_progress_lines(capsys, monkeypatch):
    """render.rs:101-105, 216-263: the Render line, Progress lines ending at 100%, Done."""
    w, h = 40, 24
    cfg = T.RenderConfig(width=w, height=h, spp=2, tile=8, seed=7)
    calls = []
    monkeypatch.setattr(T, "_TILE_BATCH_PATHS", 200)  # several batches
    monkeypatch.setattr(T, "render_pixels", _fake_render_pixels(calls))

    class _Ctx:
        device = 0

    ds = object.__new__(T.DeviceScene)
    ds.ctx = _Ctx()
    out = T.render(ds, cfg, progress=True)
    err = capsys.readouterr().err
    assert err.startswith("Render: 40x24, spp=2, tile=8, device=hip:0, morton=true, seed=7\n")
    assert "\rProgress: 0/15 tiles (0.0%)" in err
    assert "\rProgress: 15/15 tiles (100.0%)\n" in err
    assert err.rstrip().splitlines()[-1].startswith("Done in ")
    assert len(calls) > 1 and np.array_equal(out.acc_w, np.full((h, w), 2.0))
    idx = np.arange(w * h, dtype=np.float64).reshape(h, w)
    assert np.array_equal(out.acc[..., 2], 3 * idx)
