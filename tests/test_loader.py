"""Scene files: the Mitsuba-XML subset loader, OBJ and Radiance RGBE reading (CPU).

The first group restates the reference loader's own tests (mitsuba.rs:600-839); the
reference's envmap test needs sample/env.exr, which the reference does not ship, so a Radiance
.hdr map written here stands in for it.  The sample scenes (tests/golden/scenes, the reference's
sample/*.xml and cube.obj without comments) are then loaded and checked against scenes built
from the same parameters in Python, transform by transform, and the loaded Cornell box's
BVHs and light table against the oracle's.
"""
import ctypes as C
import math
import os
import struct

import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import _abi, scenes
from tinypt_amd.loader import load_obj_triangles, load_scene
from tinypt_amd.scene import Transform, color_from_srgb

SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def _cfg():
    c = T.RenderConfig()
    c.width, c.height = 16, 9
    return c


def _load(tmp_path, xml, name="scene.xml", warnings=None):
    p = tmp_path / name
    p.write_text(xml)
    return load_scene(str(p), _cfg(), warnings)


# ---------------------------------------------------------------- reference tests, restated
def test_parses_shapes_and_materials(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0">
      <sensor type="perspective">
        <float name="fov" value="40"/><string name="fov_axis" value="y"/>
        <transform name="to_world"><lookat origin="0,1.2,4" target="0,0.5,0" up="0,1,0"/></transform>
      </sensor>
      <shape type="sphere"><point name="center" x="-1" y="0.5" z="0"/><float name="radius" value="0.5"/>
        <bsdf type="diffuse"><srgb name="reflectance" value="0.8,0.3,0.3"/></bsdf></shape>
      <shape type="sphere"><point name="center" x="0" y="0.5" z="0"/><float name="radius" value="0.5"/>
        <bsdf type="roughconductor"><string name="distribution" value="ggx"/><float name="alpha" value="0.25"/>
          <rgb name="specular_reflectance" value="0.9,0.7,0.3"/></bsdf></shape>
      <shape type="sphere"><point name="center" x="0" y="3" z="-1"/><float name="radius" value="0.8"/>
        <emitter type="area"><rgb name="radiance" value="8,7,5"/></emitter></shape>
    </scene>""")
    assert len(sc.world.spheres) == 3 and len(sc.mats) == 3
    assert [m.kind for m in sc.mats] == [_abi.MAT_LAMBERT, _abi.MAT_GGX, _abi.MAT_DIFFUSE_LIGHT]
    assert abs(sc.mats[1].param[3] - 0.25) < 1e-12
    assert len(T.HostScene(sc).lights()) == 1  # the area emitter becomes a light


def test_srgb_tag_gamma_decodes_but_rgb_is_linear(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0">
      <shape type="sphere"><bsdf type="diffuse"><srgb name="reflectance" value="0.8,0.8,0.8"/></bsdf></shape>
      <shape type="sphere"><bsdf type="diffuse"><rgb name="reflectance" value="0.8,0.8,0.8"/></bsdf></shape>
    </scene>""")
    srgb, lin = sc.mats[0].param[0], sc.mats[1].param[0]
    assert abs(srgb - ((0.8 + 0.055) / 1.055) ** 2.4) < 1e-12
    assert abs(lin - 0.8) < 1e-12 and srgb < lin


def test_reads_film_sampler_integrator_into_config(tmp_path):
    p = tmp_path / "cfg.xml"
    p.write_text("""<scene version="3.0.0">
      <integrator type="path"><integer name="max_depth" value="12"/><integer name="rr_depth" value="5"/></integrator>
      <sampler type="independent"><integer name="sample_count" value="256"/></sampler>
      <film type="hdrfilm"><integer name="width" value="800"/><integer name="height" value="600"/></film>
      <shape type="sphere"><bsdf type="diffuse"/></shape>
    </scene>""")
    cfg = _cfg()
    load_scene(str(p), cfg)
    assert (cfg.width, cfg.height, cfg.spp, cfg.max_bounces, cfg.rr_start) == (800, 600, 256, 12, 5)


def test_parses_parametric_shapes(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0">
      <shape type="rectangle"><bsdf type="diffuse"/></shape>
      <shape type="cube"><bsdf type="diffuse"/></shape>
      <shape type="disk"><bsdf type="diffuse"/></shape>
    </scene>""")
    assert len(sc.world.meshes) == 3 and len(sc.world.instances) == 3
    assert [len(m.tris) for m in sc.world.meshes] == [2, 12, 64]
    np.testing.assert_array_equal(sc.world.meshes[0].tris, scenes.unit_rectangle_tris())
    np.testing.assert_array_equal(sc.world.meshes[1].tris, scenes.unit_cube_tris())
    np.testing.assert_array_equal(sc.world.meshes[2].tris, scenes.unit_disk_tris(64))


def test_parses_obj_mesh_with_transform(tmp_path):
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    sc = _load(tmp_path, """<scene version="3.0.0">
      <shape type="obj"><string name="filename" value="tri.obj"/>
        <transform name="to_world"><translate x="10" y="0" z="0"/></transform>
        <bsdf type="diffuse"><rgb name="reflectance" value="0.5,0.5,0.5"/></bsdf></shape>
    </scene>""")
    assert len(sc.world.meshes) == 1 and len(sc.world.instances) == 1 and len(sc.mats) == 1
    assert len(sc.world.meshes[0].tris) == 1
    p = sc.world.instances[0].xform.apply_point((0.0, 0.0, 0.0))
    assert np.linalg.norm(np.subtract(p, (10.0, 0.0, 0.0))) < 1e-9


def test_constant_emitter_sets_uniform_env(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0">
      <emitter type="constant"><rgb name="radiance" value="0.1,0.2,0.4"/></emitter>
      <shape type="sphere"><bsdf type="diffuse"/></shape></scene>""")
    assert (sc.env.width, sc.env.height) == (1, 1)
    np.testing.assert_allclose(sc.env.data[0, 0], (0.1, 0.2, 0.4), atol=1e-12)
    rgb, _ = oracle.OracleScene(sc).env_lookup(np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0]]))
    np.testing.assert_allclose(rgb, [[0.1, 0.2, 0.4]] * 2, atol=1e-9)


def test_no_env_emitter_defaults_to_black_background(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0"><shape type="sphere"><bsdf type="diffuse"/></shape></scene>""")
    assert sc.env is not None and (sc.env.width, sc.env.height) == (1, 1)
    assert not sc.env.data.any()


def test_emitter_scale_multiplies_radiance(tmp_path):
    sc = _load(tmp_path, """<scene version="3.0.0">
      <emitter type="constant"><rgb name="radiance" value="1,1,1"/><float name="scale" value="3"/></emitter>
    </scene>""")
    np.testing.assert_allclose(sc.env.data[0, 0], 3.0, atol=1e-9)


def test_unknown_bsdf_falls_back_to_diffuse(tmp_path):
    warns = []
    sc = _load(tmp_path, """<scene version="3.0.0">
      <integrator type="path"><integer name="max_depth" value="8"/></integrator>
      <shape type="sphere"><bsdf type="plastic"><rgb name="reflectance" value="0.2,0.4,0.6"/></bsdf></shape>
    </scene>""", warnings=warns)
    assert len(sc.mats) == 1 and sc.mats[0].kind == _abi.MAT_LAMBERT
    assert sc.mats[0].param[:3] == (0.2, 0.4, 0.6)
    assert any("plastic" in w for w in warns)


# ---------------------------------------------------------------- environment files
def _rgbe(c):
    m = max(c)
    if m < 1e-32:
        return (0, 0, 0, 0)
    e = math.floor(math.log2(m)) + 1
    f = 2.0 ** (e - 8)
    return tuple(int(min(max(x / f, 0), 255)) for x in c) + (e + 128,)


def _write_hdr(path, img, rle):
    h, w, _ = img.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        px = [_rgbe(img[y, x]) for x in range(w)]
        if rle:
            out += bytes([2, 2, w >> 8, w & 255])
            for ch in range(4):
                vals = [p[ch] for p in px]
                x = 0
                while x < w:  # one run of up to 127 equal bytes, else a literal of 1
                    n = 1
                    while x + n < w and n < 127 and vals[x + n] == vals[x]:
                        n += 1
                    if n >= 2:
                        out += bytes([128 + n, vals[x]])
                    else:
                        out += bytes([1, vals[x]])
                    x += n
        else:
            for p in px:
                out += bytes(p)
    path.write_bytes(bytes(out))
    dec = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            r, g, b, e = _rgbe(img[y, x])
            dec[y, x] = (0, 0, 0) if e == 0 else np.array([r, g, b], float) * 2.0 ** (e - 136)
    return dec


@pytest.mark.parametrize("rle,w", [(True, 16), (False, 16), (False, 5)])
def test_envmap_hdr_file_loads(tmp_path, rle, w):
    rng = np.random.default_rng(w)
    img = rng.lognormal(0.0, 1.5, (8, w, 3))
    img[0, 0] = 0.0
    img[2, :5] = (3.0, 2.0, 1.0)  # runs
    dec = _write_hdr(tmp_path / "env.hdr", img, rle)
    sc = _load(tmp_path, """<scene version="3.0.0">
      <emitter type="envmap"><string name="filename" value="env.hdr"/><float name="scale" value="2"/></emitter>
    </scene>""")
    assert (sc.env.width, sc.env.height) == (w, 8)
    np.testing.assert_array_equal(sc.env.data, dec * 2.0)


@pytest.mark.parametrize("res", ["-Y 2251799813685248 +X 8192", "-Y 4294967296 +X 4294967296",
                                 "-Y 0 +X 16", "-Y 16 +X 0", "-Y 16385 +X 16385"])
def test_envmap_hdr_malformed_size_is_rejected(tmp_path, res):
    """Resolution lines whose texel count is zero, above 2^28 or wraps 64 bits: the map is
    refused with a warning (the scene loads with the black default) instead of being read."""
    (tmp_path / "env.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + res.encode() + b"\n" + bytes(64))
    warns = []
    sc = _load(tmp_path, """<scene version="3.0.0">
      <emitter type="envmap"><string name="filename" value="env.hdr"/></emitter></scene>""", warnings=warns)
    assert not sc.env.data.any() and any("env.hdr" in w for w in warns)


def test_exr_envmap_is_skipped_with_a_warning(tmp_path):
    warns = []
    sc = _load(tmp_path, """<scene version="3.0.0">
      <emitter type="envmap"><string name="filename" value="env.exr"/></emitter></scene>""", warnings=warns)
    assert not sc.env.data.any() and any("env.exr" in w for w in warns)


# ---------------------------------------------------------------- OBJ
def test_obj_polygons_negative_and_invalid_indices(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# c\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1\nf -4 -3 -1\n"
                 "f 1 2 9\nv 2\nf 5 1 2\n")
    t = load_obj_triangles(str(p))
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0]], float)
    expect = [v[[0, 1, 2]], v[[0, 2, 3]], v[[0, 1, 3]], v[[4, 0, 1]]]  # "f 1 2 9" keeps 2 indices: dropped
    np.testing.assert_array_equal(t, np.array(expect))


def test_errors(tmp_path):
    with pytest.raises(T.TptError):
        load_scene(str(tmp_path / "missing.xml"), _cfg())
    for bad in ("<scene><shape></scene>", "<notscene/>", "", "<scene a='1' a='2'/>", "<scene x='&bogus;'/>"):
        p = tmp_path / "bad.xml"
        p.write_text(bad)
        with pytest.raises(T.TptError):
            load_scene(str(p), _cfg())
    with pytest.raises(T.TptError):
        load_obj_triangles(str(tmp_path / "missing.obj"))
    # a buffer smaller than the file's triangle count is refused, not overrun
    p = tmp_path / "quad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    lib = T.load_library()
    n = C.c_int64(1)
    buf = np.full(3 * 9, 7.0)
    rc = lib.tpt_obj_load(str(p).encode(), C.byref(n), buf.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc != 0 and n.value == 2 and (buf == 7.0).all()
    n = C.c_int64(3)
    assert lib.tpt_obj_load(str(p).encode(), C.byref(n), buf.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert n.value == 2 and (buf[18:] == 7.0).all() and not (buf[:18] == 7.0).all()
    # no <film> and a zero film size in the configuration: the camera would get a 0/0 aspect ratio
    p = tmp_path / "nofilm.xml"
    p.write_text("<scene><sensor type='perspective'/></scene>")
    c = _cfg()
    c.width = 0
    with pytest.raises(T.TptError, match="film width and height"):
        load_scene(str(p), c)
    p.write_text("<scene><film type='hdrfilm'><integer name='width' value='8'/></film></scene>")
    assert load_scene(str(p), c).cam.aspect == 8 / 9  # <film> supplies the missing width


# ---------------------------------------------------------------- the sample scenes
def _same_xform(a: Transform, b: Transform):
    np.testing.assert_allclose(np.array(a.a), np.array(b.a), rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.array(a.t), np.array(b.t), rtol=0, atol=1e-15)


def test_sample_cornell():
    cfg = T.RenderConfig()
    sc = load_scene(os.path.join(SCENES, "cornell.xml"), cfg)
    assert (cfg.width, cfg.height, cfg.spp, cfg.max_bounces, cfg.rr_start) == (600, 600, 256, 24, 5)
    assert len(sc.world.instances) == 8 and len(sc.mats) == 8 and not sc.world.spheres
    assert sc.mats[7].kind == _abi.MAT_DIFFUSE_LIGHT and sc.mats[7].param[:3] == (4.6, 3.9, 2.0)
    assert sc.mats[3].param[:3] == (0.63, 0.065, 0.05)
    Tr, R, S = Transform.translate, Transform.rotate, Transform.scale
    expect = [
        [R((1, 0, 0), -90.0)],
        [Tr((0, 2, 0)), R((1, 0, 0), 90.0)],
        [Tr((0, 1, -1))],
        [Tr((-1, 1, 0)), R((0, 1, 0), 90.0)],
        [Tr((1, 1, 0)), R((0, 1, 0), -90.0)],
        [Tr((-0.35, 0.6, -0.2)), R((0, 1, 0), 18.0), S((0.3, 0.6, 0.3))],
        [Tr((0.35, 0.3, 0.35)), R((0, 1, 0), -18.0), S((0.3, 0.3, 0.3))],
        [Tr((0, 1.98, 0)), R((1, 0, 0), 90.0), S((0.35, 0.35, 1.0))],
    ]
    for inst, ops in zip(sc.world.instances, expect):
        _same_xform(inst.xform, scenes.compose_ops(ops))
    assert sc.cam.vfov_deg == 40.0 and sc.cam.eye == (0.0, 1.0, 3.9) and sc.cam.aspect == 1.0
    # the loaded scene builds the same BVHs and lights as the oracle's reading of it
    hs, orc = T.HostScene(sc), oracle.OracleScene(sc)
    for m in range(len(sc.world.meshes)):
        a, b = hs.bvh(m), orc.bvh(m)
        for k in ("left", "right", "start", "count", "indices"):
            np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(a["bbox"], b["bbox"])
    np.testing.assert_array_equal(hs.lights(), orc.lights())


def test_sample_default_matches_builtin_default_scene():
    cfg = T.RenderConfig()
    sc = load_scene(os.path.join(SCENES, "default.xml"), cfg)
    ref = scenes.default_scene(cfg.width, cfg.height, "black")
    assert len(sc.world.spheres) == len(ref.world.spheres) == 6
    for a, b in zip(sc.world.spheres, ref.world.spheres):
        assert a.c == b.c and a.r == b.r
    for a, b in zip(sc.mats, ref.mats):
        assert a.kind == b.kind
        np.testing.assert_allclose(a.param, b.param, rtol=1e-15, atol=1e-15)
    assert (cfg.width, cfg.height, cfg.spp, cfg.max_bounces, cfg.rr_start) == (1920, 1080, 512, 8, 3)


def test_sample_mesh_reads_obj_and_transforms():
    sc = load_scene(os.path.join(SCENES, "mesh.xml"), T.RenderConfig())
    assert len(sc.world.meshes) == 2 and len(sc.world.spheres) == 2
    cube = load_obj_triangles(os.path.join(SCENES, "cube.obj"))
    assert cube.shape == (12, 3, 3)
    np.testing.assert_array_equal(sc.world.meshes[0].tris, cube)
    Tr, R, S = Transform.translate, Transform.rotate, Transform.scale
    _same_xform(sc.world.instances[0].xform, scenes.compose_ops([Tr((-0.9, 0.4, 0)), R((0, 1, 0), 30.0), S((0.8,) * 3)]))
    _same_xform(sc.world.instances[1].xform,
                scenes.compose_ops([Tr((0.9, 0.5, 0)), R((0.3, 1, 0.2), 35.0), S((0.6, 1.0, 0.6))]))
    assert sc.mats[2].kind == _abi.MAT_GGX and abs(sc.mats[2].param[3] - 0.2) < 1e-15
    np.testing.assert_allclose(sc.mats[2].param[:3], color_from_srgb(0.95, 0.8, 0.4), rtol=1e-15)


def test_sample_env_scene_without_its_exr_falls_back_to_black():
    warns = []
    sc = load_scene(os.path.join(SCENES, "env_scene.xml"), T.RenderConfig(), warns)
    assert len(sc.world.spheres) == 4 and not sc.env.data.any()
    assert any("env.exr" in w for w in warns)
