"""The product's host-side scene build (tinypt_amd/csrc/scene_host.cpp) against the oracle:
the mesh BVH must be node-for-node the reference's Bvh::build output (bvh.rs:51-242), the
light list/CDF World::build_lights' (world.rs:127-167), and the environment alias table must
realise exactly the cell probabilities of EnvMap::from_pixels' CDFs (env.rs:61-89)."""
import numpy as np
import pytest

import oracle
import tinypt_amd as T
from tinypt_amd import scenes
from tinypt_amd.scene import Camera, EnvMap, Lambert, DiffuseLight, Mesh, Instance, Transform, Scene, World


def _mesh_scene(tris, n_inst=1):
    w = World()
    w.meshes.append(Mesh(tris, 0))
    for i in range(n_inst):
        w.instances.append(Instance(0, Transform.translate((float(i), 0.0, 0.0))))
    cam = Camera.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40.0, 1.0)
    return Scene(cam, w, [Lambert((0.5, 0.5, 0.5))], None)


def _random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    return c + 0.05 * rng.normal(size=(n, 3, 3))


@pytest.mark.parametrize("tris", [
    scenes.unit_cube_tris(),                       # n = 12: n/2 splits, no SAH
    scenes.unit_disk_tris(),                       # n = 64: SAH bins
    _random_soup(15, 1),                           # just below the SAH threshold
    _random_soup(16, 2),                           # at the SAH threshold
    _random_soup(3000, 3),
    scenes.displaced_icosphere(3, 0.05, 7)[0],     # 1280 triangles
    np.repeat(_random_soup(1, 4), 40, axis=0),     # 40 identical triangles: zero centroid extent
])
def test_bvh_identical_to_oracle(tris):
    sc = _mesh_scene(tris)
    hs = T.HostScene(sc)
    ob = oracle.OracleScene(sc).bvh(0)
    pb = hs.bvh(0)
    for k in ("left", "right", "start", "count", "indices"):
        np.testing.assert_array_equal(pb[k], ob[k], err_msg=k)
    np.testing.assert_array_equal(pb["bbox"], ob["bbox"])  # bit-identical f64 boxes


def test_bvh_structural_invariants():
    tris = _random_soup(5000, 11)
    b = T.HostScene(_mesh_scene(tris)).bvh(0)
    n = len(b["left"])
    idx = b["indices"]
    assert sorted(idx.tolist()) == list(range(len(tris)))  # permutation
    leaf = (b["left"] == -1) & (b["right"] == -1)
    assert (b["count"][leaf] >= 1).all() and (b["count"][leaf] <= 4).all()
    assert (b["count"][~leaf] == 0).all()
    inner = np.where(~leaf)[0]
    assert (b["left"][inner] == inner + 1).all()  # pre-order: left child follows its parent
    covered = np.zeros(len(tris), int)
    for i in np.where(leaf)[0]:
        covered[b["start"][i]:b["start"][i] + b["count"][i]] += 1
    assert (covered == 1).all()
    lo = tris.min(axis=1) - 1e-9
    hi = tris.max(axis=1) + 1e-9
    for i in range(n):  # every node box contains its subtree's triangles
        if leaf[i]:
            t = idx[b["start"][i]:b["start"][i] + b["count"][i]]
            assert (b["bbox"][i, :3] <= lo[t]).all() and (b["bbox"][i, 3:] >= hi[t]).all()
        else:
            for c in (b["left"][i], b["right"][i]):
                assert (b["bbox"][i, :3] <= b["bbox"][c, :3]).all() and (b["bbox"][i, 3:] >= b["bbox"][c, 3:]).all()


def test_small_nodes_split_at_half_in_input_order():
    # n < 16 (bvh.rs:193-201): the range is cut at n/2 without sorting, so indices stay 0..n-1
    tris = _random_soup(12, 5)
    b = T.HostScene(_mesh_scene(tris)).bvh(0)
    np.testing.assert_array_equal(b["indices"], np.arange(12))
    assert b["count"][b["left"] == -1].tolist() == [3, 3, 3, 3]


@pytest.mark.parametrize("make", [lambda: scenes.cornell_box(8, 8), lambda: scenes.default_scene(8, 8),
                                  lambda: scenes.inverse_scene(scenes.inverse_params(), 8, 8)])
def test_lights_identical_to_oracle(make):
    sc = make()
    pl = T.HostScene(sc).lights()
    ol = oracle.OracleScene(sc).lights()
    np.testing.assert_array_equal(pl, ol)


def test_instance_override_emitter_lights_every_triangle():
    w = World()
    w.meshes.append(Mesh(scenes.unit_cube_tris(), 0))
    w.instances.append(Instance(0, Transform.translate((0, 1, 0)), 1))
    sc = Scene(Camera.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 1), w, [Lambert((1, 1, 1)), DiffuseLight((2, 3, 4))])
    pl = T.HostScene(sc).lights()
    assert len(pl) == 12
    np.testing.assert_array_equal(pl, oracle.OracleScene(sc).lights())


def _alias_probabilities(prob, alias):
    N = len(prob)
    p = prob.astype(np.float64).copy()
    np.add.at(p, alias, 1.0 - prob.astype(np.float64))
    return p / N


@pytest.mark.parametrize("wh", [(1, 1), (7, 5), (64, 32)])
def test_env_alias_table_realises_cdf_cell_probabilities(wh):
    w, h = wh
    rng = np.random.default_rng(w * 100 + h)
    img = np.exp(rng.normal(size=(h, w, 3)))
    if h > 1:
        img[h // 2, :, :] = 0.0  # a black row: zero-probability cells
    sc = Scene(Camera.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 1), World(), [], EnvMap(w, h, img))
    t = T.HostScene(sc).env_tables()
    o = oracle.OracleScene(sc)
    assert t["total"] == o.env_total()  # FP64 total weight, bit-identical
    # reference cell probabilities from its own row/column CDF construction
    lum = np.maximum(0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2], 0)
    y = np.arange(h)
    roww = np.maximum(np.cos(np.pi * y / h) - np.cos(np.pi * (y + 1) / h), 0)
    ref_p = (lum * roww[:, None]).reshape(-1) / t["total"]
    got = _alias_probabilities(t["prob"], t["alias"])
    np.testing.assert_allclose(got, ref_p, rtol=2e-6, atol=1e-12)
    assert (got[ref_p == 0] == 0).all()
    # pdf numerator (env.rs:180-181): pdf(dir) * sin(theta) for the cell
    np.testing.assert_allclose(t["cell_pdf"], ref_p * w * h / (2 * np.pi ** 2), rtol=1e-6)
