"""Host-side product logic on the CPU: the FP64 SAH builder against the
oracle's restatement of FlatBVH::build, the scene generators, and the C ABI
library's exports.  No GPU calls."""
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
import raycer_amd as ra
from raycer_amd import api

ROOT = Path(__file__).resolve().parent.parent


SMALL_SCENES = {
    "box": ra.scene_params(api.RCH_SCENE_BOX),
    "terrain96": ra.scene_params(api.RCH_SCENE_TERRAIN, seed=1, grid=96),
    "sphere4": ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1, level=4, grid=48),
    "soup": ra.scene_params(api.RCH_SCENE_SOUP, seed=5, tri_count=30000),
}


@pytest.mark.parametrize("name", list(SMALL_SCENES))
def test_bvh_matches_oracle_node_for_node(name):
    sd = ra.generate_scene(SMALL_SCENES[name])
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    ob = orc.BVH(sd.vertices)
    on = ob.nodes()
    assert len(on) == len(nodes)
    for f in on.dtype.names:
        assert np.array_equal(on[f], nodes[f]), f
    assert np.array_equal(ob.ordered(), ordered)


def test_bvh_independent_of_thread_count():
    sd = ra.generate_scene(SMALL_SCENES["soup"])
    a = ra.build_bvh(sd.vertices, threads=1)
    b = ra.build_bvh(sd.vertices, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_bvh_leaf_size_and_degenerate_input():
    # identical triangles: every partition fails and the middle fallback splits
    v = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (40, 1))
    nodes, ordered, _, depth = ra.build_bvh(v)
    ob = orc.BVH(v)
    assert np.array_equal(ob.nodes()["right_offset"], nodes["right_offset"])
    assert np.array_equal(ob.ordered(), ordered)
    leaves = nodes[nodes["right_offset"] == 0]
    assert leaves["primitive_count"].sum() == 40 and leaves["primitive_count"].max() <= 5
    n2, _, _, _ = ra.build_bvh(v, max_leaf=2)
    assert n2[n2["right_offset"] == 0]["primitive_count"].max() <= 2


def check_bvh_invariants(nodes, ordered, vertices, max_leaf):
    """Preorder flat BVH: every primitive exactly once, children inside their
    parent, leaves within max_leaf, right child after the left subtree."""
    ntri = vertices.shape[0]
    assert np.array_equal(np.sort(ordered), np.arange(ntri))
    tb = vertices.reshape(-1, 3, 3)
    lo, hi = tb.min(1).astype(np.float64), tb.max(1).astype(np.float64)
    covered = np.zeros(ntri, np.int64)
    stack = [(0, len(nodes))]
    while stack:
        i, end = stack.pop()
        n = nodes[i]
        s, c = int(n["start_offset"]), int(n["primitive_count"])
        ids = ordered[s:s + c]
        assert (lo[ids] >= n["aabb_min"] - 1e-12).all() and (hi[ids] <= n["aabb_max"] + 1e-12).all()
        if n["right_offset"] == 0:
            assert 1 <= c <= max_leaf
            covered[ids] += 1
            continue
        left, right = i + 1, i + int(n["right_offset"])
        assert left < right < end
        for ch in (left, right):
            assert (nodes[ch]["aabb_min"] >= n["aabb_min"]).all() and (nodes[ch]["aabb_max"] <= n["aabb_max"]).all()
        stack += [(left, right), (right, end)]
    assert (covered == 1).all()


@pytest.mark.parametrize("name", list(SMALL_SCENES))
def test_binned_sah_bvh_is_valid_and_cheaper(name):
    sd = ra.generate_scene(SMALL_SCENES[name])
    ref = ra.build_bvh(sd.vertices)
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    check_bvh_invariants(nodes, ordered, sd.vertices, 5)
    check_bvh_invariants(ref[0], ref[1], sd.vertices, 5)

    def sah(nd):
        e = nd["aabb_max"] - nd["aabb_min"]
        a = 2 * (e[:, 0] * e[:, 1] + e[:, 1] * e[:, 2] + e[:, 0] * e[:, 2])
        leaf = nd["right_offset"] == 0
        return (a[~leaf].sum() + (a[leaf] * nd["primitive_count"][leaf]).sum()) / a[0]

    assert sah(nodes) <= sah(ref[0]) * 1.001
    again = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED, threads=1)
    assert np.array_equal(again[0], nodes) and np.array_equal(again[1], ordered)


@pytest.mark.parametrize("method", [api.RCH_BVH_REFERENCE, api.RCH_BVH_BINNED])
def test_empty_scene_bvh(tmp_path, method):
    """No triangles: FlatBVH::build makes one leaf holding nothing, with
    AABB()'s box (min = DBL_MAX, max = lowest); the oracle builds the same."""
    v = np.zeros((0, 9), np.float32)
    nodes, ordered, _, depth = ra.build_bvh(v, method=method)
    assert len(nodes) == 1 and len(ordered) == 0 and depth == 1
    n = nodes[0]
    assert (n["right_offset"], n["start_offset"], n["primitive_count"]) == (0, 0, 0)
    assert np.all(n["aabb_min"] == np.finfo(np.float64).max) and np.all(n["aabb_max"] == np.finfo(np.float64).min)
    assert nodes.tobytes() == orc.BVH(v).nodes().tobytes()
    path = tmp_path / "empty.bvh"
    ra.save_bvh(path, v, nodes, ordered)
    n2, o2 = ra.load_bvh(path, v)
    assert n2.tobytes() == nodes.tobytes() and len(o2) == 0
    with pytest.raises(ra.RaycerError):
        ra.load_bvh(path, np.zeros((1, 9), np.float32))


def test_workload_triangle_counts():
    assert ra.describe_scene(ra.scene_params(api.RCH_SCENE_BOX)).ntri == 36
    assert ra.describe_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1)).ntri == 1_179_648
    assert ra.describe_scene(ra.scene_params(api.RCH_SCENE_TERRAIN, seed=1)).ntri == 2_093_058
    d = ra.describe_scene(ra.scene_params(api.RCH_SCENE_SOUP, seed=7))
    assert 1000 <= d.ntri <= 500_000 and d.cfg.width in (128, 256, 512, 1024) and d.cfg.spp == 32


def test_box_scene_geometry():
    sd = ra.generate_scene(SMALL_SCENES["box"])
    v = sd.vertices.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    c = v.mean(1)
    # walls (first 10 triangles) face the interior
    assert np.all(np.einsum("ij,ij->i", n[:10], np.array([0.5, 0.5, 0.5]) - c[:10]) > 0)
    # inner boxes (next 24) face outward from their own centres
    for k in range(2):
        s = slice(10 + 12 * k, 22 + 12 * k)
        assert np.all(np.einsum("ij,ij->i", n[s], c[s] - c[s].mean(0)) > 0)
    # light quad faces down with Ke (17, 12, 4)
    assert np.allclose(n[34:36], [0, -1, 0])
    light = sd.materials[sd.tri_material[34]]
    assert np.array_equal(light["emittance"], np.float32([17, 12, 4])) and light["flags"] & api.RC_MAT_EMISSIVE
    # red wall is the x=1 plane (left in the image)
    red = sd.tri_material[np.all(np.isclose(v[:, :, 0], 1.0), axis=1)]
    assert set(red.tolist()) == {1}
    assert np.allclose(sd.materials[1]["diffuse"], [0.65, 0.05, 0.05])


def test_scene_generation_is_deterministic():
    a = ra.generate_scene(SMALL_SCENES["sphere4"])
    b = ra.generate_scene(SMALL_SCENES["sphere4"])
    assert np.array_equal(a.vertices, b.vertices)
    c = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=2, level=4, grid=48))
    assert not np.array_equal(a.vertices, c.vertices)


def _declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|double|uint32_t|void)\s+(r[a-z_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) > 30
    L = api.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_create_without_gpu_reports_an_error():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(ra.RaycerError) as e:
        ra.Context(0)
    assert e.value.code in (api.RC_ERR_NO_DEVICE, api.RC_ERR_HIP)
    assert api.last_error()


def test_bvh_save_and_restore(tmp_path):
    """A saved BVH restores node-for-node for the same triangles and is refused
    for others (FlatBVH serialization and restore, FlatBVH.h:99-105)."""
    sd = ra.generate_scene(SMALL_SCENES["sphere4"])
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    path = tmp_path / "scene.bvh"
    ra.save_bvh(path, sd.vertices, nodes, ordered)
    n2, o2 = ra.load_bvh(path, sd.vertices)
    assert np.array_equal(n2, nodes) and np.array_equal(o2, ordered)
    moved = sd.vertices.copy()
    moved[0, 0] += 1e-3
    with pytest.raises(ra.RaycerError):
        ra.load_bvh(path, moved)
    with pytest.raises(ra.RaycerError):
        ra.load_bvh(path, sd.vertices[:-1])
    trunc = tmp_path / "cut.bvh"
    trunc.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(ra.RaycerError):
        ra.load_bvh(trunc, sd.vertices)
    (tmp_path / "junk.bvh").write_bytes(b"not a bvh at all, no magic here")
    with pytest.raises(ra.RaycerError):
        ra.load_bvh(tmp_path / "junk.bvh", sd.vertices)
