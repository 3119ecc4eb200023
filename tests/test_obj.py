"""OBJ/MTL loader (ModelLoader::readObjFile) on the reference's own mesh files,
held as fixtures under tests/golden/meshes (CPU only).

Pins: the reference's ModelLoaderTest (src/Tests/ModelLoaderTest.cpp:19-47:
12 triangles in each of cube1..cube5.obj); the Cornell box of TestScene3
(18 quads -> 36 triangles, the light's Ke 17 12 4, per-group materials);
vertices against an independent reading of the same text; the transform
against closed-form rotations.
"""
import numpy as np
import pytest

import raycer_amd as ra
from raycer_amd import api
from tests.conftest import ROOT

MESHES = ROOT / "tests" / "golden" / "meshes"


def parse_reference_number(tok: str) -> float:
    """StringUtils::parseDouble (src/Utils/StringUtils.cpp:70-100)."""
    i, sign, acc = 0, 1.0, 0.0
    if tok[:1] == "-":
        sign, i = -1.0, 1
    while i < len(tok) and tok[i].isdigit():
        acc = acc * 10 + (ord(tok[i]) - 48)
        i += 1
    if i < len(tok) and tok[i] == ".":
        k = 0.1
        i += 1
        while i < len(tok) and tok[i].isdigit():
            acc += (ord(tok[i]) - 48) * k
            k *= 0.1
            i += 1
    return sign * acc


def independent_read(path):
    """Vertices and fan-triangulated faces of an OBJ, read with Python's own
    float() and split()."""
    verts, tris = [], []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            verts.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            idx = []
            for ref in t[1:]:
                k = int(ref.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for j in range(2, len(idx)):
                tris.append([verts[idx[0]], verts[idx[j - 1]], verts[idx[j]]])
    return np.array(tris, np.float64).reshape(-1, 9)


@pytest.mark.parametrize("name", ["cube1", "cube2", "cube3", "cube4", "cube5"])
def test_reference_cube_files_have_12_triangles(name):
    v, m, mats = ra.load_obj(MESHES / f"{name}.obj")
    assert v.shape == (12, 9)
    ref = independent_read(MESHES / f"{name}.obj")
    # float() is correctly rounded, parseDouble accumulates: within an FP32 ulp
    np.testing.assert_allclose(v, ref, rtol=2 ** -23, atol=1e-7)


def test_cornell_box_of_test_scene_3():
    v, m, mats = ra.load_obj(MESHES / "cornellbox.obj")
    assert v.shape == (36, 9) and mats.shape == (9,)  # base + 8 MTL materials
    light = mats[8]
    np.testing.assert_array_equal(light["emittance"], np.float32([17, 12, 4]))
    assert light["flags"] & api.RC_MAT_EMISSIVE
    np.testing.assert_array_equal(light["diffuse"], np.float32([0.78, 0.78, 0.78]))
    red, green = mats[1], mats[2]  # leftWall, rightWall (MTL order)
    np.testing.assert_array_equal(red["diffuse"], np.float32([0.63, 0.065, 0.05]))
    np.testing.assert_array_equal(green["ambient"], np.float32([0.14, 0.45, 0.091]))
    assert all(not (mats[k]["flags"] & api.RC_MAT_EMISSIVE) for k in range(8))
    assert np.all(mats[1:]["shininess"] == 10.0)
    # groups in file order: floor, ceiling, backWall, rightWall, leftWall, shortBox, tallBox, light
    mtl_index = {"leftWall": 1, "rightWall": 2, "floor": 3, "ceiling": 4, "backWall": 5, "shortBox": 6,
                 "tallBox": 7, "light": 8}
    order = ["floor"] * 2 + ["ceiling"] * 2 + ["backWall"] * 2 + ["rightWall"] * 2 + ["leftWall"] * 2 + \
        ["shortBox"] * 12 + ["tallBox"] * 12 + ["light"] * 2
    np.testing.assert_array_equal(m, [mtl_index[g] for g in order])
    ref = independent_read(MESHES / "cornellbox.obj")
    np.testing.assert_allclose(v, ref, rtol=2 ** -23, atol=1e-7)
    # and bit-exactly the reference's own number parsing
    exact = []
    vs = []
    for line in open(MESHES / "cornellbox.obj"):
        t = line.split()
        if t and t[0] == "v":
            vs.append([parse_reference_number(x) for x in t[1:4]])
        elif t and t[0] == "f":
            idx = [int(r) for r in t[1:]]
            idx = [k - 1 if k > 0 else len(vs) + k for k in idx]
            for j in range(2, len(idx)):
                exact.append(vs[idx[0]] + vs[idx[j - 1]] + vs[idx[j]])
    np.testing.assert_array_equal(v, np.float32(exact))


def test_transform_and_face_forms(tmp_path):
    obj = tmp_path / "t.obj"
    obj.write_text(
        "# unit triangle and a quad with every reference form\n"
        "v 1 0 0\nv 0 1 0\nv 0 0 1\nv 1 1 0\n"
        "vt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n"
        "f 1 2 3\n"
        "f 1/1 2/2 3/3\n"
        "f 1//1 2//1 3//1\n"
        "f -4/1/1 -3/2/1 -1/3/1 -2/1/1\n"
        "f 1 2 9\n"          # out of range: skipped, as the reference does
        "f 1 2\n")           # too few vertices: skipped
    v, m, mats = ra.load_obj(obj)
    assert v.shape == (5, 9) and (m == 0).all() and mats.shape == (1,)
    np.testing.assert_array_equal(v[3], np.float32([1, 0, 0, 0, 1, 0, 1, 1, 0]))   # fan (0, 1, 2)
    np.testing.assert_array_equal(v[4], np.float32([1, 0, 0, 1, 1, 0, 0, 0, 1]))   # fan (0, 2, 3)
    # translate * rotateXYZ(pitch, yaw, roll) * scale; yaw 90 maps +x to -z
    t, _, _ = ra.load_obj(obj, scale=(2, 2, 2), rotate=(0, 90, 0), translate=(0, 0, 5))
    np.testing.assert_allclose(t[0, 0:3], [0, 0, 5 - 2], atol=1e-6)   # (1,0,0) -> (0,0,-2) + (0,0,5)
    np.testing.assert_allclose(t[0, 3:6], [0, 2, 5], atol=1e-6)
    np.testing.assert_allclose(t[0, 6:9], [2, 0, 5], atol=1e-6)       # (0,0,1) -> (2,0,0)
    # pitch 90 maps +y to +z, roll 90 maps +x to +y
    p, _, _ = ra.load_obj(obj, rotate=(90, 0, 0))
    np.testing.assert_allclose(p[0, 3:6], [0, 0, 1], atol=1e-6)
    r, _, _ = ra.load_obj(obj, rotate=(0, 0, 90))
    np.testing.assert_allclose(r[0, 0:3], [0, 1, 0], atol=1e-6)


def test_missing_files_raise(tmp_path):
    with pytest.raises(ra.RaycerError):
        ra.load_obj(tmp_path / "none.obj")
    obj = tmp_path / "m.obj"
    obj.write_text("mtllib missing.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(ra.RaycerError):
        ra.load_obj(obj)
    v, _, _ = ra.load_obj(obj, ignore_materials=True)
    assert v.shape == (1, 9)
    empty = tmp_path / "e.obj"
    empty.write_text("v 0 0 0\n")
    with pytest.raises(ra.RaycerError):
        ra.load_obj(empty)


def area(v):
    t = v.reshape(-1, 3, 3).astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


@pytest.mark.parametrize("name", ["cube1", "cube2", "cube3"])
def test_reference_ply_cubes(name):
    """ModelLoaderTest's PLY half (src/Tests/ModelLoaderTest.cpp:39-49): 12
    triangles each; cube3.ply is the binary little-endian one."""
    v, m, mats = ra.load_obj(MESHES / f"{name}.ply")
    assert v.shape == (12, 9) and (m == 0).all() and mats.shape == (1,)
    # the same 2x2x2 cube as cube1.obj: surface 24, same corner set
    assert abs(area(v).sum() - 24.0) < 1e-4
    corners = np.unique(np.round(v.reshape(-1, 3), 4), axis=0)
    ref = np.unique(np.round(ra.load_obj(MESHES / "cube1.obj")[0].reshape(-1, 3), 4), axis=0)
    np.testing.assert_array_equal(corners, ref)


def test_ply_errors(tmp_path):
    bad = tmp_path / "b.ply"
    bad.write_text("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
                   "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 7\n")
    with pytest.raises(ra.RaycerError):
        ra.load_obj(bad)  # index 7 out of range
    good = tmp_path / "g.ply"
    good.write_text(bad.read_text().replace("3 0 1 7", "4 0 1 2 0"))
    v, _, _ = ra.load_obj(good, translate=(0, 0, 1))
    assert v.shape == (2, 9)
    np.testing.assert_array_equal(v[0], np.float32([0, 0, 1, 1, 0, 1, 0, 1, 1]))
    (tmp_path / "n.ply").write_text("not a ply\n")
    with pytest.raises(ra.RaycerError):
        ra.load_obj(tmp_path / "n.ply")
