"""Scene files (Scene::loadFromFile / saveToFile, src/Raytracing/Scene.cpp:67-183), CPU only.

Pins: the reference's own data/scenes/sample.xml (a fixture under
tests/golden/scenes) read field by field against its text -- camera (8, 8, 11)
pitch -30, CMJ + Mitchell multi-sampling, the linear tone mapper, eight
materials with ids 1..8, the 1.5-intensity point light at (8, 10, 5), ambient
0.01 -- and its non-triangle primitives counted as skipped; XML and JSON
round trips of a generated scene reproduce every array bit for bit, including
a built BVH, which loads back as the same tree.

The binary form (cereal's BinaryOutputArchive, Scene.cpp:88-90,144-146) is
pinned two ways: tests/golden/cereal/cereal_binary_rules.bin, written by the
cereal release the reference vendors (make_binary_rules.cpp there), is decoded
by the encoding rules the library's schema uses (int32 enums, one-byte bools,
uint64 counts, C arrays without a count, base-class members inline, no
version words); and a saved binary scene has exactly the size its member
lists give (Scene.h, Material.h, Lights.h, Triangle.h, FlatBVH.h).
"""
import json
import struct
import ctypes as C

import numpy as np
import pytest

import raycer_amd as ra
from raycer_amd import api
from tests.conftest import ROOT

SAMPLE = ROOT / "tests" / "golden" / "scenes" / "sample.xml"


def test_reference_sample_scene():
    sf = api.load_scene_file(SAMPLE)
    assert sf.tracer_type == 0                                    # TracerType::RAY
    assert list(sf.camera.position) == [8.0, 8.0, 11.0]
    assert (sf.camera.pitch, sf.camera.yaw, sf.camera.roll, sf.camera.fov) == (-30.0, 0.0, 0.0, 75.0)
    assert sf.cfg.sampler == api.RC_SAMPLER_CMJ and sf.cfg.filter == api.RC_FILTER_MITCHELL and sf.cfg.spp == 1
    assert sf.cfg.max_path_length == 3 and sf.cfg.ray_start_offset == np.float32(1e-5)
    assert sf.ray.max_ray_iterations == 3 and sf.ray.visualize_depth == 0
    assert sf.tonemap.type == api.RC_TONEMAP_LINEAR and sf.tonemap.apply_gamma and sf.tonemap.clamp
    assert sf.tonemap.gamma == np.float32(2.2) and sf.tonemap.key == np.float32(0.18)
    assert list(sf.material_ids) == list(range(1, 9))
    m0 = sf.materials[0]
    assert np.all(m0["diffuse"] == np.float32(0.3)) and np.all(m0["ambient"] == np.float32(0.3))
    assert m0["shininess"] == 2.0 and m0["flags"] == 0
    r0 = sf.ray_materials[0]
    assert r0.refl_glossiness == 1000.0 and r0.refl_sampler == api.RC_SAMPLER_CMJ and r0.refractive_index == 1.0
    assert sf.ray.ambient_intensity == np.float32(0.01) and list(sf.ray.ambient_color) == [1.0, 1.0, 1.0]
    assert len(sf.lights) == 1
    L = sf.lights[0]
    assert L.type == api.RC_LIGHT_POINT and list(L.position) == [8.0, 10.0, 5.0]
    assert (L.intensity, L.max_distance, L.attenuation, L.area_light) == (1.5, 20.0, 1.0, 0)
    # one invisible triangle + 1 plane, 2 spheres, 1 box, 1 cylinder, 1 torus, 1 blob, 7 instances
    assert len(sf.vertices) == 0 and sf.skipped_primitives == 15
    assert sf.nodes is None  # the file's tree indexes the other primitives
    assert sf.models == []


def generated_scene_file(with_bvh=True):
    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    rm = api.ray_materials(len(sd.materials), ray_reflectance=0.25, refl_n=2, refl_sampler=api.RC_SAMPLER_JITTERED)
    rm[0].fresnel_reflection, rm[0].refractive_index, rm[0].ray_transmittance = 1, 1.5, 0.5
    rm[0].attenuation_color[:] = (0.1, 0.2, 0.3)
    lights = [api.light(api.RC_LIGHT_DIRECTIONAL, direction=(0.0, -1.0, 0.5), intensity=0.7),
              api.light(api.RC_LIGHT_POINT, position=(0.5, 0.9, 0.5), area_light=1, area_n=3, area_radius=0.2),
              api.light(api.RC_LIGHT_SPOT, position=(0.1, 0.95, 0.2), direction=(0, -1, 0), angle=30.0,
                        side_attenuation=2.0)]
    tm = api.ToneMap(0.25, 2.0, 1, 1, 1, api.RC_TONEMAP_REINHARD, 0.5, 1, 0.2, 0)
    rs = api.ray_settings(max_ray_iterations=5, background=(0.1, 0.2, 0.3), ambient_intensity=0.05, ao_enabled=1,
                          ao_n=3, fog_enabled=1, fog_distance=30.0, fog_height_dispersion=1, fog_height=2.0)
    cfg = api.RenderCfg.from_buffer_copy(sd.desc.cfg)
    cfg.sampler, cfg.filter, cfg.filter_radius, cfg.spp = api.RC_SAMPLER_CMJ, api.RC_FILTER_TENT, 1.0, 16
    info = api.ObjInfo()
    lib = api.lib()
    lib.rch_obj_default_info(C.byref(info))
    info.scale[:] = (2.0, 2.0, 2.0)
    info.rotate[:] = (0.0, 90.0, 0.0)
    sf = api.SceneFile(
        tracer_type=1, camera=sd.desc.camera, cfg=cfg, tonemap=tm, ray=rs, materials=sd.materials.copy(),
        ray_materials=rm, material_ids=np.arange(10, 10 + len(sd.materials), dtype=np.uint64),
        vertices=sd.vertices.copy(), tri_material=sd.tri_material.copy(),
        tri_ids=np.arange(100, 100 + sd.ntri, dtype=np.uint64), lights=lights,
        models=[], nodes=nodes if with_bvh else None, ordered=ordered if with_bvh else None, skipped_primitives=0)
    return sf, info


def struct_eq(a, b):
    return bytes(a) == bytes(b)


def light_fields(L):
    """The members Lights.h serializes for the light's type."""
    f = [L.type, *L.color, L.intensity]
    if L.type == api.RC_LIGHT_DIRECTIONAL:
        return f + list(L.direction)
    f += [*L.position, L.max_distance, L.attenuation, L.area_light, L.area_sampler, L.area_n, L.area_radius]
    if L.type == api.RC_LIGHT_SPOT:
        f += [*L.direction, L.angle, L.side_attenuation]
    return f


@pytest.mark.parametrize("ext", ["xml", "json", "bin"])
def test_round_trip(tmp_path, ext):
    sf, _ = generated_scene_file()
    path = tmp_path / f"scene.{ext}"
    api.save_scene_file(path, sf)
    back = api.load_scene_file(path)
    assert back.tracer_type == 1
    np.testing.assert_array_equal(back.vertices, sf.vertices)
    np.testing.assert_array_equal(back.tri_material, sf.tri_material)
    np.testing.assert_array_equal(back.tri_ids, sf.tri_ids)
    np.testing.assert_array_equal(back.material_ids, sf.material_ids)
    assert back.materials.tobytes() == sf.materials.tobytes()
    assert bytes(back.ray_materials) == bytes(sf.ray_materials)
    assert struct_eq(back.camera, sf.camera)
    for f in ("sampler", "filter", "filter_radius", "spp", "max_path_length", "ray_start_offset"):
        assert getattr(back.cfg, f) == getattr(sf.cfg, f), f
    assert struct_eq(back.tonemap, sf.tonemap)
    assert struct_eq(back.ray, sf.ray)
    assert [light_fields(x) for x in back.lights] == [light_fields(x) for x in sf.lights]
    assert back.nodes.tobytes() == sf.nodes.tobytes()
    np.testing.assert_array_equal(back.ordered, sf.ordered)
    assert back.skipped_primitives == 0


def test_json_and_xml_hold_the_same_scene(tmp_path):
    sf, _ = generated_scene_file(with_bvh=False)
    api.save_scene_file(tmp_path / "a.xml", sf)
    api.save_scene_file(tmp_path / "a.json", sf)
    x = api.load_scene_file(tmp_path / "a.xml")
    j = api.load_scene_file(tmp_path / "a.json")
    assert x.vertices.tobytes() == j.vertices.tobytes() and x.materials.tobytes() == j.materials.tobytes()
    assert struct_eq(x.ray, j.ray) and struct_eq(x.tonemap, j.tonemap) and x.nodes is None and j.nodes is None
    text = (tmp_path / "a.json").read_text()
    assert text.lstrip().startswith('{') and '"scene"' in text and '"flatNodes": []' in text
    assert (tmp_path / "a.xml").read_text().startswith('<?xml version="1.0" encoding="utf-8"?>\n<cereal>')


def test_binary_holds_the_same_scene_as_xml(tmp_path):
    sf, _ = generated_scene_file()
    sf.models = [("meshes/cornellbox.obj", generated_scene_file()[1], 1)]
    for ext in ("xml", "bin"):
        api.save_scene_file(tmp_path / f"a.{ext}", sf)
    x = api.load_scene_file(tmp_path / "a.xml")
    b = api.load_scene_file(tmp_path / "a.bin")
    for f in ("vertices", "tri_material", "tri_ids", "material_ids", "materials", "nodes", "ordered"):
        assert getattr(x, f).tobytes() == getattr(b, f).tobytes(), f
    assert bytes(x.ray_materials) == bytes(b.ray_materials)
    assert struct_eq(x.ray, b.ray) and struct_eq(x.tonemap, b.tonemap) and struct_eq(x.camera, b.camera)
    assert [light_fields(q) for q in x.lights] == [light_fields(q) for q in b.lights]
    assert x.models[0][0] == b.models[0][0] and struct_eq(x.models[0][1], b.models[0][1])
    # the sample scene through the binary form
    sample = api.load_scene_file(SAMPLE)
    api.save_scene_file(tmp_path / "sample.bin", sample)
    sb = api.load_scene_file(tmp_path / "sample.bin")
    assert sb.materials.tobytes() == sample.materials.tobytes() and struct_eq(sb.camera, sample.camera)


# Encoded sizes (bytes) of the serialized types, from the member lists of
# their serialize functions under the rules the fixture pins.
COLOR, VEC3, VEC2, EULER = 32, 24, 16, 24
MATERIAL = 8 + 6 + VEC2 + COLOR + 8 + COLOR + 8 + COLOR + 8 + 8 + COLOR + 8 + 1 + 8 + 8 + 8 + 4 + 8 + 8 + 8 + 8 + 4 \
    + 8 + 8 + 1 + 8 + COLOR + 8 + 8 + 8                                        # Material.h:80-120
GENERAL = 4 + 8 * 3 + 8 + 2 * COLOR + 4 + 4 + 8 + 4 + 8 + 4 + 8 + 1 + 8 + 1     # Scene.h:108-128
CAMERA = VEC3 + EULER + 4 + 5 * 8 + 1 + VEC3 + EULER                            # Camera.h:88-103
TONE, SFOG, VFOG = 4 + 2 + 4 * 8 + 1, 1 + COLOR + 2 * 8 + 1 + 2 * 8, 1 + COLOR + 8 + 8
BUILD, PRIM, NODE = 8 + 1 + 8 + 4 + 4, 8 + 8 + 1, 2 * VEC3 + 3 * 8               # FlatBVH.h, Primitive.h
LIGHT = COLOR + 8                                                               # Lights.h:14-24
AMBIENT, DIRECTIONAL = LIGHT + 1 + 4 + 8 + 8 + 8, LIGHT + VEC3
POINT = LIGHT + VEC3 + 8 + 8 + 1 + 4 + 8 + 8
SPOT = POINT + VEC3 + 8 + 8
TRIANGLE = PRIM + 3 * VEC3 + 3 * VEC3 + 3 * VEC2                                # Triangle.h:45-51
MODEL_FIXED = 8 + VEC3 + EULER + VEC3 + MATERIAL + 6 + 4 * 8                    # ModelLoader.h:66-84


def test_cereal_binary_rules_fixture():
    raw = (ROOT / "tests" / "golden" / "cereal" / "cereal_binary_rules.bin").read_bytes()
    want = {f[0]: f[2] for f in json.loads((ROOT / "tests" / "golden" / "cereal" / "cereal_binary_rules.json")
                                            .read_text())["fields"]}
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, raw, o)
        o += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    assert take("i") == want["kind"]                      # enum class: int32
    assert take("B") == 1 and want["flag"] is True        # bool: one byte
    assert take("q") == want["offset"] and take("Q") == want["count"] and take("d") == want["value"]
    n = take("Q")                                         # std::string: uint64 count, bytes
    assert raw[o:o + n].decode() == want["path"]
    o += n
    assert take("Q") == len(want["tris"])                 # std::vector: uint64 count
    for tid, mid, inv, verts in want["tris"]:
        assert take("QQB") == (tid, mid, int(inv))        # base class members inline, no version word
        for v in verts:                                   # C array: no count
            assert list(take("ddd")) == [float(x) for x in v]
    assert take("Q") == 3 and [take("Q") for _ in range(3)] == want["ids"]
    assert take("Q") == 0 and o == len(raw)


def test_binary_scene_size_follows_the_member_lists(tmp_path):
    sf, _ = generated_scene_file()
    sf.models = [("meshes/box.obj", generated_scene_file()[1], 0)]
    api.save_scene_file(tmp_path / "s.bin", sf)
    size = (tmp_path / "s.bin").stat().st_size
    ndir = sum(1 for L in sf.lights if L.type == api.RC_LIGHT_DIRECTIONAL)
    npt = sum(1 for L in sf.lights if L.type == api.RC_LIGHT_POINT)
    nspot = sum(1 for L in sf.lights if L.type == api.RC_LIGHT_SPOT)
    want = (GENERAL + CAMERA + TONE + SFOG + VFOG
            + 1 + BUILD + PRIM + 1 + 8 + NODE * len(sf.nodes) + 8 + 8 * len(sf.ordered)   # rootBVH
            + 2 + MATERIAL                                                                # boundingBoxes
            + 12 * 8                                                                      # textures (empty)
            + 8 + MATERIAL * len(sf.materials)
            + AMBIENT + 8 + DIRECTIONAL * ndir + 8 + POINT * npt + 8 + SPOT * nspot
            + 8 + MODEL_FIXED + len("meshes/box.obj")
            + 8 + TRIANGLE * len(sf.tri_material) + 9 * 8)                                # primitives
    assert size == want


def test_binary_scene_errors(tmp_path):
    sf, _ = generated_scene_file()
    api.save_scene_file(tmp_path / "s.bin", sf)
    raw = (tmp_path / "s.bin").read_bytes()
    assert api.parse_scene_file(raw, binary=True).vertices.shape == sf.vertices.shape
    # the last nine counts are the non-triangle primitive lists: a plane is rejected by name
    planes = bytearray(raw)
    planes[-72:-64] = struct.pack("<Q", 1)
    with pytest.raises(ra.RaycerError, match="planes are not supported"):
        api.parse_scene_file(bytes(planes), binary=True)
    with pytest.raises(ra.RaycerError, match="ends early|runs past"):
        api.parse_scene_file(raw[:-9], binary=True)
    with pytest.raises(ra.RaycerError, match="bytes after the scene"):
        api.parse_scene_file(raw + b"\0", binary=True)
    bad_bool = bytearray(raw)
    bad_bool[GENERAL - 1] = 7                             # enableNormalMapping
    with pytest.raises(ra.RaycerError, match="bool"):
        api.parse_scene_file(bytes(bad_bool), binary=True)
    huge = bytearray(raw)
    off = GENERAL + CAMERA + TONE + SFOG + VFOG + 1 + BUILD + PRIM + 1            # flatNodes count
    huge[off:off + 8] = struct.pack("<Q", 1 << 60)
    with pytest.raises(ra.RaycerError, match="runs past the end"):
        api.parse_scene_file(bytes(huge), binary=True)


def test_models_round_trip(tmp_path):
    sf, info = generated_scene_file(with_bvh=False)
    sf.models = [("meshes/cornellbox.obj", info, 0)]
    api.save_scene_file(tmp_path / "m.xml", sf)
    back = api.load_scene_file(tmp_path / "m.xml")
    (path, binfo, inv) = back.models[0]
    assert path == "meshes/cornellbox.obj" and inv == 0
    assert list(binfo.scale) == [2.0, 2.0, 2.0] and list(binfo.rotate) == [0.0, 90.0, 0.0]


def test_scene_file_errors(tmp_path):
    with pytest.raises(ra.RaycerError):
        api.load_scene_file(tmp_path / "none.xml")
    with pytest.raises(ra.RaycerError):
        api.load_scene_file(SAMPLE.with_suffix(".bin"))
    bad = SAMPLE.read_text().replace("<materialId>2</materialId>\n\t\t\t\t\t\t<invisible>true",
                                     "<materialId>99</materialId>\n\t\t\t\t\t\t<invisible>false")
    with pytest.raises(ra.RaycerError, match="material id 99"):
        api.parse_scene_file(bad, json=False)
    with pytest.raises(ra.RaycerError):
        api.parse_scene_file(SAMPLE.read_text()[:5000], json=False)   # truncated
    with pytest.raises(ra.RaycerError):
        api.parse_scene_file('{"scene": {"general": [1, 2', json=True)
    # the now-visible triangle of the edited sample loads when its material exists
    ok = SAMPLE.read_text().replace("<invisible>true</invisible>\n\t\t\t\t\t</primitive>\n\t\t\t\t\t<vertices>",
                                    "<invisible>false</invisible>\n\t\t\t\t\t</primitive>\n\t\t\t\t\t<vertices>")
    sf = api.parse_scene_file(ok, json=False)
    assert sf.vertices.shape == (1, 9) and sf.skipped_primitives == 14
    np.testing.assert_array_equal(sf.vertices[0], np.float32([-1, 0, 0, 1, 0, 0, 0, 2, 0]))
    assert sf.material_ids[sf.tri_material[0]] == 2


def test_cpp_wrapper_rethrows_c_abi_errors(tmp_path):
    """The C++ interface turns a nonzero C-ABI return into raycer::RaycerError,
    as the reference's CLManager::checkError throws: the console runner, given
    a truncated or hostile scene file or an unknown option value, reports the
    library's message and exits 1 (no crash; no GPU needed, the scene is read
    before the tracer starts)."""
    import subprocess

    from raycer_amd import build

    runner = build.RUNNER
    if not runner.exists():
        pytest.skip("the console runner is not built")
    cases = {
        "truncated.json": ('{"scene": [[[[', "raycer error 1: scene file: unexpected end"),
        "deep.json": ("[" * 100000, "nesting deeper than 256 levels"),
        "deep.xml": ("<a>" * 100000, "nesting deeper than 256 levels"),
    }
    for name, (text, message) in cases.items():
        f = tmp_path / name
        f.write_text(text)
        r = subprocess.run([str(runner), "--scene-file", str(f)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and message in r.stderr, (name, r.returncode, r.stderr[-300:])
    r = subprocess.run([str(runner), "--scene", "box", "--sampler", "bogus"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "unknown sampler bogus" in r.stderr
