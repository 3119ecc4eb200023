"""The C++ drop-in interface (raycer_render, include/raycer.hpp) and the
batched multi-scene graph, on the GPU."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import raycer_amd as ra
from raycer_amd import api

pytestmark = pytest.mark.gpu
RUNNER = Path(ra.__file__).resolve().parent / "raycer_render"


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return data.reshape(-1, 3)


def test_console_runner_matches_c_abi(ctx, tmp_path):
    out = tmp_path / "box.pfm"
    ppm = tmp_path / "box.ppm"
    r = subprocess.run([str(RUNNER), "--scene", "box", "--width", "64", "--height", "48", "--spp", "4",
                        "--linear", str(out), "--tonemapped", str(ppm)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["triangles"] == 36 and info["interrupted"] is False
    img = read_pfm(out)
    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    s = ctx.upload(sd, nodes, ordered)
    cfg = sd.desc.cfg
    cfg.width, cfg.height, cfg.spp, cfg.frame = 64, 48, 4, 0
    film = ctx.film(64, 48)
    ctx.render(s, film, sd.desc.camera, cfg)
    ref = film.linear()[:, :3]
    assert np.array_equal(img, ref)
    head = ppm.read_bytes()[:15]
    assert head.startswith(b"P6\n64 48\n255\n")
    film.close()
    s.close()


def test_batch_graph_equals_individual_renders(ctx):
    import ctypes as C
    L = api.lib()
    scenes, films, cams, cfgs = [], [], [], []
    for seed in range(5):
        sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SOUP, seed=seed, tri_count=2000 + 3000 * seed))
        nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
        s = ctx.upload(sd, nodes, ordered)
        cfg = sd.desc.cfg
        cfg.width = cfg.height = [32, 48, 64, 40, 24][seed]
        cfg.spp = 8
        scenes.append(s)
        films.append(ctx.film(cfg.width, cfg.height))
        cams.append(sd.desc.camera)
        cfgs.append(cfg)
    n = len(scenes)
    sarr = (C.c_void_p * n)(*[s.h.value for s in scenes])
    farr = (C.c_void_p * n)(*[f.h.value for f in films])
    carr = (api.Camera * n)(*cams)
    garr = (api.RenderCfg * n)(*cfgs)
    tun = api.Tuning(128, 4, 64, 0)
    api.check(L.rc_render_batch(ctx.h, sarr, farr, carr, garr, n, C.byref(tun)))
    batch = [f.linear() for f in films]
    st = ctx.stats()
    assert st.samples == sum(c.width * c.height * c.spp for c in cfgs)
    ctx.set_tuning(spp_per_pass=4)
    for i in range(n):
        f2 = ctx.film(cfgs[i].width, cfgs[i].height)
        ctx.render(scenes[i], f2, cams[i], cfgs[i])
        assert np.array_equal(f2.linear(), batch[i]), i
        f2.close()
    ctx.set_tuning()
    # the context's default tuning: each scene's pass size chosen for it (all
    # 8 samples in one pass here), the same as rendering it alone
    for f in films:
        f.clear()
    api.check(L.rc_render_batch(ctx.h, sarr, farr, carr, garr, n, None))
    batch_auto = [f.linear() for f in films]
    for i in range(n):
        f2 = ctx.film(cfgs[i].width, cfgs[i].height)
        ctx.render(scenes[i], f2, cams[i], cfgs[i])
        assert np.array_equal(f2.linear(), batch_auto[i]), i
        np.testing.assert_allclose(batch_auto[i], batch[i], rtol=1e-6, atol=1e-7)  # same paths, other grouping
        f2.close()
    for f in films:
        f.clear()
    api.check(L.rc_render_batch(ctx.h, sarr, farr, carr, garr, n, C.byref(tun)))
    # the second call with the same batch relaunches the kept executable
    # graph; the films accumulate
    api.check(L.rc_render_batch(ctx.h, sarr, farr, carr, garr, n, C.byref(tun)))
    assert ctx.stats().graph_relaunched == 1
    again = films[0].sums()
    assert np.allclose(again[:, 3], 2 * 8, atol=1e-3)
    for f in films:
        f.close()
    for s in scenes:
        s.close()


def test_autotuned_batch_plan_equals_individual_renders(ctx, tmp_path):
    """rc_autotune_batch picks samples per pass per size class; the batch
    rendered with that plan equals one render per scene with the same passes,
    and a second autotune call reads the cached JSON instead of re-timing."""
    scenes, films, cams, cfgs = [], [], [], []
    sizes = [(32, 32), (48, 40), (128, 96), (30, 34), (128, 128)]
    for seed, (w, h) in enumerate(sizes):
        sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SOUP, seed=20 + seed, tri_count=3000 + 2000 * seed))
        nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
        scenes.append(ctx.upload(sd, nodes, ordered))
        cfg = sd.desc.cfg
        cfg.width, cfg.height, cfg.spp = w, h, 8
        films.append(ctx.film(w, h))
        cams.append(sd.desc.camera)
        cfgs.append(cfg)
    cache = tmp_path / "tune.json"
    plan = ctx.autotune_batch(scenes, cams, cfgs, cache)
    text = cache.read_text()
    entries = json.loads(text)
    classes = {int(np.ceil(np.log2(c.width * c.height))) for c in cfgs}
    assert len(entries) == len(classes)
    assert all(1 <= v <= 8 and (v & (v - 1)) == 0 for v in entries.values())
    assert ((plan >= 1) & (plan <= 8)).all()
    # scenes of one class share the choice
    for i, ci in enumerate(cfgs):
        for j, cj in enumerate(cfgs):
            if np.ceil(np.log2(ci.width * ci.height)) == np.ceil(np.log2(cj.width * cj.height)):
                assert plan[i] == plan[j]
    again = ctx.autotune_batch(scenes, cams, cfgs, cache)
    assert np.array_equal(again, plan) and cache.read_text() == text
    for f in films:
        f.clear()
    ctx.render_batch(scenes, films, cams, cfgs, spp_per_pass=plan)
    batch = [f.linear() for f in films]
    try:
        for i in range(len(scenes)):
            ctx.set_tuning(spp_per_pass=int(plan[i]))
            f2 = ctx.film(cfgs[i].width, cfgs[i].height)
            ctx.render(scenes[i], f2, cams[i], cfgs[i])
            assert np.array_equal(f2.linear(), batch[i]), i
            f2.close()
    finally:
        ctx.set_tuning()
        for f in films:
            f.close()
        for s in scenes:
            s.close()


def cornell_scene(w, h, spp):
    """raycer's TestScene3 (src/TestScenes/TestScene3.cpp:11-47): the Cornell
    box OBJ, camera at (0, 1, 3.5) looking down -z with the default 75 degree
    horizontal field of view, one point light at (0, 1.9, 0)."""
    from tests.test_obj import MESHES
    cam = api.Camera((0.0, 1.0, 3.5), 0.0, 0.0, 0.0, 75.0)
    cfg = api.RenderCfg()
    cfg.width, cfg.height, cfg.spp = w, h, spp
    cfg.sampler, cfg.filter, cfg.filter_radius = api.RC_SAMPLER_JITTERED, api.RC_FILTER_BOX, 0.5
    cfg.max_path_length, cfg.rr_start, cfg.estimator = 6, 3, api.RC_EST_REFERENCE
    cfg.ray_start_offset, cfg.seed = 1e-5, 3
    light = api.PointLight((0.0, 1.9, 0.0), (1.0, 0.71, 0.24), 1.0, 3.0, 1.0, (1.0, 1.0, 1.0), 0.01)
    return ra.obj_scene(MESHES / "cornellbox.obj", cam, cfg, light)


def test_reference_cornell_scene_phong_and_path_tracer(ctx):
    import oracle as orc
    from tests.conftest import ORACLE_THREADS
    from tests.test_gpu_render import oracle_stats, render_gpu, replay_check
    sd = cornell_scene(64, 64, 16)
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    s = ctx.upload(sd, nodes, ordered)
    ob = orc.BVH(sd.vertices)
    try:
        # the reference's default tracer for this scene: Phong + the point light
        w, h = 160, 120
        light = sd.desc.light
        ids, t, rgba = ctx.raycast_phong(s, sd.desc.camera, w, h, 1, 1e-5, light)
        lt = [*light.position, *light.color, light.intensity, light.max_distance, light.attenuation]
        amb = [*light.ambient_color, light.ambient_intensity]
        hit, t_o, rgb = orc.raycast_phong(ob, sd.tri_material, sd.materials, sd.desc.camera, w, h, 1, 1e-5, lt, amb,
                                          threads=ORACLE_THREADS)
        same = ids == hit
        assert same.mean() > 0.995 and (hit >= 0).mean() > 0.3  # the box is open towards the camera
        err = np.abs(rgba[same, :3] - rgb[same])
        assert (err <= 1e-3).mean() > 0.995
        # the path tracer lit by the emissive ceiling quad (Ke 17 12 4), replayed by the oracle
        cfg = sd.desc.cfg
        lin, _ = render_gpu(ctx, s, sd.desc.camera, cfg)
        st = oracle_stats(ob, sd, sd.desc.camera, cfg)
        frac = replay_check(lin, st)
        print(f"cornell: phong {same.mean():.4f} ids equal; path replay {frac:.2%} pixels equal")
        assert lin[:, :3].mean() > 0.01
    finally:
        s.close()


def test_console_runner_renders_the_reference_cornell_scene(ctx, tmp_path):
    """raycer_render --scene cornell: TestScene3 through the C++ interface
    (Scene::models -> OBJ loader -> FlatBVH::build -> HipTracer).  Its RAY
    tracer image equals rc_raytrace with the scene's ambient light and point
    light; --tracer phong equals rc_raycast_phong; the path tracer renders the
    emitter-lit box."""
    out = tmp_path / "cornell_ray.pfm"
    r = subprocess.run([str(RUNNER), "--scene", "cornell", "--width", "80", "--height", "60", "--linear", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["tracer"] == "ray" and info["triangles"] == 36
    img = read_pfm(out)
    ph = tmp_path / "cornell_phong.pfm"
    r = subprocess.run([str(RUNNER), "--scene", "cornell", "--tracer", "phong", "--width", "80", "--height", "60",
                        "--linear", str(ph)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1])["tracer"] == "phong"
    sd = cornell_scene(80, 60, 1)
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    s = ctx.upload(sd, nodes, ordered)
    try:
        _, _, rgba = ctx.raycast_phong(s, sd.desc.camera, 80, 60, 1, 1e-5, sd.desc.light)
        np.testing.assert_array_equal(read_pfm(ph), rgba[:, :3])
        # Tracer::run with the RAY tracer: TestScene3's lights, default materials,
        # one sample (the pixel centre), seed 0, frame 0
        cfg = sd.desc.cfg
        cfg.spp, cfg.seed, cfg.frame = 1, 0, 0
        rs = api.ray_settings(ambient_intensity=0.01)
        lt = api.light(api.RC_LIGHT_POINT, position=(0.0, 1.9, 0.0), color=(1.0, 0.71, 0.24), max_distance=3.0)
        film = ctx.film(80, 60)
        ctx.raytrace(s, film, sd.desc.camera, cfg, rs, api.ray_materials(len(sd.materials)), [lt])
        ref = film.linear()[:, :3]
        film.close()
    finally:
        s.close()
    np.testing.assert_array_equal(img, ref)
    assert ref.max() > 0.05
    pt = tmp_path / "cornell_path.pfm"
    r = subprocess.run([str(RUNNER), "--scene", "cornell", "--tracer", "path", "--width", "64", "--height", "48",
                        "--spp", "16", "--linear", str(pt)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lin = read_pfm(pt)
    assert np.isfinite(lin).all() and lin.mean() > 0.0


@pytest.mark.parametrize("ext", ["xml", "json"])
def test_console_runner_scene_file_round_trip(ctx, tmp_path, ext):
    """raycer_render --save-scene then --scene-file (Scene::saveToFile /
    loadFromFile): the reloaded scene, its built BVH included, renders the
    same RAY-tracer image bit for bit."""
    scene_file = tmp_path / f"sphere.{ext}"
    a, b = tmp_path / "a.pfm", tmp_path / "b.pfm"
    common = ["--width", "48", "--height", "40", "--spp", "4"]
    r = subprocess.run([str(RUNNER), "--scene", "sphere", "--level", "2", "--grid", "8", "--tracer", "ray", *common,
                        "--save-scene", str(scene_file), "--linear", str(a)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(RUNNER), "--scene-file", str(scene_file), *common, "--linear", str(b)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1])["tracer"] == "ray"
    ia, ib = read_pfm(a), read_pfm(b)
    assert ia.max() > 0 and np.array_equal(ia, ib)
    sf = api.load_scene_file(scene_file)
    assert sf.nodes is not None and len(sf.vertices) == len(sf.tri_material) > 0 and len(sf.lights) == 1


def test_bandwidth_probes_run_and_reject_bad_arguments(ctx):
    """The achievable-bandwidth probes the roofline is priced against: the
    streaming copy, a cache-resident read and the scattered 64-byte read give
    positive rates; the scattered read refuses more records than its buffer
    holds (64 MiB holds 2^20 records)."""
    assert ctx.copy_bandwidth(64 << 20, 2) > 0.0
    assert ctx.read_bandwidth(4 << 20, 4) > 0.0
    assert ctx.gather_bandwidth(64 << 20, 1 << 20, 2) > 0.0
    with pytest.raises(RuntimeError):
        ctx.gather_bandwidth(64 << 20, (1 << 20) + 1, 2)
    with pytest.raises(RuntimeError):
        ctx.gather_bandwidth(64 << 20, 0, 2)
