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
    ctx.set_tuning(spp_per_pass=1)
    # replaying the instantiated graph is not supported: a second batch call
    # re-captures; the films accumulate
    api.check(L.rc_render_batch(ctx.h, sarr, farr, carr, garr, n, C.byref(tun)))
    again = films[0].sums()
    assert np.allclose(again[:, 3], 2 * 8, atol=1e-3)
    for f in films:
        f.close()
    for s in scenes:
        s.close()
