"""GPU Whitted ray tracer (rc_raytrace) against the FP64 oracle's recursive
Raytracer::traceRecursive.

GPU and oracle key every draw by the node's place in the ray tree, so with the
same seed they trace the same rays: per-pixel colours agree to FP32 rounding
except where an FP32 and an FP64 decision part (a shadow ray grazing an edge,
a total internal reflection at the critical angle).  The closed-form scenes of
tests/test_whitted_oracle.py are rendered on the GPU too, and a level split
(more children than the buffers hold) is checked against the oracle.
"""
import math

import numpy as np
import pytest

import oracle as orc
import raycer_amd as ra
from raycer_amd import api
from tests.conftest import ORACLE_THREADS
from tests.test_gpu_render import cfg_for, replay_check
from tests.test_whitted_oracle import CENTER, DOWN, FORWARD, H, W, material, plane

pytestmark = pytest.mark.gpu


def box():
    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    return sd, nodes, ordered


def full_featured(nmat):
    """Every Whitted feature at once: mirror, glossy reflection, glass with
    Fresnel and attenuation, glossy transmission, AO, three light types with an
    area light, fog."""
    rm = api.ray_materials(nmat)
    rm[0].ray_reflectance = 0.35                     # mirror walls
    if nmat > 1:
        rm[1].ray_reflectance, rm[1].refl_n, rm[1].refl_glossiness = 0.5, 2, 40.0
    if nmat > 2:
        m = rm[2]
        m.ray_reflectance, m.ray_transmittance, m.fresnel_reflection, m.refractive_index = 1.0, 0.9, 1, 1.5
        m.enable_attenuation, m.attenuation_factor = 1, 0.8
        m.attenuation_color[:] = (0.1, 0.3, 0.2)
    if nmat > 3:
        rm[3].ray_transmittance, rm[3].trans_n, rm[3].trans_glossiness, rm[3].refractive_index = 0.6, 2, 30.0, 1.2
    rs = api.ray_settings(max_ray_iterations=3, background=(0.05, 0.07, 0.1), ambient_intensity=0.3, ao_enabled=1,
                          ao_n=2, ao_max_distance=0.3, fog_enabled=1, fog_distance=8.0, fog_color=(0.6, 0.6, 0.7))
    lights = [api.light(api.RC_LIGHT_POINT, position=(0.5, 0.9, 0.5), intensity=0.8, max_distance=5.0,
                        area_light=1, area_n=2, area_radius=0.1),
              api.light(api.RC_LIGHT_DIRECTIONAL, direction=(0.3, -0.9, -0.3), intensity=0.4),
              api.light(api.RC_LIGHT_SPOT, position=(0.2, 0.95, 0.8), direction=(0.0, -1.0, 0.0), angle=40.0,
                        intensity=0.7, max_distance=4.0)]
    return rs, rm, lights


def oracle_ray(sd, nodes_ordered_unused, cfg, rs, rm, lights, pixel_ids=None):
    ob = orc.BVH(sd.vertices)
    c = orc.render_cfg(cfg, traversal=1)
    return orc.raytrace(ob, sd.tri_material, sd.materials, sd.desc.camera, c, rs, rm, lights, pixel_ids=pixel_ids,
                        threads=ORACLE_THREADS)


@pytest.mark.parametrize("sampler", [api.RC_SAMPLER_CMJ, api.RC_SAMPLER_JITTERED])
def test_whitted_replays_the_oracle(ctx, sampler):
    sd, nodes, ordered = box()
    s = ctx.upload(sd, nodes, ordered)
    cfg = cfg_for(sd, 48, 48, 4, sampler=sampler, filter=api.RC_FILTER_TENT, filter_radius=1.0)
    rs, rm, lights = full_featured(len(sd.materials))
    film = ctx.film(48, 48)
    ctx.raytrace(s, film, sd.desc.camera, cfg, rs, rm, lights)
    lin, sums = film.linear(), film.sums()
    st = ctx.stats()
    assert st.extension_rays > 48 * 48 * 4 and st.shadow_rays > 0
    ref = oracle_ray(sd, None, cfg, rs, rm, lights)
    np.testing.assert_allclose(sums[:, 3], ref[:, 0], rtol=1e-5)
    frac = replay_check(lin, ref, min_equal=0.95)
    # the GPU traced the oracle's trees: the same number of closest-hit rays,
    # but for the few whose FP32 and FP64 hit decisions part
    assert st.extension_rays == pytest.approx(ref[:, 11].sum(), rel=2e-3)
    print(f"whitted replay: {frac:.2%} pixels equal")
    # bit-reproducible: the pixel sums run in a fixed order
    film2 = ctx.film(48, 48)
    ctx.raytrace(s, film2, sd.desc.camera, cfg, rs, rm, lights)
    assert film2.linear().tobytes() == lin.tobytes()
    # every ray re-walked by the deep-stack kernel (on-chip stack of one
    # entry): the same image, each listed ray walked and counted once
    film3 = ctx.film(48, 48)
    ctx.set_tuning(stack_limit=1)
    try:
        ctx.raytrace(s, film3, sd.desc.camera, cfg, rs, rm, lights)
        st3 = ctx.stats()
    finally:
        ctx.set_tuning()
    assert film3.linear().tobytes() == lin.tobytes()
    assert 0.2 * st3.extension_rays < st3.overflow_rays <= st3.extension_rays + st3.shadow_rays
    film.close()
    film2.close()
    film3.close()
    s.close()


def gpu_center(ctx, tris, tri_mat, mats, cam, rs, rm, lights=()):
    v = np.array(tris, np.float32).reshape(-1, 9)
    mt = np.zeros(len(mats), api.MATERIAL_DTYPE)
    for i, m in enumerate(mats):
        mt[i]["diffuse"], mt[i]["ambient"], mt[i]["specular"] = m[0:3], m[6:9], m[9:12]
        mt[i]["shininess"], mt[i]["flags"] = m[12], int(m[13])
    sd = ra.SceneData(v, np.array(tri_mat, np.uint32), mt, api.SceneDesc())
    nodes, ordered, _, _ = ra.build_bvh(v)
    s = ctx.upload(sd, nodes, ordered)
    cfg = api.RenderCfg()
    cfg.width, cfg.height, cfg.spp, cfg.sampler = W, H, 1, api.RC_SAMPLER_CENTER
    cfg.filter, cfg.filter_radius, cfg.max_path_length = api.RC_FILTER_BOX, 0.5, 1
    cfg.ray_start_offset = 1e-4
    cam_s = api.Camera()
    cam_s.position[:] = cam[0:3]
    cam_s.pitch, cam_s.yaw, cam_s.roll, cam_s.fov = cam[3:7]
    film = ctx.film(W, H)
    try:
        ctx.raytrace(s, film, cam_s, cfg, rs, rm, lights)
        return film.linear()[CENTER, :3]
    finally:
        film.close()
        s.close()


@pytest.mark.parametrize("iters", [0, 3])
def test_gpu_facing_mirrors(ctx, iters):
    tris = plane(z=-1) + plane(z=1, up=False)
    rs = api.ray_settings(max_ray_iterations=iters)
    L = gpu_center(ctx, tris, [0] * 4, [material(ambient=1.0)], FORWARD, rs, api.ray_materials(1, ray_reflectance=0.5))
    assert L == pytest.approx([2 * (1 - 0.5 ** (iters + 1))] * 3, rel=1e-6)


def test_gpu_fresnel_and_ambient_occlusion(ctx):
    rs = api.ray_settings(background=(1, 1, 1), ambient_intensity=0.0)
    rm = api.ray_materials(1, ray_reflectance=1.0, fresnel_reflection=1, refractive_index=1.5)
    L = gpu_center(ctx, plane(y=0), [0, 0], [material()], DOWN, rs, rm)
    assert L == pytest.approx([0.04] * 3, rel=1e-5)
    rs = api.ray_settings(ambient_intensity=1.0, ao_enabled=1, ao_n=4, ao_max_distance=1.0)
    tris = plane(y=0) + plane(y=0.5, size=1000, up=False)
    L = gpu_center(ctx, tris, [0, 0, 0, 0], [material(ambient=0.7)], [0, 0.2, 0, -90, 0, 0, 60], rs,
                   api.ray_materials(1))
    assert L == pytest.approx([0.7 * 0.25] * 3, rel=1e-6)


def test_level_split_matches_the_oracle(ctx):
    """128 glossy children per hit over two levels: level 2 would hold 67M rays,
    so level 1 is processed in parts; sampled pixels still equal the oracle."""
    sd, nodes, ordered = box()
    s = ctx.upload(sd, nodes, ordered)
    cfg = cfg_for(sd, 64, 64, 1)
    rm = api.ray_materials(len(sd.materials), ray_reflectance=0.4, refl_n=8, refl_glossiness=20.0,
                           ray_transmittance=0.3, trans_n=8, trans_glossiness=20.0, refractive_index=1.3)
    rs = api.ray_settings(max_ray_iterations=2, ambient_intensity=0.5)
    lights = [api.light(api.RC_LIGHT_POINT, position=(0.5, 0.9, 0.5), max_distance=5.0)]
    film = ctx.film(64, 64)
    ctx.raytrace(s, film, sd.desc.camera, cfg, rs, rm, lights)
    lin = film.linear()
    st = ctx.stats()
    assert st.extension_rays > (1 << 22)  # far more than one level's buffers hold
    ids = np.arange(0, 64 * 64, 211, dtype=np.uint32)
    ref = oracle_ray(sd, None, cfg, rs, rm, lights, pixel_ids=ids)
    mu = ref[:, 2:5] / ref[:, 0:1]
    np.testing.assert_allclose(lin[ids, :3], mu, rtol=2e-3, atol=1e-5)
    film.close()
    s.close()


def test_raytrace_rejects_bad_settings(ctx):
    sd, nodes, ordered = box()
    s = ctx.upload(sd, nodes, ordered)
    cfg = cfg_for(sd, 16, 16, 1)
    film = ctx.film(16, 16)
    rs = api.ray_settings()
    n = len(sd.materials)
    with pytest.raises(ra.RaycerError):
        ctx.raytrace(s, film, sd.desc.camera, cfg, rs, api.ray_materials(n + 1))          # material count
    with pytest.raises(ra.RaycerError):
        ctx.raytrace(s, film, sd.desc.camera, cfg, rs, api.ray_materials(n, refl_n=17))   # glossy samples
    with pytest.raises(ra.RaycerError):
        ctx.raytrace(s, film, sd.desc.camera, cfg, api.ray_settings(ao_enabled=1, ao_n=0), api.ray_materials(n))
    with pytest.raises(ra.RaycerError):
        ctx.raytrace(s, film, sd.desc.camera, cfg, rs, api.ray_materials(n), [api.light(api.RC_LIGHT_POINT)] * 65)
    with pytest.raises(ra.RaycerError):
        ctx.raytrace(s, film, sd.desc.camera, cfg, api.ray_settings(max_ray_iterations=17), api.ray_materials(n))
    film.close()
    s.close()
