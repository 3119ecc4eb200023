"""Path-tracing throughput on one MI355X.

Workload: the ridged-fBm terrain of 2,093,058 triangles at 1920x1080,
64 jittered samples per pixel with a tent filter of radius 1, at most 6 rays
per path with Russian roulette from vertex 3, sky radiance 0.5 and a
directional sun with next-event estimation.  One step = one full frame
(132,710,400 camera samples) through the wavefront path tracer; the scene is
built and uploaded before timing, so the timed region holds only GPU work on
device-resident inputs.  The BVH is the library's 32-bin SAH build (the
reference's 6-candidate SAH build gives the same hits at ~13% lower
throughput).

Prints one JSON line.  Besides the headline Msamples/s it reports the
roofline of the dominant kernel, k_extend_shade (closest-hit traversal fused
with shading).  Its algorithmic bytes per frame are
  16 * child records fetched (a 4-wide node is 64 bytes: 8-bit quantized
    child boxes, counts and links) + 48 * triangle records tested
  + 96 per extension ray (ray 32 + path state 16 + shading record 16 in,
    radiance accumulator 16 read + 16 written)
  + 48 per continuation ray written (ray 32 + state 16)
  + 48 per shadow ray written (ray 32 + contribution 16),
counted by the counting variant on one untimed frame (materials, a few
64-byte records, are left out) and divided by the kernel time measured with HIP
events on the renderer's stream during the timed frames.  The CPU baseline is
the FP64 oracle (restatement of the reference's OpenMP CPU path, built -O3
-ffast-math) on a seeded half of the pixels.
"""
from __future__ import annotations

# bytes of node data per child of a visited 4-wide node: the library's 64-byte
# nodes (the 128-byte float-box records of RC_QNODE=0 would be 32)
NODE_BYTES_PER_CHILD = 16

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s at 1080p PT; equal-time MSE vs 4096-spp ref, cached vs plain"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = ROOT / "profiles" / "extend_traffic_terrain1080p.json"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(sd, cam, cfg, frac: int = 2, seed: int = 7) -> dict:
    """FP64 oracle (-O3 -ffast-math, OpenMP) on a seeded 1/frac pixel sample."""
    import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    npix = cfg.width * cfg.height
    rng = np.random.default_rng(seed)
    ids = np.sort(rng.choice(npix, npix // frac, replace=False)).astype(np.uint32)
    ob = orc.BVH(sd.vertices, fast=True)
    ocfg = orc.render_cfg(cfg, traversal=0)
    t0 = time.perf_counter()
    st = orc.render(ob, sd.tri_material, sd.materials, cam, ocfg, pixel_ids=ids, threads=threads)
    dt = time.perf_counter() - t0
    samples = len(ids) * cfg.spp
    return {
        "value": round(samples / dt / 1e6, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(ids)} seeded pixels (1/{frac} of {cfg.width}x{cfg.height}) x {cfg.spp} spp, "
                  f"{dt:.1f} s, unordered FP64 traversal as the reference CPU path, extension rays/sample "
                  f"{st[:, 11].sum() / samples:.2f}",
    }


def mse_pair(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))


def equal_time_mse(ctx, scene, sd, base_cfg, scale: int, cache_vertex: int = 2, cached_spp: int = 1,
                   plain_spp: int = 64, ref_spp: int = 4096, warmup_frames: int = 16,
                   pretrain_frames: int = 256, train_steps: int = 1) -> dict:
    """Plain and radiance-cache renders given the same wall time, against a
    ref_spp plain render.

    The budget is the time of one plain frame of plain_spp samples.  Cold: a
    new cache spends the budget on warmup_frames training frames (discarded)
    and then on as many training frames of cached_spp samples as fit,
    accumulated.  Warm: the same cache after pretrain_frames more untimed
    training frames (a cache carried over from earlier frames, as in an
    animation) renders and trains within the budget.  Frames of one sample use
    random sub-pixel positions under the same filter (one jittered sample would
    be the unfiltered pixel centre).  Paths end into the cache
    at vertex cache_vertex (after two bounces for 2); every frame ends with
    train_steps Adam steps on its training samples.  MSE is over linear RGB
    and over the Reinhard-mapped RGB (key 0.18, gamma 2.2)."""
    from raycer_amd import api

    cfg = api.RenderCfg.from_buffer_copy(base_cfg)
    cfg.width, cfg.height = base_cfg.width // scale, base_cfg.height // scale
    cam = sd.desc.camera
    film = ctx.film(cfg.width, cfg.height)

    def image():
        return film.linear()[:, :3].astype(np.float64), film.tonemap()[0][:, :3].astype(np.float64)

    def score(lin, tm, ref_lin, ref_tm):
        return {"mse_linear": mse_pair(lin, ref_lin), "mse_tonemapped": mse_pair(tm, ref_tm),
                "mean_ratio": round(float(lin.mean() / ref_lin.mean()), 5)}

    out = {"resolution": f"{cfg.width}x{cfg.height}", "reference_spp": ref_spp, "plain_spp": plain_spp,
           "cache_vertex": cache_vertex, "train_steps_per_frame": train_steps}
    try:
        # reference: ref_spp / plain_spp frames of plain_spp samples
        cfg.spp = plain_spp
        film.clear()
        for f in range(ref_spp // plain_spp):
            cfg.frame = 100000 + f
            ctx.render(scene, film, cam, cfg)
        ref_lin, ref_tm = image()
        # plain at the budget
        film.clear()
        cfg.frame = 7
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render(scene, film, cam, cfg)
        ctx.synchronize()
        budget = time.perf_counter() - t0
        out["budget_ms"] = round(budget * 1e3, 2)
        out["plain"] = score(*image(), ref_lin, ref_tm)
        cache = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, lr=1e-3, seed=1, max_batch=65536)
        try:
            ccfg = api.RenderCfg.from_buffer_copy(cfg)
            ccfg.spp = cached_spp
            ccfg.cache_vertex = cache_vertex
            if cached_spp == 1:
                # one jittered sample is the pixel centre with weight 1 (no
                # filtering); random positions keep the filter's footprint
                ccfg.sampler = api.RC_SAMPLER_RANDOM
            frame = [200000]

            def run(warmup):
                frames = accumulated = 0
                ctx.synchronize()
                t0 = time.perf_counter()
                film.clear()
                while True:
                    if frames == warmup:
                        film.clear()
                    ccfg.frame = frame[0]
                    frame[0] += 1
                    loss = cache.render(scene, film, cam, ccfg, train=train_steps)
                    frames += 1
                    if frames > warmup:
                        accumulated += 1
                    elapsed = time.perf_counter() - t0
                    if frames > warmup and elapsed * (frames + 1) / frames > budget:
                        break
                ctx.synchronize()
                st = ctx.stats()
                r = score(*image(), ref_lin, ref_tm)
                r.update({"spp": accumulated * cached_spp, "ms": round((time.perf_counter() - t0) * 1e3, 2),
                          "train_steps": frames, "train_samples_last_frame": int(st.train_samples),
                          "cache_queries_last_frame": int(st.cache_queries), "last_loss": round(loss, 5)})
                return r

            out["cached_cold"] = run(warmup_frames)
            out["cached_cold"]["warmup_frames"] = warmup_frames
            for f in range(pretrain_frames):
                ccfg.frame = frame[0]
                frame[0] += 1
                film.clear()
                cache.render(scene, film, cam, ccfg, train=train_steps)
            out["cached_warm"] = run(0)
            out["cached_warm"]["pretrain_frames"] = pretrain_frames
        finally:
            cache.close()
    finally:
        film.close()
    for k in ("cached_cold", "cached_warm"):
        out[k]["ratio_linear"] = round(out[k]["mse_linear"] / out["plain"]["mse_linear"], 4)
        out[k]["ratio_tonemapped"] = round(out[k]["mse_tonemapped"] / out["plain"]["mse_tonemapped"], 4)
    return out


def phong_c2(ctx, reps: int = 10) -> dict:
    """Primary ray + one point-light shadow ray per pixel with Phong shading
    (rc_raycast_phong) on the seeded sphere + ground scene (655,360 + 524,288
    triangles), 1920x1080, 1 jittered sample.  The device time is the sum of
    the kernels' HIP-event times; the wall time includes copying the hit
    indices, distances and RGBA image back to the host (the entry point's
    contract)."""
    import raycer_amd as ra
    from raycer_amd import api

    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    s = ctx.upload(sd, nodes, ordered)
    w, h = sd.desc.cfg.width, sd.desc.cfg.height
    try:
        ctx.raycast_phong(s, sd.desc.camera, w, h, 1, sd.desc.cfg.ray_start_offset, sd.desc.light)  # warm-up
        ctx.set_profiling(True)
        ctx.stats()
        dev_ms = 0.0
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.raycast_phong(s, sd.desc.camera, w, h, 1, sd.desc.cfg.ray_start_offset, sd.desc.light)
            st = ctx.stats()
            dev_ms += st.extend_ms + st.shadow_ms + st.shade_ms + st.other_ms
        wall = (time.perf_counter() - t0) / reps
        ctx.set_profiling(False)
    finally:
        s.close()
    rays = 2 * w * h
    return {"workload": f"sphere-ground-{sd.ntri}tri-{w}x{h}-1spp-primary+shadow", "triangles": sd.ntri,
            "device_ms": round(dev_ms / reps, 3), "wall_ms_with_host_copies": round(wall * 1e3, 3),
            "grays_per_s_device": round(rays / (dev_ms / reps * 1e-3) / 1e9, 3),
            "mpixels_per_s_device": round(w * h / (dev_ms / reps * 1e-3) / 1e6, 1)}


def whitted_line(ctx, reps: int = 5) -> dict:
    """The Whitted ray tracer (rc_raytrace) on the seeded sphere + ground scene
    (655,360 + 524,288 triangles), 1920x1080, one sample per pixel: every
    material reflects 0.5 with up to three reflection levels, one point light
    with its shadow rays, ambient 0.1.  Wall time of the synchronous call (ray
    tree levels, scans, shading, the film add; no copies to the host)."""
    import raycer_amd as ra
    from raycer_amd import api

    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    s = ctx.upload(sd, nodes, ordered)
    cfg = api.RenderCfg.from_buffer_copy(sd.desc.cfg)
    cfg.spp = 1
    L = sd.desc.light
    lights = [api.light(api.RC_LIGHT_POINT, position=tuple(L.position), intensity=L.intensity,
                        max_distance=L.max_distance, attenuation=L.attenuation)]
    rs = api.ray_settings(max_ray_iterations=3, ambient_intensity=0.1)
    rm = api.ray_materials(len(sd.materials), ray_reflectance=0.5)
    film = ctx.film(cfg.width, cfg.height)
    try:
        ctx.raytrace(s, film, sd.desc.camera, cfg, rs, rm, lights)  # warm-up: sizes the level buffers
        ext = shadow = 0
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.raytrace(s, film, sd.desc.camera, cfg, rs, rm, lights)
            st = ctx.stats()
            ext += st.extension_rays
            shadow += st.shadow_rays
        dt = (time.perf_counter() - t0) / reps
    finally:
        film.close()
        s.close()
    return {"workload": f"sphere-ground-{sd.ntri}tri-{cfg.width}x{cfg.height}-1spp-mirror0.5-depth3-pointlight",
            "ms_per_frame": round(dt * 1e3, 3), "closest_hit_rays_per_frame": ext // reps,
            "shadow_rays_per_frame": shadow // reps,
            "grays_per_s": round((ext + shadow) / reps / dt / 1e9, 3)}


def batch_c4(ctx, n_scenes: int = 64) -> dict:
    """Many-scene workload: soup scenes with seeds 0..n-1 (1k-500k triangles,
    log-uniform; 128^2 to 1024^2; 32 jittered spp; depth 5), autotuned samples
    per pass per size class, all frames in one captured HIP graph."""
    import tempfile

    import raycer_amd as ra
    from raycer_amd import api

    scenes, films, cams, cfgs = [], [], [], []
    tris = 0
    t0 = time.perf_counter()
    try:
        for seed in range(n_scenes):
            sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SOUP, seed=seed))
            nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
            scenes.append(ctx.upload(sd, nodes, ordered))
            films.append(ctx.film(sd.desc.cfg.width, sd.desc.cfg.height))
            cams.append(sd.desc.camera)
            cfgs.append(sd.desc.cfg)
            tris += sd.ntri
        setup = time.perf_counter() - t0
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            plan = ctx.autotune_batch(scenes, cams, cfgs, os.path.join(d, "tune.json"))
            tune = time.perf_counter() - t0
        for f in films:
            f.clear()
        ctx.render_batch(scenes, films, cams, cfgs, spp_per_pass=plan)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render_batch(scenes, films, cams, cfgs, spp_per_pass=plan)
        ctx.synchronize()
        dt = time.perf_counter() - t0
    finally:
        for f in films:
            f.close()
        for s in scenes:
            s.close()
    samples = sum(c.width * c.height * c.spp for c in cfgs)
    by_res = {}
    for c, p in zip(cfgs, plan):
        by_res[f"{c.width}x{c.height}"] = int(p)
    return {"scenes": n_scenes, "triangles": tris, "samples": samples, "setup_s": round(setup, 2),
            "autotune_ms": round(tune * 1e3, 1), "render_ms": round(dt * 1e3, 1),
            "msamples_per_s": round(samples / dt / 1e6, 1), "spp_per_pass": by_res}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=1024, help="terrain vertices per side (1024 = 2,093,058 triangles)")
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--spp-per-pass", type=int, default=64,
                    help="samples per pixel per wavefront pass (64: the whole frame at once, 132.7 M paths, "
                         "~27 GB of path queues; 0: the library's automatic choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c2", choices=("on", "off"), default="on", help="the 1080p Phong raycast line")
    ap.add_argument("--whitted", choices=("on", "off"), default="on", help="the 1080p Whitted ray tracer line")
    ap.add_argument("--c4", choices=("on", "off"), default="on",
                    help="also time the 64-scene batch (autotuned, one captured graph)")
    ap.add_argument("--mse", choices=("full", "small", "off"), default="full",
                    help="equal-time MSE of the radiance-cache renderer: at the bench resolution (full), "
                         "at a quarter of it per side (small), or not at all")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}): frames shard by pixel rows across GPUs "
              f"with no exchange but a final film gather, so the single-GPU figure is the per-GPU figure",
              file=sys.stderr)
        return 2

    import raycer_amd as ra
    from raycer_amd import api

    t0 = time.perf_counter()
    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_TERRAIN, seed=1, grid=a.grid))
    nodes, ordered, build_ms, depth = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    log(f"terrain {sd.ntri} triangles, binned-SAH BVH {len(nodes)} nodes depth {depth} in {build_ms:.0f} ms")
    cfg = sd.desc.cfg
    cfg.spp = a.spp
    cam = sd.desc.camera
    ctx = ra.Context(0)
    name, cus = ctx.device_name()
    scene = ctx.upload(sd, nodes, ordered)
    film = ctx.film(cfg.width, cfg.height)
    log(f"device {name} ({cus} CUs), setup {time.perf_counter() - t0:.1f} s")

    ctx.set_tuning(spp_per_pass=a.spp_per_pass)
    copy_gbs = ctx.copy_bandwidth(1 << 30, 20)
    log(f"streaming copy {copy_gbs:.0f} GB/s")

    # untimed counting frame: node/triangle fetches of the same rays
    ctx.set_counting(True)
    ctx.render(scene, film, cam, cfg)
    cst = ctx.stats()
    ctx.set_counting(False)
    ext_rays = cst.extension_rays
    camera_rays = cfg.width * cfg.height * cfg.spp
    # a camera ray is made in the kernel: no ray and state read (-48), its weight written (+4)
    alg_bytes = (NODE_BYTES_PER_CHILD * cst.node_records + 48 * cst.tri_tests + 96 * ext_rays
                 - 44 * camera_rays + 48 * (ext_rays - camera_rays) + 48 * cst.shadow_rays)
    log(f"counting frame: {ext_rays} extension rays, {cst.node_records / ext_rays:.1f} records/ray, "
        f"{cst.tri_tests / ext_rays:.2f} triangles/ray, {cst.shadow_rays} shadow rays")

    for i in range(a.warmup):
        cfg.frame = 1000 + i
        ctx.render(scene, film, cam, cfg)
    ctx.set_profiling(True)
    ctx.stats()  # drain pending timings
    ctx.synchronize()
    ext_ms = shd_ms = shade_ms = other_ms = 0.0
    ext_launches = 0
    total_ext = total_shadow = 0
    t_start = time.perf_counter()
    for i in range(a.steps):
        cfg.frame = i
        ctx.render(scene, film, cam, cfg)
        st = ctx.stats()
        ext_ms += st.extend_ms
        shd_ms += st.shadow_ms
        shade_ms += st.shade_ms
        other_ms += st.other_ms
        ext_launches += st.extend_launches
        total_ext += st.extension_rays
        total_shadow += st.shadow_rays
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    ctx.set_profiling(False)
    samples = cfg.width * cfg.height * cfg.spp * a.steps
    value = samples / elapsed / 1e6
    # algorithmic bytes per frame scale with the frame's rays; the counting
    # frame used frame index 0 like the first timed frame
    bytes_per_ray = alg_bytes / ext_rays
    achieved = bytes_per_ray * total_ext / (ext_ms * 1e-3) / 1e9
    traffic = None
    if PROFILE_TRAFFIC.exists():
        try:
            traffic = json.loads(PROFILE_TRAFFIC.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"terrain-{sd.ntri}tri-{cfg.width}x{cfg.height}-{cfg.spp}spp-tent1-depth6-rr3-sun-sky",
            "triangles": sd.ntri,
            "width": cfg.width,
            "height": cfg.height,
            "spp": cfg.spp,
            "max_path_length": cfg.max_path_length,
            "estimator": "reference (2 rho |cos|)",
            "bvh": "32-bin SAH, collapsed to 4-wide",
            "samples_per_pass": min(a.spp_per_pass, cfg.spp) if a.spp_per_pass else "auto",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend_shade<false> (closest-hit traversal fused with shading)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_ray": round(bytes_per_ray, 1),
            "avg_launch_ms": round(ext_ms / max(ext_launches, 1), 4),
            "achievable_copy_GBs": round(copy_gbs, 1),
            "frac_of_achievable_copy": round(achieved / copy_gbs, 4),
            "hbm_measured_GBs": (round(traffic / (ext_ms / max(ext_launches, 1) * 1e-3) / 1e9, 1)
                                 if traffic else None),
            "note": "achieved counts the algorithmic node, triangle and queue bytes; L2 and the 256 MB "
                    "Infinity Cache serve most node and triangle fetches, so it can exceed the HBM peak; "
                    "hbm_measured_GBs is the rocprofv3 FETCH_SIZE + WRITE_SIZE traffic per launch over "
                    "the same launch time",
        },
        "gsample_bounces_per_s": round(total_ext / elapsed / 1e9, 4),
        "shadow_rays_per_s_G": round(total_shadow / elapsed / 1e9, 4),
        "kernel_ms_per_step": {
            "extend": round(ext_ms / a.steps, 2),
            "shadow": round(shd_ms / a.steps, 2),
            "shade": round(shade_ms / a.steps, 2),
            "other": round(other_ms / a.steps, 2),
        },
        "bvh_build_ms": round(build_ms, 1),
    }
    film.close()
    ctx.set_tuning()  # the other lines use the library's defaults
    if a.c2 == "on":
        log("1080p primary + shadow Phong raycast (1.18M triangles)")
        out["phong_1080p"] = phong_c2(ctx)
        log(f"phong: {json.dumps(out['phong_1080p'])}")
    if a.whitted == "on":
        log("1080p Whitted ray tracer (1.18M triangles, mirrors, point light)")
        out["whitted_1080p"] = whitted_line(ctx)
        log(f"whitted: {json.dumps(out['whitted_1080p'])}")
    if a.c4 == "on":
        log("64-scene batch (autotuned, one captured graph)")
        out["batch_64_scenes"] = batch_c4(ctx)
        log(f"batch: {json.dumps(out['batch_64_scenes'])}")
    if a.mse != "off":
        log(f"equal-time MSE of the radiance cache ({a.mse})")
        out["equal_time_mse"] = equal_time_mse(ctx, scene, sd, cfg, 1 if a.mse == "full" else 4)
        log(f"equal-time MSE: {json.dumps(out['equal_time_mse'])}")
        # the same comparison on the closed box (emitter-lit, up to 8 rays per
        # path), where every path reaches the cache vertex
        bd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
        bn, bo, _, _ = ra.build_bvh(bd.vertices, method=api.RCH_BVH_BINNED)
        box = ctx.upload(bd, bn, bo)
        try:
            out["equal_time_mse_box"] = equal_time_mse(ctx, box, bd, bd.desc.cfg, 1 if a.mse == "full" else 2,
                                                       cached_spp=16, warmup_frames=4, pretrain_frames=64)
        finally:
            box.close()
        log(f"equal-time MSE (box): {json.dumps(out['equal_time_mse_box'])}")
    scene.close()
    ctx.close()
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle, half of the pixels)")
        out["cpu_baseline"] = cpu_baseline(sd, cam, cfg)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
