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
with shading): `achieved` is the HBM bytes rocprofv3 counts per launch
(FETCH_SIZE + WRITE_SIZE, calibrated on a copy of known size, kept in
profiles/extend_traffic_terrain1080p.json) over the average launch time
measured here with HIP events on the renderer's stream.  `demand` holds the
algorithmic bytes per frame,
  16 * child records fetched (a 4-wide node is 64 bytes: 8-bit quantized
    child boxes, counts and links) + 48 * triangle records tested
  + 96 per extension ray (ray 32 + path state 16 + shading record 16 in,
    radiance accumulator 16 read + 16 written)
  + 48 per continuation ray written (ray 32 + state 16)
  + 48 per shadow ray written (ray 32 + contribution 16),
counted by the counting variant on one untimed frame (materials, a few
64-byte records, are left out), against the library's L2, Infinity Cache and
HBM read probes.  Further lines: the 1080p Phong raycast (C2), the Whitted
tracer, the 64-scene batch graph (C4) and the radiance-cached renderer at
equal time (C5).  The CPU leg runs the FP64 oracle (restatement of the
reference's OpenMP CPU path, built -O3 -ffast-math) on the host cores: a
4096-spp reference on a seeded 1/64 of the 1080p pixels, timed as the CPU
baseline and used to score the plain and cached images, the split-tail
variance bound of the cache, and the C1 image against its own 4096-spp
reference.
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
PROFILE_MFMA = ROOT / "profiles" / "mfma_cache_terrain1080p.json"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_reference(sd, cam, cfg, subset: int = 64, seed: int = 7, ref_spp: int = 4096, split_spp: int = 64,
                  cache_vertices=(1, 2)) -> dict:
    """The CPU leg: the FP64 oracle (restatement of the reference's OpenMP CPU
    path, built -O3 -ffast-math, unordered traversal as the reference) renders
    a seeded 1/subset of the pixels at ref_spp samples (n x n jittered, n =
    sqrt(ref_spp)).  Timed, that is
    the CPU baseline; the image is the reference the GPU images are scored
    against.  Then, for each of cache_vertices, the split-tail estimate on the
    same pixels at split_spp: the share of a split_spp-sample pixel
    estimate's variance that the light gathered after that vertex's
    next-event ray carries, which is what a perfect radiance cache at that
    vertex would remove."""
    import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    npix = cfg.width * cfg.height
    rng = np.random.default_rng(seed)
    ids = np.sort(rng.choice(npix, npix // subset, replace=False)).astype(np.uint32)
    ob = orc.BVH(sd.vertices, fast=True)
    ocfg = orc.render_cfg(cfg, traversal=0)
    ocfg.spp = ref_spp
    ocfg.frame = 500000
    t0 = time.perf_counter()
    st = orc.render(ob, sd.tri_material, sd.materials, cam, ocfg, pixel_ids=ids, threads=threads)
    dt = time.perf_counter() - t0
    samples = len(ids) * ref_spp
    ref = st[:, 2:5] / st[:, 0:1]
    tail_var = {}
    t1 = time.perf_counter()
    for cv in cache_vertices:
        scfg = orc.render_cfg(cfg, traversal=0)
        scfg.spp = split_spp
        scfg.frame = 600000
        scfg.cache_vertex = cv
        sp = orc.render_split_tail(ob, sd.tri_material, sd.materials, cam, scfg, ids, threads=threads)
        tail_var[cv] = float(np.mean(sp[:, 4:7] / (2.0 * sp[:, 0:1] ** 2)))
    dt_split = time.perf_counter() - t1
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), cpu_model)
    except OSError:
        pass
    base = {
        "value": round(samples / dt / 1e6, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "cpu_model": cpu_model,
        "kind": "port",
        "sample": f"{len(ids)} seeded pixels (1/{subset} of {cfg.width}x{cfg.height}) x {ref_spp} spp "
                  f"(the 4096-spp reference image), {dt:.1f} s, unordered FP64 traversal as the reference CPU path, "
                  f"extension rays/sample {st[:, 11].sum() / samples:.2f}",
        # the reference's unordered FP64 traversal of the same paths' extension rays
        "reference_work_per_extension_ray": {
            "binary_node_visits": round(float(st[:, 12].sum() / st[:, 11].sum()), 2),
            "triangle_tests": round(float(st[:, 13].sum() / st[:, 11].sum()), 2),
        },
    }
    return {"baseline": base, "ids": ids, "ref": ref, "tail_var": tail_var, "split_spp": split_spp,
            "split_s": round(dt_split, 1)}


def c1_gpu_image(ctx) -> tuple:
    """C1: the closed box at 256 x 256, 16 jittered samples, box filter, up to
    8 rays with Russian roulette from vertex 3, one frame on the GPU."""
    import raycer_amd as ra
    from raycer_amd import api

    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    s = ctx.upload(sd, nodes, ordered)
    cfg = api.RenderCfg.from_buffer_copy(sd.desc.cfg)
    cfg.width, cfg.height, cfg.spp, cfg.frame = 256, 256, 16, 0
    film = ctx.film(cfg.width, cfg.height)
    try:
        ctx.render(s, film, sd.desc.camera, cfg)
        img = film.linear()[:, :3].astype(np.float64)
    finally:
        film.close()
        s.close()
    return sd, cfg, img


def c1_vs_cpu(sd, cfg, img, subset: int = 4, ref_spp: int = 4096, seed: int = 11) -> dict:
    """The C1 image's MSE against a ref_spp-sample CPU oracle reference on a
    seeded 1/subset of its pixels, and the MSE the two images' noise alone
    predicts (from the oracle's per-sample moments; an upper estimate for
    jittered sampling)."""
    import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    npix = cfg.width * cfg.height
    ids = np.sort(np.random.default_rng(seed).choice(npix, npix // subset, replace=False)).astype(np.uint32)
    ocfg = orc.render_cfg(cfg, traversal=0)
    ocfg.spp = ref_spp
    ocfg.frame = 700000
    t0 = time.perf_counter()
    st = orc.render(orc.BVH(sd.vertices), sd.tri_material, sd.materials, sd.desc.camera, ocfg, pixel_ids=ids,
                    threads=threads)
    dt = time.perf_counter() - t0
    sw, sw2 = st[:, 0:1], st[:, 1:2]
    ref = st[:, 2:5] / sw
    var_ref = np.maximum((st[:, 8:11] - 2 * ref * st[:, 5:8] + ref * ref * sw2) / (sw * sw), 0)
    mse = mse_pair(img[ids], ref)
    expected = float(np.mean(var_ref * (1.0 + ref_spp / cfg.spp)))
    return {"workload": "box-256x256-16spp-boxfilter-depth8-rr3", "reference": f"FP64 CPU oracle, {len(ids)} seeded "
            f"pixels, {ref_spp} spp, {dt:.1f} s", "gpu_mse_vs_cpu_ref": mse, "expected_from_noise": expected,
            "ratio": round(mse / expected, 4)}


def mse_pair(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))


# FLOPs of the radiance-cache MLP (81 -> 64 x 4 -> 3): 2 x 17,664 weights per
# inference query; a training sample adds the backward pass through the
# hidden layers (2 x 12,480) and the weight gradients (2 x 17,664)
MLP_INFER_FLOP = 35328
MLP_TRAIN_FLOP = 35328 + 24960 + 35328
BF16_PEAK_TFLOPS = 2500.0


def cached_forms(ctx, scene, sd, base_cfg, scale: int = 1, cache_vertex: int = 2, plain_spp: int = 64,
                 ref_spp: int = 4096, pretrain_frames: int = 128, pretrain_spp: int = 4, pretrain_steps: int = 4,
                 train_steps: int = 1, cached_spp: int = 0, gpu_ref=None) -> dict:
    """The plain renderer and the radiance-cached renderer given the same wall
    time.  The budget is one plain frame of plain_spp jittered samples.  A
    cached frame is a jittered cached_spp-sample pattern (0: plain_spp) as one
    wavefront pass whose paths end into the cache at vertex cache_vertex (after
    two bounces for 2), followed by train_steps Adam steps on its training
    samples, which overlap the next frame's traversal.  Cached frames are
    rendered and accumulated while the projected time of the next one stays
    within the budget (at least one frame).  Warm: the cache was trained on
    pretrain_frames earlier frames (as in an animation); cold: a new cache.
    Returns the linear images (scored against the CPU reference later), the
    times, and a full-image MSE against a ref_spp GPU render (gpu_ref: that
    image from an earlier call, not rendered again)."""
    from raycer_amd import api

    cfg = api.RenderCfg.from_buffer_copy(base_cfg)
    cfg.width, cfg.height = base_cfg.width // scale, base_cfg.height // scale
    cfg.spp = plain_spp
    cam = sd.desc.camera
    film = ctx.film(cfg.width, cfg.height)

    def lin():
        return film.linear()[:, :3].astype(np.float32)

    out = {"resolution": f"{cfg.width}x{cfg.height}", "plain_spp": plain_spp, "cache_vertex": cache_vertex,
           "train_steps_per_frame": train_steps}
    images = {}
    try:
        if gpu_ref is None:
            film.clear()
            for f in range(ref_spp // plain_spp):
                cfg.frame = 100000 + f
                ctx.render(scene, film, cam, cfg)
            gpu_ref = lin()
        images["gpu_ref"] = gpu_ref
        film.clear()
        cfg.frame = 7
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render(scene, film, cam, cfg)
        ctx.synchronize()
        budget = time.perf_counter() - t0
        out["budget_ms"] = round(budget * 1e3, 2)
        images["plain"] = lin()
        ccfg = api.RenderCfg.from_buffer_copy(cfg)
        ccfg.cache_vertex = cache_vertex
        ccfg.spp = cached_spp or plain_spp
        out["cached_spp_per_frame"] = int(ccfg.spp)
        frame = [200000]

        def run(cache):
            frames = 0
            ctx.synchronize()
            film.clear()
            t0 = time.perf_counter()
            while True:
                ccfg.frame = frame[0]
                frame[0] += 1
                cache.render(scene, film, cam, ccfg, train=train_steps, want_loss=False)
                frames += 1
                elapsed = time.perf_counter() - t0
                if elapsed * (frames + 1) / frames > budget:
                    break
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            st = ctx.stats()
            return {"frames": frames, "spp": frames * int(ccfg.spp), "ms": round(ms, 2),
                    "ms_per_frame": round(ms / frames, 2), "train_samples_last_frame": int(st.train_samples),
                    "cache_queries_last_frame": int(st.cache_queries)}, lin()

        # one untimed cached frame with a throwaway cache first: the context's
        # query and record buffers are allocated on first use, which is not
        # what a cold (untrained) cache costs per frame
        scratch = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, lr=1e-3, seed=2, max_batch=65536)
        try:
            ccfg.frame = 150000
            scratch.render(scene, film, cam, ccfg, train=train_steps, want_loss=False)
            ctx.synchronize()
        finally:
            scratch.close()
        cold = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, lr=1e-3, seed=1, max_batch=65536)
        try:
            out["cached_cold"], images["cached_cold"] = run(cold)
        finally:
            cold.close()
        warm = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, lr=1e-3, seed=1, max_batch=65536)
        try:
            pcfg = api.RenderCfg.from_buffer_copy(ccfg)
            pcfg.spp = pretrain_spp
            pcfg.sampler = api.RC_SAMPLER_RANDOM
            t0 = time.perf_counter()
            for f in range(pretrain_frames):
                pcfg.frame = 300000 + f
                film.clear()
                loss = warm.render(scene, film, cam, pcfg, train=pretrain_steps, want_loss=f == pretrain_frames - 1)
            out["pretrain"] = {"frames": pretrain_frames, "spp": pretrain_spp, "steps_per_frame": pretrain_steps,
                               "s": round(time.perf_counter() - t0, 2), "last_loss": round(loss, 5)}
            out["cached_warm"], images["cached_warm"] = run(warm)
            # one more warm frame with the kernel timers on: the learned
            # kernels' time and FLOP rates (BF16 MFMA, FP32 accumulation)
            ctx.set_profiling(True)
            ctx.stats()
            film.clear()
            ccfg.frame = frame[0]
            warm.render(scene, film, cam, ccfg, train=train_steps, want_loss=False)
            st = ctx.stats()
            ctx.set_profiling(False)
            infer_flop = st.cache_queries * MLP_INFER_FLOP
            train_flop = train_steps * st.train_samples * MLP_TRAIN_FLOP
            out["mlp_kernels"] = {
                "infer_queries": int(st.cache_queries), "infer_ms": round(st.cache_ms, 4),
                "infer_tflops": round(infer_flop / (st.cache_ms * 1e-3) / 1e12, 2) if st.cache_ms else None,
                "infer_frac_of_bf16_peak": round(infer_flop / (st.cache_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)
                if st.cache_ms else None,
                "train_samples": int(st.train_samples), "train_ms": round(st.train_ms, 4),
                "train_tflops": round(train_flop / (st.train_ms * 1e-3) / 1e12, 2) if st.train_ms else None,
                "train_frac_of_bf16_peak": round(train_flop / (st.train_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)
                if st.train_ms else None,
                "peak_bf16_tflops": BF16_PEAK_TFLOPS,
                "frame_ms": round(st.render_ms, 2),
                "note": "algorithmic FLOPs (35,328 per query; 95,616 per training sample) over the kernels' HIP-event "
                        "time; MFMA busy and BF16 MOPS counters are in profiles/"}
        finally:
            warm.close()
    finally:
        film.close()
    ref = images["gpu_ref"].astype(np.float64)
    for k in ("plain", "cached_cold", "cached_warm"):
        out.setdefault(k, {})
        out[k]["mse_linear_vs_gpu_ref"] = mse_pair(images[k], ref)
    for k in ("cached_cold", "cached_warm"):
        out[k]["ratio_vs_gpu_ref"] = round(out[k]["mse_linear_vs_gpu_ref"] / out["plain"]["mse_linear_vs_gpu_ref"], 4)
    return {"summary": out, "images": images, "budget_s": budget}


def score_against_cpu(forms: dict, cpu: dict) -> None:
    """Adds the MSE of each image against the CPU reference on its pixels, the
    cached/plain ratios, and the bound the split-tail variance puts on them."""
    ids, ref = cpu["ids"], cpu["ref"]
    out = forms["summary"]
    m = {k: mse_pair(img[ids], ref) for k, img in forms["images"].items()}
    plain = m["plain"]
    out["mse_vs_cpu_ref"] = {
        "reference": f"FP64 CPU oracle, {len(ids)} seeded pixels, 4096 spp",
        "plain": plain, "cached_warm": m["cached_warm"], "cached_cold": m["cached_cold"],
        "gpu_4096spp_vs_cpu_4096spp": m["gpu_ref"],
        "gpu_4096spp_expected_from_noise": 2.0 * plain * out["plain_spp"] / 4096.0,
        "ratio_warm": round(m["cached_warm"] / plain, 4),
        "ratio_cold": round(m["cached_cold"] / plain, 4),
    }
    # the same per-sample variance at plain_spp
    tail = cpu["tail_var"][out["cache_vertex"]] * cpu["split_spp"] / out["plain_spp"]
    warm = out["cached_warm"]
    time_ratio = warm["ms_per_frame"] / out["budget_ms"]
    out["variance_bound"] = {
        "tail_variance": tail,
        "tail_share_of_plain_mse": round(tail / plain, 4),
        "ratio_bound_equal_spp": round(1.0 - tail / plain, 4),
        "cached_frame_time_ratio": round(time_ratio, 4),
        "ratio_bound_equal_time": round((1.0 - tail / plain) * time_ratio, 4),
        "ratio_warm_time_scaled": round(m["cached_warm"] / plain * time_ratio, 4),
        "note": "a perfect cache removes the variance of the light gathered after the cache vertex's next-event "
                "ray (split-tail estimate on the CPU reference's pixels); at equal samples the cached MSE is at "
                "least (1 - share) x plain; at equal time the cached form renders 1 / frame-time-ratio as many "
                "samples",
    }


def phong_c2(ctx, reps: int = 10) -> dict:
    """Primary ray + one point-light shadow ray per pixel with Phong shading
    (rc_raycast_phong) on the seeded sphere + ground scene (655,360 + 524,288
    triangles), 1920x1080, 1 jittered sample.  The device time is the sum of
    the kernels' HIP-event times; the wall time includes copying the hit
    indices, distances and RGBA image back to page-locked host arrays (the
    entry point's contract; the index and distance copies overlap the shadow
    rays' traversal)."""
    import raycer_amd as ra
    from raycer_amd import api

    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    s = ctx.upload(sd, nodes, ordered)
    w, h = sd.desc.cfg.width, sd.desc.cfg.height
    npix = w * h
    # the outputs in page-locked host memory, as a caller that wants the host
    # link's full rate passes them
    bufs = [ctx.pinned((npix,), np.int32), ctx.pinned((npix,), np.float32), ctx.pinned((npix, 4), np.float32)]
    out = [b.array for b in bufs]
    try:
        ctx.raycast_phong(s, sd.desc.camera, w, h, 1, sd.desc.cfg.ray_start_offset, sd.desc.light, out=out)  # warm-up
        ctx.set_profiling(True)
        ctx.stats()
        dev_ms = 0.0
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.raycast_phong(s, sd.desc.camera, w, h, 1, sd.desc.cfg.ray_start_offset, sd.desc.light, out=out)
            st = ctx.stats()
            dev_ms += st.extend_ms + st.shadow_ms + st.shade_ms + st.other_ms
        wall = (time.perf_counter() - t0) / reps
        ctx.set_profiling(False)
    finally:
        for b in bufs:
            b.close()
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
    log-uniform; 128^2 to 1024^2; 32 jittered spp; depth 5) in one captured
    HIP graph.  The autotuner picks, per size class, samples per pass (the
    wavefront queue length) and the kernels' block shapes, then the number of
    concurrent graph branches; the first batch call captures and instantiates
    the graph (graph_ms), the timed call relaunches it."""
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
            plans, branches = ctx.autotune_batch_plan(scenes, cams, cfgs, os.path.join(d, "tune.json"))
            tune = time.perf_counter() - t0
        for f in films:
            f.clear()
        ctx.render_batch_tuned(scenes, films, cams, cfgs, plans, branches)  # captures and instantiates
        first = ctx.stats()
        for f in films:
            f.clear()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.render_batch_tuned(scenes, films, cams, cfgs, plans, branches)  # relaunches the graph
        ctx.synchronize()
        dt = time.perf_counter() - t0
        st = ctx.stats()
    finally:
        for f in films:
            f.close()
        for s in scenes:
            s.close()
    samples = sum(c.width * c.height * c.spp for c in cfgs)
    by_res = {}
    for c, p in zip(cfgs, plans):
        by_res[f"{c.width}x{c.height}"] = {"spp_per_pass": int(p.spp_per_pass), "ext_block": int(p.ext_block),
                                          "shadow_block": int(p.shadow_block)}
    return {"scenes": n_scenes, "triangles": tris, "samples": samples, "setup_s": round(setup, 2),
            "autotune_ms": round(tune * 1e3, 1), "branches": int(branches),
            "graph_capture_instantiate_ms": round(first.graph_ms, 1),
            "first_launch_ms": round(first.render_ms, 1), "relaunched": bool(st.graph_relaunched),
            "render_ms": round(dt * 1e3, 1), "msamples_per_s": round(samples / dt / 1e6, 1), "plans": by_res}


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
    ap.add_argument("--park-steps", type=int, default=None,
                    help="traversal steps before an extension ray is parked for the resume kernel (0: never; "
                         "default: the library's)")
    ap.add_argument("--hit", choices=("fp32", "fp64"), default="fp32",
                    help="triangle decision: FP32 Möller–Trumbore (default) or FP32 filter + FP64 decision")
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
    ctx.set_hit_decision(a.hit)
    scene = ctx.upload(sd, nodes, ordered)
    film = ctx.film(cfg.width, cfg.height)
    log(f"device {name} ({cus} CUs), setup {time.perf_counter() - t0:.1f} s")

    ctx.set_tuning(spp_per_pass=a.spp_per_pass, park_steps=a.park_steps)
    copy_gbs = ctx.copy_bandwidth(1 << 30, 20)
    probes = {"l2_4MiB": ctx.read_bandwidth(4 << 20, 512), "infinity_cache_128MiB": ctx.read_bandwidth(128 << 20, 16),
              "hbm_4GiB": ctx.read_bandwidth(4 << 30, 1),
              # 64-byte records at distinct random addresses of 4 GiB (node-sized, no reuse)
              "gather_64B_4GiB": ctx.gather_bandwidth(4 << 30, 1 << 24, 3)}
    log(f"streaming copy {copy_gbs:.0f} GB/s, reads {json.dumps({k: round(v) for k, v in probes.items()})} GB/s")

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
    total_ext = total_shadow = overflow = parked = 0
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
        overflow += st.overflow_rays
        parked += st.parked_rays
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    ctx.set_profiling(False)
    samples = cfg.width * cfg.height * cfg.spp * a.steps
    value = samples / elapsed / 1e6
    # algorithmic bytes per frame scale with the frame's rays; the counting
    # frame used frame index 0 like the first timed frame
    bytes_per_ray = alg_bytes / ext_rays
    avg_launch_s = ext_ms / max(ext_launches, 1) * 1e-3
    demand = bytes_per_ray * total_ext / (ext_ms * 1e-3) / 1e9
    traffic = None
    if PROFILE_TRAFFIC.exists():
        try:
            traffic = json.loads(PROFILE_TRAFFIC.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    # achieved: the HBM bytes rocprofv3 counts per launch of this kernel
    # (FETCH_SIZE + WRITE_SIZE, calibrated on a copy of known size) over the
    # launch time measured here; without that profile, the algorithmic bytes
    achieved = traffic / avg_launch_s / 1e9 if traffic else demand
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
            "hit_decision": a.hit,
            "samples_per_pass": min(a.spp_per_pass, cfg.spp) if a.spp_per_pass else "auto",
            "park_steps": a.park_steps if a.park_steps is not None else api.PARK_STEPS_DEFAULT,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "extension step: k_extend_shade<false> (closest-hit traversal fused with shading) and the "
                      "k_extend_resume kernels that finish its parked traversals; one step per bounce",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "achieved_basis": "rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per extension step (profiles/) over the "
                              "step's live time (HIP events)" if traffic else "algorithmic bytes (no counter profile "
                                                                              "found)",
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "achievable_copy_GBs": round(copy_gbs, 1),
            "frac_of_achievable_copy": round(achieved / copy_gbs, 4),
            # every HBM read is a 128-byte line request (rocprofv3 TCC_EA0_RDREQ_128B, profiles/), so
            # the scattered-read probe's 64-byte records move two bytes of lines per byte requested
            "achievable_gather_lines_GBs": round(2.0 * probes["gather_64B_4GiB"], 1),
            "frac_of_achievable_gather": round(achieved / (2.0 * probes["gather_64B_4GiB"]), 4),
            "demand": {
                "bytes_per_ray": round(bytes_per_ray, 1),
                "work_per_ray": {"wide_node_visits": round(cst.node_records / 4 / ext_rays, 2),
                                 "triangle_tests": round(cst.tri_tests / ext_rays, 2)},
                "bytes_per_launch": round(bytes_per_ray * total_ext / max(ext_launches, 1)),
                "GBs": round(demand, 1),
                "frac_of_hbm_peak": round(demand / HBM_PEAK_GBS, 4),
                "read_probe_GBs": {k: round(v, 1) for k, v in probes.items()},
                "frac_of_infinity_cache_read": round(demand / probes["infinity_cache_128MiB"], 4),
                "note": "node, triangle and queue bytes the traversal asks for (the algorithmic bytes per ray); "
                        "L2 and the 256 MB Infinity Cache, which holds the whole scene, serve most of them",
            },
        },
        "gsample_bounces_per_s": round(total_ext / elapsed / 1e9, 4),
        "shadow_rays_per_s_G": round(total_shadow / elapsed / 1e9, 4),
        "overflow_rays_per_step": overflow // a.steps,
        "parked_rays_per_step": parked // a.steps,
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
        log("64-scene batch (autotuned plans and branches, one captured graph, relaunched)")
        out["batch_64_scenes"] = batch_c4(ctx)
        log(f"batch: {json.dumps(out['batch_64_scenes'])}")
    forms = forms2 = None
    if a.mse != "off":
        log(f"plain and radiance-cached forms at equal time ({a.mse})")
        scale = 1 if a.mse == "full" else 4
        # paths end into the cache after two bounces (vertex 2); beside it the
        # same cache one bounce earlier (vertex 1, the first bounce's hit)
        forms = cached_forms(ctx, scene, sd, cfg, scale=scale, cache_vertex=2)
        forms2 = cached_forms(ctx, scene, sd, cfg, scale=scale, cache_vertex=1, gpu_ref=forms["images"]["gpu_ref"])
        out["equal_time_mse"] = forms["summary"]
        out["equal_time_mse"]["cache_vertex_1"] = forms2["summary"]
        if PROFILE_MFMA.exists():
            try:  # the MFMA counters of the same kernels (rocprofv3, profiles/)
                forms["summary"]["mlp_kernels"]["mfma_counters"] = json.loads(PROFILE_MFMA.read_text())["kernels"]
            except (OSError, ValueError, KeyError):
                pass
        log(f"cached forms: {json.dumps(forms['summary'])}")
        # the same comparison in the closed box (emitter-lit, up to 8 rays per
        # path), where every path reaches the cache vertex
        bd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_BOX))
        bn, bo, _, _ = ra.build_bvh(bd.vertices, method=api.RCH_BVH_BINNED)
        box = ctx.upload(bd, bn, bo)
        try:
            out["equal_time_mse_box"] = cached_forms(ctx, box, bd, bd.desc.cfg, scale=1 if a.mse == "full" else 2,
                                                     plain_spp=16, pretrain_frames=64)["summary"]
        finally:
            box.close()
        log(f"cached forms (box): {json.dumps(out['equal_time_mse_box'])}")
    scene.close()
    c1 = c1_gpu_image(ctx) if not a.no_cpu_baseline else None
    ctx.close()
    if not a.no_cpu_baseline:
        log("cpu leg: oracle 4096-spp reference on 1/64 of the pixels (the CPU baseline) and the split-tail bound")
        ocfg = api.RenderCfg.from_buffer_copy(cfg)
        if forms is not None:
            ocfg.width, ocfg.height = (cfg.width // (1 if a.mse == "full" else 4),
                                       cfg.height // (1 if a.mse == "full" else 4))
        cpu = cpu_reference(sd, cam, ocfg)
        out["cpu_baseline"] = cpu["baseline"]
        if forms is not None:
            for f in (forms, forms2):
                score_against_cpu(f, cpu)
                log(f"cache vertex {f['summary']['cache_vertex']} vs CPU reference: "
                    f"{json.dumps(f['summary']['mse_vs_cpu_ref'])}")
                log(f"variance bound: {json.dumps(f['summary']['variance_bound'])}")
        out["c1_mse_vs_cpu_ref"] = c1_vs_cpu(*c1)
        log(f"C1 vs CPU reference: {json.dumps(out['c1_mse_vs_cpu_ref'])}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
