"""Radiance cache on the GPU: BF16 MFMA inference and training against the
FP64 oracle and the NumPy model of the kernels' BF16 roundings, and the
cached path tracer.

Tolerances: the GPU and the BF16 model round the same values to bfloat16 and
differ only by FP32 vs FP64 accumulation (a few rounding flips), so they are
compared tightly; against the FP64 oracle the BF16 error itself is allowed
(output log-radiance within 0.03, gradient blocks within 6 % relative L2 —
the measured BF16 model error is 0.011 and 3.2 %).
"""
import numpy as np
import pytest

import oracle as orc
from raycer_amd import api
from tests import cache_ref as cr
from tests.test_gpu_render import cfg_for, render_gpu, scene

pytestmark = pytest.mark.gpu

LO, HI = np.array([-1.0, 0.0, -2.0]), np.array([3.0, 1.5, 2.0])


def make_cache(ctx, **kw):
    return api.RadianceCache(ctx, LO, HI, **kw)


@pytest.mark.parametrize("n", [1, 15, 4093])
def test_infer_matches_bf16_model_and_oracle(ctx, n):
    p = cr.random_params(5)
    q = cr.random_queries(n, LO, HI, 1)
    cache = make_cache(ctx)
    try:
        cache.set_params(p)
        np.testing.assert_array_equal(cache.params(), p)
        z = np.log(cache.infer(q).astype(np.float64))
    finally:
        cache.close()
    enc = cr.encode(q, LO, HI)
    zb = cr.forward_bf16(p, enc)
    z64 = np.log(orc.mlp_forward(p.astype(np.float64), enc, 81, 64, 4, 3))
    d_model, d_oracle = np.abs(z - zb), np.abs(z - z64)
    print(f"n={n}: |z - bf16 model| max {d_model.max():.2e} mean {d_model.mean():.2e}; "
          f"|z - fp64 oracle| max {d_oracle.max():.2e} median {np.median(d_oracle):.2e}")
    assert d_model.max() < 5e-3 and d_model.mean() < 2e-4
    assert d_oracle.max() < 0.03 and np.median(d_oracle) < 0.005


def test_default_initialisation_starts_near_half(ctx):
    cache = make_cache(ctx, seed=7)
    try:
        y = cache.infer(cr.random_queries(1000, LO, HI, 3))
        p = cache.params()
    finally:
        cache.close()
    assert p.shape == (cr.N_PARAMS,) and np.isfinite(p).all()
    np.testing.assert_allclose(cr.bias(p, 4), np.log(0.5), rtol=1e-6)
    assert np.abs(cr.weights(p, 4)).max() <= 0.01
    assert ((y > 0.35) & (y < 0.7)).all()


def test_gradient_matches_bf16_model_and_oracle(ctx):
    p = cr.random_params(5)
    n = 4093
    q = cr.random_queries(n, LO, HI, 1)
    t = (np.exp(np.random.default_rng(9).normal(size=(n, 3))) * 0.5).astype(np.float32)
    cache = make_cache(ctx)
    try:
        cache.set_params(p)
        g, loss = cache.gradient(q, t)
        g2, _ = cache.gradient(q, t)
    finally:
        cache.close()
    np.testing.assert_array_equal(g, g2)  # fixed-order reductions
    enc = cr.encode(q, LO, HI)
    lb, gb = cr.loss_grad_bf16(p, enc, t)
    l64, g64 = orc.mlp_loss_grad(p.astype(np.float64), enc, t.astype(np.float64), 81, 64, 4, 3, rel=True, eps=1e-2)
    assert abs(loss - lb) < 2e-3 * lb and abs(loss - l64) < 1e-2 * l64
    for l in range(5):
        for name, s in (("W", slice(cr.OFF_W[l], cr.OFF_B[l])), ("b", slice(cr.OFF_B[l], cr.OFF_B[l] + cr.FAN_OUT[l]))):
            rm = np.linalg.norm(g[s] - gb[s]) / np.linalg.norm(gb[s])
            ro = np.linalg.norm(g[s] - g64[s]) / np.linalg.norm(g64[s])
            cos = float(g[s] @ g64[s] / (np.linalg.norm(g[s]) * np.linalg.norm(g64[s])))
            print(f"layer {l} {name}: rel L2 vs bf16 model {rm:.2e}, vs fp64 {ro:.2e}, cos {cos:.6f}")
            assert rm < 1e-2, (l, name, rm)
            assert ro < 0.06 and cos > 0.998, (l, name, ro, cos)


def test_adam_steps_match_oracle(ctx):
    p = cr.random_params(6)
    q = cr.random_queries(2048, LO, HI, 4)
    t = cr.smooth_target(q, LO, HI)
    cache = make_cache(ctx, lr=1e-3)
    p64 = p.astype(np.float64)
    m = np.zeros_like(p64)
    v = np.zeros_like(p64)
    try:
        cache.set_params(p)
        for step in range(1, 4):
            g, _ = cache.gradient(q, t)
            cache.train_step(q, t)
            p64, m, v = orc.adam_step(p64, m, v, g.astype(np.float64), lr=1e-3, b1=0.9, b2=0.99, eps=1e-8, t=step)
            got = cache.params()
            err = np.abs(got - p64).max()
            print(f"adam step {step}: max |p_gpu - p_oracle| {err:.2e}")
            assert err < 2e-6
    finally:
        cache.close()


def test_training_fits_a_smooth_target(ctx):
    cache = make_cache(ctx, seed=11, lr=1e-3)
    qv = cr.random_queries(4096, LO, HI, 99)
    tv = cr.smooth_target(qv, LO, HI)
    try:
        _, l0 = cache.gradient(qv, tv)
        for step in range(300):
            q = cr.random_queries(8192, LO, HI, 1000 + step)
            cache.train_step(q, cr.smooth_target(q, LO, HI))
        _, l1 = cache.gradient(qv, tv)
    finally:
        cache.close()
    print(f"validation loss {l0:.4f} -> {l1:.4f} after 300 Adam steps")
    assert l1 < 0.1 * l0


def terrain_cfg(sd, w=96, h=96, spp=16, **kw):
    return cfg_for(sd, w, h, spp, **kw)


@pytest.mark.parametrize("cv", [1, 2])
def test_cached_render_with_a_zero_cache_equals_the_truncated_render(ctx, cv):
    """A cache that returns exactly 0: every path ends after its vertex-cv
    next-event ray, which (without sky) is what a plain render with at most
    cv + 2 rays per path gathers — bit for bit."""
    sd, s, _ = scene(ctx, "terrain")
    b0, b1 = np.asarray(sd.desc.bounds_min), np.asarray(sd.desc.bounds_max)
    cache = api.RadianceCache(ctx, b0, b1, train_stride=0xFFFFFFFF)
    p = cache.params()
    p[cr.OFF_W[4]:cr.OFF_B[4]] = 0.0
    p[cr.OFF_B[4]:] = -200.0  # exp underflows to exactly 0 in FP32
    cache.set_params(p)
    cfg = terrain_cfg(sd, cache_vertex=cv)
    cfg.sky[0] = cfg.sky[1] = cfg.sky[2] = 0.0
    film = ctx.film(cfg.width, cfg.height)
    try:
        cache.render(s, film, sd.desc.camera, cfg, train=False)
        st = ctx.stats()
        cached = film.sums()
    finally:
        film.close()
        cache.close()
    plain_cfg = terrain_cfg(sd, max_path_length=cv + 2)
    plain_cfg.sky[0] = plain_cfg.sky[1] = plain_cfg.sky[2] = 0.0
    _, plain = render_gpu(ctx, s, sd.desc.camera, plain_cfg)
    print(f"cache queries {st.cache_queries}, training samples {st.train_samples}")
    assert st.cache_queries > 0 and st.train_samples == 0
    np.testing.assert_array_equal(cached, plain)


@pytest.fixture(scope="module")
def workload_terrain(ctx):
    """The 2,093,058-triangle terrain of the 1080p workload (32-bin SAH BVH)."""
    import raycer_amd as ra

    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_TERRAIN, seed=1, grid=1024))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, method=api.RCH_BVH_BINNED)
    s = ctx.upload(sd, nodes, ordered)
    yield sd, s
    s.close()


@pytest.mark.slow
def test_zero_cache_identity_at_the_workload(ctx, workload_terrain):
    """The zero-cache identity on the cached workload itself: the
    2,093,058-triangle terrain at 1920 x 1080, cache vertex 2, the 65,536-sample
    training batch.  Without training paths the cached film equals the plain
    render truncated after 4 rays bit for bit; with them (one Adam step of
    65,536 samples per frame; a default-initialised cache, since the zero
    cache's saturated exp has no gradient) the frame takes 65,536 training
    samples, its loss is finite and the cache has moved."""
    sd, s = workload_terrain
    cfg = cfg_for(sd, 1920, 1080, 16, cache_vertex=2)
    cfg.sky[0] = cfg.sky[1] = cfg.sky[2] = 0.0
    assert sd.ntri == 2093058
    b0, b1 = np.asarray(sd.desc.bounds_min), np.asarray(sd.desc.bounds_max)
    film = ctx.film(cfg.width, cfg.height)
    zero = api.RadianceCache(ctx, b0, b1, train_stride=0xFFFFFFFF, max_batch=65536)
    # one path in 16 trains (2,073,600 at 1080p x 16 spp, of which ~11% reach the cache vertex; most leave for
    # the sky first): the batch fills to its 65,536 samples
    trained = api.RadianceCache(ctx, b0, b1, max_batch=65536, train_stride=16)
    try:
        p = zero.params()
        p[cr.OFF_W[4]:cr.OFF_B[4]] = 0.0
        p[cr.OFF_B[4]:] = -200.0  # exp underflows to exactly 0 in FP32
        zero.set_params(p)
        zero.render(s, film, sd.desc.camera, cfg, train=False)
        st = ctx.stats()
        cached = film.sums()
        film.clear()
        p0 = trained.params()
        loss = trained.render(s, film, sd.desc.camera, cfg, train=1)
        st2 = ctx.stats()
        p1 = trained.params()
    finally:
        film.close()
        zero.close()
        trained.close()
    plain_cfg = cfg_for(sd, 1920, 1080, 16, max_path_length=4)
    plain_cfg.sky[0] = plain_cfg.sky[1] = plain_cfg.sky[2] = 0.0
    _, plain = render_gpu(ctx, s, sd.desc.camera, plain_cfg)
    print(f"1080p x 16 spp: {st.cache_queries} cache queries; trained frame: {st2.train_samples} samples, "
          f"loss {loss:.4f}")
    assert st.cache_queries > 1920 * 1080 and st.train_samples == 0
    np.testing.assert_array_equal(cached, plain)
    assert st2.train_samples == 65536 and np.isfinite(loss) and loss > 0
    assert np.isfinite(p1).all() and (p1 != p0).any()


@pytest.mark.slow
def test_gradient_at_the_training_batch(ctx):
    """The gradient at the cached workload's 65,536-sample training batch (the
    weight-gradient kernel's full split-K depth) against the BF16 model and
    the FP64 oracle, as at 4,093 samples."""
    p = cr.random_params(7)
    n = 65536
    q = cr.random_queries(n, LO, HI, 2)
    t = (np.exp(np.random.default_rng(10).normal(size=(n, 3))) * 0.5).astype(np.float32)
    cache = make_cache(ctx, max_batch=n)
    try:
        cache.set_params(p)
        g, loss = cache.gradient(q, t)
        g2, _ = cache.gradient(q, t)
    finally:
        cache.close()
    np.testing.assert_array_equal(g, g2)
    enc = cr.encode(q, LO, HI)
    lb, gb = cr.loss_grad_bf16(p, enc, t)
    l64, g64 = orc.mlp_loss_grad(p.astype(np.float64), enc, t.astype(np.float64), 81, 64, 4, 3, rel=True, eps=1e-2)
    assert abs(loss - lb) < 2e-3 * lb and abs(loss - l64) < 1e-2 * l64
    for l in range(5):
        for name, sl in (("W", slice(cr.OFF_W[l], cr.OFF_B[l])),
                         ("b", slice(cr.OFF_B[l], cr.OFF_B[l] + cr.FAN_OUT[l]))):
            rm = np.linalg.norm(g[sl] - gb[sl]) / np.linalg.norm(gb[sl])
            ro = np.linalg.norm(g[sl] - g64[sl]) / np.linalg.norm(g64[sl])
            cos = float(g[sl] @ g64[sl] / (np.linalg.norm(g[sl]) * np.linalg.norm(g64[sl])))
            print(f"n = {n}, layer {l} {name}: rel L2 vs bf16 model {rm:.2e}, vs fp64 {ro:.2e}, cos {cos:.6f}")
            assert rm < 1e-2, (l, name, rm)
            assert ro < 0.06 and cos > 0.998, (l, name, ro, cos)


def test_cached_render_rejects_bad_cache_vertex(ctx):
    sd, s, _ = scene(ctx, "terrain")
    cache = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max)
    film = ctx.film(32, 32)
    try:
        for cv in (0, 5, 6):  # max_path_length 6: needs 1 <= cv < 5
            cfg = terrain_cfg(sd, 32, 32, 4, cache_vertex=cv)
            with pytest.raises(api.RaycerError):
                cache.render(s, film, sd.desc.camera, cfg)
    finally:
        film.close()
        cache.close()


def block_means(img, w, h, b=8):
    return img.reshape(h // b, b, w // b, b, -1).mean(axis=(1, 3))


def test_online_training_converges_to_the_path_traced_image(ctx):
    """Paths end into the cache after 2 bounces; the cache trains one step per
    frame on the suffix radiance of the training paths.  After training, the
    cached image must agree with a converged plain render in mean, and its
    8x8-block means (noise averaged out, bias left) must be closer to the
    reference's than with the untrained cache.  (The per-frame training loss is
    not monotone: its targets are one-path estimates.)"""
    sd, s, _ = scene(ctx, "terrain")
    cache = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, seed=3, max_batch=16384)
    cfg = terrain_cfg(sd, cache_vertex=2)
    ev = terrain_cfg(sd, spp=256, cache_vertex=2)
    ev.frame = 5000
    film = ctx.film(cfg.width, cfg.height)

    def evaluate():
        film.clear()
        cache.render(s, film, sd.desc.camera, ev, train=False)
        return film.linear()[:, :3].astype(np.float64)

    try:
        untrained = evaluate()
        for f in range(300):
            cfg.frame = f
            film.clear()
            cache.render(s, film, sd.desc.camera, cfg, train=True)
            if f == 0:
                st = ctx.stats()
                assert st.cache_queries > 0 and 0 < st.train_samples <= 16384
        trained = evaluate()
    finally:
        film.close()
        cache.close()
    ref, _ = render_gpu(ctx, s, sd.desc.camera, terrain_cfg(sd, spp=4096))
    ref = ref[:, :3].astype(np.float64)
    w, h = cfg.width, cfg.height
    br = block_means(ref, w, h)
    b0 = np.mean((block_means(untrained, w, h) - br) ** 2)
    b1 = np.mean((block_means(trained, w, h) - br) ** 2)
    r1 = trained.mean() / ref.mean()
    print(f"trained image mean / reference {r1:.4f}; 8x8-block MSE untrained {b0:.3e} -> trained {b1:.3e}")
    assert np.isfinite(trained).all()
    assert abs(r1 - 1.0) < 0.03
    assert b1 < b0


def test_several_adam_steps_per_frame(ctx):
    """train = k runs k Adam steps on interleaved slices of the frame's
    training samples: parameters move k times, the loss is finite."""
    sd, s, _ = scene(ctx, "terrain")
    cfg = terrain_cfg(sd, cache_vertex=2)
    film = ctx.film(cfg.width, cfg.height)
    caches = [api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, seed=5, max_batch=8192)
              for _ in range(2)]
    try:
        p0 = caches[0].params()
        l1 = caches[0].render(s, film, sd.desc.camera, cfg, train=1)
        l4 = caches[1].render(s, film, sd.desc.camera, cfg, train=4)
        p1, p4 = caches[0].params(), caches[1].params()
    finally:
        film.close()
        for c in caches:
            c.close()
    assert np.isfinite(l1) and np.isfinite(l4) and l1 > 0 and l4 > 0
    # Adam's first step moves every parameter with a non-zero gradient by
    # lr |g| / (|g| + eps) ~ lr; four steps move most of them further
    d1, d4 = np.abs(p1 - p0), np.abs(p4 - p0)
    m1, m4 = np.median(d1[d1 > 0]), np.median(d4[d4 > 0])
    print(f"median non-zero |dp|: 1 step {m1:.2e}, 4 steps {m4:.2e}")
    assert 0.9e-3 < m1 <= 1.0001e-3
    assert m4 > 1.5e-3


def test_trained_cache_beats_the_plain_render_at_equal_samples(ctx):
    """What the trained network gains over the method without it: in the
    closed box (emitter-lit, up to 8 rays, every path reaches vertex 2) a
    16-spp frame whose paths end into a cache trained on 64 earlier 4-spp
    frames has a lower MSE against a 4096-spp plain render than a 16-spp
    plain frame (the bench measures 0.78x at this size, 256 x 256; at 128 x
    128 the 4-spp frames hold too few training paths and the cache loses,
    1.09x)."""
    sd, s, _ = scene(ctx, "box")
    cam = sd.desc.camera
    ref, _ = render_gpu(ctx, s, cam, cfg_for(sd, 256, 256, 16), frames=256)
    plain_cfg = cfg_for(sd, 256, 256, 16)
    plain_cfg.frame = 9000
    film = ctx.film(256, 256)
    cache = api.RadianceCache(ctx, sd.desc.bounds_min, sd.desc.bounds_max, lr=1e-3, seed=1, max_batch=65536)
    try:
        film.clear()
        ctx.render(s, film, cam, plain_cfg)
        plain = film.linear()[:, :3].astype(np.float64)
        pre = cfg_for(sd, 256, 256, 4, cache_vertex=2, sampler=api.RC_SAMPLER_RANDOM)
        for f in range(64):
            pre.frame = 300000 + f
            film.clear()
            cache.render(s, film, cam, pre, train=4, want_loss=False)
        ccfg = cfg_for(sd, 256, 256, 16, cache_vertex=2)
        ccfg.frame = 200000
        film.clear()
        cache.render(s, film, cam, ccfg, train=1, want_loss=False)
        cached = film.linear()[:, :3].astype(np.float64)
    finally:
        film.close()
        cache.close()
    ref = ref[:, :3].astype(np.float64)
    m_plain = float(np.mean((plain - ref) ** 2))
    m_cached = float(np.mean((cached - ref) ** 2))
    print(f"MSE vs 4096-spp plain: plain 16 spp {m_plain:.4e}, trained cache 16 spp {m_cached:.4e} "
          f"(ratio {m_cached / m_plain:.3f})")
    assert np.isfinite(cached).all()
    assert m_cached < 0.9 * m_plain
