"""Oracle pins for the reference's filters, samplers and tone mappers (CPU only).

None of these has a golden vector in the reference (FilterTest.cpp and
SamplerTest.cpp only write plots), so each is pinned by what the mathematics
fixes:
  * filters (src/Rendering/Filters/): partition of unity of the bell
    (quadratic B-spline) and of Mitchell-Netravali for any B, C; unit integral;
    the Gaussian's radius is its full width at a thousandth of the maximum; the
    Lanczos window vanishes at the integers and outside its radius;
  * CMJ (CMJSampler.cpp:12-77): permute is a bijection of [0, l) for any l and
    any key p <= 2^64 - l (above that the reference's final (i + p) % l wraps
    in 64 bits and two indices can meet, as they do there; a random key lands
    there with probability below 2^-56), and the n x n pattern is multi-jittered (one point per cell, and one
    per fine row and fine column) -- swapping the two permutations breaks it;
  * samplers as generateMultiSamples uses them: the weights of the centre,
    regular and Poisson-disc samples are closed forms;
  * tone mappers (src/Rendering/ToneMappers/): linear scales by
    fastPow(2, exposure) (fastPow(2, 0) = 0.97100782394409180, computed from
    MathUtils.cpp:56-68), simple is c / (1 + c), passthrough copies, Reinhard's
    moving average follows MovingAverage.cpp:28-31.
"""
import math

import numpy as np
import pytest

import oracle as orc

BOX, TENT, BELL, GAUSS, MITCHELL, LANCZOS = range(6)


def w1(f, r, x, p0=0.0, p1=0.0):
    """1-D weight: the 2-D weight at (x, 0) divided by the weight at (0, 0)
    times the 1-D weight at 0, i.e. wx(x) * wy(0) / wy(0)."""
    return orc.filter_weight(f, r, x, 0.0, p0, p1) / math.sqrt(orc.filter_weight(f, r, 0.0, 0.0, p0, p1))


def integral(f, r, p0=0.0, p1=0.0, lo=-8.0, hi=8.0, n=16001):
    xs = np.linspace(lo, hi, n)
    ys = np.array([w1(f, r, x, p0, p1) for x in xs])
    return np.trapezoid(ys, xs)


def test_bell_closed_forms():
    assert w1(BELL, 0, 0.0) == 0.75
    assert w1(BELL, 0, 1.0) == pytest.approx(0.125)
    assert w1(BELL, 0, 1.5) == 0.0 and w1(BELL, 0, 2.0) == 0.0
    for x in (0.0, 0.1, 0.37, 0.5, 0.93):
        assert sum(w1(BELL, 0, x + k) for k in range(-3, 4)) == pytest.approx(1.0, abs=1e-12)
    assert integral(BELL, 0) == pytest.approx(1.0, abs=1e-6)
    assert orc.filter_radius(BELL, 0.5) == 1.5


@pytest.mark.parametrize("B,C", [(1 / 3, 1 / 3), (0.0, 0.5), (1.0, 0.0), (0.2, 0.7)])
def test_mitchell_partition_of_unity(B, C):
    for x in (0.0, 0.1, 0.37, 0.5, 0.93):
        assert sum(w1(MITCHELL, 0, x + k, B, C) for k in range(-3, 4)) == pytest.approx(1.0, abs=1e-12)
    assert w1(MITCHELL, 0, 0.0, B, C) == pytest.approx((6 - 2 * B) / 6)
    assert w1(MITCHELL, 0, 1.0, B, C) == pytest.approx(B / 6)
    assert w1(MITCHELL, 0, 2.0, B, C) == pytest.approx(0.0, abs=1e-13)
    assert w1(MITCHELL, 0, 2.01, B, C) == 0.0
    assert integral(MITCHELL, 0, B, C) == pytest.approx(1.0, abs=1e-6)
    assert orc.filter_radius(MITCHELL, 0.5, B, C) == 2.0


@pytest.mark.parametrize("sd", [0.1, 0.5, 1.0])
def test_gaussian_normalisation_and_radius(sd):
    assert w1(GAUSS, 0, 0.0, sd) == pytest.approx(1 / (math.sqrt(2 * math.pi) * sd))
    r = orc.filter_radius(GAUSS, 0.5, sd)
    assert r == pytest.approx(7.43384 * sd / 2)
    # the radius is the full width at a thousandth of the maximum
    assert w1(GAUSS, 0, r, sd) / w1(GAUSS, 0, 0.0, sd) == pytest.approx(1e-3, rel=1e-4)
    assert integral(GAUSS, 0, sd, lo=-12 * sd, hi=12 * sd) == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("a", [1, 2, 3, 6])
def test_lanczos_window(a):
    assert w1(LANCZOS, a, 0.0) == 1.0
    for k in range(1, a + 1):
        assert abs(w1(LANCZOS, a, float(k))) < 1e-15
    assert w1(LANCZOS, a, a + 0.01) == 0.0
    # sinc(x) sinc(x / a) at x = 1/2: (2 / pi) * sin(pi / 2a) / (pi / 2a)
    assert w1(LANCZOS, a, 0.5) == pytest.approx((2 / math.pi) * math.sin(math.pi / (2 * a)) / (math.pi / (2 * a)))
    assert orc.filter_radius(LANCZOS, float(a)) == a


def test_two_d_weight_is_the_product():
    for f, r, p0, p1 in [(BELL, 0, 0, 0), (GAUSS, 0, 0.4, 0), (MITCHELL, 0, 1 / 3, 1 / 3), (LANCZOS, 2, 0, 0)]:
        x, y = 0.3, -0.7
        assert orc.filter_weight(f, r, x, y, p0, p1) == pytest.approx(
            orc.filter_weight(f, r, x, 0, p0, p1) * orc.filter_weight(f, r, 0, y, p0, p1)
            / orc.filter_weight(f, r, 0, 0, p0, p1))


@pytest.mark.parametrize("p", [0, 1, 0xDEADBEEFCAFEBABE, 12345678901234567, 2 ** 64 - 256])
def test_cmj_permute_is_a_bijection(p):
    for n in (1, 2, 3, 5, 7, 8, 9, 16, 17, 64, 100, 256):
        assert sorted(orc.cmj_permute(i, n, p) for i in range(n)) == list(range(n))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8])
def test_cmj_pattern_is_multi_jittered(n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        p = int(rng.integers(0, 2 ** 63)) * 2 + int(rng.integers(0, 2))
        pts = np.array([orc.cmj_sample(x, y, n, p, rng.random(), rng.random()) for y in range(n) for x in range(n)])
        cells = [(int(px * n), int(py * n)) for px, py in pts]
        assert cells == [(x, y) for y in range(n) for x in range(n)]      # jittered: one per cell
        assert sorted(int(px * n * n) for px in pts[:, 0]) == list(range(n * n))  # one per fine column
        assert sorted(int(py * n * n) for py in pts[:, 1]) == list(range(n * n))  # one per fine row


def _one_pixel_weight(sampler, filt, radius, spp, sample, p0=0.0, p1=0.0):
    """The oracle renderer's filter weight of one camera sample (or_trace_path)."""
    v = np.array([[-5, -5, 3, 5, -5, 3, 0, 5, 3]], np.float32)
    b = orc.BVH(v)
    mats = np.zeros((1, 16))
    cfg = orc.RenderCfg()
    cfg.width, cfg.height, cfg.spp, cfg.sampler = 8, 8, spp, sampler
    cfg.max_path_length, cfg.filter, cfg.filter_radius, cfg.traversal = 1, filt, radius, 1
    cfg.filter_param[0], cfg.filter_param[1] = p0, p1
    cfg.ray_start_offset = 1e-5
    cam = np.array([0, 0, 0, 0, 0, 0, 60.0])
    _, w, _ = orc.trace_path(b, np.zeros(1, np.uint32), mats, cam, cfg, 9, sample)
    return w


def test_sampler_weights_closed_forms():
    CENTER, REGULAR, POISSON = 2, 3, 5
    # centre: offset 0 -> tent(0)^2 = 1 / r^2
    assert _one_pixel_weight(CENTER, TENT, 1.0, 16, 7) == pytest.approx(1.0)
    assert _one_pixel_weight(CENTER, TENT, 2.0, 16, 7) == pytest.approx(0.25)
    # regular: stratum (1, 1) of 4 x 4 -> offset ((1.5 / 4) - 0.5) * 2 = -0.25 -> tent 0.75 each
    assert _one_pixel_weight(REGULAR, TENT, 1.0, 16, 5) == pytest.approx(0.75 * 0.75)
    # stratum (3, 0): offsets 0.75, -0.75
    assert _one_pixel_weight(REGULAR, TENT, 1.0, 16, 3) == pytest.approx(0.25 * 0.25)
    # Poisson disc: getSample2D gives (0, 0) -> offset (-r, -r): box keeps it, tent weighs 0
    assert _one_pixel_weight(POISSON, BOX, 0.5, 16, 11) == pytest.approx(1.0)
    assert _one_pixel_weight(POISSON, TENT, 1.0, 16, 11) == 0.0
    # one sample per pixel: the centre with weight 1 whatever the filter
    assert _one_pixel_weight(REGULAR, GAUSS, 0.5, 1, 0, 0.3) == 1.0
    # bell with the centre sampler: bell(0)^2
    assert _one_pixel_weight(CENTER, BELL, 0.5, 4, 2) == pytest.approx(0.75 ** 2)


def _img(n=64, seed=0):
    rng = np.random.default_rng(seed)
    img = np.ones((n, 4), np.float32)
    img[:, :3] = rng.random((n, 3)).astype(np.float32) * 3
    return img


def test_linear_and_simple_tone_mappers():
    img = _img()
    out, la, ml = orc.tonemap(img, type=1, clamp=False, apply_gamma=False)
    scale = 0.97100782394409180  # fastPow(2, 0)
    assert orc.fastpow(2.0, 0.0) == scale
    np.testing.assert_array_equal(out[:, :3], (img[:, :3].astype(np.float64) * scale).astype(np.float32))
    assert la == 0.0 and ml == 0.0 and (out[:, 3] == 1).all()
    out, _, _ = orc.tonemap(img, type=1, clamp=False, apply_gamma=False, fast_pow=False, exposure=1.0)
    np.testing.assert_array_equal(out[:, :3], (img[:, :3].astype(np.float64) * 2).astype(np.float32))
    out, _, _ = orc.tonemap(img, type=2, clamp=False, apply_gamma=False, fast_pow=False)
    c = img[:, :3].astype(np.float64)
    np.testing.assert_array_equal(out[:, :3], (c / (1 + c)).astype(np.float32))
    out, _, _ = orc.tonemap(img, type=2, fast_pow=True)  # clamp + fastPow gamma
    cc = np.clip(c * scale / (1 + c * scale), 0, 1)
    ref = np.vectorize(lambda v: orc.fastpow(v, 1 / 2.2))(cc)
    np.testing.assert_array_equal(out[:, :3], ref.astype(np.float32))


def test_passthrough_copies():
    img = _img()
    img[3, 3] = 0.25
    out, _, _ = orc.tonemap(img, type=3)
    np.testing.assert_array_equal(out, img)


def test_reinhard_general_equals_reinhard_and_moving_average():
    img = _img(seed=2)
    a, la, ml = orc.tonemap(img, type=0)
    b, lb, mb = orc.reinhard(img)
    np.testing.assert_array_equal(a, b)
    assert (la, ml) == (lb, mb)
    # moving average of the maximum, starting at 1: avg' = alpha max + (1 - alpha) avg
    st = orc.ToneMapState()
    d = img.astype(np.float64)
    maxl = max(1.0, float((0.2126 * d[:, 0] + 0.7152 * d[:, 1] + 0.0722 * d[:, 2]).max()))
    _, _, m1 = orc.tonemap(img, type=0, enable_averaging=True, averaging_alpha=0.25, state=st)
    assert m1 == pytest.approx(0.25 * maxl + 0.75 * 1.0, rel=1e-15)
    _, _, m2 = orc.tonemap(img, type=0, enable_averaging=True, averaging_alpha=0.25, state=st)
    assert m2 == pytest.approx(0.25 * maxl + 0.75 * m1, rel=1e-15)
    assert st.max_lum_average.value == m2
