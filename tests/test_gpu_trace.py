"""GPU closest-hit / any-hit parity against the FP64 oracle (through the C ABI).

Bar: the closest-hit triangle index is bit-exact except rays the oracle lists
as near ties (another triangle within 1e-6 relative distance, an edge inside
the barycentric band, or a grazing determinant); the GPU's distance is within
1e-5 relative of the FP64 Möller–Trumbore distance to the triangle it reports.
Any-hit decisions are exact except rays the oracle lists as fragile.  Each
parity test runs under both triangle decision modes: FP64 (the bar above
as stated) and FP32 (the default), where the listing band is its rounding
band, 1e-4, and the distance bound max(1e-5, 16 * 2^-24 / cond) for a hit of
conditioning cond (tests/helpers.py).
"""
import numpy as np
import pytest

import oracle as orc
import raycer_amd as ra
from raycer_amd import api
from tests.conftest import ORACLE_THREADS
from tests.helpers import FP32_BAND, camera_rays, check_closest, closest_diag, mt_fp64, random_rays

pytestmark = pytest.mark.gpu

_cache = {}


def scene(ctx, name):
    if name not in _cache:
        p = {
            "box": ra.scene_params(api.RCH_SCENE_BOX),
            "sphere5": ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1, level=5, grid=96),
            "terrain128": ra.scene_params(api.RCH_SCENE_TERRAIN, seed=1, grid=128),
            "soup": ra.scene_params(api.RCH_SCENE_SOUP, seed=11, tri_count=40000),
        }[name]
        sd = ra.generate_scene(p)
        nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
        _cache[name] = (sd, ctx.upload(sd, nodes, ordered), orc.BVH(sd.vertices))
    return _cache[name]


@pytest.mark.parametrize("name", ["box", "sphere5", "terrain128", "soup"])
def test_closest_hit_camera_rays(ctx, decision, name):
    sd, s, ob = scene(ctx, name)
    w, h = 320, 240
    rng = np.random.default_rng(3)
    rays = camera_rays(sd.desc.camera, w, h, jitter=rng.random((w * h, 2)))
    hits = ctx.trace_closest(s, rays)
    diag = closest_diag(ob, rays, decision, ORACLE_THREADS)
    mism, listed = check_closest(hits, sd.vertices, rays, diag, fp32=decision)
    print(f"{name}: {mism} mismatches, {listed} listed near ties of {len(rays)}")
    assert (hits["tri"] >= 0).mean() > 0.3


@pytest.mark.parametrize("name", ["box", "sphere5", "soup"])
def test_closest_hit_random_rays(ctx, decision, name):
    sd, s, ob = scene(ctx, name)
    v = sd.vertices.reshape(-1, 3)
    rays = random_rays(60000, 7, v.min(0), v.max(0))
    hits = ctx.trace_closest(s, rays)
    # in the box the inner boxes' bottom faces are coplanar with the floor:
    # rays from below hit both at the same t, exact ties the oracle lists
    check_closest(hits, sd.vertices, rays, closest_diag(ob, rays, decision, ORACLE_THREADS),
                  max_listed=0.05 if name == "box" else 0.01, fp32=decision)


@pytest.mark.parametrize("name,method,max_leaf", [("sphere5", api.RCH_BVH_BINNED, 5), ("soup", api.RCH_BVH_BINNED, 5),
                                                  ("soup", api.RCH_BVH_REFERENCE, 100)])
def test_results_do_not_depend_on_the_bvh(ctx, name, method, max_leaf):
    """Another BVH visits triangles in another order; with the canonical tie
    rule (lower original id) the hits must be bit-identical.  Leaves of up to
    100 triangles exercise the upload's splitting of long leaves (a leaf
    record holds at most 31)."""
    sd, s, _ = scene(ctx, name)
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices, max_leaf=max_leaf, method=method)
    if max_leaf > 31:
        leaves = nodes[nodes["right_offset"] == 0]
        assert leaves["primitive_count"].max() >= 32
    s2 = ctx.upload(sd, nodes, ordered)
    try:
        v = sd.vertices.reshape(-1, 3)
        rays = random_rays(60000, 13, v.min(0), v.max(0))
        rng = np.random.default_rng(4)
        rays = np.concatenate([rays, camera_rays(sd.desc.camera, 160, 120, jitter=rng.random((160 * 120, 2)))])
        a, b = ctx.trace_closest(s, rays), ctx.trace_closest(s2, rays)
        assert np.array_equal(a, b)
        assert np.array_equal(ctx.trace_any(s, rays), ctx.trace_any(s2, rays))
    finally:
        s2.close()


def test_closest_hit_matches_brute_force_without_bvh(ctx, decision):
    sd, s, _ = scene(ctx, "box")
    rays = random_rays(5000, 9, [0, 0, 0], [1, 1, 1])
    hits = ctx.trace_closest(s, rays)
    t_bf, id_bf = orc.trace_brute(sd.vertices, rays, threads=ORACLE_THREADS)
    ob = orc.BVH(sd.vertices)
    flags = closest_diag(ob, rays, decision, ORACLE_THREADS)[2]
    bad = (hits["tri"] != id_bf) & (flags == 0)
    assert not bad.any()


def test_tmin_tmax_window(ctx, decision):
    sd, s, ob = scene(ctx, "box")
    rays = camera_rays(sd.desc.camera, 64, 64)
    full = ctx.trace_closest(s, rays)
    win = rays.copy()
    win["tmax"] = np.where(full["tri"] >= 0, full["t"] * 0.999, np.inf)
    assert np.all(ctx.trace_closest(s, win)["t"][full["tri"] >= 0] > full["t"][full["tri"] >= 0] * 0.999)
    win2 = rays.copy()
    win2["tmin"] = np.where(full["tri"] >= 0, full["t"] * 1.001, 0.0)
    h2 = ctx.trace_closest(s, win2)
    hit = (full["tri"] >= 0) & (h2["tri"] >= 0)
    assert np.all(h2["t"][hit] >= full["t"][hit] * 1.001)
    # past the first hit many rays run along the floor under the inner boxes,
    # whose bottom faces are coplanar with it (listed exact ties)
    check_closest(h2, sd.vertices, win2, closest_diag(ob, win2, decision, ORACLE_THREADS), max_listed=0.05,
                  fp32=decision)


def test_axis_aligned_rays_and_miss(ctx, decision):
    sd, s, ob = scene(ctx, "box")
    g = np.linspace(0.05, 0.95, 17)
    org = np.array([[x, y, -1.0] for x in g for y in g], np.float32)
    rays = api.make_rays(org, np.tile(np.float32([0, 0, 1]), (len(org), 1)))
    rays["org"] = org
    hits = ctx.trace_closest(s, rays)
    # the 17 rays with x == y hit the back wall exactly on the diagonal shared
    # by its two triangles: exact ties, resolved to the lower id
    check_closest(hits, sd.vertices, rays, closest_diag(ob, rays, decision, ORACLE_THREADS), max_listed=0.1,
                  fp32=decision)
    up =api.make_rays(np.float32([0.5, 2.0, 0.5]), np.float32([[0, 1, 0]]))
    assert ctx.trace_closest(s, up)["tri"][0] == -1
    assert np.isinf(ctx.trace_closest(s, up)["t"][0])


def test_single_leaf_scene_and_empty_batch(ctx):
    v = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)
    mats = np.zeros(1, api.MATERIAL_DTYPE)
    sd = ra.SceneData(v, np.zeros(2, np.uint32), mats, None)
    nodes, ordered, _, _ = ra.build_bvh(v)
    assert len(nodes) == 1
    s = ctx.upload(sd, nodes, ordered)
    rays = api.make_rays(np.float32([0.2, 0.2, -1]), np.float32([[0, 0, 1], [0, 0, -1]]))
    h = ctx.trace_closest(s, rays)
    assert h["tri"].tolist() == [0, -1] and h["t"][0] == 1.0
    assert len(ctx.trace_closest(s, rays[:0])) == 0
    s.close()


def test_exact_tie_goes_to_lower_original_id(ctx):
    # the same triangle twice: both at the same t; the canonical rule picks id 0
    v = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]] * 7, np.float32)
    sd = ra.SceneData(v, np.zeros(7, np.uint32), np.zeros(1, api.MATERIAL_DTYPE), None)
    nodes, ordered, _, _ = ra.build_bvh(v)
    s = ctx.upload(sd, nodes, ordered)
    rays = api.make_rays(np.float32([0.2, 0.2, 1]), np.float32([[0, 0, -1]]))
    assert ctx.trace_closest(s, rays)["tri"][0] == 0
    s.close()


def test_any_hit_parity(ctx, decision):
    sd, s, ob = scene(ctx, "sphere5")
    w, h = 200, 150
    rays = camera_rays(sd.desc.camera, w, h)
    hits = ctx.trace_closest(s, rays)
    hit = hits["tri"] >= 0
    p = rays["org"][hit] + hits["t"][hit][:, None] * rays["dir"][hit]
    L = np.array(sd.desc.light.position, np.float32)
    d = L - p
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    sr = api.make_rays(p + d * 1e-4, d.astype(np.float32), tmax=dist.astype(np.float32))
    sr["org"] = (p + d * 1e-4).astype(np.float32)
    occ = ctx.trace_any(s, sr)
    o_occ, fragile = ob.trace_any(sr, band=FP32_BAND if decision else 2e-5, threads=ORACLE_THREADS)
    bad = (occ != o_occ) & (fragile == 0)
    assert not bad.any(), np.nonzero(bad)[0][:10]
    assert 0.05 < occ.mean() < 0.95
    # a shorter tmax can only remove occluders
    sr2 = sr.copy()
    sr2["tmax"] = sr["tmax"] * 0.5
    assert np.all(ctx.trace_any(s, sr2) <= occ | (fragile != 0))


def test_any_hit_skips_non_shadowing(ctx):
    v = np.array([[-1, -1, 0, 1, -1, 0, -1, 1, 0], [-1, -1, 1, 1, -1, 1, -1, 1, 1]], np.float32)
    mats = np.zeros(2, api.MATERIAL_DTYPE)
    mats[1]["flags"] = api.RC_MAT_NON_SHADOWING
    sd = ra.SceneData(v, np.array([1, 0], np.uint32), mats, None)
    nodes, ordered, _, _ = ra.build_bvh(v)
    s = ctx.upload(sd, nodes, ordered)
    r = api.make_rays(np.float32([-0.5, -0.5, -1]), np.float32([[0, 0, 1]]), tmax=1.5)
    assert ctx.trace_any(s, r)[0] == 0     # z=0 triangle is nonShadowing, z=1 is beyond tmax
    r["tmax"] = 2.5
    assert ctx.trace_any(s, r)[0] == 1
    assert ctx.trace_closest(s, r)["tri"][0] == 0   # closest hit still sees it
    s.close()


def test_determinism_and_counting(ctx):
    sd, s, _ = scene(ctx, "terrain128")
    rays = camera_rays(sd.desc.camera, 256, 192)
    a = ctx.trace_closest(s, rays)
    b = ctx.trace_closest(s, rays)
    assert a.tobytes() == b.tobytes()
    ctx.set_counting(True)
    try:
        c = ctx.trace_closest(s, rays)
        st = ctx.stats()
    finally:
        ctx.set_counting(False)
    assert c.tobytes() == a.tobytes()
    assert st.node_records >= 2 * len(rays) and st.tri_tests > 0
    print("records/ray", st.node_records / len(rays), "tris/ray", st.tri_tests / len(rays))


def test_answers_do_not_depend_on_the_lanes_a_ray_shares_a_wave_with(ctx, decision):
    """Leaf postponing makes the order in which a ray's leaves are tested depend
    on the other rays of its wave; the closest hit (exact ties to the lower
    id) and the any-hit answer must not.  The same rays in a shuffled order
    (camera rays mixed with random ones, so waves mix long and short
    traversals) give bit-identical results."""
    sd, s, _ = scene(ctx, "terrain128")
    cam = camera_rays(sd.desc.camera, 160, 120)
    lo, hi = sd.vertices.reshape(-1, 3).min(axis=0), sd.vertices.reshape(-1, 3).max(axis=0)
    rays = np.concatenate([cam, random_rays(len(cam), 5, lo, hi)])
    perm = np.random.default_rng(3).permutation(len(rays))
    a = ctx.trace_closest(s, rays)
    b = ctx.trace_closest(s, rays[perm])
    back = np.empty_like(b)
    back[perm] = b
    assert a.tobytes() == back.tobytes()
    assert (a["tri"] >= 0).mean() > 0.3
    occ_a = ctx.trace_any(s, rays)
    occ_b = ctx.trace_any(s, rays[perm])
    assert (occ_a == occ_b[np.argsort(perm)]).all()


def test_upload_rejects_corrupt_bvh(ctx):
    sd, _, _ = scene(ctx, "box")
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    bad = ordered.copy()
    bad[0] = bad[1]
    with pytest.raises(ra.RaycerError):
        ctx.upload(sd, nodes, bad)
    n2 = nodes.copy()
    n2["right_offset"][0] = 10 ** 6
    with pytest.raises(ra.RaycerError):
        ctx.upload(sd, n2, ordered)


@pytest.mark.slow
def test_full_size_sphere_ground_ray_set(ctx, decision):
    """The 1,179,648-triangle scene at 1920x1080 with per-pixel jitter: every
    primary ray checked against the FP64 oracle."""
    p = ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1)
    sd = ra.generate_scene(p)
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    s = ctx.upload(sd, nodes, ordered)
    w, h = 1920, 1080
    rng = np.random.default_rng(1)
    rays = camera_rays(sd.desc.camera, w, h, jitter=rng.random((w * h, 2)))
    hits = ctx.trace_closest(s, rays)
    ob = orc.BVH(sd.vertices)
    mism, listed = check_closest(hits, sd.vertices, rays, closest_diag(ob, rays, decision, ORACLE_THREADS),
                                 max_listed=0.005 if not decision else 0.02, fp32=decision)
    print(f"full C2 ray set: {mism} mismatches, {listed} listed of {w * h}")
    s.close()
