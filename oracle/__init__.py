"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference results are compared
with, never as the thing measured or shipped.  See oracle.h for the reference functions each entry
point restates.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
NODE_DTYPE = np.dtype([("aabb_min", "<f8", 3), ("aabb_max", "<f8", 3), ("right_offset", "<i8"),
                       ("start_offset", "<u8"), ("primitive_count", "<u8")])
MAT_STRIDE = 16


class RenderCfg(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("sampler", C.c_uint32),
                ("max_path_length", C.c_uint32), ("rr_start", C.c_uint32), ("filter", C.c_uint32),
                ("estimator", C.c_uint32), ("traversal", C.c_uint32), ("filter_radius", C.c_double),
                ("ray_start_offset", C.c_double), ("sky", C.c_double * 3), ("sun_dir", C.c_double * 3),
                ("sun_irradiance", C.c_double * 3), ("seed", C.c_uint64), ("frame", C.c_uint32),
                ("cache_vertex", C.c_uint32), ("filter_param", C.c_double * 2)]


_libs: dict[str, C.CDLL] = {}


def lib(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = HERE / name
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(str(path))
        vp, sz, d = C.c_void_p, C.c_size_t, C.c_double
        L.or_fastpow.restype = d
        L.or_fastpow.argtypes = [d, d]
        L.or_uniform.restype = d
        L.or_uniform.argtypes = [C.c_uint32]
        L.or_filter_weight.restype = d
        L.or_filter_weight.argtypes = [C.c_uint32, d, d, d, d, d]
        L.or_filter_radius.restype = d
        L.or_filter_radius.argtypes = [C.c_uint32, d, d, d]
        L.or_cmj_permute.restype = C.c_uint64
        L.or_cmj_permute.argtypes = [C.c_uint64] * 3
        L.or_cmj_sample.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, d, d, vp]
        L.or_tonemap.argtypes = [vp, sz, vp, vp, C.POINTER(d), C.POINTER(d), C.POINTER(d)]
        L.or_raytrace.argtypes = [vp, vp, sz, vp, vp, vp, C.POINTER(RenderCfg), vp, vp, vp, C.c_uint32, vp, sz, vp,
                                  C.c_int]
        L.or_bvh_build.argtypes = [vp, sz, C.c_uint64, C.POINTER(vp)]
        L.or_bvh_node_count.restype = sz
        L.or_bvh_node_count.argtypes = [vp]
        L.or_bvh_nodes.argtypes = [vp, vp]
        L.or_bvh_ordered_ids.argtypes = [vp, vp]
        L.or_bvh_free.argtypes = [vp]
        L.or_intersect_triangle.argtypes = [vp, vp, vp, d, d, vp, vp, vp]
        L.or_trace_closest_ref.argtypes = [vp, vp, vp, sz, vp, vp, C.c_int]
        L.or_trace_closest_diag.argtypes = [vp, vp, vp, sz, d, d, vp, vp, vp, vp, C.c_int]
        L.or_trace_closest_brute.argtypes = [vp, sz, vp, sz, vp, vp, C.c_int]
        L.or_trace_any.argtypes = [vp, vp, vp, sz, d, vp, vp, C.c_int]
        L.or_camera_ray.argtypes = [vp, C.c_uint32, C.c_uint32, d, d, vp, vp]
        L.or_render_split_tail.argtypes = [vp, vp, sz, vp, vp, vp, C.POINTER(RenderCfg), vp, sz, vp, C.c_int]
        L.or_render_split_tail.restype = C.c_int
        L.or_render.argtypes = [vp, vp, sz, vp, vp, sz, vp, C.POINTER(RenderCfg), vp, sz, vp, C.c_int,
                                C.POINTER(C.c_int)]
        L.or_trace_path.argtypes = [vp, vp, sz, vp, vp, vp, C.POINTER(RenderCfg), C.c_uint32, C.c_uint32, vp,
                                    C.POINTER(d), vp]
        L.or_raycast_phong.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint64, d, vp, vp, vp, sz,
                                       vp, vp, vp, C.c_int]
        L.or_reinhard.argtypes = [vp, sz, vp, vp, C.POINTER(d), C.POINTER(d)]
        L.or_hemisphere.argtypes = [vp, d, d, vp]
        L.or_mlp_param_count.restype = sz
        L.or_mlp_param_count.argtypes = [C.c_uint32] * 4
        L.or_mlp_forward.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, sz, vp]
        L.or_mlp_loss_grad.restype = d
        L.or_mlp_loss_grad.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, sz, C.c_int, d,
                                       vp]
        L.or_adam_step.argtypes = [vp, vp, vp, vp, sz, d, d, d, d, C.c_uint32]
        L.or_cache_encode.argtypes = [vp, vp, vp, vp, vp]
        _libs[name] = L
    return _libs[name]


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f64(x):
    return np.ascontiguousarray(x, np.float64)


def fastpow(a: float, b: float) -> float:
    return lib().or_fastpow(a, b)


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def normalize(v):
    i, o = _f64(v), np.empty(3)
    lib().or_normalize(_p(i), _p(o))
    return o


def cross(a, b):
    a, b, o = _f64(a), _f64(b), np.empty(3)
    lib().or_cross(_p(a), _p(b), _p(o))
    return o


def intersect_triangle(tri9, org, d, tmin=-np.inf, tmax=np.inf):
    t, u, v = C.c_double(), C.c_double(), C.c_double()
    tri9, org, d = _f64(tri9), _f64(org), _f64(d)
    hit = lib().or_intersect_triangle(_p(tri9), _p(org), _p(d), tmin, tmax, C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def camera_array(cam) -> np.ndarray:
    """rc_camera-like object or 7-sequence -> [px, py, pz, pitch, yaw, roll, fov]."""
    if hasattr(cam, "position"):
        return np.array([*cam.position, cam.pitch, cam.yaw, cam.roll, cam.fov], np.float64)
    return _f64(cam)


def camera_ray(cam, w, h, px, py):
    o, d, ca = np.empty(3), np.empty(3), camera_array(cam)
    lib().or_camera_ray(_p(ca), w, h, px, py, _p(o), _p(d))
    return o, d


def materials_array(mats) -> np.ndarray:
    """rc_material structured array -> (nmat, 16) float64 in the oracle layout."""
    out = np.zeros((len(mats), MAT_STRIDE))
    out[:, 0:3] = mats["diffuse"]
    out[:, 3:6] = mats["emittance"]
    out[:, 6:9] = mats["ambient"]
    out[:, 9:12] = mats["specular"]
    out[:, 12] = mats["shininess"]
    out[:, 13] = mats["flags"]
    return out


class BVH:
    """FlatBVH restatement; fast=True builds it in the -ffast-math library
    (used only for CPU-baseline timing)."""

    def __init__(self, vertices: np.ndarray, max_leaf: int = 5, fast: bool = False):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 9)
        self.fast = fast
        self.h = C.c_void_p()
        if lib(fast).or_bvh_build(_p(self.vertices), self.vertices.shape[0], max_leaf, C.byref(self.h)) != 0:
            raise ValueError("or_bvh_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib(self.fast).or_bvh_free(self.h)
            self.h = None

    def nodes(self) -> np.ndarray:
        n = lib().or_bvh_node_count(self.h)
        out = np.empty(n, NODE_DTYPE)
        lib().or_bvh_nodes(self.h, _p(out))
        return out

    def ordered(self) -> np.ndarray:
        out = np.empty(self.vertices.shape[0], np.uint64)
        lib().or_bvh_ordered_ids(self.h, _p(out))
        return out

    def trace_ref(self, rays: np.ndarray, threads: int = 0):
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t, i = np.empty(n), np.empty(n, np.int64)
        lib().or_trace_closest_ref(self.h, _p(self.vertices), _p(rays), n, _p(t), _p(i), threads)
        return t, i

    def trace_diag(self, rays: np.ndarray, rel_tie=1e-6, band=2e-5, threads: int = 0, with_cond=False):
        """Closest hit (t, id) with near-tie flags; with_cond also returns the
        conditioning |det| / (|e1| |e2| |d|) of each best hit (1 for a miss)."""
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t, i, f, c = np.empty(n), np.empty(n, np.int64), np.empty(n, np.uint32), np.empty(n)
        lib().or_trace_closest_diag(self.h, _p(self.vertices), _p(rays), n, rel_tie, band, _p(t), _p(i), _p(f), _p(c),
                                    threads)
        return (t, i, f, c) if with_cond else (t, i, f)

    def trace_any(self, rays: np.ndarray, band=2e-5, threads: int = 0):
        rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = rays.shape[0]
        o, f = np.empty(n, np.uint8), np.empty(n, np.uint8)
        lib().or_trace_any(self.h, _p(self.vertices), _p(rays), n, band, _p(o), _p(f), threads)
        return o, f


def trace_brute(vertices: np.ndarray, rays: np.ndarray, threads: int = 0):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
    n = rays.shape[0]
    t, i = np.empty(n), np.empty(n, np.int64)
    lib().or_trace_closest_brute(_p(v), v.shape[0], _p(rays), n, _p(t), _p(i), threads)
    return t, i


def render_cfg(cfg, traversal: int = 1) -> RenderCfg:
    """Product RenderCfg (ctypes) -> oracle RenderCfg."""
    o = RenderCfg()
    for f in ("width", "height", "spp", "sampler", "max_path_length", "rr_start", "filter", "estimator", "seed",
              "frame", "cache_vertex"):
        setattr(o, f, getattr(cfg, f))
    o.filter_radius = cfg.filter_radius
    o.filter_param[0], o.filter_param[1] = cfg.filter_param[0], cfg.filter_param[1]
    o.ray_start_offset = cfg.ray_start_offset
    for k in range(3):
        o.sky[k] = cfg.sky[k]
        o.sun_dir[k] = cfg.sun_dir[k]
        o.sun_irradiance[k] = cfg.sun_irradiance[k]
    o.traversal = traversal
    return o


def render(bvh: BVH, tri_mat, mats, cam, cfg: RenderCfg, pixel_ids=None, threads: int = 0):
    """Per-pixel stats (npix, 14): sum_w, sum_w2, sum_wL[3], sum_w2L[3], sum_w2L2[3], rays, and
    the reference traversal's node visits and triangle tests for those rays (traversal 0).
    Runs in the library the BVH was built with."""
    L = lib(bvh.fast)
    tm = np.ascontiguousarray(tri_mat, np.uint32)
    ma = np.ascontiguousarray(materials_array(mats) if mats.dtype.names else mats, np.float64)
    ids = None if pixel_ids is None else np.ascontiguousarray(pixel_ids, np.uint32)
    npix = cfg.width * cfg.height if ids is None else ids.shape[0]
    out = np.empty((npix, 14))
    stop = C.c_int(0)
    ca = camera_array(cam)
    L.or_render(bvh.h, _p(bvh.vertices), bvh.vertices.shape[0], _p(tm), _p(ma), ma.shape[0],
                _p(ca), C.byref(cfg), _p(ids), npix, _p(out), threads, C.byref(stop))
    return out


def render_split_tail(bvh: BVH, tri_mat, mats, cam, cfg: RenderCfg, pixel_ids, threads: int = 0):
    """or_render_split_tail over the listed pixels (cfg.cache_vertex > 0):
    per pixel (sum_w, sum_w La[3], sum_w^2 (La - Lb)^2 [3]) for two paths of
    each sample that share everything up to that vertex's next-event ray."""
    L = lib(bvh.fast)
    tm = np.ascontiguousarray(tri_mat, np.uint32)
    ma = np.ascontiguousarray(materials_array(mats) if mats.dtype.names else mats, np.float64)
    ids = np.ascontiguousarray(pixel_ids, np.uint32)
    out = np.empty((ids.shape[0], 7))
    rc = L.or_render_split_tail(bvh.h, _p(bvh.vertices), bvh.vertices.shape[0], _p(tm), _p(ma), _p(camera_array(cam)),
                                C.byref(cfg), _p(ids), ids.shape[0], _p(out), threads)
    if rc != 0:
        raise ValueError("render_split_tail needs cfg.cache_vertex > 0")
    return out


def raytrace(bvh: BVH, tri_mat, mats, cam, cfg: RenderCfg, settings, ray_mats, lights=(), pixel_ids=None,
             threads: int = 0):
    """The Whitted tracer's per-pixel stats (npix, 12) like render().  settings,
    ray_mats and lights are the product's ctypes structures (same layouts)."""
    L = lib(bvh.fast)
    tm = np.ascontiguousarray(tri_mat, np.uint32)
    ma = np.ascontiguousarray(materials_array(mats) if mats.dtype.names else mats, np.float64)
    ids = None if pixel_ids is None else np.ascontiguousarray(pixel_ids, np.uint32)
    npix = cfg.width * cfg.height if ids is None else ids.shape[0]
    out = np.empty((npix, 12))
    ca = camera_array(cam)
    lts = lights if isinstance(lights, C.Array) or len(lights) == 0 else (type(lights[0]) * len(lights))(*lights)
    L.or_raytrace(bvh.h, _p(bvh.vertices), bvh.vertices.shape[0], _p(tm), _p(ma), _p(ca), C.byref(cfg),
                  C.addressof(settings), C.addressof(ray_mats), C.addressof(lts) if len(lights) else None, len(lights),
                  _p(ids), npix, _p(out), threads)
    return out


def trace_path(bvh: BVH, tri_mat, mats, cam, cfg: RenderCfg, pixel: int, sample: int):
    tm = np.ascontiguousarray(tri_mat, np.uint32)
    ma = np.ascontiguousarray(materials_array(mats) if mats.dtype.names else mats, np.float64)
    L = np.empty(3)
    w = C.c_double()
    verts = np.zeros((cfg.max_path_length, 8))
    ca = camera_array(cam)
    n = lib().or_trace_path(bvh.h, _p(bvh.vertices), bvh.vertices.shape[0], _p(tm), _p(ma), _p(ca),
                            C.byref(cfg), pixel, sample, _p(L), C.byref(w), _p(verts))
    return L, w.value, verts[:n]


def raycast_phong(bvh: BVH, tri_mat, mats, cam, w, h, seed, offset, light, ambient, pixel_ids=None, threads=0):
    tm = np.ascontiguousarray(tri_mat, np.uint32)
    ma = np.ascontiguousarray(materials_array(mats) if mats.dtype.names else mats, np.float64)
    ids = None if pixel_ids is None else np.ascontiguousarray(pixel_ids, np.uint32)
    n = w * h if ids is None else ids.shape[0]
    hit, t, rgb = np.empty(n, np.int64), np.empty(n), np.empty((n, 3))
    ca, lg, am = camera_array(cam), _f64(light), _f64(ambient)
    lib().or_raycast_phong(bvh.h, _p(bvh.vertices), _p(tm), _p(ma), _p(ca), w, h, seed, offset, _p(lg), _p(am),
                           _p(ids), n, _p(hit), _p(t), _p(rgb), threads)
    return hit, t, rgb


def reinhard(linear_rgba: np.ndarray, key=0.18, gamma=2.2, clamp=True, apply_gamma=True, fast_pow=True):
    lin = np.ascontiguousarray(linear_rgba, np.float32).reshape(-1, 4)
    out = np.empty_like(lin)
    prm = np.array([key, gamma, float(clamp), float(apply_gamma), float(fast_pow)])
    la, ml = C.c_double(), C.c_double()
    lib().or_reinhard(_p(lin), lin.shape[0], _p(prm), _p(out), C.byref(la), C.byref(ml))
    return out, la.value, ml.value


def filter_weight(f: int, r: float, x: float, y: float, p0: float = 0.0, p1: float = 0.0) -> float:
    """2-D filter weight; p0, p1 = Gaussian standard deviation, or Mitchell's B and C."""
    return lib().or_filter_weight(f, r, p0, p1, x, y)


def filter_radius(f: int, r: float, p0: float = 0.0, p1: float = 0.0) -> float:
    return lib().or_filter_radius(f, r, p0, p1)


def cmj_permute(i: int, n: int, p: int) -> int:
    return lib().or_cmj_permute(i, n, p)


def cmj_sample(x: int, y: int, n: int, p: int, r0: float, r1: float):
    o = np.empty(2)
    lib().or_cmj_sample(x, y, n, p, r0, r1, _p(o))
    return o


class ToneMapState:
    """ReinhardToneMapper's moving average of the maximum luminance (starts at 1)."""

    def __init__(self):
        self.max_lum_average = C.c_double(1.0)


def tonemap(linear_rgba: np.ndarray, type=0, key=0.18, gamma=2.2, clamp=True, apply_gamma=True, fast_pow=True,
            exposure=0.0, enable_averaging=False, averaging_alpha=0.1, state: ToneMapState | None = None):
    """type: 0 Reinhard, 1 linear, 2 simple, 3 passthrough.  Returns (rgba, log_avg, max_lum)."""
    lin = np.ascontiguousarray(linear_rgba, np.float32).reshape(-1, 4)
    out = np.empty_like(lin)
    prm = np.array([type, key, gamma, float(clamp), float(apply_gamma), float(fast_pow), exposure,
                    float(enable_averaging), averaging_alpha])
    st = state or ToneMapState()
    la, ml = C.c_double(), C.c_double()
    lib().or_tonemap(_p(lin), lin.shape[0], _p(prm), _p(out), C.byref(la), C.byref(ml), C.byref(st.max_lum_average))
    return out, la.value, ml.value


def hemisphere(n, u1, u2):
    o, n = np.empty(3), _f64(n)
    lib().or_hemisphere(_p(n), u1, u2, _p(o))
    return o


# ------------------------------------------------------------------ MLP
def mlp_param_count(i, h, nh, o) -> int:
    return lib().or_mlp_param_count(i, h, nh, o)


def mlp_forward(params, x, i, h, nh, o):
    x, params = _f64(x).reshape(-1, i), _f64(params)
    y = np.empty((x.shape[0], o))
    lib().or_mlp_forward(_p(params), i, h, nh, o, _p(x), x.shape[0], _p(y))
    return y


def mlp_loss_grad(params, x, target, i, h, nh, o, rel=True, eps=1e-2):
    x = _f64(x).reshape(-1, i)
    t = _f64(target).reshape(-1, o)
    g, params = np.empty(len(params)), _f64(params)
    loss = lib().or_mlp_loss_grad(_p(params), i, h, nh, o, _p(x), _p(t), x.shape[0], int(rel), eps, _p(g))
    return loss, g


def adam_step(params, m, v, g, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8, t=1):
    p, m, v, g = _f64(params).copy(), _f64(m).copy(), _f64(v).copy(), _f64(g)
    lib().or_adam_step(_p(p), _p(m), _p(v), _p(g), p.shape[0], lr, b1, b2, eps, t)
    return p, m, v


def cache_encode(pos01, d, n, albedo):
    o = np.empty(81)
    a, b, c, e = _f64(pos01), _f64(d), _f64(n), _f64(albedo)
    lib().or_cache_encode(_p(a), _p(b), _p(c), _p(e), _p(o))
    return o
