"""ctypes binding of libraycer.so (include/raycer_hip.h, include/raycer_host.h).

The structures below mirror the C headers field for field.  Every call goes
through the C ABI; a nonzero status raises RaycerError with the library's
message (the counterpart of CLManager::checkError throwing
std::runtime_error, src/OpenCL/CLManager.cpp:220-226).  There is no fallback:
if libraycer.so is missing or no gfx950 device is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("RAYCER_LIBRARY") or str(PKG / "libraycer.so"))

RC_OK, RC_ERR_INVALID, RC_ERR_HIP, RC_ERR_ALLOC, RC_ERR_INTERRUPTED, RC_ERR_NO_DEVICE = range(6)
RC_MAT_EMISSIVE, RC_MAT_NON_SHADOWING = 1, 2
RC_FILTER_BOX, RC_FILTER_TENT, RC_FILTER_BELL, RC_FILTER_GAUSSIAN, RC_FILTER_MITCHELL, RC_FILTER_LANCZOS_SINC = range(6)
(RC_SAMPLER_JITTERED, RC_SAMPLER_RANDOM, RC_SAMPLER_CENTER, RC_SAMPLER_REGULAR, RC_SAMPLER_CMJ,
 RC_SAMPLER_POISSON_DISC) = range(6)
RC_TONEMAP_REINHARD, RC_TONEMAP_LINEAR, RC_TONEMAP_SIMPLE, RC_TONEMAP_PASSTHROUGH = range(4)
RC_EST_REFERENCE, RC_EST_LAMBERT = 0, 1
RC_HIT_FP32, RC_HIT_FP64 = 0, 1
RCH_SCENE_BOX, RCH_SCENE_SPHERE_GROUND, RCH_SCENE_TERRAIN, RCH_SCENE_SOUP = 1, 2, 3, 4
RCH_BVH_REFERENCE, RCH_BVH_BINNED = 0, 1

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<i4")])
NODE_DTYPE = np.dtype([("aabb_min", "<f8", 3), ("aabb_max", "<f8", 3), ("right_offset", "<i8"),
                       ("start_offset", "<u8"), ("primitive_count", "<u8")])
MATERIAL_DTYPE = np.dtype([("diffuse", "<f4", 3), ("emittance", "<f4", 3), ("ambient", "<f4", 3),
                           ("specular", "<f4", 3), ("shininess", "<f4"), ("flags", "<u4"), ("pad", "<u4", 2)])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and NODE_DTYPE.itemsize == 72
assert MATERIAL_DTYPE.itemsize == 64


class RaycerError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"raycer error {code}: {msg}")
        self.code = code


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("pitch", C.c_float), ("yaw", C.c_float), ("roll", C.c_float),
                ("fov", C.c_float)]


class RenderCfg(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("sampler", C.c_uint32),
                ("max_path_length", C.c_uint32), ("rr_start", C.c_uint32), ("filter", C.c_uint32),
                ("estimator", C.c_uint32), ("cache_vertex", C.c_uint32), ("filter_radius", C.c_float),
                ("filter_param", C.c_float * 2), ("ray_start_offset", C.c_float), ("sky", C.c_float * 3),
                ("sun_dir", C.c_float * 3), ("sun_irradiance", C.c_float * 3), ("frame", C.c_uint32),
                ("seed", C.c_uint64)]


class ToneMap(C.Structure):
    _fields_ = [("key", C.c_float), ("gamma", C.c_float), ("clamp", C.c_uint32), ("apply_gamma", C.c_uint32),
                ("fast_pow", C.c_uint32), ("type", C.c_uint32), ("exposure", C.c_float),
                ("enable_averaging", C.c_uint32), ("averaging_alpha", C.c_float), ("pad", C.c_uint32)]


RC_LIGHT_DIRECTIONAL, RC_LIGHT_POINT, RC_LIGHT_SPOT = range(3)


class Light(C.Structure):
    """rc_light: a directional, point or spot light of the Whitted tracer (Lights.h)."""
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("intensity", C.c_float), ("position", C.c_float * 3),
                ("direction", C.c_float * 3), ("max_distance", C.c_float), ("attenuation", C.c_float),
                ("area_light", C.c_uint32), ("area_sampler", C.c_uint32), ("area_n", C.c_uint32),
                ("area_radius", C.c_float), ("angle", C.c_float), ("side_attenuation", C.c_float),
                ("pad", C.c_uint32)]


class RayMaterial(C.Structure):
    """rc_ray_material: Material's ray-tracing fields (Material.h:20-66)."""
    _fields_ = [("ray_reflectance", C.c_float), ("ray_transmittance", C.c_float), ("refractive_index", C.c_float),
                ("fresnel_reflection", C.c_uint32), ("skip_lighting", C.c_uint32), ("refl_sampler", C.c_uint32),
                ("refl_n", C.c_uint32), ("refl_glossiness", C.c_float), ("trans_sampler", C.c_uint32),
                ("trans_n", C.c_uint32), ("trans_glossiness", C.c_float), ("enable_attenuation", C.c_uint32),
                ("attenuation_factor", C.c_float), ("attenuation_color", C.c_float * 3)]


class RaySettings(C.Structure):
    """rc_ray_settings: Scene::General's ray-tracer fields, the ambient light and SimpleFog."""
    _fields_ = [("background", C.c_float * 3), ("max_ray_iterations", C.c_uint32), ("ambient_color", C.c_float * 3),
                ("ambient_intensity", C.c_float), ("ao_enabled", C.c_uint32), ("ao_sampler", C.c_uint32),
                ("ao_n", C.c_uint32), ("ao_max_distance", C.c_float), ("ao_distribution", C.c_float),
                ("visualize_depth", C.c_uint32), ("visualize_depth_max", C.c_float), ("fog_enabled", C.c_uint32),
                ("fog_color", C.c_float * 3), ("fog_distance", C.c_float), ("fog_steepness", C.c_float),
                ("fog_height_dispersion", C.c_uint32), ("fog_height", C.c_float),
                ("fog_height_steepness", C.c_float)]


def ray_materials(n: int, **defaults):
    """n RayMaterial with the reference's defaults (Material.h: glossiness 1000,
    refractive index 1, CMJ samplers), then `defaults` applied to each."""
    arr = (RayMaterial * n)()
    for m in arr:
        m.refractive_index = 1.0
        m.refl_glossiness = m.trans_glossiness = 1000.0
        m.refl_sampler = m.trans_sampler = RC_SAMPLER_CMJ
        m.attenuation_factor = 1.0
        for k, v in defaults.items():
            setattr(m, k, v)
    return arr


def ray_settings(**kw) -> RaySettings:
    """RaySettings with the reference's defaults (Scene.h, Lights.h), then kw."""
    r = RaySettings()
    r.max_ray_iterations = 3
    r.ambient_color[:] = (1.0, 1.0, 1.0)
    r.ambient_intensity = 1.0
    r.ao_sampler = RC_SAMPLER_CMJ
    r.ao_max_distance = 1.0
    r.ao_distribution = 1.0
    r.visualize_depth_max = 25.0
    r.fog_color[:] = (1.0, 1.0, 1.0)
    r.fog_distance = r.fog_height = 100.0
    r.fog_steepness = r.fog_height_steepness = 1.0
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(r, k)[:] = v
        else:
            setattr(r, k, v)
    return r


def light(type: int, **kw) -> Light:
    """A Light with the reference's defaults (Lights.h: white, intensity 1,
    maxDistance 1000, attenuation 1, CMJ area sampler, radius 1, angle 45)."""
    L = Light()
    L.type = type
    L.color[:] = (1.0, 1.0, 1.0)
    L.intensity = 1.0
    L.max_distance = 1000.0
    L.attenuation = 1.0
    L.area_sampler = RC_SAMPLER_CMJ
    L.area_radius = 1.0
    L.angle = 45.0
    L.side_attenuation = 1.0
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(L, k)[:] = v
        else:
            setattr(L, k, v)
    return L


class PointLight(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("max_distance", C.c_float), ("attenuation", C.c_float), ("ambient_color", C.c_float * 3),
                ("ambient_intensity", C.c_float)]


PARK_STEPS_DEFAULT = 20  # RC_PARK_STEPS_DEFAULT in include/raycer_hip.h


class Tuning(C.Structure):
    _fields_ = [("block", C.c_uint32), ("spp_per_pass", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("stack_limit", C.c_uint32), ("park_steps", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("samples", C.c_uint64),
                ("node_records", C.c_uint64), ("tri_tests", C.c_uint64), ("overflow_rays", C.c_uint64),
                ("extend_launches", C.c_uint64), ("extend_ms", C.c_double), ("shadow_ms", C.c_double),
                ("shade_ms", C.c_double), ("other_ms", C.c_double), ("render_ms", C.c_double),
                ("shadow_launches", C.c_uint64), ("cache_ms", C.c_double), ("train_ms", C.c_double),
                ("cache_queries", C.c_uint64), ("train_samples", C.c_uint64), ("graph_ms", C.c_double),
                ("graph_relaunched", C.c_uint64), ("parked_rays", C.c_uint64)]


class BatchPlan(C.Structure):
    """rc_batch_plan: one scene's samples per pass and kernel block shapes."""
    _fields_ = [("spp_per_pass", C.c_uint32), ("ext_block", C.c_uint32), ("shadow_block", C.c_uint32),
                ("pad", C.c_uint32)]


class CacheCfg(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("adam_eps", C.c_float),
                ("loss_eps", C.c_float), ("seed", C.c_uint32), ("max_batch", C.c_uint32),
                ("train_stride", C.c_uint32)]


CACHE_PARAMS = 17923


class BVHBuildInfo(C.Structure):
    _fields_ = [("max_leaf_size", C.c_uint64), ("threads", C.c_uint32), ("method", C.c_uint32),
                ("tri_cost", C.c_float), ("pad", C.c_uint32)]


class SceneParams(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("seed", C.c_uint32), ("level", C.c_uint32), ("grid", C.c_uint32),
                ("tri_count", C.c_uint64)]


class Material(C.Structure):
    _fields_ = [("diffuse", C.c_float * 3), ("emittance", C.c_float * 3), ("ambient", C.c_float * 3),
                ("specular", C.c_float * 3), ("shininess", C.c_float), ("flags", C.c_uint32), ("pad", C.c_uint32 * 2)]


class ObjInfo(C.Structure):
    _fields_ = [("scale", C.c_double * 3), ("rotate", C.c_double * 3), ("translate", C.c_double * 3),
                ("base", Material), ("ignore_materials", C.c_uint32), ("pad", C.c_uint32)]


class SceneModel(C.Structure):
    _fields_ = [("path", C.c_char * 1024), ("info", ObjInfo), ("invisible_triangles", C.c_uint32), ("pad", C.c_uint32)]


class SceneFileData(C.Structure):
    """rch_scene_file_data: a scene read from or written to a cereal XML / JSON scene file."""
    _fields_ = [("tracer_type", C.c_uint32), ("pad", C.c_uint32), ("camera", Camera), ("cfg", RenderCfg),
                ("tonemap", ToneMap), ("ray", RaySettings),
                ("nmat", C.c_size_t), ("materials", C.POINTER(Material)), ("ray_materials", C.POINTER(RayMaterial)),
                ("material_ids", C.POINTER(C.c_uint64)),
                ("ntri", C.c_size_t), ("vertices", C.POINTER(C.c_float)), ("tri_material", C.POINTER(C.c_uint32)),
                ("tri_ids", C.POINTER(C.c_uint64)),
                ("nlights", C.c_size_t), ("lights", C.POINTER(Light)),
                ("nmodels", C.c_size_t), ("models", C.POINTER(SceneModel)),
                ("nnodes", C.c_size_t), ("nodes", C.c_void_p), ("ordered", C.POINTER(C.c_uint64)),
                ("skipped_primitives", C.c_size_t)]


class SceneDesc(C.Structure):
    _fields_ = [("ntri", C.c_size_t), ("nmat", C.c_size_t), ("camera", Camera), ("cfg", RenderCfg),
                ("light", PointLight), ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3)]


_lib = None


def lib() -> C.CDLL:
    """Loads libraycer.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RaycerError(RC_ERR_NO_DEVICE, f"{LIB_PATH} is missing: run `python -m raycer_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp, sz, u32, u64, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int
        sig = {
            "rc_last_error": [C.c_char_p, sz],
            "rc_create": [i32, C.POINTER(vp)],
            "rc_destroy": [vp],
            "rc_device_name": [vp, C.c_char_p, sz, C.POINTER(i32)],
            "rc_set_profiling": [vp, i32],
            "rc_set_counting": [vp, i32],
            "rc_set_hit_decision": [vp, i32],
            "rc_set_tuning": [vp, C.POINTER(Tuning)],
            "rc_get_stats": [vp, C.POINTER(Stats)],
            "rc_synchronize": [vp],
            "rc_scene_upload": [vp, vp, sz, vp, vp, sz, vp, sz, vp, C.POINTER(vp)],
            "rc_scene_free": [vp],
            "rc_scene_info": [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)],
            "rc_trace_closest": [vp, vp, vp, sz, vp],
            "rc_trace_any": [vp, vp, vp, sz, vp],
            "rc_trace_closest_device": [vp, vp, vp, sz, vp],
            "rc_trace_any_device": [vp, vp, vp, sz, vp],
            "rc_device_alloc": [vp, sz, C.POINTER(vp)],
            "rc_device_free": [vp, vp],
            "rc_host_alloc": [vp, sz, C.POINTER(vp)],
            "rc_host_free": [vp, vp],
            "rc_memcpy_h2d": [vp, vp, vp, sz],
            "rc_memcpy_d2h": [vp, vp, vp, sz],
            "rc_raytrace": [vp, vp, vp, C.POINTER(Camera), C.POINTER(RenderCfg), C.POINTER(RaySettings), vp, u32,
                            vp, u32],
            "rc_film_create": [vp, u32, u32, C.POINTER(vp)],
            "rc_film_free": [vp],
            "rc_film_clear": [vp, vp],
            "rc_film_download_linear": [vp, vp, vp],
            "rc_film_download_sums": [vp, vp, vp],
            "rc_film_tonemap": [vp, vp, C.POINTER(ToneMap), vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
            "rc_render": [vp, vp, vp, C.POINTER(Camera), C.POINTER(RenderCfg), C.POINTER(C.c_int),
                          C.POINTER(C.c_uint64)],
            "rc_raycast_phong": [vp, vp, C.POINTER(Camera), u32, u32, u64, C.c_float, C.POINTER(PointLight), vp, vp,
                                 vp],
            "rc_render_batch": [vp, vp, vp, vp, vp, sz, C.POINTER(Tuning)],
            "rc_render_batch_plan": [vp, vp, vp, vp, vp, sz, vp],
            "rc_autotune_batch": [vp, vp, vp, vp, sz, C.c_char_p, vp],
            "rc_render_batch_tuned": [vp, vp, vp, vp, vp, sz, vp, u32],
            "rc_autotune_batch_plan": [vp, vp, vp, vp, sz, C.c_char_p, vp, C.POINTER(u32)],
            "rc_copy_bandwidth": [vp, sz, i32, C.POINTER(C.c_double)],
            "rc_read_bandwidth": [vp, sz, i32, C.POINTER(C.c_double)],
            "rc_gather_bandwidth": [vp, sz, sz, i32, C.POINTER(C.c_double)],
            "rc_render_progress": [vp, vp, vp, C.POINTER(Camera), C.POINTER(RenderCfg), vp, vp],
            "rc_cache_create": [vp, C.POINTER(CacheCfg), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(vp)],
            "rc_cache_free": [vp],
            "rc_cache_set_params": [vp, vp],
            "rc_cache_get_params": [vp, vp],
            "rc_cache_infer": [vp, vp, vp, sz, vp],
            "rc_cache_train_step": [vp, vp, vp, vp, sz, C.POINTER(C.c_float)],
            "rc_cache_gradient": [vp, vp, vp, vp, sz, vp, C.POINTER(C.c_float)],
            "rc_render_cached": [vp, vp, vp, C.POINTER(Camera), C.POINTER(RenderCfg), vp, i32, C.POINTER(C.c_float)],
            "rch_bvh_build": [vp, sz, C.POINTER(BVHBuildInfo), C.POINTER(vp)],
            "rch_bvh_node_count": [vp],
            "rch_bvh_get": [vp, vp, vp],
            "rch_bvh_build_ms": [vp],
            "rch_bvh_depth": [vp],
            "rch_bvh_free": [vp],
            "rch_scene_describe": [C.POINTER(SceneParams), C.POINTER(SceneDesc)],
            "rch_scene_generate": [C.POINTER(SceneParams), vp, vp, vp],
            "rch_bvh_save": [C.c_char_p, vp, sz, vp, sz, vp],
            "rch_bvh_file_counts": [C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
            "rch_bvh_load": [C.c_char_p, vp, sz, vp, vp],
            "rch_obj_load": [C.c_char_p, C.POINTER(ObjInfo), C.POINTER(vp)],
            "rch_scene_file_load": [C.c_char_p, C.POINTER(C.POINTER(SceneFileData))],
            "rch_scene_file_parse": [C.c_char_p, sz, i32, C.POINTER(C.POINTER(SceneFileData))],
            "rch_scene_file_save": [C.c_char_p, C.POINTER(SceneFileData)],
            "rch_ply_load": [C.c_char_p, C.POINTER(ObjInfo), C.POINTER(vp)],
            "rch_obj_triangle_count": [vp],
            "rch_obj_material_count": [vp],
            "rch_obj_get": [vp, vp, vp, vp],
            "rch_obj_free": [vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.rch_bvh_node_count.restype = C.c_size_t
        L.rch_bvh_build_ms.restype = C.c_double
        L.rch_bvh_depth.restype = C.c_uint32
        L.rch_obj_triangle_count.restype = C.c_size_t
        L.rch_obj_material_count.restype = C.c_size_t
        L.rch_obj_default_info.argtypes = [C.POINTER(ObjInfo)]
        L.rch_obj_default_info.restype = None
        L.rch_scene_file_free.argtypes = [C.POINTER(SceneFileData)]
        L.rch_scene_file_free.restype = None
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().rc_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != RC_OK:
        raise RaycerError(rc, last_error())


def _ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ host services
@dataclass
class SceneData:
    vertices: np.ndarray      # (ntri, 9) float32
    tri_material: np.ndarray  # (ntri,) uint32
    materials: np.ndarray     # (nmat,) MATERIAL_DTYPE
    desc: SceneDesc

    @property
    def ntri(self) -> int:
        return int(self.vertices.shape[0])


def scene_params(kind: int, seed: int = 0, level: int = 0, grid: int = 0, tri_count: int = 0) -> SceneParams:
    return SceneParams(kind, seed, level, grid, tri_count)


def describe_scene(p: SceneParams) -> SceneDesc:
    d = SceneDesc()
    check(lib().rch_scene_describe(C.byref(p), C.byref(d)))
    return d


def generate_scene(p: SceneParams) -> SceneData:
    d = describe_scene(p)
    v = np.empty((d.ntri, 9), np.float32)
    m = np.empty(d.ntri, np.uint32)
    mats = np.zeros(d.nmat, MATERIAL_DTYPE)
    check(lib().rch_scene_generate(C.byref(p), _ptr(v), _ptr(m), _ptr(mats)))
    return SceneData(v, m, mats, d)


def save_bvh(path, vertices, nodes, ordered) -> None:
    """rch_bvh_save: the flat BVH with a hash of the triangles it was built over."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 9)
    nodes, ordered = np.ascontiguousarray(nodes), np.ascontiguousarray(ordered, np.uint64)
    check(lib().rch_bvh_save(str(path).encode(), _ptr(v), v.shape[0], _ptr(nodes), nodes.shape[0], _ptr(ordered)))


def load_bvh(path, vertices):
    """rch_bvh_load: (nodes, ordered_ids) saved for these triangles."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 9)
    nn, nt = C.c_size_t(), C.c_size_t()
    check(lib().rch_bvh_file_counts(str(path).encode(), C.byref(nn), C.byref(nt)))
    if nt.value != v.shape[0]:
        raise RaycerError(RC_ERR_INVALID, "rch_bvh_load: the saved BVH was built over other triangles")
    nodes, ordered = np.empty(nn.value, NODE_DTYPE), np.empty(nt.value, np.uint64)
    check(lib().rch_bvh_load(str(path).encode(), _ptr(v), v.shape[0], _ptr(nodes), _ptr(ordered)))
    return nodes, ordered


def load_obj(path, scale=(1.0, 1.0, 1.0), rotate=(0.0, 0.0, 0.0), translate=(0.0, 0.0, 0.0),
             ignore_materials=False, base: Material | None = None):
    """ModelLoader::readObjFile (readPlyFile for .ply) through rch_obj_load /
    rch_ply_load: (vertices (n, 9) float32,
    tri_material (n,) uint32, materials MATERIAL_DTYPE with [0] = the base material)."""
    info = ObjInfo()
    lib().rch_obj_default_info(C.byref(info))
    for k in range(3):
        info.scale[k], info.rotate[k], info.translate[k] = scale[k], rotate[k], translate[k]
    info.ignore_materials = int(ignore_materials)
    if base is not None:
        info.base = base
    h = C.c_void_p()
    loader = lib().rch_ply_load if str(path).lower().endswith(".ply") else lib().rch_obj_load
    check(loader(str(path).encode(), C.byref(info), C.byref(h)))
    try:
        n, nm = lib().rch_obj_triangle_count(h), lib().rch_obj_material_count(h)
        v = np.empty((n, 9), np.float32)
        m = np.empty(n, np.uint32)
        mats = np.zeros(nm, MATERIAL_DTYPE)
        check(lib().rch_obj_get(h, _ptr(v), _ptr(m), _ptr(mats)))
        return v, m, mats
    finally:
        lib().rch_obj_free(h)


def obj_scene(path, camera: Camera, cfg: RenderCfg, light: PointLight | None = None, **kw) -> SceneData:
    """A SceneData for an OBJ model with the given camera, render settings and
    point light (the form of raycer's test scenes built from models)."""
    v, m, mats = load_obj(path, **kw)
    d = SceneDesc()
    d.ntri, d.nmat, d.camera, d.cfg = v.shape[0], mats.shape[0], camera, cfg
    if light is not None:
        d.light = light
    p = v.reshape(-1, 3)
    for k in range(3):
        d.bounds_min[k], d.bounds_max[k] = float(p[:, k].min()), float(p[:, k].max())
    return SceneData(v, m, mats, d)


@dataclass
class SceneFile:
    """A scene file's contents (Scene::loadFromFile): settings and arrays."""
    tracer_type: int
    camera: Camera
    cfg: RenderCfg
    tonemap: ToneMap
    ray: RaySettings
    materials: np.ndarray       # MATERIAL_DTYPE
    ray_materials: C.Array      # RayMaterial * nmat
    material_ids: np.ndarray    # uint64
    vertices: np.ndarray        # (ntri, 9) float32
    tri_material: np.ndarray    # uint32 index into materials
    tri_ids: np.ndarray         # uint64 Primitive::id
    lights: list
    models: list                # (path, ObjInfo, invisible_triangles)
    nodes: np.ndarray | None    # NODE_DTYPE, None without a built BVH
    ordered: np.ndarray | None
    skipped_primitives: int

    def scene_data(self) -> "SceneData":
        d = SceneDesc()
        d.ntri, d.nmat, d.camera, d.cfg = len(self.tri_material), len(self.materials), self.camera, self.cfg
        if len(self.vertices):
            p = self.vertices.reshape(-1, 3)
            for k in range(3):
                d.bounds_min[k], d.bounds_max[k] = float(p[:, k].min()), float(p[:, k].max())
        return SceneData(self.vertices, self.tri_material, self.materials, d)


def _scene_file(ptr) -> SceneFile:
    d = ptr.contents
    try:
        def arr(p, n, dt, shape=None):
            if n == 0:
                return np.zeros(shape or (0,), dt)
            a = np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,))
            return a.copy().reshape(shape) if shape else a.copy()
        mats = np.zeros(d.nmat, MATERIAL_DTYPE)
        if d.nmat:
            C.memmove(mats.ctypes.data, d.materials, d.nmat * MATERIAL_DTYPE.itemsize)
        rm = (RayMaterial * d.nmat)()
        if d.nmat:
            C.memmove(rm, d.ray_materials, d.nmat * C.sizeof(RayMaterial))
        nodes = ordered = None
        if d.nnodes:
            nodes = np.zeros(d.nnodes, NODE_DTYPE)
            C.memmove(nodes.ctypes.data, d.nodes, d.nnodes * NODE_DTYPE.itemsize)
            ordered = arr(d.ordered, d.ntri, np.uint64)
        return SceneFile(
            d.tracer_type, Camera.from_buffer_copy(d.camera), RenderCfg.from_buffer_copy(d.cfg),
            ToneMap.from_buffer_copy(d.tonemap), RaySettings.from_buffer_copy(d.ray), mats, rm,
            arr(d.material_ids, d.nmat, np.uint64), arr(d.vertices, d.ntri * 9, np.float32, (d.ntri, 9)),
            arr(d.tri_material, d.ntri, np.uint32), arr(d.tri_ids, d.ntri, np.uint64),
            [Light.from_buffer_copy(d.lights[i]) for i in range(d.nlights)],
            [(d.models[i].path.decode(), ObjInfo.from_buffer_copy(d.models[i].info), d.models[i].invisible_triangles)
             for i in range(d.nmodels)],
            nodes, ordered, int(d.skipped_primitives))
    finally:
        lib().rch_scene_file_free(ptr)


RCH_SCENE_XML, RCH_SCENE_JSON, RCH_SCENE_BINARY = 0, 1, 2


def load_scene_file(path) -> SceneFile:
    """rch_scene_file_load: a cereal XML, JSON or binary (.bin) scene file (Scene::loadFromFile)."""
    p = C.POINTER(SceneFileData)()
    check(lib().rch_scene_file_load(str(path).encode(), C.byref(p)))
    return _scene_file(p)


def parse_scene_file(text, json: bool = False, binary: bool = False) -> SceneFile:
    """rch_scene_file_parse from memory: text (XML or JSON) or bytes (cereal binary)."""
    b = text if isinstance(text, (bytes, bytearray)) else text.encode()
    fmt = RCH_SCENE_BINARY if binary else RCH_SCENE_JSON if json else RCH_SCENE_XML
    p = C.POINTER(SceneFileData)()
    check(lib().rch_scene_file_parse(bytes(b), len(b), fmt, C.byref(p)))
    return _scene_file(p)


def save_scene_file(path, sf: SceneFile) -> None:
    """rch_scene_file_save: writes the scene as cereal XML, JSON or binary (by extension)."""
    d = SceneFileData()
    d.tracer_type, d.camera, d.cfg, d.tonemap, d.ray = sf.tracer_type, sf.camera, sf.cfg, sf.tonemap, sf.ray
    mats = np.ascontiguousarray(sf.materials)
    rm = sf.ray_materials
    ids = np.ascontiguousarray(sf.material_ids, np.uint64)
    v = np.ascontiguousarray(sf.vertices, np.float32)
    tm = np.ascontiguousarray(sf.tri_material, np.uint32)
    tid = np.ascontiguousarray(sf.tri_ids, np.uint64)
    lts = (Light * max(1, len(sf.lights)))(*sf.lights)
    models = (SceneModel * max(1, len(sf.models)))()
    for i, (path_, info, inv) in enumerate(sf.models):
        models[i].path, models[i].info, models[i].invisible_triangles = path_.encode(), info, inv
    d.nmat = len(mats)
    d.materials = C.cast(mats.ctypes.data, C.POINTER(Material))
    d.ray_materials = C.cast(rm, C.POINTER(RayMaterial))
    d.material_ids = C.cast(ids.ctypes.data, C.POINTER(C.c_uint64))
    d.ntri = len(tm)
    d.vertices = C.cast(v.ctypes.data, C.POINTER(C.c_float))
    d.tri_material = C.cast(tm.ctypes.data, C.POINTER(C.c_uint32))
    d.tri_ids = C.cast(tid.ctypes.data, C.POINTER(C.c_uint64))
    d.nlights = len(sf.lights)
    d.lights = C.cast(lts, C.POINTER(Light))
    d.nmodels = len(sf.models)
    d.models = C.cast(models, C.POINTER(SceneModel))
    nodes = ordered = None
    if sf.nodes is not None and len(sf.nodes):
        nodes = np.ascontiguousarray(sf.nodes)
        ordered = np.ascontiguousarray(sf.ordered, np.uint64)
        d.nnodes = len(nodes)
        d.nodes = nodes.ctypes.data
        d.ordered = C.cast(ordered.ctypes.data, C.POINTER(C.c_uint64))
    check(lib().rch_scene_file_save(str(path).encode(), C.byref(d)))


def build_bvh(vertices: np.ndarray, max_leaf: int = 5, threads: int = 0, method: int = 0, tri_cost: float = 0.0):
    """FlatBVH::build through rch_bvh_build (method RCH_BVH_REFERENCE, or
    RCH_BVH_BINNED); returns (nodes, ordered_ids, build_ms, depth)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 9)
    h = C.c_void_p()
    info = BVHBuildInfo(max_leaf, threads, method, tri_cost, 0)
    check(lib().rch_bvh_build(_ptr(v), v.shape[0], C.byref(info), C.byref(h)))
    try:
        n = lib().rch_bvh_node_count(h)
        nodes = np.empty(n, NODE_DTYPE)
        ordered = np.empty(v.shape[0], np.uint64)
        check(lib().rch_bvh_get(h, _ptr(nodes), _ptr(ordered)))
        return nodes, ordered, lib().rch_bvh_build_ms(h), lib().rch_bvh_depth(h)
    finally:
        lib().rch_bvh_free(h)


# ------------------------------------------------------------------ device objects
class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().rc_create(device, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().rc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self) -> tuple[str, int]:
        buf = C.create_string_buffer(256)
        cu = C.c_int()
        check(lib().rc_device_name(self.h, buf, len(buf), C.byref(cu)))
        return buf.value.decode(), cu.value

    def set_profiling(self, on: bool) -> None:
        check(lib().rc_set_profiling(self.h, int(on)))

    def set_counting(self, on: bool) -> None:
        check(lib().rc_set_counting(self.h, int(on)))

    def set_hit_decision(self, mode: str) -> None:
        """"fp32" (RC_HIT_FP32, the default) or "fp64" (RC_HIT_FP64)."""
        check(lib().rc_set_hit_decision(self.h, {"fp32": RC_HIT_FP32, "fp64": RC_HIT_FP64}[mode]))

    def set_tuning(self, block=256, spp_per_pass=0, blocks_per_cu=64, stack_limit=0, park_steps=None) -> None:
        """rc_set_tuning; the defaults are the context's own (spp_per_pass 0:
        each frame chooses its pass size; stack_limit 0: the whole on-chip
        traversal stack before the deep-stack kernel takes a ray over;
        park_steps None: the library's default)."""
        t = Tuning(block, spp_per_pass, blocks_per_cu, stack_limit,
                   PARK_STEPS_DEFAULT if park_steps is None else park_steps)
        check(lib().rc_set_tuning(self.h, C.byref(t)))

    def stats(self) -> Stats:
        s = Stats()
        check(lib().rc_get_stats(self.h, C.byref(s)))
        return s

    def synchronize(self) -> None:
        check(lib().rc_synchronize(self.h))

    def copy_bandwidth(self, nbytes: int = 1 << 30, iters: int = 20) -> float:
        g = C.c_double()
        check(lib().rc_copy_bandwidth(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def gather_bandwidth(self, nbytes: int = 4 << 30, records: int = 1 << 24, iters: int = 3) -> float:
        """rc_gather_bandwidth: GB/s of 64-byte reads at distinct random records."""
        g = C.c_double()
        check(lib().rc_gather_bandwidth(self.h, nbytes, records, iters, C.byref(g)))
        return g.value

    def read_bandwidth(self, nbytes: int, iters: int) -> float:
        """rc_read_bandwidth: read GB/s of a working set of nbytes."""
        g = C.c_double()
        check(lib().rc_read_bandwidth(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def upload(self, sd: SceneData, nodes=None, ordered=None) -> "Scene":
        if nodes is None:
            nodes, ordered, _, _ = build_bvh(sd.vertices)
        return Scene(self, sd, nodes, ordered)

    def trace_closest(self, scene: "Scene", rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        hits = np.empty(rays.shape[0], HIT_DTYPE)
        check(lib().rc_trace_closest(self.h, scene.h, _ptr(rays), rays.shape[0], _ptr(hits)))
        return hits

    def trace_any(self, scene: "Scene", rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        occ = np.empty(rays.shape[0], np.uint8)
        check(lib().rc_trace_any(self.h, scene.h, _ptr(rays), rays.shape[0], _ptr(occ)))
        return occ

    def film(self, width: int, height: int) -> "Film":
        return Film(self, width, height)

    def render(self, scene: "Scene", film: "Film", cam: Camera, cfg: RenderCfg, interrupt=None,
               pixels_processed=None) -> None:
        check(lib().rc_render(self.h, scene.h, film.h, C.byref(cam), C.byref(cfg), interrupt, pixels_processed))

    def _batch_arrays(self, scenes, cams, cfgs, films=None):
        n = len(scenes)
        sarr = (C.c_void_p * n)(*[s.h.value for s in scenes])
        farr = (C.c_void_p * n)(*[f.h.value for f in films]) if films is not None else None
        return n, sarr, farr, (Camera * n)(*cams), (RenderCfg * n)(*cfgs)

    def render_batch(self, scenes, films, cams, cfgs, spp_per_pass=None, tuning: Tuning | None = None) -> None:
        """All scenes in one captured graph: one samples-per-pass for all
        (tuning, default the context's) or one per scene (spp_per_pass)."""
        n, sarr, farr, carr, garr = self._batch_arrays(scenes, cams, cfgs, films)
        if spp_per_pass is not None:
            sp = np.ascontiguousarray(spp_per_pass, np.uint32)
            check(lib().rc_render_batch_plan(self.h, sarr, farr, carr, garr, n, _ptr(sp)))
        else:
            check(lib().rc_render_batch(self.h, sarr, farr, carr, garr, n, C.byref(tuning) if tuning else None))

    def render_batch_tuned(self, scenes, films, cams, cfgs, plans, branches: int = 1) -> None:
        """rc_render_batch_tuned: one BatchPlan per scene, frames over
        `branches` concurrent graph branches."""
        n, sarr, farr, carr, garr = self._batch_arrays(scenes, cams, cfgs, films)
        parr = plans if isinstance(plans, C.Array) else (BatchPlan * n)(*plans)
        check(lib().rc_render_batch_tuned(self.h, sarr, farr, carr, garr, n, C.addressof(parr), int(branches)))

    def autotune_batch_plan(self, scenes, cams, cfgs, cache_path=None):
        """rc_autotune_batch_plan: (plans, branches)."""
        n, sarr, _, carr, garr = self._batch_arrays(scenes, cams, cfgs)
        plans = (BatchPlan * n)()
        br = C.c_uint32()
        check(lib().rc_autotune_batch_plan(self.h, sarr, carr, garr, n,
                                           str(cache_path).encode() if cache_path else None, C.addressof(plans),
                                           C.byref(br)))
        return plans, br.value

    def autotune_batch(self, scenes, cams, cfgs, cache_path=None) -> np.ndarray:
        """Samples per pass for each scene from rc_autotune_batch."""
        n, sarr, _, carr, garr = self._batch_arrays(scenes, cams, cfgs)
        out = np.zeros(n, np.uint32)
        check(lib().rc_autotune_batch(self.h, sarr, carr, garr, n, str(cache_path).encode() if cache_path else None,
                                      _ptr(out)))
        return out

    def raytrace(self, scene: "Scene", film: "Film", cam: Camera, cfg: RenderCfg, settings: RaySettings,
                 materials, lights=()) -> None:
        """rc_raytrace: the Whitted tracer's frame, accumulated into the film."""
        mats = materials if isinstance(materials, C.Array) else (RayMaterial * len(materials))(*materials)
        lts = lights if isinstance(lights, C.Array) else (Light * max(1, len(lights)))(*lights)
        check(lib().rc_raytrace(self.h, scene.h, film.h, C.byref(cam), C.byref(cfg), C.byref(settings),
                                C.cast(mats, C.c_void_p), len(mats), C.cast(lts, C.c_void_p), len(lights)))

    def pinned(self, shape, dtype) -> "PinnedArray":
        """A numpy array in page-locked host memory (rc_host_alloc)."""
        return PinnedArray(self, shape, dtype)

    def raycast_phong(self, scene: "Scene", cam: Camera, width: int, height: int, seed: int, offset: float,
                      light: PointLight, out=None):
        """Returns (hit ids, t, rgba); out: the three arrays to fill (numpy
        arrays of those shapes and types, pinned ones for the full copy rate)."""
        n = width * height
        if out is None:
            ids = np.empty(n, np.int32)
            t = np.empty(n, np.float32)
            rgba = np.empty((n, 4), np.float32)
        else:
            ids, t, rgba = (np.asarray(a) for a in out)
            if (ids.shape, ids.dtype, t.shape, t.dtype, rgba.shape, rgba.dtype) != (
                    (n,), np.int32, (n,), np.float32, (n, 4), np.float32) or not all(
                    a.flags.c_contiguous and a.flags.writeable for a in (ids, t, rgba)):
                raise ValueError("raycast_phong: out must be writable contiguous int32 (n,), float32 (n,), "
                                 "float32 (n, 4)")
        check(lib().rc_raycast_phong(self.h, scene.h, C.byref(cam), width, height, seed, offset, C.byref(light),
                                     _ptr(ids), _ptr(t), _ptr(rgba)))
        return ids, t, rgba


class PinnedArray:
    """Page-locked host memory (rc_host_alloc) viewed as a numpy array; freed
    by close()."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        check(lib().rc_host_alloc(ctx.h, max(nbytes, 1), C.byref(p)))
        self.p = p
        buf = (C.c_uint8 * max(nbytes, 1)).from_address(p.value)
        self.array = np.frombuffer(buf, np.uint8, nbytes).view(dt).reshape(shape)

    def close(self) -> None:
        if self.p:
            self.array = None
            check(lib().rc_host_free(self.ctx.h, self.p))
            self.p = None


class RadianceCache:
    """The MLP radiance cache (rc_cache_*)."""

    def __init__(self, ctx: Context, bounds_min, bounds_max, lr=1e-3, seed=1, max_batch=65536, train_stride=0,
                 loss_eps=1e-2):
        self.ctx = ctx
        cfg = CacheCfg(lr, 0.9, 0.99, 1e-8, loss_eps, seed, max_batch, train_stride)
        lo = (C.c_float * 3)(*bounds_min)
        hi = (C.c_float * 3)(*bounds_max)
        h = C.c_void_p()
        check(lib().rc_cache_create(ctx.h, C.byref(cfg), lo, hi, C.byref(h)))
        self.h = h

    def params(self) -> np.ndarray:
        p = np.empty(CACHE_PARAMS, np.float32)
        check(lib().rc_cache_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(lib().rc_cache_set_params(self.h, _ptr(p)))

    def infer(self, queries) -> np.ndarray:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 12)
        out = np.empty((q.shape[0], 3), np.float32)
        check(lib().rc_cache_infer(self.ctx.h, self.h, _ptr(q), q.shape[0], _ptr(out)))
        return out

    def gradient(self, queries, targets):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.empty(CACHE_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().rc_cache_gradient(self.ctx.h, self.h, _ptr(q), _ptr(t), q.shape[0], _ptr(g), C.byref(loss)))
        return g, loss.value

    def train_step(self, queries, targets) -> float:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = C.c_float()
        check(lib().rc_cache_train_step(self.ctx.h, self.h, _ptr(q), _ptr(t), q.shape[0], C.byref(loss)))
        return loss.value

    def render(self, scene: "Scene", film: "Film", cam: Camera, cfg: RenderCfg, train: int = 1,
               want_loss: bool = True) -> float:
        """One frame with paths ending into the cache at cfg.cache_vertex; then
        `train` Adam steps (0 = none) on interleaved slices of the frame's
        training samples, queued behind the frame.  With want_loss the call
        waits for them and returns their mean loss; without, it returns 0 and
        the steps overlap whatever the caller does next."""
        loss = C.c_float()
        check(lib().rc_render_cached(self.ctx.h, scene.h, film.h, C.byref(cam), C.byref(cfg), self.h, int(train),
                                     C.byref(loss) if want_loss else None))
        return loss.value

    def close(self) -> None:
        if self.h:
            lib().rc_cache_free(self.h)
            self.h = None


class Scene:
    def __init__(self, ctx: Context, sd: SceneData, nodes: np.ndarray, ordered: np.ndarray):
        self.ctx = ctx
        h = C.c_void_p()
        v = np.ascontiguousarray(sd.vertices, np.float32)
        m = np.ascontiguousarray(sd.tri_material, np.uint32)
        mats = np.ascontiguousarray(sd.materials)
        nodes = np.ascontiguousarray(nodes)
        ordered = np.ascontiguousarray(ordered, np.uint64)
        check(lib().rc_scene_upload(ctx.h, _ptr(v), v.shape[0], _ptr(m), _ptr(mats), mats.shape[0], _ptr(nodes),
                                    nodes.shape[0], _ptr(ordered), C.byref(h)))
        self.h = h

    def info(self) -> tuple[int, int, int]:
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(lib().rc_scene_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def close(self) -> None:
        if self.h:
            lib().rc_scene_free(self.h)
            self.h = None


class Film:
    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx, self.width, self.height = ctx, width, height
        h = C.c_void_p()
        check(lib().rc_film_create(ctx.h, width, height, C.byref(h)))
        self.h = h

    def clear(self) -> None:
        check(lib().rc_film_clear(self.ctx.h, self.h))

    def linear(self) -> np.ndarray:
        out = np.empty((self.width * self.height, 4), np.float32)
        check(lib().rc_film_download_linear(self.ctx.h, self.h, _ptr(out)))
        return out

    def sums(self) -> np.ndarray:
        out = np.empty((self.width * self.height, 4), np.float32)
        check(lib().rc_film_download_sums(self.ctx.h, self.h, _ptr(out)))
        return out

    def tonemap(self, key=0.18, gamma=2.2, clamp=True, apply_gamma=True, fast_pow=True, type=RC_TONEMAP_REINHARD,
                exposure=0.0, enable_averaging=False, averaging_alpha=0.1):
        """(rgba, log_avg, max_lum); log_avg and max_lum are Reinhard's (0 for the other mappers)."""
        out = np.empty((self.width * self.height, 4), np.float32)
        tm = ToneMap(key, gamma, int(clamp), int(apply_gamma), int(fast_pow), type, exposure, int(enable_averaging),
                     averaging_alpha, 0)
        la, ml = C.c_double(), C.c_double()
        check(lib().rc_film_tonemap(self.ctx.h, self.h, C.byref(tm), _ptr(out), C.byref(la), C.byref(ml)))
        return out, la.value, ml.value

    def close(self) -> None:
        if self.h:
            lib().rc_film_free(self.h)
            self.h = None


def make_rays(org: np.ndarray, dirs: np.ndarray, tmin=0.0, tmax=np.inf) -> np.ndarray:
    n = dirs.shape[0]
    r = np.empty(n, RAY_DTYPE)
    r["org"] = np.broadcast_to(org, (n, 3))
    r["dir"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r
