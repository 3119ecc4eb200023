"""MI355X (gfx950) triangle path tracer with raycer's Scene/Tracer/Film interface.

The product is libraycer.so (HIP kernels + C ABI, include/raycer_hip.h and
include/raycer_host.h) with the C++ interface in include/raycer.hpp.  This
package holds the build script and a ctypes binding used by the tests and
the benchmark.
"""
from .api import (  # noqa: F401
    Camera, Context, Film, RaycerError, RenderCfg, Scene, SceneData, build_bvh, describe_scene, generate_scene,
    lib, load_bvh, load_obj, make_rays, obj_scene, save_bvh, scene_params,
)
