"""The allocation-free closest-hit entry point (rc_trace_closest_device_noalloc)
against rc_trace_closest_device, through the C ABI, on device-resident rays.

Two calls in a row (the second with fewer rays), each bit-identical to the
allocating entry point on the same rays; once more with the on-chip stack
limited to one entry, so every traversal goes through the deep-stack kernel
and its scratch; and the size limit rejected with a clear error.
"""
import ctypes as C

import numpy as np
import pytest

import raycer_amd as ra
from raycer_amd import api
from tests.helpers import camera_rays, random_rays

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20


def _lib():
    L = api.lib()
    f = L.rc_trace_closest_device_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    f.restype = C.c_int
    return L


class DevArray:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.p = C.c_void_p()
        api.check(api.lib().rc_device_alloc(ctx.h, max(nbytes, 16), C.byref(self.p)))

    def close(self):
        api.check(api.lib().rc_device_free(self.ctx.h, self.p))


def _trace_both(ctx, s, rays):
    L = _lib()
    n = len(rays)
    d_rays = DevArray(ctx, rays.nbytes)
    d_a = DevArray(ctx, n * 16)
    d_b = DevArray(ctx, n * 16)
    try:
        api.check(L.rc_memcpy_h2d(ctx.h, d_rays.p, rays.ctypes.data, rays.nbytes))
        api.check(L.rc_trace_closest_device(ctx.h, s.h, d_rays.p, n, d_a.p))
        api.check(L.rc_trace_closest_device_noalloc(ctx.h, s.h, d_rays.p, n, d_b.p))
        a = np.empty(n, api.HIT_DTYPE)
        b = np.empty(n, api.HIT_DTYPE)
        api.check(L.rc_memcpy_d2h(ctx.h, a.ctypes.data, d_a.p, a.nbytes))
        api.check(L.rc_memcpy_d2h(ctx.h, b.ctypes.data, d_b.p, b.nbytes))
        return a, b
    finally:
        for d in (d_rays, d_a, d_b):
            d.close()


def _scene(ctx):
    sd = ra.generate_scene(ra.scene_params(api.RCH_SCENE_SPHERE_GROUND, seed=1, level=5, grid=96))
    nodes, ordered, _, _ = ra.build_bvh(sd.vertices)
    return sd, ctx.upload(sd, nodes, ordered)


def test_noalloc_equals_the_allocating_entry_point(ctx):
    sd, s = _scene(ctx)
    try:
        rng = np.random.default_rng(5)
        w, h = 512, 384
        rays = camera_rays(sd.desc.camera, w, h, jitter=rng.random((w * h, 2)))
        v = sd.vertices.reshape(-1, 3)
        rays = np.concatenate([rays, random_rays(60000, 3, v.min(0), v.max(0))])
        a, b = _trace_both(ctx, s, rays)
        assert a.tobytes() == b.tobytes()
        assert (a["tri"] >= 0).mean() > 0.3
        # a second call with fewer rays reuses the same scratch
        a2, b2 = _trace_both(ctx, s, rays[:12345])
        assert a2.tobytes() == b2.tobytes()
        assert a2.tobytes() == a[:12345].tobytes()
        # every ray through the deep-stack kernel and the scratch's deep stack
        ctx.set_tuning(stack_limit=1)
        try:
            a3, b3 = _trace_both(ctx, s, rays[:50000])
        finally:
            ctx.set_tuning()
        assert a3.tobytes() == b3.tobytes()
        assert a3.tobytes() == a[:50000].tobytes()
    finally:
        s.close()


def test_noalloc_rejects_batches_beyond_its_scratch(ctx):
    sd, s = _scene(ctx)
    L = _lib()
    try:
        rc = L.rc_trace_closest_device_noalloc(ctx.h, s.h, C.c_void_p(16), MAX_RAYS + 1, C.c_void_p(16))
        assert rc == api.RC_ERR_INVALID
        assert "exceed the limit" in api.last_error()
        assert L.rc_trace_closest_device_noalloc(ctx.h, s.h, None, 0, None) == api.RC_OK
    finally:
        s.close()
