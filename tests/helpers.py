"""Input generators and FP64 recomputations shared by the tests."""
import math

import numpy as np

from raycer_amd import api


def camera_rays(cam, w, h, jitter=None, pixels=None):
    """Camera::getRay for pixel coordinates (x + jx, y + jy) in FP64 numpy,
    rounded to the rc_ray layout.  jitter: (npix, 2) offsets or None."""
    pitch, yaw = math.radians(cam.pitch), math.radians(cam.yaw)
    f = np.array([-math.sin(yaw) * math.cos(pitch), math.sin(pitch), -math.cos(yaw) * math.cos(pitch)])
    f /= np.linalg.norm(f)
    r = np.cross(f, [0.0001, 1.0, 0.0001])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    u /= np.linalg.norm(u)
    dist = 0.5 / math.tan(math.radians(min(max(cam.fov, 1.0), 180.0)) / 2.0)
    pos = np.array(cam.position, np.float64)
    centre = pos + f * dist
    idx = np.arange(w * h) if pixels is None else np.asarray(pixels)
    px = (idx % w).astype(np.float64)
    py = (idx // w).astype(np.float64)
    if jitter is not None:
        px = px + jitter[:, 0]
        py = py + jitter[:, 1]
    dx = px / (w - 1) - 0.5
    dy = py / (h - 1) - 0.5
    p = centre + dx[:, None] * r + (dy * (h / w))[:, None] * u
    d = p - pos
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return api.make_rays(pos.astype(np.float32), d.astype(np.float32))


def mt_fp64(vertices, tri_ids, rays):
    """FP64 Möller–Trumbore distance of each ray to the given triangle."""
    v = vertices[tri_ids].astype(np.float64).reshape(-1, 3, 3)
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    tv = o - v[:, 0]
    q = np.cross(tv, e1)
    return np.einsum("ij,ij->i", e2, q) / det


def mt_fp64_uv(vertices, tri_ids, rays):
    """FP64 Möller–Trumbore barycentrics (u, v) of each ray on the given triangle."""
    v = vertices[tri_ids].astype(np.float64).reshape(-1, 3, 3)
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    tv = o - v[:, 0]
    q = np.cross(tv, e1)
    return np.einsum("ij,ij->i", tv, p) / det, np.einsum("ij,ij->i", d, q) / det


def fp32_uv_tolerance(vertices, tri_ids, rays):
    """Absolute bound on the FP32 decision's barycentrics: FP32_ERR times the
    largest coordinate or edge over the shorter edge, over cond (the rounding
    of o - v0 and of the edges, carried through the division by det)."""
    v = vertices[tri_ids].astype(np.float64).reshape(-1, 3, 3)
    o = rays["org"].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    l1, l2 = np.linalg.norm(e1, axis=1), np.linalg.norm(e2, axis=1)
    big = np.max(np.stack([np.abs(o).max(axis=1), np.abs(v).max(axis=(1, 2)), l1, l2]), axis=0)
    amp = big / np.maximum(np.minimum(l1, l2), 1e-30)
    return np.maximum(1e-5, FP32_ERR * amp / np.maximum(tri_cond(vertices, tri_ids, rays), 1e-30))


def random_rays(n, seed, lo, hi, tmax=np.inf):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ext = hi - lo
    org = lo - 0.5 * ext + rng.random((n, 3)) * 2 * ext
    tgt = lo + rng.random((n, 3)) * ext
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty(n, api.RAY_DTYPE)
    r["org"] = org
    r["dir"] = d
    r["tmin"] = 0.0
    r["tmax"] = tmax
    return r


# The FP32 decision (the default): Möller–Trumbore's numerators and
# determinant each lose a few ulps relative to |e1| |e2| |o - v0| and
# |e1| |e2| |d|, so the error in t is at most about
#   10 * 2^-24 * max(1, |o - v0| / (t |d|)) / cond
# relative to t, for a hit of conditioning cond = |det| / (|e1| |e2| |d|)
# (the second factor matters when the ray starts close to the triangle's
# plane but far from its first vertex).  Measured on 54k soup hits the
# constant is at most 4.3; with a factor 16 for margin:
FP32_ERR = 16.0 * 2.0 ** -24
# ties and edge grazes within this relative band are listed for the FP32
# decision (the FP64 decision's band is 1e-6 / 2e-5) ...
FP32_BAND = 1e-4
# ... and so are hits whose conditioning allows an error above the band
FP32_COND = FP32_ERR / FP32_BAND


def closest_diag(ob, rays, fp32, threads):
    """Oracle closest hits with the near-tie listing of the decision mode:
    (t, id, flags, cond)."""
    if fp32:
        t, i, f, c = ob.trace_diag(rays, rel_tie=FP32_BAND, band=FP32_BAND, threads=threads, with_cond=True)
        f = f | np.where(c < FP32_COND, 8, 0).astype(np.uint32)
        return t, i, f, c
    return ob.trace_diag(rays, threads=threads, with_cond=True)


def tri_cond(vertices, tri_ids, rays):
    """|det| / (|e1| |e2| |d|) of each ray against the given triangle (FP64)."""
    v = vertices[tri_ids].astype(np.float64).reshape(-1, 3, 3)
    d = rays["dir"].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    det = np.einsum("ij,ij->i", e1, np.cross(d, e2))
    return np.abs(det) / (np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1) * np.linalg.norm(d, axis=1))


def fp32_t_tolerance(vertices, tri_ids, rays, t, rel_t=1e-5):
    """Relative bound on the FP32 decision's distance to the given triangles
    (FP32_ERR * max(1, |o - v0| / (t |d|)) / cond), at least rel_t."""
    v0 = vertices[tri_ids].astype(np.float64).reshape(-1, 3, 3)[:, 0]
    o = rays["org"].astype(np.float64)
    dn = np.linalg.norm(rays["dir"].astype(np.float64), axis=1)
    amp = np.maximum(1.0, np.linalg.norm(o - v0, axis=1) / np.maximum(np.abs(t) * dn, 1e-30))
    return np.maximum(rel_t, FP32_ERR * amp / np.maximum(tri_cond(vertices, tri_ids, rays), 1e-30))


def check_closest(gpu_hits, vertices, rays, diag, max_listed=0.01, rel_t=1e-5, fp32=False):
    """Hit index bit-exact against the oracle except listed near-ties; the GPU's
    t within rel_t of the FP64 distance to the triangle it reports (for the
    FP32 decision: within its rounding bound, fp32_t_tolerance)."""
    t_o, id_o, flags = diag[:3]
    gid = gpu_hits["tri"].astype(np.int64)
    mismatch = gid != id_o
    listed = flags != 0
    unlisted = mismatch & ~listed
    assert listed.mean() <= max_listed, f"{listed.mean():.4%} of rays listed as near ties"
    assert not unlisted.any(), (
        f"{unlisted.sum()} unlisted mismatches, e.g. rays {np.nonzero(unlisted)[0][:8].tolist()} "
        f"gpu {gid[unlisted][:8].tolist()} oracle {id_o[unlisted][:8].tolist()}")
    hit = gid >= 0
    if hit.any():
        t64 = mt_fp64(vertices, gid[hit], rays[hit])
        err = np.abs(gpu_hits["t"][hit].astype(np.float64) - t64) / np.maximum(np.abs(t64), 1e-30)
        tol = np.full(err.shape, rel_t)
        if fp32:
            tol = fp32_t_tolerance(vertices, gid[hit], rays[hit], t64, rel_t)
        assert (err <= tol).all(), f"max relative t error {err.max():.3g}, worst excess {(err / tol).max():.3g}x"
        # barycentrics of the reported triangle: inside it, and near the FP64 values
        u, v = gpu_hits["u"][hit].astype(np.float64), gpu_hits["v"][hit].astype(np.float64)
        assert (u >= 0).all() and (v >= 0).all() and (u + v <= 1 + 2.0 ** -22).all()
        u64, v64 = mt_fp64_uv(vertices, gid[hit], rays[hit])
        tol_uv = fp32_uv_tolerance(vertices, gid[hit], rays[hit]) if fp32 else np.full(u.shape, rel_t)
        err_uv = np.maximum(np.abs(u - u64), np.abs(v - v64))
        assert (err_uv <= tol_uv).all(), (
            f"max barycentric error {err_uv.max():.3g}, worst excess {(err_uv / tol_uv).max():.3g}x")
    agree = hit & ~mismatch
    if agree.any() and not fp32:
        err = np.abs(gpu_hits["t"][agree] - t_o[agree]) / t_o[agree]
        assert err.max() <= rel_t
    return mismatch.sum(), listed.sum()
