// FP64 restatement of raycer's camera, sampling, path tracer, film, Phong
// raycast and Reinhard tone mapper.  TEST INFRASTRUCTURE ONLY.
#include <omp.h>

#include <vector>

#include "core.hpp"
#include "oracle.h"

using namespace orc;

extern "C" size_t or_bvh_ordered_count(const or_bvh* b);

// ---------------------------------------------------------------- small pins
// MathUtils::fastPow (src/Math/MathUtils.cpp:56-68): the high 32-bit word of the
// double is treated as an integer, scaled about 1072632447, the low word zeroed.
extern "C" double or_fastpow(double a, double b) {
    uint64_t bits;
    std::memcpy(&bits, &a, 8);
    int32_t hi = int32_t(uint32_t(bits >> 32));
    int32_t nhi = static_cast<int32_t>(b * (hi - 1072632447) + 1072632447);
    uint64_t ob = uint64_t(uint32_t(nhi)) << 32;
    double r;
    std::memcpy(&r, &ob, 8);
    return r;
}
extern "C" void or_normalize(const double in[3], double out[3]) {
    V3 r = normalized(V3(in[0], in[1], in[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
extern "C" void or_cross(const double a[3], const double b[3], double out[3]) {
    V3 r = cross(V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
extern "C" void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr, key, out);
}
extern "C" double or_uniform(uint32_t x) { return u01(x); }

// The filters of Filter::getFilter (src/Rendering/Filters/Filter.cpp:17-29);
// the 2-D weight is the product wx * wy (Filter.cpp:31-39).  p0, p1: the
// Gaussian's standard deviation, or Mitchell's B and C.
namespace {
struct FilterShape {
    uint32_t f;
    double radius, a, b;
};
FilterShape filter_shape(uint32_t f, double r, double p0, double p1) {
    FilterShape s{f, r, 0.0, 0.0};
    switch (f) {
    case 2: s.radius = 1.5; break;                     // BellFilter.cpp:24-28
    case 3:                                            // GaussianFilter.cpp:15-23
        s.a = 1.0 / (std::sqrt(2.0 * M_PI) * p0);
        s.b = -1.0 / (2.0 * p0 * p0);
        s.radius = (7.43384 * p0) / 2.0;
        break;
    case 4: s.radius = 2.0; s.a = p0; s.b = p1; break;  // MitchellFilter.cpp:30-37
    default: break;                                    // box, tent, lanczos: r
    }
    return s;
}
double filter_1d(const FilterShape& F, double x) {
    const double r = F.radius;
    switch (F.f) {
    case 0: return (x >= -r && x < r) ? 1.0 / (2.0 * r) : 0.0;  // BoxFilter.cpp:22-37
    case 1: {                                                  // TentFilter.cpp:12-44
        double rinv = 1.0 / r;
        double s = std::abs(x * rinv);
        return (s < 1.0 ? 1.0 - s : 0.0) * rinv;
    }
    case 2: {  // BellFilter.cpp:12-22
        double s = std::abs(x);
        if (s < 0.5) return 0.75 - s * s;
        if (s <= 1.5) return 0.5 * std::pow(s - 1.5, 2.0);
        return 0.0;
    }
    case 3: return F.a * std::exp(x * x * F.b);  // GaussianFilter.cpp:25-33
    case 4: {                                    // MitchellFilter.cpp:12-23
        const double B = F.a, C = F.b, s = std::abs(x);
        if (s <= 1.0)
            return ((12.0 - 9.0 * B - 6.0 * C) * (s * s * s) + (-18.0 + 12.0 * B + 6.0 * C) * (s * s) + (6.0 - 2.0 * B)) *
                   (1.0 / 6.0);
        if (s <= 2.0)
            return ((-B - 6.0 * C) * (s * s * s) + (6.0 * B + 30.0 * C) * (s * s) + (-12.0 * B - 48.0 * C) * s +
                    (8.0 * B + 24.0 * C)) *
                   (1.0 / 6.0);
        return 0.0;
    }
    default: {  // LanczosSincFilter.cpp:12-30
        double s = std::abs(x);
        if (s == 0.0) return 1.0;
        if (s > r) return 0.0;
        auto sinc = [](double v) { return std::sin(M_PI * v) / (M_PI * v); };
        return sinc(s) * sinc(s / r);
    }
    }
}

// CMJSampler's permute (CMJSampler.cpp:12-44), Kensler's hash permutation
// ("Correlated Multi-Jittered Sampling", Pixar Technical Memo 13-01) in the
// reference's 64-bit arithmetic.
uint64_t cmj_permute(uint64_t i, uint64_t l, uint64_t p) {
    uint64_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893dull;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3full;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1 | p >> 27;
        i *= 0x6935fa69ull;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303ull;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3ull;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfull;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}
}  // namespace

extern "C" double or_filter_weight(uint32_t f, double r, double p0, double p1, double x, double y) {
    FilterShape F = filter_shape(f, r, p0, p1);
    return filter_1d(F, x) * filter_1d(F, y);
}
extern "C" double or_filter_radius(uint32_t f, double r, double p0, double p1) {
    return filter_shape(f, r, p0, p1).radius;
}
extern "C" uint64_t or_cmj_permute(uint64_t i, uint64_t l, uint64_t p) { return cmj_permute(i, l, p); }
/* The CMJ sample (x, y) of an n x n pattern (CMJSampler.cpp:57-77) for
 * permutation p and jitter (r0, r1). */
extern "C" void or_cmj_sample(uint64_t x, uint64_t y, uint64_t n, uint64_t p, double r0, double r1, double out[2]) {
    const uint64_t sx = cmj_permute(x, n, p * 0x68bc21ebull);
    const uint64_t sy = cmj_permute(y, n, p * 0x02e5be93ull);
    out[0] = (double(x) + (double(sy) + r0) / double(n)) / double(n);
    out[1] = (double(y) + (double(sx) + r1) / double(n)) / double(n);
}
extern "C" void or_hemisphere(const double n[3], double u1, double u2, double out[3]) {
    V3 r = hemisphere(V3(n[0], n[1], n[2]), u1, u2);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// ---------------------------------------------------------------- camera
namespace {
struct Cam {
    V3 pos, fwd, right, up, center;
    double ipw, iph, aspect;
};
// EulerAngle::getDirection (EulerAngle.cpp:77-87), ONB::fromNormal (ONB.cpp:32-39),
// Camera::setImagePlaneSize (Camera.cpp:29-34), imagePlaneDistance (Camera.cpp:160),
// cameraState (Camera.cpp:279-283).
Cam make_cam(const double c[7], uint32_t w, uint32_t h) {
    Cam k;
    k.pos = V3(c[0], c[1], c[2]);
    double pitch = c[3], yaw = c[4];
    double fov = std::max(1.0, std::min(c[6], 180.0));
    V3 d(-std::sin(deg2rad(yaw)) * std::cos(deg2rad(pitch)), std::sin(deg2rad(pitch)),
         -std::cos(deg2rad(yaw)) * std::cos(deg2rad(pitch)));
    d = normalized(d);
    k.right = normalized(cross(d, ALMOST_UP));
    k.up = normalized(cross(k.right, d));
    k.fwd = d;
    double dist = 0.5 / std::tan(deg2rad(fov / 2.0));
    k.center = k.pos + k.fwd * dist;
    k.ipw = double(w - 1);
    k.iph = double(h - 1);
    k.aspect = double(h) / double(w);
    return k;
}
// Camera::getRay, perspective branch (Camera.cpp:308-331)
void cam_ray(const Cam& k, double px, double py, V3& o, V3& d) {
    double dx = (px / k.ipw) - 0.5;
    double dy = (py / k.iph) - 0.5;
    V3 p = k.center + (dx * k.right) + ((dy * k.aspect) * k.up);
    o = k.pos;
    d = normalized(p - k.pos);
}
RayD make_ray(V3 o, V3 d, double tmin, double tmax) {
    RayD r;
    r.o = o; r.d = d; r.tmin = tmin; r.tmax = tmax;
    r.inv = V3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    return r;
}
}  // namespace

extern "C" void or_camera_ray(const double cam[7], uint32_t w, uint32_t h, double px, double py,
                              double org[3], double dir[3]) {
    Cam k = make_cam(cam, w, h);
    V3 o, d;
    cam_ray(k, px, py, o, d);
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

// ---------------------------------------------------------------- path tracer
namespace {
struct Scene {
    const or_bvh* bvh;
    const float* verts;
    const uint32_t* tri_mat;
    const double* mats;
    std::vector<uint32_t> tri_flags;  // material flags per triangle
};

// nodes / tris: the reference traversal's node visits and triangle tests
// (the unordered traversal only; the ordered one is not counted).
bool closest(const Scene& s, const RayD& r, uint32_t traversal, double& t, int64_t& id, uint64_t* nodes = nullptr,
             uint64_t* tris = nullptr) {
    if (traversal == 0) return trace_ref(s.bvh, s.verts, r, false, t, id, nullptr, nodes, tris);
    Diag D;
    trace_ordered(s.bvh, s.verts, r, 0.0, 0.0, D);
    t = D.t;
    id = D.id;
    return id >= 0;
}
bool occluded(const Scene& s, const RayD& r) {
    double t;
    int64_t id;
    return trace_ref(s.bvh, s.verts, r, true, t, id, s.tri_flags.data(), nullptr);
}

constexpr int kRenderStats = 14;  // doubles per pixel of or_render

struct PathOut {
    V3 L;
    int rays = 0;
    uint64_t nodes = 0, tris = 0;  // the reference traversal's work for the path's extension rays
};

// PathTracer::traceRecursive (PathTracer.cpp:25-75) unrolled into a loop, plus
// the terms the reference lacks: sky on miss, sun next-event estimation,
// Russian roulette, and the optional unbiased Lambert weight.
//   reference estimator: throughput *= 2 * rho * |n . w|   (PathTracer.cpp:70-71)
//   lambert estimator:   throughput *= rho
// Sun NEE uses the BRDF each estimator implies: (2 rho / pi) cos_i (reference)
// or rho / pi (lambert), times E cos_s; it happens at vertex k only when a
// continuation ray k + 1 would still be allowed.
// tail_vertex / tail_frame: the draws of vertices k >= tail_vertex (the
// bounce from vertex tail_vertex on) use frame tail_frame instead of c.frame,
// so two calls that differ only in tail_frame share the path up to and
// including that vertex's next-event ray and continue independently.
PathOut trace_path(const Scene& s, const or_render_cfg& c, uint32_t pixel, uint32_t sample, V3 o, V3 d,
                   double* trace_out, uint32_t tail_vertex = 0xFFFFFFFFu, uint32_t tail_frame = 0) {
    PathOut P;
    V3 thr(1, 1, 1);
    V3 L(0, 0, 0);
    V3 sky(c.sky[0], c.sky[1], c.sky[2]);
    V3 sun(c.sun_dir[0], c.sun_dir[1], c.sun_dir[2]);
    V3 E(c.sun_irradiance[0], c.sun_irradiance[1], c.sun_irradiance[2]);
    bool has_sun = maxc(E) > 0.0;
    for (uint32_t k = 0; k < c.max_path_length; ++k) {
        RayD r = make_ray(o, d, 0.0, std::numeric_limits<double>::max());
        double t;
        int64_t id;
        bool hit = closest(s, r, c.traversal, t, id, &P.nodes, &P.tris);
        P.rays++;
        if (trace_out) {
            double* q = trace_out + 8 * k;
            q[0] = hit ? t : -1.0; q[1] = double(id);
        }
        if (!hit) {
            L = L + mul(thr, sky);
            if (trace_out) { double* q = trace_out + 8 * k; q[2] = thr.x; q[3] = thr.y; q[4] = thr.z; q[5] = L.x; q[6] = L.y; q[7] = L.z; }
            break;
        }
        const double* m = s.mats + OR_MAT_STRIDE * s.tri_mat[id];
        uint32_t flags = uint32_t(m[13]);
        if (flags & 1u) {
            L = L + mul(thr, V3(m[3], m[4], m[5]));
            if (trace_out) { double* q = trace_out + 8 * k; q[2] = thr.x; q[3] = thr.y; q[4] = thr.z; q[5] = L.x; q[6] = L.y; q[7] = L.z; }
            break;
        }
        Tri tr = load_tri(s.verts, uint64_t(id));
        V3 n = tri_normal(tr);
        V3 p = r.o + (t * r.d);
        V3 rho(m[0], m[1], m[2]);
        double u[4];
        draw4(c.seed, pixel, sample, k >= tail_vertex ? tail_frame : c.frame, k, u);
        bool more = (k + 1) < c.max_path_length;
        if (has_sun && more) {
            double cs = dot(n, sun);
            if (cs > 0.0) {
                RayD sr = make_ray(p + sun * c.ray_start_offset, sun, 0.0, std::numeric_limits<double>::max());
                if (!occluded(s, sr)) {
                    double f = c.estimator == 0 ? (2.0 / M_PI) * cs * cs : (1.0 / M_PI) * cs;
                    L = L + mul(thr, mul(rho, E)) * f;
                }
            }
        }
        if (trace_out) { double* q = trace_out + 8 * k; q[2] = thr.x; q[3] = thr.y; q[4] = thr.z; q[5] = L.x; q[6] = L.y; q[7] = L.z; }
        if (!more) break;
        V3 w = hemisphere(n, u[0], u[1]);
        if (c.estimator == 0) thr = mul(thr, rho * (2.0 * std::abs(dot(w, n))));
        else thr = mul(thr, rho);
        if (c.rr_start > 0 && k >= c.rr_start) {
            double q = std::min(1.0, maxc(thr));
            if (u[2] >= q) break;
            thr = thr * (1.0 / q);
        }
        o = p + w * c.ray_start_offset;
        d = w;
    }
    P.L = L;
    return P;
}

uint32_t isqrt_exact(uint32_t v) {
    uint32_t r = uint32_t(std::sqrt(double(v)));
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

void sample_camera(const Cam& k, const or_render_cfg& c, uint32_t pixel, uint32_t s, V3& o, V3& d, double& w) {
    double px = double(pixel % c.width), py = double(pixel / c.width);
    // Tracer::generateMultiSamples (Tracer.cpp:115-144): one jittered sample is
    // the pixel centre with weight 1; otherwise the sample point (x + r) / n of
    // stratum (x, y) of an n x n grid (JitteredSampler.cpp:21-34) or a uniform
    // point (RandomSampler.cpp:23-35), offset (s - 0.5) * 2 * radius, weighted by
    // the filter.
    // Samplers (src/Rendering/Samplers/*.cpp): 0 jittered, 1 random, 2 center,
    // 3 regular, 4 CMJ with one 64-bit permutation per pixel and frame, 5
    // Poisson disc, whose getSample2D returns (0, 0) (PoissonDiscSampler.cpp:22-32).
    if (c.sampler != 1 && c.spp == 1) {
        w = 1.0;
    } else {
        double u[4];
        draw4(c.seed, pixel, s, c.frame, CAM_DIM, u);
        const uint32_t n = isqrt_exact(c.spp);
        const uint32_t sx = s % n, sy = s / n;
        double ox, oy;
        switch (c.sampler) {
        case 0:
            ox = (double(sx) + u[0]) / double(n);
            oy = (double(sy) + u[1]) / double(n);
            break;
        case 1:
            ox = u[0];
            oy = u[1];
            break;
        case 2: ox = oy = 0.5; break;
        case 3:
            ox = (double(sx) + 0.5) / double(n);
            oy = (double(sy) + 0.5) / double(n);
            break;
        case 4: {
            double o[2];
            or_cmj_sample(sx, sy, n, pixel_permutation(c.seed, pixel, c.frame), u[0], u[1], o);
            ox = o[0];
            oy = o[1];
            break;
        }
        default: ox = oy = 0.0; break;
        }
        const FilterShape F = filter_shape(c.filter, c.filter_radius, c.filter_param[0], c.filter_param[1]);
        ox = (ox - 0.5) * 2.0 * F.radius;
        oy = (oy - 0.5) * 2.0 * F.radius;
        w = filter_1d(F, ox) * filter_1d(F, oy);
        px += ox;
        py += oy;
    }
    cam_ray(k, px, py, o, d);
}
}  // namespace

extern "C" int or_render(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                         const double* mats, size_t nmat, const double cam[7], const or_render_cfg* cfg,
                         const uint32_t* pixel_ids, size_t npix, double* stats, int nthreads,
                         volatile int* interrupt) {
    Scene s{b, vertices, tri_mat, mats, {}};
    s.tri_flags.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) s.tri_flags[i] = uint32_t(mats[OR_MAT_STRIDE * tri_mat[i] + 13]);
    (void)nmat;
    const or_render_cfg c = *cfg;
    Cam k = make_cam(cam, c.width, c.height);
    size_t count = pixel_ids ? npix : size_t(c.width) * c.height;
    uint32_t spp = c.spp;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int stopped = 0;
    // Tracer::run (Tracer.cpp:58-113): dynamic schedule in chunks of 1000 pixels.
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t i = 0; i < int64_t(count); ++i) {
        if (interrupt && *interrupt) { stopped = 1; continue; }
        uint32_t pixel = pixel_ids ? pixel_ids[i] : uint32_t(i);
        double* st = stats + kRenderStats * i;
        for (int q = 0; q < kRenderStats; ++q) st[q] = 0.0;
        for (uint32_t smp = 0; smp < spp; ++smp) {
            V3 o, d;
            double w;
            sample_camera(k, c, pixel, smp, o, d, w);
            PathOut P = trace_path(s, c, pixel, smp, o, d, nullptr);
            double L[3] = {P.L.x, P.L.y, P.L.z};
            st[0] += w;
            st[1] += w * w;
            for (int ch = 0; ch < 3; ++ch) {
                st[2 + ch] += w * L[ch];
                st[5 + ch] += w * w * L[ch];
                st[8 + ch] += w * w * L[ch] * L[ch];
            }
            st[11] += P.rays;
            st[12] += double(P.nodes);
            st[13] += double(P.tris);
        }
    }
    return stopped;
}

// Variance a radiance cache at vertex c.cache_vertex could remove: each
// sample's path is traced twice, sharing everything up to that vertex's
// next-event ray and with independent tails after it (tail frames frame' =
// frame ^ 0x9E3779B9 for the second).  Per listed pixel, 7 doubles:
// sum_w, sum_w La[3] (the first path's estimate), sum_w^2 (La - Lb)^2 [3].
// E[(La - Lb)^2] = 2 Var(tail | prefix), so sum_w^2 (La - Lb)^2 / (2 sum_w^2)
// is the tail's share of the pixel estimate's variance (law of total
// variance), which a perfect cache replaces by its expectation.
extern "C" int or_render_split_tail(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                                    const double* mats, const double cam[7], const or_render_cfg* cfg,
                                    const uint32_t* pixel_ids, size_t npix, double* out, int nthreads) {
    Scene s{b, vertices, tri_mat, mats, {}};
    s.tri_flags.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) s.tri_flags[i] = uint32_t(mats[OR_MAT_STRIDE * tri_mat[i] + 13]);
    const or_render_cfg c = *cfg;
    if (c.cache_vertex == 0) return 1;
    Cam k = make_cam(cam, c.width, c.height);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const uint32_t tf = c.frame ^ 0x9E3779B9u;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < int64_t(npix); ++i) {
        uint32_t pixel = pixel_ids[i];
        double* q = out + 7 * i;
        for (int j = 0; j < 7; ++j) q[j] = 0.0;
        for (uint32_t smp = 0; smp < c.spp; ++smp) {
            V3 o, d;
            double w;
            sample_camera(k, c, pixel, smp, o, d, w);
            PathOut A = trace_path(s, c, pixel, smp, o, d, nullptr, c.cache_vertex, c.frame);
            PathOut B = trace_path(s, c, pixel, smp, o, d, nullptr, c.cache_vertex, tf);
            const double la[3] = {A.L.x, A.L.y, A.L.z}, lb[3] = {B.L.x, B.L.y, B.L.z};
            q[0] += w;
            for (int ch = 0; ch < 3; ++ch) {
                q[1 + ch] += w * la[ch];
                q[4 + ch] += w * w * (la[ch] - lb[ch]) * (la[ch] - lb[ch]);
            }
        }
    }
    return 0;
}

extern "C" int or_trace_path(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                             const double* mats, const double cam[7], const or_render_cfg* cfg, uint32_t pixel,
                             uint32_t sample, double L_out[3], double* weight_out, double* vertices_out) {
    Scene s{b, vertices, tri_mat, mats, {}};
    s.tri_flags.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) s.tri_flags[i] = uint32_t(mats[OR_MAT_STRIDE * tri_mat[i] + 13]);
    Cam k = make_cam(cam, cfg->width, cfg->height);
    V3 o, d;
    double w;
    sample_camera(k, *cfg, pixel, sample, o, d, w);
    PathOut P = trace_path(s, *cfg, pixel, sample, o, d, vertices_out);
    L_out[0] = P.L.x; L_out[1] = P.L.y; L_out[2] = P.L.z;
    *weight_out = w;
    return P.rays;
}

// ---------------------------------------------------------------- Phong raycast
// Raytracer::calculateLightColor for one point light (Raytracer.cpp:312-384),
// calculatePhongShadingColor (:386-407) and the point-light hard shadow
// (:524-551).  The primary ray goes through pixel (x + u0, y + u1) with
// (u0, u1) the CAM_DIM draw of (pixel, sample 0, frame 0).
extern "C" void or_raycast_phong(const or_bvh* b, const float* vertices, const uint32_t* tri_mat,
                                 const double* mats, const double cam[7], uint32_t w, uint32_t h, uint64_t seed,
                                 double offset, const double light[9], const double amb[4],
                                 const uint32_t* pixel_ids, size_t npix, int64_t* hit_out, double* t_out,
                                 double* rgb_out, int nthreads) {
    Cam k = make_cam(cam, w, h);
    size_t count = pixel_ids ? npix : size_t(w) * h;
    V3 lp(light[0], light[1], light[2]);
    V3 lc = V3(light[3], light[4], light[5]) * light[6];
    double maxd = light[7], att = light[8];
    V3 ambient = V3(amb[0], amb[1], amb[2]) * amb[3];
    size_t ntri = or_bvh_ordered_count(b);
    std::vector<uint32_t> tri_flags(ntri);
    for (size_t i = 0; i < ntri; ++i) tri_flags[i] = uint32_t(mats[OR_MAT_STRIDE * tri_mat[i] + 13]);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t i = 0; i < int64_t(count); ++i) {
        uint32_t pixel = pixel_ids ? pixel_ids[i] : uint32_t(i);
        double u[4];
        draw4(seed, pixel, 0, 0, CAM_DIM, u);
        V3 o, d;
        cam_ray(k, double(pixel % w) + u[0], double(pixel / w) + u[1], o, d);
        RayD r = make_ray(o, d, 0.0, std::numeric_limits<double>::max());
        Diag D;
        trace_ordered(b, vertices, r, 0.0, 0.0, D);
        hit_out[i] = D.id;
        t_out[i] = D.t;
        V3 col(0, 0, 0);
        if (D.id >= 0) {
            const double* m = mats + OR_MAT_STRIDE * tri_mat[D.id];
            V3 kd(m[0], m[1], m[2]), ka(m[6], m[7], m[8]), ks(m[9], m[10], m[11]);
            double shin = m[12];
            Tri tr = load_tri(vertices, uint64_t(D.id));
            V3 n = tri_normal(tr);
            V3 p = o + (D.t * d);
            col = mul(ambient, ka);
            V3 tl = lp - p;
            double dist = length(tl);
            V3 l = normalized(tl);
            V3 ph(0, 0, 0);
            double diff = dot(l, n);
            if (diff > 0.0) {
                ph = mul(lc * diff, kd);
                if (!(ks.x == 0.0 && ks.y == 0.0 && ks.z == 0.0)) {
                    V3 refl = normalized((2.0 * diff * n) - l);
                    double sa = dot(refl, V3(-d.x, -d.y, -d.z));
                    if (sa > 0.0) ph = ph + mul(lc * std::pow(sa, shin), ks);
                }
            }
            V3 ldir = normalized(lp - p);
            RayD sr = make_ray(p + ldir * offset, ldir, std::numeric_limits<double>::lowest(), length(lp - p));
            double t2;
            int64_t id2;
            // shadow rays skip nonShadowing materials (Triangle.cpp:55-56) and
            // stop at the first hit (fastOcclusion, FlatBVH.cpp:49-50)
            bool shadow = trace_ref(b, vertices, sr, true, t2, id2, tri_flags.data(), nullptr);
            double da = std::min(1.0, dist / maxd);
            da = 1.0 - std::pow(da, att);
            col = col + ph * (da * (shadow ? 0.0 : 1.0));
        }
        rgb_out[3 * i] = col.x; rgb_out[3 * i + 1] = col.y; rgb_out[3 * i + 2] = col.z;
    }
}

// ---------------------------------------------------------------- tone mappers
// ToneMapper::getToneMapper's four mappers (ToneMapper.cpp:13-23) on the linear
// image: Reinhard below (with MovingAverage::addMeasurement on the maximum,
// MovingAverage.cpp:28-31, when averaging), Linear c * fastPow(2, exposure)
// (LinearToneMapper.cpp:12-35), Simple c' / (1 + c') of that
// (SimpleToneMapper.cpp:12-37), Passthrough a copy (PassthroughToneMapper.cpp:12-22).
extern "C" void or_tonemap(const float* lin, size_t npix, const double prm[9], float* out, double* log_avg_out,
                           double* max_lum_out, double* max_lum_average) {
    const uint32_t type = uint32_t(prm[0]);
    const double key = prm[1], gamma = prm[2], exposure = prm[6], alpha = prm[8];
    const bool do_clamp = prm[3] != 0.0, do_gamma = prm[4] != 0.0, fast = prm[5] != 0.0, avg = prm[7] != 0.0;
    if (log_avg_out) *log_avg_out = 0.0;
    if (max_lum_out) *max_lum_out = 0.0;
    if (type == 3) {
        for (size_t i = 0; i < 4 * npix; ++i) out[i] = lin[i];
        return;
    }
    double scale = 0.0, inv2 = 0.0;
    if (type == 0) {
        double logsum = 0.0, maxl = 1.0;
        for (size_t i = 0; i < npix; ++i) {
            double l = 0.2126 * double(lin[4 * i]) + 0.7152 * double(lin[4 * i + 1]) + 0.0722 * double(lin[4 * i + 2]);
            logsum += std::log(0.01 + l);
            if (l > maxl) maxl = l;
        }
        if (avg) {
            *max_lum_average = alpha * maxl + (1.0 - alpha) * *max_lum_average;
            maxl = *max_lum_average;
        }
        double logavg = std::exp(logsum / double(npix));
        scale = key / logavg;
        inv2 = 1.0 / (maxl * maxl);
        if (log_avg_out) *log_avg_out = logavg;
        if (max_lum_out) *max_lum_out = maxl;
    }
    const double expo = fast ? or_fastpow(2.0, exposure) : std::pow(2.0, exposure);
    const double ig = 1.0 / gamma;
    for (size_t i = 0; i < npix; ++i) {
        double c[3] = {lin[4 * i], lin[4 * i + 1], lin[4 * i + 2]};
        if (type == 0) {
            double l = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
            double sl = scale * l;
            double cs = ((sl * (1.0 + (sl * inv2))) / (1.0 + sl)) / l;
            for (double& v : c) v *= cs;
        } else {
            for (double& v : c) v *= expo;
            if (type == 2)
                for (double& v : c) v = v / (1.0 + v);
        }
        for (int k = 0; k < 3; ++k) {
            double v = c[k];
            if (do_clamp) v = std::max(0.0, std::min(v, 1.0));
            if (do_gamma) v = fast ? or_fastpow(v, ig) : std::pow(v, ig);
            out[4 * i + k] = float(v);
        }
        out[4 * i + 3] = 1.0f;
    }
}

// ---------------------------------------------------------------- Reinhard
// ReinhardToneMapper::apply (ReinhardToneMapper.cpp:18-86) without the moving
// average (enableAveraging = false, Scene.h default).  Luminance weights from
// Color::getLuminance (Color.cpp:130-133); clamp as Color::clamped
// (Color.cpp:166-176, std::max(0, std::min(c, 1)) so NaN becomes 0); gamma via
// fastPow (Color.cpp:260-270) or std::pow.
extern "C" void or_reinhard(const float* lin, size_t npix, const double prm[5], float* out,
                            double* log_avg_out, double* max_lum_out) {
    const double eps = 0.01;
    double key = prm[0], gamma = prm[1];
    bool do_clamp = prm[2] != 0.0, do_gamma = prm[3] != 0.0, fast = prm[4] != 0.0;
    double logsum = 0.0, maxl = 1.0;
    for (size_t i = 0; i < npix; ++i) {
        double r = lin[4 * i], g = lin[4 * i + 1], b = lin[4 * i + 2];
        double l = 0.2126 * r + 0.7152 * g + 0.0722 * b;
        logsum += std::log(eps + l);
        if (l > maxl) maxl = l;
    }
    double logavg = std::exp(logsum / double(npix));
    double scale = key / logavg;
    double inv2 = 1.0 / (maxl * maxl);
    double ig = 1.0 / gamma;
    for (size_t i = 0; i < npix; ++i) {
        double c[3] = {lin[4 * i], lin[4 * i + 1], lin[4 * i + 2]};
        double l = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
        double sl = scale * l;
        double ml = (sl * (1.0 + (sl * inv2))) / (1.0 + sl);
        double cs = ml / l;
        for (int k = 0; k < 3; ++k) {
            double v = c[k] * cs;
            if (do_clamp) v = std::max(0.0, std::min(v, 1.0));
            if (do_gamma) v = fast ? or_fastpow(v, ig) : std::pow(v, ig);
            out[4 * i + k] = float(v);
        }
        out[4 * i + 3] = 1.0f;
    }
    if (log_avg_out) *log_avg_out = logavg;
    if (max_lum_out) *max_lum_out = maxl;
}

// ---------------------------------------------------------------- Whitted ray tracer
// Raytracer::traceRecursive (Raytracer.cpp:27-310) and its helpers
// (:312-597), recursive as the reference is, in FP64.  The per-hit draws are
// keyed by the node's place in the ray tree (root key 0, child key =
// hash(key + golden * (kind << 16 | s) + golden)), so the GPU's level-by-level
// wavefront reproduces the same rays.
namespace {
uint32_t wt_hash(uint32_t x) {  // lowbias32 (C. Wellons)
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
uint32_t wt_child(uint32_t key, uint32_t kind, uint32_t s) { return wt_hash(key + 0x9E3779B9u * (1u + ((kind << 16) | s))); }
const uint32_t P_REFL = 1, P_TRANS = 2, P_AO = 3, P_LIGHT = 16, PERM_S = 0xFFFFu;

struct WS {
    Scene sc;
    const or_ray_settings* rs;
    const or_ray_material* rm;
    const or_light* lights;
    uint32_t nlights;
    double offset;
    uint64_t seed;
    uint32_t frame;
};
struct WNode {
    uint32_t pixel, sample, key;
};

void node_bits(const WS& S, const WNode& N, uint32_t p, uint32_t s, uint32_t o[4]) {
    uint32_t ctr[4] = {N.pixel, N.sample, S.frame, wt_hash(N.key ^ wt_hash((p << 16) | s))};
    uint32_t key[2] = {uint32_t(S.seed), uint32_t(S.seed >> 32)};
    philox4x32_10(ctr, key, o);
}
void node_uv(const WS& S, const WNode& N, uint32_t p, uint32_t s, double& u, double& v) {
    uint32_t o[4];
    node_bits(S, N, p, s, o);
    u = u01(o[0]);
    v = u01(o[1]);
}
uint64_t node_perm(const WS& S, const WNode& N, uint32_t p) {
    uint32_t o[4];
    node_bits(S, N, p, PERM_S, o);
    return uint64_t(o[0]) | (uint64_t(o[1]) << 32);
}
// Sampler::getSample2D of each SamplerType
void sample_point(uint32_t type, uint32_t x, uint32_t y, uint32_t n, uint64_t perm, double r0, double r1, double& px,
                  double& py) {
    switch (type) {
    case 0: px = (x + r0) / n; py = (y + r1) / n; break;
    case 1: px = r0; py = r1; break;
    case 2: px = py = 0.5; break;
    case 3: px = (x + 0.5) / n; py = (y + 0.5) / n; break;
    case 4: {
        double o[2];
        or_cmj_sample(x, y, n, perm, r0, r1, o);
        px = o[0];
        py = o[1];
        break;
    }
    default: px = py = 0.0; break;
    }
}
// ONB::fromNormal (ONB.cpp:32-39)
void onb(V3 w, V3& u, V3& v) {
    u = normalized(cross(w, ALMOST_UP));
    v = normalized(cross(u, w));
}
// mapToHemisphere (Sampler.cpp:50-64)
V3 map_hemisphere(V3 u, V3 v, V3 w, double e, double px, double py) {
    double phi = 2.0 * M_PI * px;
    double ct = std::pow(1.0 - py, 1.0 / (e + 1.0));
    double st = std::sqrt(1.0 - ct * ct);
    return (st * std::cos(phi)) * u + (st * std::sin(phi)) * v + ct * w;
}
V3 reflect(V3 a, V3 n) { return a - (2.0 * dot(a, n)) * n; }
V3 lerp(V3 a, V3 b, double t) { return a + (b - a) * t; }
V3 f3(const float* a) { return V3(a[0], a[1], a[2]); }

// calculatePhongShadingColor (Raytracer.cpp:386-407)
V3 phong(V3 n, V3 l, V3 v, V3 lc, V3 kd, V3 ks, double shin) {
    V3 c(0, 0, 0);
    double diff = dot(l, n);
    if (diff > 0.0) {
        c = mul(lc * diff, kd);
        if (!(ks.x == 0.0 && ks.y == 0.0 && ks.z == 0.0)) {
            V3 r = normalized((2.0 * diff * n) - l);
            double sa = dot(r, v);
            if (sa > 0.0) c = c + mul(lc * std::pow(sa, shin), ks);
        }
    }
    return c;
}

bool wt_closest(const WS& S, V3 o, V3 d, double tmax, double& t, int64_t& id) {
    return closest(S.sc, make_ray(o, d, 0.0, tmax), 1, t, id);
}
bool wt_shadowed(const WS& S, V3 o, V3 d, double tmax) { return occluded(S.sc, make_ray(o, d, 0.0, tmax)); }

// calculateSimpleFogColor (Raytracer.cpp:409-425)
V3 fog(const WS& S, double dist, V3 p, V3 c) {
    const or_ray_settings& R = *S.rs;
    double t1 = std::max(0.0, std::min(dist / R.fog_distance, 1.0));
    t1 = std::pow(t1, R.fog_steepness);
    if (R.fog_height_dispersion && p.y > 0.0) {
        double t2 = std::max(0.0, std::min(p.y / R.fog_height, 1.0));
        t2 = std::pow(t2, R.fog_height_steepness);
        t1 *= 1.0 - t2;
    }
    return lerp(c, f3(R.fog_color), t1);
}

// calculateShadowAmount for point and spot lights (Raytracer.cpp:524-597)
double shadow_amount(const WS& S, const WNode& N, uint32_t li, V3 p) {
    const or_light& L = S.lights[li];
    V3 lp = f3(L.position);
    V3 l = normalized(lp - p);
    if (!L.area_light) return wt_shadowed(S, p + l * S.offset, l, length(lp - p)) ? 1.0 : 0.0;
    V3 right = normalized(cross(l, ALMOST_UP));
    V3 up = normalized(cross(right, l));
    uint64_t perm = node_perm(S, N, P_LIGHT + li);
    uint32_t n = L.area_n;
    double amount = 0.0;
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            double r0, r1, px, py;
            node_uv(S, N, P_LIGHT + li, y * n + x, r0, r1);
            sample_point(L.area_sampler, x, y, n, perm, r0, r1, px, py);
            // getDiscSample: mapToDisc (Sampler.cpp:36-48) times the radius
            double th = 2.0 * M_PI * px, rr = std::sqrt(py);
            double jx = rr * std::cos(th) * L.area_radius, jy = rr * std::sin(th) * L.area_radius;
            V3 np = lp + (jx * right + jy * up);
            V3 dl = normalized(np - p);
            if (wt_shadowed(S, p + dl * S.offset, dl, length(np - p))) amount += 1.0;
        }
    return amount / (double(n) * double(n));
}

// calculateAmbientOcclusionAmount (Raytracer.cpp:453-494); AO rays are not
// shadow rays, so nonShadowing triangles occlude them
double ao_amount(const WS& S, const WNode& N, V3 p, V3 nrm) {
    const or_ray_settings& R = *S.rs;
    V3 u, v;
    onb(nrm, u, v);
    uint64_t perm = node_perm(S, N, P_AO);
    uint32_t n = R.ao_n;
    double occ = 0.0;
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            double r0, r1, px, py;
            node_uv(S, N, P_AO, y * n + x, r0, r1);
            sample_point(R.ao_sampler, x, y, n, perm, r0, r1, px, py);
            V3 sd = map_hemisphere(u, v, nrm, R.ao_distribution, px, py);
            double t;
            int64_t id;
            if (wt_closest(S, p + sd * S.offset, sd, R.ao_max_distance, t, id)) occ += 1.0;
        }
    return 1.0 - occ / (double(n) * double(n));
}

// calculateLightColor (Raytracer.cpp:312-384): ambient (with occlusion), then
// the directional, point and spot lights, each group in list order.
V3 light_color(const WS& S, const WNode& N, V3 d, V3 p, V3 nrm, const double* m) {
    const or_ray_settings& R = *S.rs;
    V3 kd(m[0], m[1], m[2]), ka(m[6], m[7], m[8]), ks(m[9], m[10], m[11]);
    double shin = m[12];
    V3 v(-d.x, -d.y, -d.z);
    double ao = R.ao_enabled ? ao_amount(S, N, p, nrm) : 1.0;
    V3 c = mul(f3(R.ambient_color) * double(R.ambient_intensity) * ao, ka);
    for (uint32_t type = 0; type < 3; ++type)
        for (uint32_t li = 0; li < S.nlights; ++li) {
            const or_light& L = S.lights[li];
            if (L.type != type) continue;
            V3 lc = f3(L.color) * double(L.intensity);
            if (type == 0) {
                V3 l = f3(L.direction) * -1.0;
                V3 pc = phong(nrm, l, v, lc, kd, ks, shin);
                double sh = wt_shadowed(S, p + l * S.offset, l, std::numeric_limits<double>::max()) ? 1.0 : 0.0;
                c = c + pc * (1.0 - sh);
                continue;
            }
            V3 tl = f3(L.position) - p;
            double dist = length(tl);
            V3 l = normalized(tl);
            V3 pc = phong(nrm, l, v, lc, kd, ks, shin);
            double sh = shadow_amount(S, N, li, p);
            double da = std::min(1.0, dist / L.max_distance);
            da = 1.0 - std::pow(da, double(L.attenuation));
            if (type == 1) {
                c = c + pc * (da * (1.0 - sh));
            } else {
                double sa = dot(f3(L.direction), l * -1.0);
                if (sa > 0.0) {
                    sa = std::min(1.0, (1.0 - sa) / (L.angle / 180.0));
                    sa = 1.0 - std::pow(sa, double(L.side_attenuation));
                    c = c + pc * (da * sa * (1.0 - sh));
                }
            }
        }
    return c;
}

V3 wt_trace(const WS& S, V3 o, V3 d, uint32_t iter, const WNode& N, bool& found, double& dist, uint64_t& rays);

// calculateReflectedColor / calculateTransmittedColor (Raytracer.cpp:160-310)
V3 child_color(const WS& S, const or_ray_material& M, uint32_t kind, V3 p, V3 nrm, bool outside, V3 dir, double k,
               uint32_t n, uint32_t sampler, double glossiness, uint32_t iter, const WNode& N, uint64_t& rays) {
    V3 attc = f3(M.attenuation_color);
    if (n == 0) {
        bool cf = false;
        double cd = 0.0;
        WNode C{N.pixel, N.sample, wt_child(N.key, kind, 0)};
        V3 c = wt_trace(S, p + dir * S.offset, dir, iter + 1, C, cf, cd, rays) * k;
        // only attenuate if the ray has travelled inside a primitive
        bool inside_path = kind == P_REFL ? !outside : outside;
        if (inside_path && cf && M.enable_attenuation) c = lerp(attc, c, std::exp(-M.attenuation_factor * cd));
        return c;
    }
    V3 u, v;
    onb(dir, u, v);
    uint64_t perm = node_perm(S, N, kind);
    V3 sum(0, 0, 0);
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x = 0; x < n; ++x) {
            uint32_t s = y * n + x;
            double r0, r1, px, py;
            node_uv(S, N, kind, s, r0, r1);
            sample_point(sampler, x, y, n, perm, r0, r1, px, py);
            V3 sd = map_hemisphere(u, v, dir, glossiness, px, py);
            double sn = dot(sd, nrm);
            bool flip = kind == P_REFL ? ((outside && sn < 0.0) || (!outside && sn > 0.0))
                                       : ((outside && sn > 0.0) || (!outside && sn < 0.0));
            if (flip) sd = reflect(sd, dir);
            bool cf = false;
            double cd = 0.0;
            WNode C{N.pixel, N.sample, wt_child(N.key, kind, s)};
            V3 c = wt_trace(S, p + sd * S.offset, sd, iter + 1, C, cf, cd, rays) * k;
            if (!outside && cf && M.enable_attenuation) c = lerp(attc, c, std::exp(-M.attenuation_factor * cd));
            sum = sum + c;
        }
    return sum * (1.0 / (double(n) * double(n)));
}

V3 wt_trace(const WS& S, V3 o, V3 d, uint32_t iter, const WNode& N, bool& found, double& dist, uint64_t& rays) {
    const or_ray_settings& R = *S.rs;
    double t;
    int64_t id;
    ++rays;
    found = wt_closest(S, o, d, std::numeric_limits<double>::max(), t, id);
    if (!found) return f3(R.background);
    dist = t;
    if (R.visualize_depth) {
        double z = 1.0 - std::min(t, double(R.visualize_depth_max)) / R.visualize_depth_max;
        z = std::pow(z, 2.0);
        return V3(z, z, z);
    }
    const uint32_t mid = S.sc.tri_mat[id];
    const double* m = S.sc.mats + OR_MAT_STRIDE * mid;
    const or_ray_material& M = S.rm[mid];
    V3 p = o + t * d;
    V3 nrm = tri_normal(load_tri(S.sc.verts, uint64_t(id)));
    bool outside = dot(d, nrm) < 0.0;
    if (M.skip_lighting) {
        V3 c(m[0], m[1], m[2]);
        if (R.fog_enabled && outside) c = fog(S, t, p, c);
        return c;
    }
    // calculateRayReflectanceAndTransmittance (Raytracer.cpp:126-158)
    double fr = 1.0, ft = 1.0;
    double eta = M.refractive_index;
    if (M.fresnel_reflection) {
        double cs = dot(d, nrm);
        bool out = cs < 0.0;
        double n1 = out ? 1.0 : eta, n2 = out ? eta : 1.0;
        double rf0 = (n2 - n1) / (n2 + n1);
        rf0 = rf0 * rf0;
        fr = rf0 + (1.0 - rf0) * std::pow(1.0 - std::abs(cs), 5.0);
        ft = 1.0 - fr;
    }
    double Rr = M.ray_reflectance * fr, Tt = M.ray_transmittance * ft;
    V3 refl(0, 0, 0), trans(0, 0, 0);
    if (Rr > 0.0 && iter < R.max_ray_iterations) {
        V3 rd = normalized(d + (2.0 * -dot(d, nrm)) * nrm);
        refl = child_color(S, M, P_REFL, p, nrm, outside, rd, Rr, M.refl_n, M.refl_sampler, M.refl_glossiness, iter, N,
                           rays);
    }
    if (Tt > 0.0 && iter < R.max_ray_iterations) {
        double c1 = dot(d, nrm);
        bool out = c1 < 0.0;
        double n1 = out ? 1.0 : eta, n2 = out ? eta : 1.0;
        double n3 = n1 / n2;
        double c2 = 1.0 - (n3 * n3) * (1.0 - c1 * c1);
        if (c2 > 0.0) {
            V3 td = normalized(d * n3 + (std::abs(c1) * n3 - std::sqrt(c2)) * nrm);
            trans = child_color(S, M, P_TRANS, p, nrm, outside, td, Tt, M.trans_n, M.trans_sampler,
                                M.trans_glossiness, iter, N, rays);
        }
    }
    V3 c = light_color(S, N, d, p, nrm, m) + refl + trans;
    if (outside && R.fog_enabled) c = fog(S, t, p, c);
    return c;
}
}  // namespace

extern "C" void or_raytrace(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                            const double* mats, const double cam[7], const or_render_cfg* cfg,
                            const or_ray_settings* settings, const or_ray_material* rmats, const or_light* lights,
                            uint32_t nlights, const uint32_t* pixel_ids, size_t npix, double* stats, int nthreads) {
    WS S{Scene{b, vertices, tri_mat, mats, {}}, settings, rmats, lights, nlights, cfg->ray_start_offset, cfg->seed,
         cfg->frame};
    S.sc.tri_flags.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) S.sc.tri_flags[i] = uint32_t(mats[OR_MAT_STRIDE * tri_mat[i] + 13]);
    const or_render_cfg c = *cfg;
    Cam k = make_cam(cam, c.width, c.height);
    size_t count = pixel_ids ? npix : size_t(c.width) * c.height;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(count); ++i) {
        uint32_t pixel = pixel_ids ? pixel_ids[i] : uint32_t(i);
        double* st = stats + 12 * i;
        for (int q = 0; q < 12; ++q) st[q] = 0.0;
        for (uint32_t smp = 0; smp < c.spp; ++smp) {
            V3 o, d;
            double w;
            sample_camera(k, c, pixel, smp, o, d, w);
            bool found = false;
            double dist = 0.0;
            uint64_t rays = 0;
            V3 L = wt_trace(S, o, d, 0, WNode{pixel, smp, 0u}, found, dist, rays);
            double Lc[3] = {L.x, L.y, L.z};
            st[0] += w;
            st[1] += w * w;
            for (int ch = 0; ch < 3; ++ch) {
                st[2 + ch] += w * Lc[ch];
                st[5 + ch] += w * w * Lc[ch];
                st[8 + ch] += w * w * Lc[ch] * Lc[ch];
            }
            st[11] += double(rays);
        }
    }
}
