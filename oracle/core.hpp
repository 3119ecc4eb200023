// FP64 math shared by the oracle translation units.  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double get(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vector3::cross component order (src/Math/Vector3.cpp:205-213)
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3::normalized = v / length, and operator/(v, s) multiplies by 1/s
// (src/Math/Vector3.cpp:57-61, 162-165)
inline V3 normalized(V3 a) { double inv = 1.0 / length(a); return {a.x * inv, a.y * inv, a.z * inv}; }
inline double maxc(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }

// Vector3::ALMOST_UP (src/Math/Vector3.cpp:15)
static const V3 ALMOST_UP(0.0001, 1.0, 0.0001);

inline double deg2rad(double d) { return d * (M_PI / 180.0); }

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Round multipliers and Weyl key increments are the
// published constants of that paper.
inline void philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = uint64_t(M0) * c0;
        uint64_t p1 = uint64_t(M1) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 24-bit uniform in [0,1): exactly representable in FP32, so the GPU's float
// and this double agree bit for bit.
inline double u01(uint32_t x) { return double(x >> 8) * (1.0 / 16777216.0); }

// Counter layout shared with the GPU renderer: (pixel, sample, frame, dim),
// key = seed.  dim = path vertex index k for bounce draws, CAM_DIM for the
// sub-pixel jitter.
static const uint32_t CAM_DIM = 0xFFFFFFFFu;
inline void draw4(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t frame, uint32_t dim, double u[4]) {
    uint32_t ctr[4] = {pixel, sample, frame, dim};
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = u01(o[i]);
}

// The CMJ sampler's per-pixel 64-bit permutation: counter (pixel, 0, frame,
// PERM_DIM), low word first.
static const uint32_t PERM_DIM = 0xFFFFFFFEu;
inline uint64_t pixel_permutation(uint64_t seed, uint32_t pixel, uint32_t frame) {
    uint32_t ctr[4] = {pixel, 0u, frame, PERM_DIM};
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    return uint64_t(o[0]) | (uint64_t(o[1]) << 32);
}

// ---------------------------------------------------------------- geometry
struct Tri {
    V3 v0, v1, v2;
};
inline Tri load_tri(const float* vertices, uint64_t i) {
    const float* p = vertices + 9 * i;
    return {V3(p[0], p[1], p[2]), V3(p[3], p[4], p[5]), V3(p[6], p[7], p[8])};
}

// Triangle::initialize normal (src/Raytracing/Primitives/Triangle.cpp:27)
inline V3 tri_normal(const Tri& t) { return normalized(cross(t.v1 - t.v0, t.v2 - t.v0)); }

// Möller–Trumbore as Triangle::intersect (Triangle.cpp:61-91).
// Returns true and fills t,u,v,det on a hit.
inline bool mt_intersect(const Tri& tr, V3 org, V3 dir, double min_d, double max_d,
                         double& t, double& u, double& v, double& det) {
    V3 e1 = tr.v1 - tr.v0;
    V3 e2 = tr.v2 - tr.v0;
    V3 pvec = cross(dir, e2);
    det = dot(e1, pvec);
    if (std::abs(det) < std::numeric_limits<double>::epsilon()) return false;
    double inv = 1.0 / det;
    V3 tvec = org - tr.v0;
    u = dot(tvec, pvec) * inv;
    if (u < 0.0 || u > 1.0) return false;
    V3 qvec = cross(tvec, e1);
    v = dot(dir, qvec) * inv;
    if (v < 0.0 || (u + v) > 1.0) return false;
    t = dot(e2, qvec) * inv;
    if (t < 0.0) return false;
    if (t < min_d || t > max_d) return false;
    return true;
}

// Same arithmetic with the barycentric bounds widened by `band` (used to list
// rays whose FP32 decision can legitimately differ from FP64).
inline bool mt_intersect_band(const Tri& tr, V3 org, V3 dir, double min_d, double max_d, double band,
                              double& t, double& u, double& v, double& det) {
    V3 e1 = tr.v1 - tr.v0;
    V3 e2 = tr.v2 - tr.v0;
    V3 pvec = cross(dir, e2);
    det = dot(e1, pvec);
    if (det == 0.0) return false;
    double inv = 1.0 / det;
    V3 tvec = org - tr.v0;
    u = dot(tvec, pvec) * inv;
    if (u < -band || u > 1.0 + band) return false;
    V3 qvec = cross(tvec, e1);
    v = dot(dir, qvec) * inv;
    if (v < -band || (u + v) > 1.0 + band) return false;
    t = dot(e2, qvec) * inv;
    if (t < 0.0) return false;
    if (t < min_d || t > max_d) return false;
    return true;
}

struct Box {
    V3 mn{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
    V3 mx{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest()};
    void expand(const Box& o) {
        mn.x = std::min(mn.x, o.mn.x); mn.y = std::min(mn.y, o.mn.y); mn.z = std::min(mn.z, o.mn.z);
        mx.x = std::max(mx.x, o.mx.x); mx.y = std::max(mx.y, o.mx.y); mx.z = std::max(mx.z, o.mx.z);
    }
    V3 center() const { return (mn + mx) * 0.5; }
    V3 extent() const { return mx - mn; }
    // AABB::getSurfaceArea (src/Raytracing/AABB.cpp:166-170)
    double area() const {
        V3 e = extent();
        return 2.0 * (e.x * e.y + e.z * e.y + e.x * e.z);
    }
};

inline Box tri_box(const Tri& t) {
    Box b;
    b.mn = V3(std::min(t.v0.x, std::min(t.v1.x, t.v2.x)), std::min(t.v0.y, std::min(t.v1.y, t.v2.y)),
              std::min(t.v0.z, std::min(t.v1.z, t.v2.z)));
    b.mx = V3(std::max(t.v0.x, std::max(t.v1.x, t.v2.x)), std::max(t.v0.y, std::max(t.v1.y, t.v2.y)),
              std::max(t.v0.z, std::max(t.v1.z, t.v2.z)));
    return b;
}

// AABB::intersects (src/Raytracing/AABB.cpp:56-80): slab test, no t-culling.
inline bool box_hit_ref(const double mn[3], const double mx[3], V3 o, V3 inv) {
    double tx0 = (mn[0] - o.x) * inv.x;
    double tx1 = (mx[0] - o.x) * inv.x;
    double tmin = std::min(tx0, tx1);
    double tmax = std::max(tx0, tx1);
    double ty0 = (mn[1] - o.y) * inv.y;
    double ty1 = (mx[1] - o.y) * inv.y;
    tmin = std::max(tmin, std::min(ty0, ty1));
    tmax = std::min(tmax, std::max(ty0, ty1));
    double tz0 = (mn[2] - o.z) * inv.z;
    double tz1 = (mx[2] - o.z) * inv.z;
    tmin = std::max(tmin, std::min(tz0, tz1));
    tmax = std::min(tmax, std::max(tz0, tz1));
    if (tmax < std::max(tmin, 0.0)) return false;
    return true;
}

// Slab interval for the ordered traversal (entry distance, or +inf on miss).
inline double box_entry(const double mn[3], const double mx[3], V3 o, V3 inv, double tmax_ray) {
    double tx0 = (mn[0] - o.x) * inv.x, tx1 = (mx[0] - o.x) * inv.x;
    double ty0 = (mn[1] - o.y) * inv.y, ty1 = (mx[1] - o.y) * inv.y;
    double tz0 = (mn[2] - o.z) * inv.z, tz1 = (mx[2] - o.z) * inv.z;
    // NaN (0 * inf) is dropped by fmin/fmax, which keeps the interval conservative.
    double tmin = std::fmax(std::fmax(std::fmin(tx0, tx1), std::fmin(ty0, ty1)), std::fmax(std::fmin(tz0, tz1), 0.0));
    double tmax = std::fmin(std::fmin(std::fmax(tx0, tx1), std::fmax(ty0, ty1)), std::fmin(std::fmax(tz0, tz1), tmax_ray));
    tmax *= 1.0 + 1e-12;
    if (tmin > tmax) return std::numeric_limits<double>::infinity();
    return tmin;
}

// ONB of a triangle without texcoords: Triangle::initialize (Triangle.cpp:40-44)
inline void tri_onb(V3 n, V3& T, V3& B) {
    T = normalized(cross(n, ALMOST_UP));
    B = normalized(cross(T, n));
}

// mapToHemisphere with distribution 1 (src/Rendering/Samplers/Sampler.cpp:50-64)
inline V3 hemisphere(V3 n, double u1, double u2) {
    V3 T, B;
    tri_onb(n, T, B);
    double phi = 2.0 * M_PI * u1;
    double cp = std::cos(phi), sp = std::sin(phi);
    double ct = std::pow(1.0 - u2, 1.0 / (1.0 + 1.0));
    double st = std::sqrt(1.0 - ct * ct);
    return (st * cp) * T + (st * sp) * B + ct * n;
}

struct RayD {
    V3 o, d, inv;
    double tmin, tmax;
};
struct Diag {
    double t = std::numeric_limits<double>::infinity();
    int64_t id = -1;
    double u = 0, v = 0, det = 0, cond = 0;
    uint32_t flags = 0;
};

}  // namespace orc

struct or_bvh;
namespace orc {
bool trace_ref(const or_bvh* b, const float* vertices, const RayD& r, bool any, double& best_t, int64_t& best_id,
               const uint32_t* tri_flags, uint64_t* nodes_visited, uint64_t* tris_tested = nullptr);
void trace_ordered(const or_bvh* b, const float* vertices, const RayD& r, double rel_tie, double band, Diag& D);
}  // namespace orc
