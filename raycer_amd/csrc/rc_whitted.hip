// Whitted ray tracer on the device: Raytracer::traceRecursive
// (src/Raytracing/Tracers/Raytracer.cpp:27-310) with calculateLightColor,
// calculatePhongShadingColor, calculateSimpleFogColor, ambient occlusion and
// hard or area-light shadows (:312-597), driven by Tracer::run's camera
// sampling into the film.
//
// The recursion becomes a ray tree traced level by level.  Level 0 holds the
// pass's camera rays; a hit at level k spawns its reflection and transmission
// rays (one each, or n*n glossy ones) into level k + 1 while k <
// maxRayIterations, plus its shadow and ambient-occlusion rays.  Each node
// carries the product of the factors above it (ray reflectance or
// transmittance, fog, the parent's attenuation), so its own terms go straight
// into its pixel sample's radiance L[slot]:
//   * a miss adds w * background;
//   * the parent's attenuation lerp(c_att, R * C, a), a = exp(-f t), becomes
//     w *= a plus (1 - a) * c_att * w_parent;
//   * simple fog lerp(C, fog, t1) becomes w *= 1 - t1 plus t1 * fog * w;
//   * every light adds w * phong * attenuation once per unoccluded shadow ray
//     (divided by n*n for area lights), ambient likewise per unoccluded AO ray.
//
// Per level: count (each node's children, shadow rays and AO rays, and the
// decisions behind them, stored so the write pass cannot disagree), exclusive
// scans of the counts, write, then the sums.  The scans keep parent order, so
// the nodes of one pixel sample stay contiguous and in tree order at every
// level; each sample's terms are then added by the first node of its run, in a
// fixed order, which makes renders bit-reproducible.  A level whose children
// or shadow rays would outgrow their buffers is processed in halves (depth
// first), so memory stays bounded for any glossy sample count.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "rc_internal.hpp"

using namespace rcd;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxLights = 64;
constexpr uint32_t kMaxN = 16;         // glossy / AO / area-light samples per axis
constexpr uint32_t kMaxIterations = 16;

// purposes of the per-node draws, and kinds of children
constexpr uint32_t kPurposeReflect = 1, kPurposeTransmit = 2, kPurposeAO = 3, kPurposeLight = 16;
constexpr uint32_t kPermSample = 0xFFFFu;

struct alignas(16) WLight {
    float4 color;  // rgb = color * intensity, w = type (bits)
    float4 pos;    // xyz position, w max distance
    float4 dir;    // xyz direction, w attenuation
    float4 area;   // x radius, y angle, z side attenuation, w = area flag | sampler << 1 | n << 8 (bits)
};

struct alignas(16) WMaterial {
    float4 a;          // ray reflectance, ray transmittance, refractive index, reflection glossiness
    float4 b;          // transmission glossiness, attenuation factor, flags (bits), samplers (bits)
    float4 att_color;  // rgb
};
constexpr uint32_t kFresnel = 1u, kSkipLighting = 2u, kAttenuate = 4u;

struct WParams {
    float3 background;
    float3 ambient;  // colour * intensity
    float offset;
    uint32_t max_iter;
    uint32_t ao, ao_sampler, ao_n;
    float ao_maxd, ao_dist;
    uint32_t vis_depth;
    float vis_max;
    uint32_t fog, fog_hd;
    float3 fog_color;
    float fog_dist, fog_steep, fog_h, fog_hsteep;
    uint32_t nlights;
    PixelOrder order;
    uint32_t npix, sample0, spp_pass, frame;
    uint64_t seed;
};

// ---------------------------------------------------------------- per-node draws
// lowbias32 (C. Wellons' integer hash), for the keys of the ray tree: the root
// key is 0 and child c of a node with key k has key hash(k + golden * (c + 1)).
__device__ __forceinline__ uint32_t wt_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t child_key(uint32_t key, uint32_t kind, uint32_t s) {
    return wt_hash(key + 0x9E3779B9u * (1u + ((kind << 16) | s)));
}
// Draw s of purpose p of a node: Philox counter (pixel, sample, frame, dim)
// with dim = hash(key ^ hash(p << 16 | s)).
__device__ __forceinline__ uint4 node_bits(const WParams& P, uint32_t slot, uint32_t key, uint32_t p, uint32_t s) {
    const uint32_t pixel = order_to_pixel(P.order, slot_index(slot, P.npix, P.spp_pass));
    const uint32_t sample = P.sample0 + slot_sample_in_pass(slot, P.npix, P.spp_pass);
    const uint32_t dim = wt_hash(key ^ wt_hash((p << 16) | s));
    return philox(make_uint4(pixel, sample, P.frame, dim), make_uint2(uint32_t(P.seed), uint32_t(P.seed >> 32)));
}
__device__ __forceinline__ float2 node_uv(const WParams& P, uint32_t slot, uint32_t key, uint32_t p, uint32_t s) {
    const uint4 o = node_bits(P, slot, key, p, s);
    return make_float2(u01(o.x), u01(o.y));
}
__device__ __forceinline__ uint64_t node_perm(const WParams& P, uint32_t slot, uint32_t key, uint32_t p) {
    const uint4 o = node_bits(P, slot, key, p, kPermSample);
    return uint64_t(o.x) | (uint64_t(o.y) << 32);
}

// ---------------------------------------------------------------- geometry helpers
__device__ __forceinline__ bool nonzero(float3 a) { return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f; }

// ONB::fromNormal (ONB.cpp:32-39); also the frame of a triangle without
// texcoords (Triangle.cpp:40-44).
__device__ __forceinline__ void onb_from(float3 w, float3& u, float3& v) {
    u = normalize3(cross3(w, make_float3(0.0001f, 1.0f, 0.0001f)));
    v = normalize3(cross3(u, w));
}

// mapToHemisphere (Sampler.cpp:50-64): cos(theta) = (1 - y)^(1 / (e + 1)).
__device__ __forceinline__ float3 map_hemisphere(float3 u, float3 v, float3 w, float e, float2 pt) {
    const float phi = 2.0f * 3.14159265358979f * pt.x;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float ct = powf(1.0f - pt.y, 1.0f / (e + 1.0f));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    return ((st * cp) * u + (st * sp) * v) + ct * w;
}

// Vector3::reflect (Vector3.cpp:216-219): a - 2 (a . n) n.
__device__ __forceinline__ float3 reflect3(float3 a, float3 n) { return a - (2.0f * dot3(a, n)) * n; }

// calculatePhongShadingColor (Raytracer.cpp:386-407).
__device__ __forceinline__ float3 phong(float3 n, float3 l, float3 v, float3 lc, float3 kd, float3 ks, float shin) {
    float3 col = make_float3(0.0f, 0.0f, 0.0f);
    const float diff = dot3(l, n);
    if (diff > 0.0f) {
        col = mul3(lc * diff, kd);
        if (nonzero(ks)) {
            const float3 r = normalize3((2.0f * diff * n) - l);
            const float sa = dot3(r, v);
            if (sa > 0.0f) col = col + mul3(lc * powf(sa, shin), ks);
        }
    }
    return col;
}

// ---------------------------------------------------------------- one node
struct LevelIn {
    const rc_ray* rays;
    const float4* A;       // w.rgb, slot (bits)
    const float4* B;       // attenuation colour * parent weight, attenuation factor (0 = none)
    const uint32_t* keys;
    const rc_hit* hits;
};
struct Counts {
    uint32_t* nc;
    uint32_t* ns;
    uint32_t* na;
    uint4* dec;  // light mask lo, light mask hi, flags, unused
};
constexpr uint32_t kDecReflect = 1u, kDecTransmit = 2u, kDecAO = 4u;
struct Outs {
    const uint32_t* oc;
    const uint32_t* os;
    const uint32_t* oa;
    rc_ray* crays;
    float4* cA;
    float4* cB;
    uint32_t* ckeys;
    rc_ray* srays;
    float4* sval;  // contribution rgb, slot (bits)
    rc_ray* arays;
    float4* aval;
    float4* contrib;  // the node's own terms, slot (bits)
};

struct Geo {
    float3 o, d, p, n, w;
    float t;
    uint32_t slot, key, mid;
    bool outside;
};

__device__ __forceinline__ void put_ray(rc_ray* r, float3 o, float3 d, float tmax) {
    float4* q = reinterpret_cast<float4*>(r);
    q[0] = make_float4(o.x, o.y, o.z, 0.0f);
    q[1] = make_float4(d.x, d.y, d.z, tmax);
}

// The light's Phong term at the hit times its distance and cone attenuation
// (calculateLightColor, Raytracer.cpp:331-381); zero when it adds nothing.
__device__ __forceinline__ float3 light_term(const WLight& L, const Geo& g, const DevMaterial& m, float3& l,
                                             float& dist) {
    const uint32_t type = __float_as_uint(L.color.w);
    const float3 lc = f3(L.color), kd = f3(m.diffuse), ks = f3(m.specular);
    const float shin = m.emittance.w;
    const float3 v = make_float3(-g.d.x, -g.d.y, -g.d.z);
    if (type == RC_LIGHT_DIRECTIONAL) {
        l = make_float3(-L.dir.x, -L.dir.y, -L.dir.z);
        dist = __builtin_inff();
        return phong(g.n, l, v, lc, kd, ks, shin);
    }
    const float3 tl = f3(L.pos) - g.p;
    dist = sqrtf(dot3(tl, tl));
    l = normalize3(tl);
    float3 c = phong(g.n, l, v, lc, kd, ks, shin);
    float da = fminf(1.0f, dist / L.pos.w);
    da = 1.0f - powf(da, L.dir.w);
    c = c * da;
    if (type == RC_LIGHT_SPOT) {
        float sa = dot3(f3(L.dir), make_float3(-l.x, -l.y, -l.z));
        if (!(sa > 0.0f)) return make_float3(0.0f, 0.0f, 0.0f);
        sa = fminf(1.0f, (1.0f - sa) / (L.area.y / 180.0f));
        sa = 1.0f - powf(sa, L.area.z);
        c = c * sa;
    }
    return c;
}

__device__ __forceinline__ uint32_t light_rays(const WLight& L) {
    const uint32_t bits = __float_as_uint(L.area.w);
    if (__float_as_uint(L.color.w) == RC_LIGHT_DIRECTIONAL || !(bits & 1u)) return 1u;
    const uint32_t n = bits >> 8;
    return n * n;
}

// Fresnel-weighted ray reflectance and transmittance
// (calculateRayReflectanceAndTransmittance, Raytracer.cpp:126-158).
__device__ __forceinline__ void reflect_transmit(const WMaterial& M, const Geo& g, float& R, float& T) {
    float fr = 1.0f, ft = 1.0f;
    if (__float_as_uint(M.b.z) & kFresnel) {
        const float c = dot3(g.d, g.n);
        const bool out = c < 0.0f;
        const float n1 = out ? 1.0f : M.a.z, n2 = out ? M.a.z : 1.0f;
        float rf0 = (n2 - n1) / (n2 + n1);
        rf0 = rf0 * rf0;
        fr = rf0 + (1.0f - rf0) * powf(1.0f - fabsf(c), 5.0f);
        ft = 1.0f - fr;
    }
    R = M.a.x * fr;
    T = M.a.y * ft;
}

// The refraction direction of calculateTransmittedColor (Raytracer.cpp:242-258);
// false on total internal reflection.
__device__ __forceinline__ bool refract_dir(const WMaterial& M, const Geo& g, float3& t) {
    const float c1 = dot3(g.d, g.n);
    const bool out = c1 < 0.0f;
    const float n1 = out ? 1.0f : M.a.z, n2 = out ? M.a.z : 1.0f;
    const float n3 = n1 / n2;
    const float c2 = 1.0f - (n3 * n3) * (1.0f - c1 * c1);
    if (c2 <= 0.0f) return false;
    t = normalize3(g.d * n3 + (fabsf(c1) * n3 - sqrtf(c2)) * g.n);
    return true;
}

// Counts a node's children, shadow rays and AO rays and records why.
__global__ __launch_bounds__(kBlock) void k_wt_count(WParams P, uint32_t level, uint32_t n, LevelIn in,
                                                     const float4* __restrict__ shade,
                                                     const DevMaterial* __restrict__ mats,
                                                     const WMaterial* __restrict__ wm, const WLight* __restrict__ lights,
                                                     Counts C) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t nc = 0, ns = 0, na = 0, flags = 0;
    uint64_t mask = 0;
    const rc_hit h = in.hits[i];
    if (h.tri >= 0 && !P.vis_depth) {
        const float4 sh = shade[h.tri];
        const uint32_t mid = __float_as_uint(sh.w);
        const WMaterial M = wm[mid];
        if (!(__float_as_uint(M.b.z) & kSkipLighting)) {
            const float4* rp = reinterpret_cast<const float4*>(in.rays + i);
            Geo g;
            g.o = f3(rp[0]);
            g.d = f3(rp[1]);
            g.t = h.t;
            g.p = g.o + (g.t * g.d);
            g.n = f3(sh);
            g.w = f3(in.A[i]);
            const DevMaterial m = mats[mid];
            const uint32_t sb = __float_as_uint(M.b.w);
            if (level < P.max_iter) {
                float R, T;
                reflect_transmit(M, g, R, T);
                if (R > 0.0f) {
                    const uint32_t rn = (sb >> 8) & 0xFFu;
                    nc += rn ? rn * rn : 1u;
                    flags |= kDecReflect;
                }
                float3 td;
                if (T > 0.0f && refract_dir(M, g, td)) {
                    const uint32_t tn = (sb >> 24) & 0xFFu;
                    nc += tn ? tn * tn : 1u;
                    flags |= kDecTransmit;
                }
            }
            const float3 amb = mul3(P.ambient, f3(m.ambient));
            if (P.ao && nonzero(amb)) {
                na = P.ao_n * P.ao_n;
                flags |= kDecAO;
            }
            for (uint32_t li = 0; li < P.nlights; ++li) {
                const WLight L = lights[li];
                float3 l;
                float dist;
                if (nonzero(light_term(L, g, m, l, dist))) {
                    mask |= 1ull << li;
                    ns += light_rays(L);
                }
            }
        }
    }
    C.nc[i] = nc;
    C.ns[i] = ns;
    C.na[i] = na;
    C.dec[i] = make_uint4(uint32_t(mask), uint32_t(mask >> 32), flags, 0u);
}

// The node's terms, children, shadow rays and AO rays, placed at the scanned
// offsets (relative to the first node of the range).
__global__ __launch_bounds__(kBlock) void k_wt_write(WParams P, uint32_t level, uint32_t base, uint32_t n, LevelIn in,
                                                     const float4* __restrict__ shade,
                                                     const DevMaterial* __restrict__ mats,
                                                     const WMaterial* __restrict__ wm, const WLight* __restrict__ lights,
                                                     const uint4* __restrict__ dec, Outs O) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const uint32_t i = base + r;
    const float4 A = in.A[i];
    const uint32_t slot = __float_as_uint(A.w);
    float3 w = f3(A);
    float3 c = make_float3(0.0f, 0.0f, 0.0f);
    const rc_hit h = in.hits[i];
    if (h.tri < 0) {
        c = mul3(w, P.background);  // a miss returns the background (Raytracer.cpp:29-43)
    } else {
        const float4 Bt = in.B[i];
        if (Bt.w != 0.0f) {
            // the parent's lerp(attenuation colour, C, exp(-f t)) (Raytracer.cpp:178-183, 228-233, 266-271)
            const float a = expf(-Bt.w * h.t);
            c = c + f3(Bt) * (1.0f - a);
            w = w * a;
        }
        const float4* rp = reinterpret_cast<const float4*>(in.rays + i);
        Geo g;
        g.o = f3(rp[0]);
        g.d = f3(rp[1]);
        g.t = h.t;
        if (P.vis_depth) {  // Raytracer.cpp:45-50
            float dd = 1.0f - fminf(g.t, P.vis_max) / P.vis_max;
            dd = dd * dd;
            c = c + w * dd;
        } else {
            const float4 sh = shade[h.tri];
            const uint32_t mid = __float_as_uint(sh.w);
            const WMaterial M = wm[mid];
            const DevMaterial m = mats[mid];
            g.p = g.o + (g.t * g.d);
            g.n = f3(sh);
            g.outside = dot3(g.d, g.n) < 0.0f;
            g.slot = slot;
            g.key = in.keys[i];
            const uint32_t mflags = __float_as_uint(M.b.z);
            // calculateSimpleFogColor (Raytracer.cpp:409-425) on the node's colour when outside
            float fog_t = 0.0f;
            if (P.fog && g.outside) {
                float t1 = fminf(fmaxf(g.t / P.fog_dist, 0.0f), 1.0f);
                t1 = powf(t1, P.fog_steep);
                if (P.fog_hd && g.p.y > 0.0f) {
                    float t2 = fminf(fmaxf(g.p.y / P.fog_h, 0.0f), 1.0f);
                    t2 = powf(t2, P.fog_hsteep);
                    t1 *= 1.0f - t2;
                }
                fog_t = t1;
            }
            if (mflags & kSkipLighting) {  // Raytracer.cpp:54-67
                const float3 kd = f3(m.diffuse);
                const float3 col = kd + (P.fog_color - kd) * fog_t;
                c = c + mul3(w, col);
            } else {
                if (fog_t != 0.0f) {
                    c = c + (w * fog_t) * P.fog_color;
                    w = w * (1.0f - fog_t);
                }
                g.w = w;
                const uint4 D = dec[i];
                const uint32_t sb = __float_as_uint(M.b.w);
                float R, T;
                reflect_transmit(M, g, R, T);
                uint32_t co = O.oc[i] - O.oc[base];
                // calculateReflectedColor (Raytracer.cpp:160-226)
                if (D.z & kDecReflect) {
                    float3 rd = normalize3(g.d + (2.0f * -dot3(g.d, g.n)) * g.n);
                    const uint32_t rn = (sb >> 8) & 0xFFu;
                    const bool att = !g.outside && (mflags & kAttenuate);
                    if (rn == 0) {
                        put_ray(O.crays + co, g.p + rd * P.offset, rd, __builtin_inff());
                        const float3 cw = w * R;
                        O.cA[co] = make_float4(cw.x, cw.y, cw.z, __uint_as_float(slot));
                        const float3 ac = mul3(w, f3(M.att_color));
                        O.cB[co] = att ? make_float4(ac.x, ac.y, ac.z, M.b.y) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        O.ckeys[co] = child_key(g.key, kPurposeReflect, 0);
                        ++co;
                    } else {
                        float3 u, v;
                        onb_from(rd, u, v);
                        const uint64_t perm = node_perm(P, slot, g.key, kPurposeReflect);
                        const uint32_t rsamp = sb & 0xFFu;
                        const float inv = 1.0f / float(rn * rn);
                        for (uint32_t y = 0; y < rn; ++y)
                            for (uint32_t x = 0; x < rn; ++x) {
                                const uint32_t s = y * rn + x;
                                const float2 uv = node_uv(P, slot, g.key, kPurposeReflect, s);
                                float3 sd = map_hemisphere(u, v, rd, M.a.w, sample_point(rsamp, x, y, rn, perm, uv.x, uv.y));
                                const float sn = dot3(sd, g.n);
                                if ((g.outside && sn < 0.0f) || (!g.outside && sn > 0.0f)) sd = reflect3(sd, rd);
                                put_ray(O.crays + co, g.p + sd * P.offset, sd, __builtin_inff());
                                const float3 cw = w * (R * inv);
                                O.cA[co] = make_float4(cw.x, cw.y, cw.z, __uint_as_float(slot));
                                const float3 ac = mul3(w, f3(M.att_color)) * inv;
                                O.cB[co] = att ? make_float4(ac.x, ac.y, ac.z, M.b.y) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                                O.ckeys[co] = child_key(g.key, kPurposeReflect, s);
                                ++co;
                            }
                    }
                }
                // calculateTransmittedColor (Raytracer.cpp:228-310)
                if (D.z & kDecTransmit) {
                    float3 td;
                    if (!refract_dir(M, g, td)) td = g.d;  // decided by the count pass; rounding cannot undo it
                    const uint32_t tn = (sb >> 24) & 0xFFu;
                    if (tn == 0) {
                        const bool att = g.outside && (mflags & kAttenuate);
                        put_ray(O.crays + co, g.p + td * P.offset, td, __builtin_inff());
                        const float3 cw = w * T;
                        O.cA[co] = make_float4(cw.x, cw.y, cw.z, __uint_as_float(slot));
                        const float3 ac = mul3(w, f3(M.att_color));
                        O.cB[co] = att ? make_float4(ac.x, ac.y, ac.z, M.b.y) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        O.ckeys[co] = child_key(g.key, kPurposeTransmit, 0);
                        ++co;
                    } else {
                        const bool att = !g.outside && (mflags & kAttenuate);
                        float3 u, v;
                        onb_from(td, u, v);
                        const uint64_t perm = node_perm(P, slot, g.key, kPurposeTransmit);
                        const uint32_t tsamp = (sb >> 16) & 0xFFu;
                        const float inv = 1.0f / float(tn * tn);
                        for (uint32_t y = 0; y < tn; ++y)
                            for (uint32_t x = 0; x < tn; ++x) {
                                const uint32_t s = y * tn + x;
                                const float2 uv = node_uv(P, slot, g.key, kPurposeTransmit, s);
                                float3 sd = map_hemisphere(u, v, td, M.b.x, sample_point(tsamp, x, y, tn, perm, uv.x, uv.y));
                                const float sn = dot3(sd, g.n);
                                if ((g.outside && sn > 0.0f) || (!g.outside && sn < 0.0f)) sd = reflect3(sd, td);
                                put_ray(O.crays + co, g.p + sd * P.offset, sd, __builtin_inff());
                                const float3 cw = w * (T * inv);
                                O.cA[co] = make_float4(cw.x, cw.y, cw.z, __uint_as_float(slot));
                                const float3 ac = mul3(w, f3(M.att_color)) * inv;
                                O.cB[co] = att ? make_float4(ac.x, ac.y, ac.z, M.b.y) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                                O.ckeys[co] = child_key(g.key, kPurposeTransmit, s);
                                ++co;
                            }
                    }
                }
                // calculateLightColor (Raytracer.cpp:312-384): ambient, with occlusion
                const float3 amb = mul3(P.ambient, f3(m.ambient));
                if (D.z & kDecAO) {
                    // calculateAmbientOcclusionAmount (Raytracer.cpp:453-494) in the triangle's frame
                    float3 u, v;
                    onb_from(g.n, u, v);
                    const uint64_t perm = node_perm(P, slot, g.key, kPurposeAO);
                    const uint32_t an = P.ao_n;
                    const float3 val = mul3(w, amb) * (1.0f / float(an * an));
                    uint32_t ao = O.oa[i] - O.oa[base];
                    for (uint32_t y = 0; y < an; ++y)
                        for (uint32_t x = 0; x < an; ++x) {
                            const uint32_t s = y * an + x;
                            const float2 uv = node_uv(P, slot, g.key, kPurposeAO, s);
                            const float3 sd =
                                map_hemisphere(u, v, g.n, P.ao_dist, sample_point(P.ao_sampler, x, y, an, perm, uv.x, uv.y));
                            put_ray(O.arays + ao, g.p + sd * P.offset, sd, P.ao_maxd);
                            O.aval[ao] = make_float4(val.x, val.y, val.z, __uint_as_float(slot));
                            ++ao;
                        }
                } else {
                    c = c + mul3(w, amb);
                }
                // directional, point and spot lights with their shadow rays
                // (calculateShadowAmount, Raytracer.cpp:496-597)
                const uint64_t mask = uint64_t(D.x) | (uint64_t(D.y) << 32);
                uint32_t so = O.os[i] - O.os[base];
                for (uint32_t li = 0; li < P.nlights; ++li) {
                    if (!((mask >> li) & 1ull)) continue;
                    const WLight L = lights[li];
                    float3 l;
                    float dist;
                    const float3 lt = mul3(w, light_term(L, g, m, l, dist));
                    const uint32_t bits = __float_as_uint(L.area.w);
                    const bool area = __float_as_uint(L.color.w) != RC_LIGHT_DIRECTIONAL && (bits & 1u);
                    if (!area) {
                        put_ray(O.srays + so, g.p + l * P.offset, l, dist);
                        O.sval[so] = make_float4(lt.x, lt.y, lt.z, __uint_as_float(slot));
                        ++so;
                        continue;
                    }
                    const uint32_t an = bits >> 8, asamp = (bits >> 1) & 0x7Fu;
                    const float3 right = normalize3(cross3(l, make_float3(0.0001f, 1.0f, 0.0001f)));
                    const float3 up = normalize3(cross3(right, l));
                    const uint64_t perm = node_perm(P, slot, g.key, kPurposeLight + li);
                    const float3 val = lt * (1.0f / float(an * an));
                    for (uint32_t y = 0; y < an; ++y)
                        for (uint32_t x = 0; x < an; ++x) {
                            const uint32_t s = y * an + x;
                            const float2 uv = node_uv(P, slot, g.key, kPurposeLight + li, s);
                            // getDiscSample: mapToDisc (Sampler.cpp:36-48) times the radius
                            const float2 pt = sample_point(asamp, x, y, an, perm, uv.x, uv.y);
                            float sp, cp;
                            sincosf(2.0f * 3.14159265358979f * pt.x, &sp, &cp);
                            const float rr = sqrtf(pt.y);
                            const float jx = rr * cp * L.area.x, jy = rr * sp * L.area.x;
                            const float3 np = f3(L.pos) + (jx * right + jy * up);
                            const float3 tl = np - g.p;
                            const float nd = sqrtf(dot3(tl, tl));
                            const float3 dl = normalize3(tl);
                            put_ray(O.srays + so, g.p + dl * P.offset, dl, nd);
                            O.sval[so] = make_float4(val.x, val.y, val.z, __uint_as_float(slot));
                            ++so;
                        }
                }
            }
        }
    }
    O.contrib[i] = make_float4(c.x, c.y, c.z, __uint_as_float(slot));
}

// Adds runs of terms to their pixel samples: the first entry of each run of
// equal slots sums the run in order.  mode 0: every entry; 1: entries whose
// shadow ray is not occluded; 2: entries whose AO ray hit nothing.
__global__ __launch_bounds__(kBlock) void k_wt_sum(uint32_t n, const float4* __restrict__ val,
                                                   const uint8_t* __restrict__ occ, const rc_hit* __restrict__ hits,
                                                   uint32_t mode, float4* __restrict__ L) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = __float_as_uint(val[i].w);
    if (i > 0 && __float_as_uint(val[i - 1].w) == s) return;
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    for (uint32_t j = i; j < n; ++j) {
        const float4 v = val[j];
        if (__float_as_uint(v.w) != s) break;
        if (mode == 1 && occ[j]) continue;
        if (mode == 2 && hits[j].tri >= 0) continue;
        acc = acc + f3(v);
    }
    const float4 l = L[s];
    L[s] = make_float4(l.x + acc.x, l.y + acc.y, l.z + acc.z, l.w);
}

__global__ void k_wt_total(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, uint32_t last,
                           uint32_t* __restrict__ out) {
    *out = off[last] + cnt[last];
}

__global__ void k_wt_root(uint32_t n, uint32_t* __restrict__ keys, float4* __restrict__ B) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    keys[i] = 0u;
    B[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

// ---------------------------------------------------------------- host side
struct WhittedBufs {
    struct Level {
        DevBuf rays, A, B, keys, hits, nc, ns, na, oc, os, oa, dec, contrib;
        size_t cap = 0;
    };
    std::vector<Level> levels;
    DevBuf srays, sval, socc, arays, aval, ahits, scan_tmp, totals, lights, mats;
    size_t scap = 0, acap = 0;
};

void rc_whitted_release(rc_ctx* c) {
    if (!c || !c->whitted) return;
    WhittedBufs* W = c->whitted;
    for (auto& L : W->levels) {
        DevBuf* b[] = {&L.rays, &L.A, &L.B, &L.keys, &L.hits, &L.nc, &L.ns, &L.na, &L.oc, &L.os, &L.oa, &L.dec, &L.contrib};
        for (DevBuf* d : b) d->release();
    }
    DevBuf* b[] = {&W->srays, &W->sval, &W->socc, &W->arays, &W->aval, &W->ahits, &W->scan_tmp, &W->totals,
                   &W->lights, &W->mats};
    for (DevBuf* d : b) d->release();
    delete W;
    c->whitted = nullptr;
}

namespace {

int ensure_level(WhittedBufs::Level& L, size_t cap) {
    RC_TRY(L.rays.ensure(cap * sizeof(rc_ray)));
    RC_TRY(L.A.ensure(cap * sizeof(float4)));
    RC_TRY(L.B.ensure(cap * sizeof(float4)));
    RC_TRY(L.keys.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.hits.ensure(cap * sizeof(rc_hit)));
    RC_TRY(L.nc.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.ns.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.na.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.oc.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.os.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.oa.ensure(cap * sizeof(uint32_t)));
    RC_TRY(L.dec.ensure(cap * sizeof(uint4)));
    RC_TRY(L.contrib.ensure(cap * sizeof(float4)));
    L.cap = std::max(L.cap, cap);
    return RC_OK;
}

struct Run {
    rc_ctx* c;
    const rc_scene* s;
    WhittedBufs* W;
    WParams P;
    const WMaterial* wm;
    const WLight* wl;
    float4* L;
    uint32_t nlev;  // levels in use: maxRayIterations + 1
    uint64_t ext_rays = 0, shadow_rays = 0;

    int scan(const uint32_t* in, uint32_t* out, uint32_t n) {
        size_t bytes = 0;
        RC_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c->stream));
        RC_TRY(W->scan_tmp.ensure(std::max<size_t>(bytes, 256)));
        RC_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(W->scan_tmp.p, bytes, in, out, n, c->stream));
        return RC_OK;
    }

    LevelIn level_in(uint32_t k) {
        WhittedBufs::Level& Lv = W->levels[k];
        return LevelIn{Lv.rays.as<rc_ray>(), Lv.A.as<float4>(), Lv.B.as<float4>(), Lv.keys.as<uint32_t>(),
                       Lv.hits.as<rc_hit>()};
    }

    // traces level k's n nodes and counts them
    int trace_and_count(uint32_t k, uint32_t n) {
        WhittedBufs::Level& Lv = W->levels[k];
        RC_TRY(rc_launch_trace(c, s, false, Lv.rays.as<rc_ray>(), nullptr, n, Lv.hits.as<rc_hit>(), nullptr));
        ext_rays += n;
        Counts C{Lv.nc.as<uint32_t>(), Lv.ns.as<uint32_t>(), Lv.na.as<uint32_t>(), Lv.dec.as<uint4>()};
        k_wt_count<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(P, k, n, level_in(k), s->shade, s->mats, wm,
                                                                       wl, C);
        RC_TRY_HIP(hipGetLastError());
        return RC_OK;
    }

    // children, shadow rays and AO rays of nodes [base, base + n) of level k
    int totals(uint32_t k, uint32_t base, uint32_t n, uint32_t out[3]) {
        WhittedBufs::Level& Lv = W->levels[k];
        const uint32_t* cnt[3] = {Lv.nc.as<uint32_t>(), Lv.ns.as<uint32_t>(), Lv.na.as<uint32_t>()};
        uint32_t* off[3] = {Lv.oc.as<uint32_t>(), Lv.os.as<uint32_t>(), Lv.oa.as<uint32_t>()};
        uint32_t* dev = W->totals.as<uint32_t>();
        for (int q = 0; q < 3; ++q) {
            RC_TRY(scan(cnt[q] + base, off[q] + base, n));
            k_wt_total<<<1, 1, 0, c->stream>>>(cnt[q], off[q], base + n - 1, dev + q);
        }
        RC_TRY_HIP(hipGetLastError());
        RC_TRY_HIP(hipMemcpyAsync(out, dev, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        RC_TRY_HIP(hipStreamSynchronize(c->stream));
        return RC_OK;
    }

    // Offsets are scanned over [base, base + n) only, so the write kernel
    // subtracts the range's first offset (zero by construction).
    int process(uint32_t k, uint32_t base, uint32_t n) {
        if (n == 0) return RC_OK;
        uint32_t tot[3];
        RC_TRY(totals(k, base, n, tot));
        const size_t ccap = k + 1 < nlev ? W->levels[k + 1].cap : 0;
        if ((k + 1 < nlev && tot[0] > ccap) || tot[1] > W->scap || tot[2] > W->acap) {
            if (n == 1)
                return rc_host_fail(RC_ERR_INVALID, "rc_raytrace: one hit spawns more rays than the buffers hold");
            const uint32_t h = n / 2;
            RC_TRY(process(k, base, h));
            return process(k, base + h, n - h);
        }
        WhittedBufs::Level& Lv = W->levels[k];
        Outs O{};
        O.oc = Lv.oc.as<uint32_t>();
        O.os = Lv.os.as<uint32_t>();
        O.oa = Lv.oa.as<uint32_t>();
        if (k + 1 < nlev) {
            WhittedBufs::Level& Nx = W->levels[k + 1];
            O.crays = Nx.rays.as<rc_ray>();
            O.cA = Nx.A.as<float4>();
            O.cB = Nx.B.as<float4>();
            O.ckeys = Nx.keys.as<uint32_t>();
        }
        O.srays = W->srays.as<rc_ray>();
        O.sval = W->sval.as<float4>();
        O.arays = W->arays.as<rc_ray>();
        O.aval = W->aval.as<float4>();
        O.contrib = Lv.contrib.as<float4>();
        k_wt_write<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(P, k, base, n, level_in(k), s->shade, s->mats,
                                                                       wm, wl, Lv.dec.as<uint4>(), O);
        k_wt_sum<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(n, Lv.contrib.as<float4>() + base, nullptr,
                                                                     nullptr, 0, L);
        RC_TRY_HIP(hipGetLastError());
        if (tot[1]) {
            RC_TRY(rc_launch_trace(c, s, true, W->srays.as<rc_ray>(), nullptr, tot[1], nullptr, W->socc.as<uint8_t>()));
            k_wt_sum<<<(tot[1] + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(tot[1], W->sval.as<float4>(),
                                                                               W->socc.as<uint8_t>(), nullptr, 1, L);
            shadow_rays += tot[1];
        }
        if (tot[2]) {
            // AO rays are not shadow rays: nonShadowing materials occlude them (Raytracer.cpp:470-489)
            RC_TRY(rc_launch_trace(c, s, false, W->arays.as<rc_ray>(), nullptr, tot[2], W->ahits.as<rc_hit>(), nullptr));
            k_wt_sum<<<(tot[2] + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(tot[2], W->aval.as<float4>(), nullptr,
                                                                               W->ahits.as<rc_hit>(), 2, L);
            shadow_rays += tot[2];
        }
        RC_TRY_HIP(hipGetLastError());
        if (tot[0] && k + 1 < nlev) {
            RC_TRY(trace_and_count(k + 1, tot[0]));
            RC_TRY(process(k + 1, 0, tot[0]));
        }
        return RC_OK;
    }
};

float bits_as_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, sizeof f);
    return f;
}

int validate(const rc_scene* s, const rc_ray_settings* rs, const rc_ray_material* mats, uint32_t nmat,
             const rc_light* lights, uint32_t nlights) {
    auto bad = [](const std::string& m) { return rc_host_fail(RC_ERR_INVALID, "rc_raytrace: " + m); };
    if (!rs) return bad("null settings");
    if (nmat != s->nmat || (nmat && !mats)) return bad("one ray material per scene material is needed");
    if (nlights > kMaxLights || (nlights && !lights)) return bad("at most 64 lights");
    if (rs->max_ray_iterations > kMaxIterations) return bad("max_ray_iterations above 16");
    if (rs->ao_enabled && (rs->ao_n < 1 || rs->ao_n > kMaxN || rs->ao_sampler > RC_SAMPLER_POISSON_DISC))
        return bad("ambient occlusion needs 1..16 samples per axis and a known sampler");
    if (rs->visualize_depth && !(rs->visualize_depth_max > 0.0f)) return bad("visualize_depth_max must be positive");
    if (rs->fog_enabled && !(rs->fog_distance > 0.0f && (!rs->fog_height_dispersion || rs->fog_height > 0.0f)))
        return bad("fog distance and height must be positive");
    for (uint32_t i = 0; i < nmat; ++i) {
        const rc_ray_material& m = mats[i];
        if (m.refl_n > kMaxN || m.trans_n > kMaxN || m.refl_sampler > RC_SAMPLER_POISSON_DISC ||
            m.trans_sampler > RC_SAMPLER_POISSON_DISC)
            return bad("glossy sample counts are 0..16 per axis with a known sampler");
        if (!(m.refractive_index > 0.0f) && (m.fresnel_reflection || m.ray_transmittance > 0.0f))
            return bad("refractive_index must be positive");
    }
    for (uint32_t i = 0; i < nlights; ++i) {
        const rc_light& L = lights[i];
        if (L.type > RC_LIGHT_SPOT) return bad("unknown light type");
        if (L.type != RC_LIGHT_DIRECTIONAL) {
            if (!(L.max_distance > 0.0f)) return bad("max_distance must be positive");
            if (L.area_light && (L.area_n < 1 || L.area_n > kMaxN || L.area_sampler > RC_SAMPLER_POISSON_DISC))
                return bad("area lights need 1..16 samples per axis and a known sampler");
            if (L.type == RC_LIGHT_SPOT && !(L.angle > 0.0f)) return bad("spot angle must be positive");
        }
    }
    return RC_OK;
}

}  // namespace

extern "C" int rc_raytrace(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                           const rc_ray_settings* rs, const rc_ray_material* mats, uint32_t nmat,
                           const rc_light* lights, uint32_t nlights) {
    RC_TRY(rc_check_scene(c, s, "rc_raytrace"));
    RC_TRY(rc_validate_cfg(cfg, "rc_raytrace"));
    if (!cam) return rc_host_fail(RC_ERR_INVALID, "rc_raytrace: null camera");
    if (!f || f->ctx != c || f->width != cfg->width || f->height != cfg->height)
        return rc_host_fail(RC_ERR_INVALID, "rc_raytrace: film does not match the config");
    RC_TRY(validate(s, rs, mats, nmat, lights, nlights));
    RC_TRY_HIP(hipSetDevice(c->device));
    if (!c->whitted) c->whitted = new WhittedBufs();
    WhittedBufs* W = c->whitted;
    auto t0 = std::chrono::steady_clock::now();
    rc_zero_counters(c->counters, c->stream);

    // device copies of the materials and lights
    std::vector<WMaterial> wm(std::max<uint32_t>(nmat, 1));
    for (uint32_t i = 0; i < nmat; ++i) {
        const rc_ray_material& m = mats[i];
        const uint32_t flags = (m.fresnel_reflection ? kFresnel : 0u) | (m.skip_lighting ? kSkipLighting : 0u) |
                               (m.enable_attenuation ? kAttenuate : 0u);
        const uint32_t samp = (m.refl_sampler & 0xFFu) | ((m.refl_n & 0xFFu) << 8) | ((m.trans_sampler & 0xFFu) << 16) |
                              ((m.trans_n & 0xFFu) << 24);
        wm[i].a = make_float4(m.ray_reflectance, m.ray_transmittance, m.refractive_index, m.refl_glossiness);
        wm[i].b = make_float4(m.trans_glossiness, m.attenuation_factor, bits_as_float(flags), bits_as_float(samp));
        wm[i].att_color = make_float4(m.attenuation_color[0], m.attenuation_color[1], m.attenuation_color[2], 0.0f);
    }
    std::vector<WLight> wl(std::max<uint32_t>(nlights, 1));
    for (uint32_t i = 0; i < nlights; ++i) {
        const rc_light& L = lights[i];
        const uint32_t bits = (L.area_light ? 1u : 0u) | ((L.area_sampler & 0x7Fu) << 1) | ((L.area_n & 0xFFu) << 8);
        wl[i].color = make_float4(L.color[0] * L.intensity, L.color[1] * L.intensity, L.color[2] * L.intensity,
                                  bits_as_float(L.type));
        wl[i].pos = make_float4(L.position[0], L.position[1], L.position[2], L.max_distance);
        wl[i].dir = make_float4(L.direction[0], L.direction[1], L.direction[2], L.attenuation);
        wl[i].area = make_float4(L.area_radius, L.angle, L.side_attenuation, bits_as_float(bits));
    }
    RC_TRY(W->mats.ensure(wm.size() * sizeof(WMaterial)));
    RC_TRY(W->lights.ensure(wl.size() * sizeof(WLight)));
    RC_TRY_HIP(hipMemcpyAsync(W->mats.p, wm.data(), wm.size() * sizeof(WMaterial), hipMemcpyHostToDevice, c->stream));
    RC_TRY_HIP(hipMemcpyAsync(W->lights.p, wl.data(), wl.size() * sizeof(WLight), hipMemcpyHostToDevice, c->stream));

    Run run{c, s, W, {}, W->mats.as<WMaterial>(), W->lights.as<WLight>(), nullptr, rs->max_ray_iterations + 1};
    WParams& P = run.P;
    P.background = make_float3(rs->background[0], rs->background[1], rs->background[2]);
    P.ambient = make_float3(rs->ambient_color[0] * rs->ambient_intensity, rs->ambient_color[1] * rs->ambient_intensity,
                            rs->ambient_color[2] * rs->ambient_intensity);
    P.offset = cfg->ray_start_offset;
    P.max_iter = rs->max_ray_iterations;
    P.ao = rs->ao_enabled;
    P.ao_sampler = rs->ao_sampler;
    P.ao_n = rs->ao_n;
    P.ao_maxd = rs->ao_max_distance;
    P.ao_dist = rs->ao_distribution;
    P.vis_depth = rs->visualize_depth;
    P.vis_max = rs->visualize_depth_max;
    P.fog = rs->fog_enabled;
    P.fog_hd = rs->fog_height_dispersion;
    P.fog_color = make_float3(rs->fog_color[0], rs->fog_color[1], rs->fog_color[2]);
    P.fog_dist = rs->fog_distance;
    P.fog_steep = rs->fog_steepness;
    P.fog_h = rs->fog_height;
    P.fog_hsteep = rs->fog_height_steepness;
    P.nlights = nlights;
    P.order = make_order(cfg->width, cfg->height);
    P.npix = cfg->width * cfg->height;
    P.frame = cfg->frame;
    P.seed = cfg->seed;

    // one pass per tuning.spp_per_pass samples of every pixel
    const uint32_t spp_pass = pass_samples(c->tuning, P.npix, cfg->spp, kWhittedPassPaths);
    const size_t n0max = size_t(P.npix) * spp_pass;
    if (n0max > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_raytrace: pass too large");
    const size_t cap = std::max<size_t>(2 * n0max, size_t(1) << 20);
    if (W->levels.size() < run.nlev) W->levels.resize(run.nlev);  // grow only: DevBuf frees nothing by itself
    RC_TRY(ensure_level(W->levels[0], n0max));
    for (uint32_t k = 1; k < run.nlev; ++k) RC_TRY(ensure_level(W->levels[k], cap));
    W->scap = W->acap = cap;
    RC_TRY(W->srays.ensure(cap * sizeof(rc_ray)));
    RC_TRY(W->sval.ensure(cap * sizeof(float4)));
    RC_TRY(W->socc.ensure(cap));
    RC_TRY(W->arays.ensure(cap * sizeof(rc_ray)));
    RC_TRY(W->aval.ensure(cap * sizeof(float4)));
    RC_TRY(W->ahits.ensure(cap * sizeof(rc_hit)));
    RC_TRY(W->totals.ensure(4 * sizeof(uint32_t)));
    RC_TRY(c->fb.slotL.ensure(n0max * sizeof(float4)));
    RC_TRY(c->fb.slotW.ensure(n0max * sizeof(float)));
    RC_TRY(c->fb.ovf_list.ensure(cap * sizeof(uint32_t)));
    RC_TRY(c->fb.deep_stack.ensure(rc_deep_stack_bytes()));
    run.L = c->fb.slotL.as<float4>();

    for (uint32_t s0 = 0; s0 < cfg->spp; s0 += spp_pass) {
        const uint32_t this_pass = std::min(spp_pass, cfg->spp - s0);
        const uint32_t n0 = P.npix * this_pass;
        P.sample0 = s0;
        P.spp_pass = this_pass;
        WhittedBufs::Level& L0 = W->levels[0];
        RC_TRY(rc_launch_camera(c, cam, cfg, s0, this_pass, L0.rays.as<rc_ray>(), L0.A.as<float4>(), run.L,
                                c->fb.slotW.as<float>()));
        k_wt_root<<<(n0 + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(n0, L0.keys.as<uint32_t>(), L0.B.as<float4>());
        RC_TRY_HIP(hipGetLastError());
        RC_TRY(run.trace_and_count(0, n0));
        RC_TRY(run.process(0, 0, n0));
        RC_TRY(rc_launch_film_add(c, f, cfg, this_pass, run.L, c->fb.slotW.as<float>()));
    }
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    DevCounters h;
    RC_TRY_HIP(hipMemcpy(&h, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost));
    if (h.error) return rc_host_fail(RC_ERR_HIP, "rc_raytrace: traversal stack exceeded the deep-stack limit");
    c->stats = rc_stats{};
    c->stats.extension_rays = run.ext_rays;
    c->stats.shadow_rays = run.shadow_rays;
    c->stats.samples = uint64_t(P.npix) * cfg->spp;
    c->stats.overflow_rays = h.ovf_total;
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RC_OK;
}
