// Device-side data layouts and inline helpers of the gfx950 path tracer.
//
// Scene layout in HBM (all read-only during a render):
//   NodeRec  32 B  child-pair records: an interior node of the host FlatBVH is
//                  stored as the two records of its children, side by side
//                  (64 B, 64-B aligned), so one visit reads both child boxes
//                  with four 128-bit loads.  Record fields: child box
//                  lo.xyz/hi.xyz, link, count.  count > 0: leaf child whose
//                  triangles are tris[link, link + count); count == 0:
//                  interior child whose own pair starts at record `link`.
//                  Pairs are laid out in the host tree's preorder, so links
//                  always point forward.
//   TriRec   48 B  v0.xyz + original triangle id, e1.xyz + flags, e2.xyz
//                  (e1 = v1 - v0, e2 = v2 - v0), in BVH leaf order.
//   ShadeRec 16 B  geometric normal + material index, in original order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "raycer_hip.h"

namespace rcd {

struct alignas(32) NodeRec {
    float lo[3];
    float hi[3];
    uint32_t link;
    uint32_t count;
};
static_assert(sizeof(NodeRec) == 32, "NodeRec must be 32 bytes");

struct alignas(16) TriRec {
    float4 v0;  // w: original triangle id (bits)
    float4 e1;  // w: flags (bit1 nonShadowing)
    float4 e2;  // w: unused
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

struct alignas(16) DevMaterial {
    float4 diffuse;   // w: flags (bits)
    float4 emittance; // w: shininess
    float4 ambient;
    float4 specular;
};

constexpr uint32_t kNoHit = 0x7fffffffu;
constexpr int kDeepStack = 256;  // deepest tree the overflow kernel can walk

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64 products.
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}
// 24-bit uniform in [0,1), exact in FP32.
__device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }
constexpr uint32_t kCamDim = 0xFFFFFFFFu;
__device__ __forceinline__ float4 draw4(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t frame, uint32_t dim) {
    uint4 o = philox(make_uint4(pixel, sample, frame, dim), make_uint2(uint32_t(seed), uint32_t(seed >> 32)));
    return make_float4(u01(o.x), u01(o.y), u01(o.z), u01(o.w));
}

// ---------------------------------------------------------------- vec math
__device__ __forceinline__ float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 mul3(float3 a, float3 b) { return make_float3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize3(float3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }

// Direction with no zero component, so 1/d is finite and the slab products
// never form 0 * inf.
__device__ __forceinline__ float3 safe_inv(float3 d) {
    const float e = 1e-20f;
    float x = fabsf(d.x) < e ? copysignf(e, d.x) : d.x;
    float y = fabsf(d.y) < e ? copysignf(e, d.y) : d.y;
    float z = fabsf(d.z) < e ? copysignf(e, d.z) : d.z;
    return make_float3(1.0f / x, 1.0f / y, 1.0f / z);
}

// Slab test of one child record against [tmin, tmax]; entry distance or +inf.
// The exit distance is widened by a few ulps so rounding never culls a box the
// triangle test could still hit.
__device__ __forceinline__ float slab(float4 r0, float4 r1, float3 o, float3 inv, float tmin, float tmax) {
    float tx0 = (r0.x - o.x) * inv.x, tx1 = (r0.w - o.x) * inv.x;
    float ty0 = (r0.y - o.y) * inv.y, ty1 = (r1.x - o.y) * inv.y;
    float tz0 = (r0.z - o.z) * inv.z, tz1 = (r1.y - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tn <= tf * 1.0000005f ? tn : __builtin_inff();
}

// FP32 Möller–Trumbore with the reference's structure (Triangle.cpp:61-91):
// double-sided, inclusive barycentric bounds; the determinant test is the
// scale-free det == 0 instead of |det| < epsilon.
__device__ __forceinline__ bool mt_hit(const TriRec& T, float3 o, float3 d, float& t, float& u, float& v) {
    float3 e1 = f3(T.e1), e2 = f3(T.e2);
    float3 pv = cross3(d, e2);
    float det = dot3(e1, pv);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    float3 tv = o - f3(T.v0);
    u = dot3(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    float3 qv = cross3(tv, e1);
    v = dot3(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot3(e2, qv) * inv;
    return true;
}

struct TraceResult {
    float t, u, v;
    uint32_t id;
    bool overflow;
};

// Ordered, t-culled traversal over child-pair records.  The nearer child is
// visited first, the farther one pushed on a per-thread stack whose entries
// live at stack[k * stride] (LDS for the main kernels, global memory for the
// deep-stack kernel).  Leaves are tested as soon as their box is hit.
// Closest hit keeps the smallest t; an exact tie goes to the lower original
// triangle id, so the answer does not depend on traversal order.
// ANY: stop at the first triangle (not nonShadowing) with t in [tmin, tmax].
template <bool ANY, bool COUNT>
__device__ __forceinline__ TraceResult traverse(const NodeRec* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                float3 o, float3 d, float tmin, float tmax, uint32_t* stack,
                                                int stride, int max_depth, uint32_t& n_rec, uint32_t& n_tri) {
    TraceResult R{tmax, 0.0f, 0.0f, kNoHit, false};
    tmin = fmaxf(tmin, 0.0f);
    float3 inv = safe_inv(d);
    uint32_t cur = 0;
    int sp = 0;
    while (true) {
        const float4* p = reinterpret_cast<const float4*>(nodes + cur);
        float4 a0 = p[0], a1 = p[1], b0 = p[2], b1 = p[3];
        if (COUNT) n_rec += 2;
        float ta = slab(a0, a1, o, inv, tmin, R.t);
        float tb = slab(b0, b1, o, inv, tmin, R.t);
        uint32_t la = __float_as_uint(a1.z), ca = __float_as_uint(a1.w);
        uint32_t lb = __float_as_uint(b1.z), cb = __float_as_uint(b1.w);
        if (tb < ta) {  // make `a` the nearer child
            float tt = ta; ta = tb; tb = tt;
            uint32_t x = la; la = lb; lb = x;
            x = ca; ca = cb; cb = x;
        }
        bool ha = ta != __builtin_inff(), hb = tb != __builtin_inff();
#pragma unroll 1
        for (int leaf = 0; leaf < 2; ++leaf) {
            bool h = leaf == 0 ? ha : hb;
            uint32_t l = leaf == 0 ? la : lb, c = leaf == 0 ? ca : cb;
            float te = leaf == 0 ? ta : tb;
            if (!(h && c > 0u)) continue;
            if (leaf == 0) ha = false; else hb = false;
            if (te > R.t) continue;
            for (uint32_t j = l; j < l + c; ++j) {
                const float4* q = reinterpret_cast<const float4*>(tris + j);
                TriRec T;
                T.v0 = q[0];
                T.e1 = q[1];
                T.e2 = q[2];
                if (COUNT) n_tri += 1;
                float t, u, v;
                if (!mt_hit(T, o, d, t, u, v)) continue;
                if (t < tmin) continue;
                uint32_t id = __float_as_uint(T.v0.w);
                if (ANY) {
                    if (t <= tmax && !(__float_as_uint(T.e1.w) & RC_MAT_NON_SHADOWING)) {
                        R.t = t; R.id = id;
                        return R;
                    }
                } else if (t < R.t || (t == R.t && id < R.id)) {
                    R.t = t; R.u = u; R.v = v; R.id = id;
                }
            }
        }
        if (ha && ta > R.t) ha = false;
        if (hb && tb > R.t) hb = false;
        if (ha && hb) {
            if (sp >= max_depth) { R.overflow = true; return R; }
            stack[sp * stride] = lb;
            ++sp;
            cur = la;
        } else if (ha) {
            cur = la;
        } else if (hb) {
            cur = lb;
        } else {
            if (sp == 0) break;
            --sp;
            cur = stack[sp * stride];
        }
    }
    return R;
}

}  // namespace rcd
