// Device-side data layouts and inline helpers of the gfx950 path tracer.
//
// Scene layout in HBM (all read-only during a render):
//   QNode    64 B  per 4-wide node (the default, below): origin, three scale
//                  exponents, 8-bit child bounds, child counts and links.
//   NodeRec  32 B  child records (RC_QNODE = 0).  The host FlatBVH is collapsed into 4-wide
//                  nodes; a wide node is stored as the 4 records of its
//                  children, side by side (128 B, one cache line), so one visit
//                  reads all child boxes with eight 128-bit loads.  Record
//                  fields: child box lo.xyz/hi.xyz, link, count.  count in
//                  1..kEmptyChild-1: leaf child whose triangles are
//                  tris[link, link + count); count == 0: interior child whose
//                  own wide node starts at record `link`; count == kEmptyChild:
//                  unused slot.  Wide nodes are laid out in preorder, so links
//                  always point forward.
//   TriRec   48 B  v0.xyz + original triangle id, v1.xyz + flags, v2.xyz,
//                  in BVH leaf order (the vertices themselves, so the FP64
//                  refinement sees exactly the input edges).
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

// RC_QNODE = 1 (default) stores each wide node in 64 bytes instead of 128:
// the child boxes as 8-bit offsets from the node's origin in units of a power
// of two per axis, widened by one unit on each side so the decoded boxes
// contain the exact ones.  Links are node indices.  2.6% faster on the
// 2M-triangle terrain than the 128-byte records (RC_QNODE = 0).
#ifndef RC_QNODE
#define RC_QNODE 1
#endif
struct alignas(64) QNode {
    float origin[3];
    uint32_t exps;     // biased float exponents of the three scales: ex | ey << 8 | ez << 16
    uint32_t qlo[3];   // per axis, child k's lower bound in byte k
    uint32_t qhi[3];
    uint32_t meta;     // per child, byte k: 0xFF unused, 0 interior, else leaf triangle count (1..31)
    uint32_t pad;      // the root's: the number of top nodes kept in LDS (nodes [0, pad)); others 0
    uint32_t link[4];  // child k's stack entry: interior, node index; leaf, kLeafTag | count << 26 | first
                       // triangle; unused, kEmptyChild
};
// RC_LDS_TOP = 1 keeps the first kTopNodes wide nodes (the upload puts the
// tree's breadth-first top there) in LDS for the traversal kernels.  Off: on
// the 2M-triangle terrain it was 1% slower than fetching those nodes through
// the caches (1,848 against 1,868 Msamples/s).
// RC_NODE_ORDER = 1 stores each wide node's interior children side by side
// (blocks in depth-first order); 0 (the default) keeps the collapse's
// preorder, a node followed by its first child's subtree.
#ifndef RC_NODE_ORDER
#define RC_NODE_ORDER 0
#endif
#ifndef RC_LDS_TOP
#define RC_LDS_TOP 0
#endif
#ifndef RC_TOP_NODES
#define RC_TOP_NODES 21
#endif
constexpr uint32_t kTopNodes = RC_TOP_NODES;  // 21: levels 0..2 of a 4-wide tree, 1,344 bytes
static_assert(sizeof(QNode) == 64, "QNode must be 64 bytes");
#if RC_QNODE
using NodeT = QNode;
#else
using NodeT = NodeRec;
#endif

// RC_TRI_EDGES 1: the FP32 decision's records hold v0 and the edges
// e1 = v1 - v0, e2 = v2 - v0 (rounded in FP32 on the host exactly as the
// kernel would round them), saving six subtractions per triangle test; the FP64
// decision keeps a second array with the vertices.
#ifndef RC_TRI_EDGES
#define RC_TRI_EDGES 0
#endif
struct alignas(16) TriRec {
    float4 v0;  // w: original triangle id (bits)
    float4 v1;  // w: flags (bit1 nonShadowing); e1 in the FP32 decision's records with RC_TRI_EDGES
    float4 v2;  // w: unused; e2 likewise
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

struct alignas(16) DevMaterial {
    float4 diffuse;   // w: flags (bits)
    float4 emittance; // w: shininess
    float4 ambient;
    float4 specular;
};

constexpr uint32_t kNoHit = 0x7fffffffu;
constexpr int kDeepStack = 256;  // stack entries of the overflow kernel (memory part)

// Traversal stack of one lane: the top kRegStack entries in registers, as a
// shift register (no dynamic register indexing), deeper entries in LDS laid
// out [depth][lane] so a wave's accesses hit 64 distinct banks.  A pop takes
// its entry from a register, so the next node fetch never waits on an LDS
// read; the LDS refill of the bottom register runs off the critical path.
// A ray whose stack would outgrow kOnChipStack entries is handed to the
// deep-stack kernel.
#ifndef RC_REG_STACK
#define RC_REG_STACK 4
#endif
#ifndef RC_LDS_DEPTH
#define RC_LDS_DEPTH 16
#endif
constexpr int kRegStack = RC_REG_STACK;
constexpr int kLdsStack = RC_LDS_DEPTH;
constexpr int kOnChipStack = kRegStack + kLdsStack;
constexpr int kWidth = 4;        // children per wide node
constexpr uint32_t kEmptyChild = 0xffffffffu;

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

// Pixel order of the slots of a pass: the 8x8 tiles of the largest tiled
// region first (tiles row by row, pixels row by row inside a tile), then the
// right strip and the bottom strip row by row.  A wave of 64 consecutive slots
// is one 8x8 tile, and since the queues keep slot order from bounce to bounce
// its rays start from nearby points at every bounce.
struct PixelOrder {
    uint32_t width, tiles_x, x0, y0, a;  // x0 = 8 tiles_x, y0 = 8 tiles_y, a = x0 y0
};

inline PixelOrder make_order(uint32_t w, uint32_t h) {
    PixelOrder o;
    o.width = w;
    o.tiles_x = w / 8;
    o.x0 = 8 * (w / 8);
    o.y0 = 8 * (h / 8);
    o.a = o.x0 * o.y0;
    return o;
}

__host__ __device__ __forceinline__ uint32_t order_to_pixel(const PixelOrder& o, uint32_t q) {
    if (q < o.a) {
        const uint32_t t = q >> 6, w = q & 63u, ty = t / o.tiles_x, tx = t - ty * o.tiles_x;
        return (ty * 8 + (w >> 3)) * o.width + tx * 8 + (w & 7u);
    }
    uint32_t r = q - o.a;
    const uint32_t rw = o.width - o.x0;
    if (r < rw * o.y0) {
        const uint32_t y = r / rw;
        return y * o.width + o.x0 + (r - y * rw);
    }
    r -= rw * o.y0;
    return o.y0 * o.width + r;
}

__host__ __device__ __forceinline__ uint32_t pixel_to_order(const PixelOrder& o, uint32_t p) {
    const uint32_t y = p / o.width, x = p - y * o.width;
    if (x < o.x0 && y < o.y0) return ((((y >> 3) * o.tiles_x + (x >> 3)) << 6) | ((y & 7u) << 3) | (x & 7u));
    if (y < o.y0) return o.a + y * (o.width - o.x0) + (x - o.x0);
    return o.a + (o.width - o.x0) * o.y0 + (p - o.y0 * o.width);
}

// A pass's queue slot <-> (tile-order index, sample within the pass).
// Pixel-major (RC_PIXEL_MAJOR = 1, the default): slot = index * spp_pass + j,
// the pass's samples of one pixel side by side, so a wave's rays start from
// nearly one point (its sun shadow rays nearly coincide): 4% faster on the
// 2M-triangle terrain at 64 samples per pass than sample-major (0: slot =
// j * npix + index, a wave holding 64 neighbouring pixels of one sample).
#ifndef RC_PIXEL_MAJOR
#define RC_PIXEL_MAJOR 1
#endif
__host__ __device__ __forceinline__ uint32_t slot_index(uint32_t slot, uint32_t npix, uint32_t spp_pass) {
    return RC_PIXEL_MAJOR ? slot / spp_pass : slot % npix;
}
__host__ __device__ __forceinline__ uint32_t slot_sample_in_pass(uint32_t slot, uint32_t npix, uint32_t spp_pass) {
    return RC_PIXEL_MAJOR ? slot % spp_pass : slot / npix;
}

// CMJSampler's permute (CMJSampler.cpp:12-44): Kensler's hash-based
// permutation of [0, l) ("Correlated Multi-Jittered Sampling", Pixar Technical
// Memo 13-01, 2013) in the reference's 64-bit arithmetic.  The hash is a
// bijection of [0, 2^k) with 2^k >= l, so cycle walking ends within 2^k steps.
__device__ __forceinline__ uint64_t cmj_permute(uint64_t i, uint64_t l, uint64_t p) {
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

// The per-pixel permutation of generateMultiSamples (Tracer.cpp:129-130), one
// 64-bit draw per pixel and frame, shared by all the pixel's samples.
constexpr uint32_t kPermDim = 0xFFFFFFFEu;
__device__ __forceinline__ uint64_t pixel_permutation(uint64_t seed, uint32_t pixel, uint32_t frame) {
    uint4 o = philox(make_uint4(pixel, 0u, frame, kPermDim), make_uint2(uint32_t(seed), uint32_t(seed >> 32)));
    return uint64_t(o.x) | (uint64_t(o.y) << 32);
}

// Sampler::getSample2D of each SamplerType (src/Rendering/Samplers/) for
// stratum (x, y) of an n x n pattern, with the jitter (r0, r1) and the
// permutation perm (CMJ only).
__device__ __forceinline__ float2 sample_point(uint32_t type, uint32_t x, uint32_t y, uint32_t n, uint64_t perm,
                                               float r0, float r1) {
    switch (type) {
    case RC_SAMPLER_JITTERED:  // JitteredSampler.cpp:21-34
        return make_float2((float(x) + r0) / float(n), (float(y) + r1) / float(n));
    case RC_SAMPLER_RANDOM:  // RandomSampler.cpp:23-35
        return make_float2(r0, r1);
    case RC_SAMPLER_CENTER:  // CenterSampler.cpp:19-33
        return make_float2(0.5f, 0.5f);
    case RC_SAMPLER_REGULAR:  // RegularSampler.cpp:19-34
        return make_float2((float(x) + 0.5f) / float(n), (float(y) + 0.5f) / float(n));
    case RC_SAMPLER_CMJ: {  // CMJSampler.cpp:57-77
        const float sx = float(cmj_permute(x, n, perm * 0x68bc21ebull));
        const float sy = float(cmj_permute(y, n, perm * 0x02e5be93ull));
        return make_float2((float(x) + (sy + r0) / float(n)) / float(n), (float(y) + (sx + r1) / float(n)) / float(n));
    }
    default:  // PoissonDiscSampler::getSample2D returns (0, 0) (PoissonDiscSampler.cpp:22-32)
        return make_float2(0.0f, 0.0f);
    }
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
    return tn <= tf * 1.000002f ? tn : __builtin_inff();
}

// FP32 Möller–Trumbore with the reference's structure (Triangle.cpp:61-91):
// double-sided; the determinant test is the scale-free det == 0 instead of
// |det| < epsilon.  This is a conservative filter: the barycentric bounds are
// widened by kFilterBand so FP32 rounding never rejects a triangle the FP64
// decision (mt_hit64, inclusive bounds) would accept.
constexpr float kFilterBand = 1e-3f;
// RC_FAST_RCP 1: the FP32 filter's 1/det is the hardware reciprocal (1 ulp;
// the filter bands absorb it and FP64 decides).  RC_SLAB_SELECT 1: the four
// child slab tests run unconditionally and empty slots are masked by a select
// (no branch).  Together 10% faster on the 2M-triangle terrain.
#ifndef RC_FAST_RCP
#define RC_FAST_RCP 1
#endif
#ifndef RC_SLAB_SELECT
#define RC_SLAB_SELECT 1
#endif
__device__ __forceinline__ bool mt_hit(const TriRec& T, float3 o, float3 d, float& t) {
    float3 v0 = f3(T.v0);
    float3 e1 = f3(T.v1) - v0, e2 = f3(T.v2) - v0;
    float3 pv = cross3(d, e2);
    float det = dot3(e1, pv);
    if (det == 0.0f) return false;
#if RC_FAST_RCP
    // hardware reciprocal (1 ulp): the bands above absorb it, FP64 decides
    float inv = __builtin_amdgcn_rcpf(det);
#else
    float inv = 1.0f / det;
#endif
    float3 tv = o - v0;
    float u = dot3(tv, pv) * inv;
    if (u < -kFilterBand || u > 1.0f + kFilterBand) return false;
    float3 qv = cross3(tv, e1);
    float v = dot3(d, qv) * inv;
    if (v < -kFilterBand || u + v > 1.0f + kFilterBand) return false;
    t = dot3(e2, qv) * inv;
    return true;
}

// The same test in FP64 on the exact FP32 inputs, for candidates the filter
// accepted: it decides the winner and gives t to FP64 accuracy, so sliver
// triangles at grazing angles (where FP32 t loses ~1e-4) still resolve like
// the FP64 reference.
__device__ __forceinline__ bool mt_hit64(const TriRec& T, float3 of, float3 df, double& t, double& u, double& v) {
    const double v0x = T.v0.x, v0y = T.v0.y, v0z = T.v0.z;
    const double e1x = double(T.v1.x) - v0x, e1y = double(T.v1.y) - v0y, e1z = double(T.v1.z) - v0z;
    const double e2x = double(T.v2.x) - v0x, e2y = double(T.v2.y) - v0y, e2z = double(T.v2.z) - v0z;
    const double dx = df.x, dy = df.y, dz = df.z;
    const double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
    const double det = e1x * px + e1y * py + e1z * pz;
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    const double tx = double(of.x) - v0x, ty = double(of.y) - v0y, tz = double(of.z) - v0z;
    u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const double qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    v = (dx * qx + dy * qy + dz * qz) * inv;
    if (v < 0.0 || (u + v) > 1.0) return false;
    t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    return true;
}

// The FP32 decision (the default, RC_HIT_FP32): Triangle::intersect's
// Möller–Trumbore (Triangle.cpp:61-91) and its OpenCL form
// (data/opencl/intersections.cl:27-68) in FP32, double-sided, inclusive
// barycentric bounds, with the scale-free det == 0 instead of |det| <
// epsilon.  The bounds are tested without a division: u = U/det and v =
// V/det lie in [0, 1] with u + v <= 1 exactly when sign(det) U, sign(det) V
// >= 0 and their sum <= |det|; t = T/det is one IEEE division for the
// triangles that pass.  u and v are returned undivided (U, V and adet = |det|).
__device__ __forceinline__ bool mt_dec32(const TriRec& T, float3 o, float3 d, float& t, float& U, float& V,
                                         float& adet) {
    const float3 v0 = f3(T.v0);
    const float3 e1 = RC_TRI_EDGES ? f3(T.v1) : f3(T.v1) - v0, e2 = RC_TRI_EDGES ? f3(T.v2) : f3(T.v2) - v0;
    const float3 pv = cross3(d, e2);
    const float det = dot3(e1, pv);
    const float sg = det < 0.0f ? -1.0f : 1.0f;
    adet = fabsf(det);
    const float3 tv = o - v0;
    U = sg * dot3(tv, pv);
    // one branch for both tests, so the load of v0 is not sunk behind the first
    if (det == 0.0f || U < 0.0f || U > adet) return false;
    const float3 qv = cross3(tv, e1);
    V = sg * dot3(d, qv);
    if (V < 0.0f || U + V > adet) return false;
    t = dot3(e2, qv) / det;
    return true;
}

// Relative window inside which an FP32 distance defers to the FP64 one.
constexpr float kRefineWindow = 1e-3f;
// FP32 decision: nodes are culled a hair beyond the best distance so a
// triangle at exactly that distance (an exact tie) is still reached.
constexpr float kTieWindow = 4.0f / 16777216.0f;

// Stack push of x (false: the stack already holds `limit` entries).  sp
// counts all entries; the first kRegStack live in reg[0..] (reg[0] the top),
// entry k >= kRegStack at mem[(k - kRegStack) * stride].
__device__ __forceinline__ bool stack_push(uint32_t* reg, int& sp, uint32_t x, uint32_t* mem, int stride, int limit) {
    if (sp >= limit) return false;
    if (kRegStack == 0) {
        mem[sp * stride] = x;
    } else {
        if (sp >= kRegStack) mem[(sp - kRegStack) * stride] = reg[kRegStack - 1];
#pragma unroll
        for (int k = kRegStack - 1; k > 0; --k) reg[k] = reg[k - 1];
        reg[0] = x;
    }
    ++sp;
    return true;
}

// Stack pop (sp > 0): the top register, the registers shifted up and the
// bottom one refilled from memory while entries remain there.
__device__ __forceinline__ uint32_t stack_pop(uint32_t* reg, int& sp, const uint32_t* mem, int stride) {
    --sp;
    if (kRegStack == 0) return mem[sp * stride];
    const uint32_t x = reg[0];
#pragma unroll
    for (int k = 0; k < kRegStack - 1; ++k) reg[k] = reg[k + 1];
    if (sp >= kRegStack) reg[kRegStack - 1] = mem[(sp - kRegStack) * stride];
    return x;
}

// The barycentrics of the hit are u / w and v / w: the FP32 decision keeps
// them undivided (w = |det|) so only a reader of u and v pays the divisions
// (bary_u / bary_v); the FP64 decision stores them divided, with w = 1.
// Streamed accesses: queue entries, path states, slot radiance and parked
// traversals are written once and read once per bounce.  RC_NT_QUEUES 1 makes
// them nontemporal loads and stores, so they do not displace nodes and
// triangles from L2 and the Infinity Cache.
#ifndef RC_NT_QUEUES
#define RC_NT_QUEUES 0
#endif
typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream(const float4* p) {
    if constexpr (RC_NT_QUEUES) {
        const v4f_t v = __builtin_nontemporal_load(reinterpret_cast<const v4f_t*>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}
__device__ __forceinline__ void st_stream(float4* p, float4 v) {
    if constexpr (RC_NT_QUEUES) {
        __builtin_nontemporal_store(v4f_t{v.x, v.y, v.z, v.w}, reinterpret_cast<v4f_t*>(p));
    } else {
        *p = v;
    }
}
__device__ __forceinline__ uint4 ld_stream(const uint4* p) {
    if constexpr (RC_NT_QUEUES) {
        const v4u_t v = __builtin_nontemporal_load(reinterpret_cast<const v4u_t*>(p));
        return make_uint4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}
__device__ __forceinline__ void st_stream(uint4* p, uint4 v) {
    if constexpr (RC_NT_QUEUES) {
        __builtin_nontemporal_store(v4u_t{v.x, v.y, v.z, v.w}, reinterpret_cast<v4u_t*>(p));
    } else {
        *p = v;
    }
}

struct TraceResult {
    float t, u, v;
    uint32_t id;
    bool overflow;
    float w;
};
__device__ __forceinline__ float bary_u(const TraceResult& R) { return R.u / R.w; }
__device__ __forceinline__ float bary_v(const TraceResult& R) { return R.v / R.w; }

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w runs on XCD
// w % 8).  This bijection on [0, n) gives XCD x the contiguous range of
// logical blocks [x n / 8, (x + 1) n / 8), so each XCD's private L2 sees a
// contiguous, spatially coherent slice of the ray queue.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t w, uint32_t n) {
    uint32_t x = w % kXcds, local = w / kXcds;
    uint32_t per = n / kXcds, rem = n % kXcds;
    return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

// Wave-wide sum, one 64-bit atomic per wave (every lane must call it).
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
    unsigned long long x = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(dst, x);
}

// Tests the triangles tris[first, first + count).  DEC64: candidates pass
// the FP32 filter within kRefineWindow of the current best and are then
// decided in FP64 on the exact FP32 inputs (RC_HIT_FP64); otherwise the FP32
// test decides (mt_dec32).  Closest hit keeps the smallest t; an exact tie
// goes to the lower original triangle id, so the answer does not depend on
// traversal order.  Returns true when an any-hit query is answered (tmax_any:
// its tmax).
template <bool DEC64>
struct BestT {
    using T = float;
};
template <>
struct BestT<true> {
    using T = double;
};

template <bool ANY, bool COUNT, bool DEC64>
__device__ __forceinline__ bool test_leaf(const TriRec* __restrict__ tris, uint32_t first, uint32_t count, float3 o,
                                          float3 d, float tmin, float tmax_any, typename BestT<DEC64>::T& best,
                                          float& cull, TraceResult& R, uint32_t& n_tri) {
    for (uint32_t j = first; j < first + count; ++j) {
        const float4* q = reinterpret_cast<const float4*>(tris + j);
        TriRec T;
        T.v0 = q[0];
        T.v1 = q[1];
        T.v2 = q[2];
        if (COUNT) n_tri += 1;
        if (DEC64) {
            float t32;
            if (!mt_hit(T, o, d, t32)) continue;
            if (t32 > cull || t32 < tmin * (1.0f - kRefineWindow) - 1e-30f) continue;
            double t64, u64, v64;
            if (!mt_hit64(T, o, d, t64, u64, v64)) continue;
            if (t64 < double(tmin)) continue;
            uint32_t id = __float_as_uint(T.v0.w);
            if (ANY) {
                if (t64 <= double(tmax_any) && !(__float_as_uint(T.v1.w) & RC_MAT_NON_SHADOWING)) {
                    R.t = float(t64);
                    R.id = id;
                    return true;
                }
            } else if (t64 < best || (t64 == best && id < R.id)) {
                best = t64;
                R.t = float(t64);
                R.u = float(u64);
                R.v = float(v64);
                R.w = 1.0f;
                R.id = id;
                cull = R.t * (1.0f + kRefineWindow);
            }
        } else {
            float t, U, V, adet;
            if (!mt_dec32(T, o, d, t, U, V, adet)) continue;
            if (t < tmin || t > cull) continue;
            uint32_t id = __float_as_uint(T.v0.w);
            if (ANY) {
                if (t <= tmax_any && !(__float_as_uint(T.v1.w) & RC_MAT_NON_SHADOWING)) {
                    R.t = t;
                    R.id = id;
                    return true;
                }
            } else if (t < best || (t == best && id < R.id)) {
                best = t;
                R.t = t;
                R.u = U;
                R.v = V;
                R.w = adet;
                R.id = id;
                cull = t * (1.0f + kTieWindow);
            }
        }
    }
    return false;
}

// Cull distance of a query with no hit yet.
template <bool DEC64>
__device__ __forceinline__ float initial_cull(float tmax) {
    return tmax * (1.0f + (DEC64 ? kRefineWindow : kTieWindow));
}

// Compare-exchange of two children (entry distance, stack entry) by distance,
// as selects (no branch).
__device__ __forceinline__ void order2(float& ta, uint32_t& ea, float& tb, uint32_t& eb) {
    const bool sw = tb < ta;
    const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
    const uint32_t e0 = sw ? eb : ea, e1 = sw ? ea : eb;
    ta = t0; tb = t1;
    ea = e0; eb = e1;
}

// The same with the link and count of a child record carried separately.
__device__ __forceinline__ void order2(float& ta, uint32_t& la, uint32_t& ca, float& tb, uint32_t& lb, uint32_t& cb) {
    const bool sw = tb < ta;
    const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
    const uint32_t l0 = sw ? lb : la, l1 = sw ? la : lb;
    const uint32_t c0 = sw ? cb : ca, c1 = sw ? ca : cb;
    ta = t0; tb = t1;
    la = l0; lb = l1;
    ca = c0; cb = c1;
}

// 1: the if-if traversal (leaves postponed onto the stack), 2% faster on the
// 2M-triangle terrain than testing leaf children at once (0).
#ifndef RC_TRAV_IFIF
#define RC_TRAV_IFIF 1
#endif
// 1: any-hit queries visit children nearest first too; 0: in slot order (an
// any-hit answer does not depend on the order, only its cost does).
#ifndef RC_ANY_SORT
#define RC_ANY_SORT 0
#endif
// 0: off; n > 0: leaves are set aside and tested when at least n / 64 of the
// wave's lanes hold one (or none has a node), the lanes meanwhile going on
// with their stacks.
#ifndef RC_LEAF_POSTPONE
#define RC_LEAF_POSTPONE 20
#endif
// The same threshold for any-hit queries.
#ifndef RC_LEAF_POSTPONE_ANY
#define RC_LEAF_POSTPONE_ANY 32
#endif
// Stack entries of the if-if traversal: an interior child's record index, or
// a leaf tagged with kLeafTag, its triangle count (1..31) in bits 26..30 and
// its first triangle (< 2^26) below.  The upload guarantees both: scenes hold
// at most 2^26 triangles and longer leaves are split into wide nodes.
constexpr uint32_t kLeafTag = 0x80000000u;
constexpr uint32_t kMaxLeafRecord = 32;           // leaf records hold fewer triangles
constexpr size_t kMaxTriangles = size_t(1) << 26;

// Ordered, t-culled traversal over 4-wide nodes.  The four child boxes are
// tested and sorted by entry distance (a 5-comparator network); leaf children
// are tested in that order at once, the nearest interior child is visited
// next and the others are pushed far-to-near on the lane's stack (registers,
// then stack[k * stride]: LDS for the main kernels, global memory for the
// deep-stack kernel), at most max_depth entries.  ANY: stop at the first
// triangle (not nonShadowing) with t in [tmin, tmax].
template <bool ANY, bool COUNT, bool DEC64>
__device__ __forceinline__ TraceResult traverse_inline(const NodeRec* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                float3 o, float3 d, float tmin, float tmax, uint32_t* stack,
                                                int stride, int max_depth, uint32_t& n_rec, uint32_t& n_tri) {
    const float INF = __builtin_inff();
    TraceResult R{tmax, 0.0f, 0.0f, kNoHit, false, 1.0f};
    tmin = fmaxf(tmin, 0.0f);
    typename BestT<DEC64>::T best = tmax;
    float cull = initial_cull<DEC64>(tmax);
    float3 inv = safe_inv(d);
    uint32_t cur = 0;
    int sp = 0;
    uint32_t reg[kRegStack > 0 ? kRegStack : 1];
    while (true) {
        const float4* p = reinterpret_cast<const float4*>(nodes + cur);
        float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5], q6 = p[6], q7 = p[7];
        if (COUNT) n_rec += kWidth;
        uint32_t l0 = __float_as_uint(q1.z), c0 = __float_as_uint(q1.w);
        uint32_t l1 = __float_as_uint(q3.z), c1 = __float_as_uint(q3.w);
        uint32_t l2 = __float_as_uint(q5.z), c2 = __float_as_uint(q5.w);
        uint32_t l3 = __float_as_uint(q7.z), c3 = __float_as_uint(q7.w);
        float t0 = c0 == kEmptyChild ? INF : slab(q0, q1, o, inv, tmin, cull);
        float t1 = c1 == kEmptyChild ? INF : slab(q2, q3, o, inv, tmin, cull);
        float t2 = c2 == kEmptyChild ? INF : slab(q4, q5, o, inv, tmin, cull);
        float t3 = c3 == kEmptyChild ? INF : slab(q6, q7, o, inv, tmin, cull);
        order2(t0, l0, c0, t1, l1, c1);
        order2(t2, l2, c2, t3, l3, c3);
        order2(t0, l0, c0, t2, l2, c2);
        order2(t1, l1, c1, t3, l3, c3);
        order2(t1, l1, c1, t2, l2, c2);
        // leaf children, nearest first
        if (c0 != 0 && t0 != INF && t0 <= cull)
            if (test_leaf<ANY, COUNT, DEC64>(tris, l0, c0, o, d, tmin, tmax, best, cull, R, n_tri)) return R;
        if (c1 != 0 && t1 != INF && t1 <= cull)
            if (test_leaf<ANY, COUNT, DEC64>(tris, l1, c1, o, d, tmin, tmax, best, cull, R, n_tri)) return R;
        if (c2 != 0 && t2 != INF && t2 <= cull)
            if (test_leaf<ANY, COUNT, DEC64>(tris, l2, c2, o, d, tmin, tmax, best, cull, R, n_tri)) return R;
        if (c3 != 0 && t3 != INF && t3 <= cull)
            if (test_leaf<ANY, COUNT, DEC64>(tris, l3, c3, o, d, tmin, tmax, best, cull, R, n_tri)) return R;
        // interior children: farthest pushed first, nearest visited next
        bool have = false;
        uint32_t next = 0;
#define RC_INTERIOR(t, l, c)                                            \
        if (c == 0 && t != INF && t <= cull) {                          \
            if (have && !stack_push(reg, sp, next, stack, stride, max_depth)) { \
                R.overflow = true;                                      \
                return R;                                               \
            }                                                           \
            next = l;                                                   \
            have = true;                                                \
        }
        RC_INTERIOR(t3, l3, c3)
        RC_INTERIOR(t2, l2, c2)
        RC_INTERIOR(t1, l1, c1)
        RC_INTERIOR(t0, l0, c0)
#undef RC_INTERIOR
        if (have) {
            cur = next;
        } else {
            if (sp == 0) break;
            cur = stack_pop(reg, sp, stack, stride);
        }
    }
    return R;
}

// The same traversal with leaves postponed ("if-if"): a node visit pushes its
// hit children, leaves included, far-to-near and keeps the nearest; each step
// then does at most one node visit and one leaf, so the lanes of a wave stay
// in step instead of waiting while one lane tests up to four leaves.  The
// state is explicit (trav_begin / trav_step).
template <bool DEC64>
struct Trav {
    float3 o, d, inv;
    float tmin, cull;
    typename BestT<DEC64>::T best;
    TraceResult R;
    uint32_t cur;
    int sp;
    uint32_t reg[kRegStack > 0 ? kRegStack : 1];  // top of the stack
    uint32_t leaf;     // a leaf set aside (leaf postponing), else kEmptyChild
    const uint4* top;  // LDS copy of nodes [0, ntop), or null
    uint32_t ntop;
};

// Copies the first `ntop` nodes (the root's count) into LDS `top` (four
// 16-byte words per node); every thread of the block calls it, and it ends in
// a barrier.  Returns the count (0 with the 128-byte node records).
template <class N>
__device__ __forceinline__ uint32_t load_top(const N* __restrict__ nodes, uint4* top) {
    if constexpr (!RC_LDS_TOP || sizeof(N) != sizeof(QNode)) {
        return 0u;
    } else {
        const uint32_t ntop = min(reinterpret_cast<const QNode*>(nodes)[0].pad, kTopNodes);
        const uint4* src = reinterpret_cast<const uint4*>(nodes);
        for (uint32_t w = threadIdx.x; w < 4 * ntop; w += blockDim.x) top[w] = src[w];
        __syncthreads();
        return ntop;
    }
}

template <bool DEC64>
__device__ __forceinline__ void trav_begin(Trav<DEC64>& S, float3 o, float3 d, float tmin, float tmax,
                                           const uint4* top = nullptr, uint32_t ntop = 0) {
    S.top = top;
    S.ntop = top ? ntop : 0u;
    S.o = o;
    S.d = d;
    S.inv = safe_inv(d);
    S.tmin = fmaxf(tmin, 0.0f);
    S.best = tmax;
    S.cull = initial_cull<DEC64>(tmax);
    S.R = TraceResult{tmax, 0.0f, 0.0f, kNoHit, false, 1.0f};
    S.cur = 0;  // root: interior record 0
    S.sp = 0;
    S.leaf = kEmptyChild;
}

// One step; true when the traversal has ended (S.R holds the result,
// S.R.overflow when the stack would have outgrown max_depth entries).
template <bool ANY, bool COUNT, bool DEC64>
__device__ __forceinline__ bool trav_step(Trav<DEC64>& S, const NodeT* __restrict__ nodes,
                                          const TriRec* __restrict__ tris, uint32_t* stack, int stride, int max_depth,
                                          uint32_t& n_rec, uint32_t& n_tri) {
    const float INF = __builtin_inff();
    if (!(S.cur & kLeafTag)) {
#if RC_QNODE
        uint4 q0, q1, q2, q3;
        if (S.cur < S.ntop) {  // the top of the tree, from LDS
            const uint4* t = S.top + 4 * S.cur;
            q0 = t[0];
            q1 = t[1];
            q2 = t[2];
            q3 = t[3];
        } else {
            const uint4* p = reinterpret_cast<const uint4*>(nodes + S.cur);
            q0 = p[0];
            q1 = p[1];
            q2 = p[2];
            q3 = p[3];
        }
        if (COUNT) n_rec += kWidth;
        // bound = origin + q * 2^e; in ray distance q * (2^e / d) + (origin - o) / d
        const float ax = __uint_as_float((q0.w & 0xFFu) << 23) * S.inv.x;
        const float ay = __uint_as_float(((q0.w >> 8) & 0xFFu) << 23) * S.inv.y;
        const float az = __uint_as_float(((q0.w >> 16) & 0xFFu) << 23) * S.inv.z;
        const float bx = (__uint_as_float(q0.x) - S.o.x) * S.inv.x;
        const float by = (__uint_as_float(q0.y) - S.o.y) * S.inv.y;
        const float bz = (__uint_as_float(q0.z) - S.o.z) * S.inv.z;
        // Per axis the entry plane is the lower bound when the direction is
        // positive and the upper one when it is negative (the decode is
        // monotonic in q): picking the two words by sign once per node takes
        // the place of a min and a max per axis and child, with the same result.
        const uint32_t nx = S.inv.x < 0.0f ? q1.w : q1.x, fx = S.inv.x < 0.0f ? q1.x : q1.w;
        const uint32_t ny = S.inv.y < 0.0f ? q2.x : q1.y, fy = S.inv.y < 0.0f ? q1.y : q2.x;
        const uint32_t nz = S.inv.z < 0.0f ? q2.y : q1.z, fz = S.inv.z < 0.0f ? q1.z : q2.y;
#define RC_QSLAB(k, e)                                                                                         \
        [&]() {                                                                                                \
            const float tnx = fmaf(float((nx >> (8 * k)) & 0xFFu), ax, bx);                                     \
            const float tfx = fmaf(float((fx >> (8 * k)) & 0xFFu), ax, bx);                                     \
            const float tny = fmaf(float((ny >> (8 * k)) & 0xFFu), ay, by);                                     \
            const float tfy = fmaf(float((fy >> (8 * k)) & 0xFFu), ay, by);                                     \
            const float tnz = fmaf(float((nz >> (8 * k)) & 0xFFu), az, bz);                                     \
            const float tfz = fmaf(float((fz >> (8 * k)) & 0xFFu), az, bz);                                     \
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, S.tmin));                                        \
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, S.cull));                                        \
            return tn <= tf * 1.000002f && e != kEmptyChild ? tn : INF;                                        \
        }()
        // links are the children's stack entries (kEmptyChild for an unused slot)
        uint32_t e0 = q3.x, e1 = q3.y, e2 = q3.z, e3 = q3.w;
        float t0 = RC_QSLAB(0, e0), t1 = RC_QSLAB(1, e1), t2 = RC_QSLAB(2, e2), t3 = RC_QSLAB(3, e3);
#undef RC_QSLAB
        if (!ANY || RC_ANY_SORT) {
            order2(t0, e0, t1, e1);
            order2(t2, e2, t3, e3);
            order2(t0, e0, t2, e2);
            order2(t1, e1, t3, e3);
            order2(t1, e1, t2, e2);
        }
#else
        const float4* p = reinterpret_cast<const float4*>(nodes + S.cur);
        float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5], q6 = p[6], q7 = p[7];
        if (COUNT) n_rec += kWidth;
        uint32_t l0 = __float_as_uint(q1.z), c0 = __float_as_uint(q1.w);
        uint32_t l1 = __float_as_uint(q3.z), c1 = __float_as_uint(q3.w);
        uint32_t l2 = __float_as_uint(q5.z), c2 = __float_as_uint(q5.w);
        uint32_t l3 = __float_as_uint(q7.z), c3 = __float_as_uint(q7.w);
#if RC_SLAB_SELECT
        float t0 = slab(q0, q1, S.o, S.inv, S.tmin, S.cull), t1 = slab(q2, q3, S.o, S.inv, S.tmin, S.cull);
        float t2 = slab(q4, q5, S.o, S.inv, S.tmin, S.cull), t3 = slab(q6, q7, S.o, S.inv, S.tmin, S.cull);
        t0 = c0 == kEmptyChild ? INF : t0;
        t1 = c1 == kEmptyChild ? INF : t1;
        t2 = c2 == kEmptyChild ? INF : t2;
        t3 = c3 == kEmptyChild ? INF : t3;
#else
        float t0 = c0 == kEmptyChild ? INF : slab(q0, q1, S.o, S.inv, S.tmin, S.cull);
        float t1 = c1 == kEmptyChild ? INF : slab(q2, q3, S.o, S.inv, S.tmin, S.cull);
        float t2 = c2 == kEmptyChild ? INF : slab(q4, q5, S.o, S.inv, S.tmin, S.cull);
        float t3 = c3 == kEmptyChild ? INF : slab(q6, q7, S.o, S.inv, S.tmin, S.cull);
#endif
        uint32_t e0 = c0 != 0 ? (kLeafTag | (c0 << 26) | l0) : l0, e1 = c1 != 0 ? (kLeafTag | (c1 << 26) | l1) : l1;
        uint32_t e2 = c2 != 0 ? (kLeafTag | (c2 << 26) | l2) : l2, e3 = c3 != 0 ? (kLeafTag | (c3 << 26) | l3) : l3;
        order2(t0, e0, t1, e1);
        order2(t2, e2, t3, e3);
        order2(t0, e0, t2, e2);
        order2(t1, e1, t3, e3);
        order2(t1, e1, t2, e2);
#endif  // RC_QNODE
        bool have = false;
        uint32_t next = 0;
#define RC_CHILD(t, e)                                                                                         \
        if (t != INF && t <= S.cull) {                                                                         \
            if (have && !stack_push(S.reg, S.sp, next, stack, stride, max_depth)) {                            \
                S.R.overflow = true;                                                                           \
                return true;                                                                                   \
            }                                                                                                  \
            next = e;                                                                                          \
            have = true;                                                                                       \
        }
        RC_CHILD(t3, e3)
        RC_CHILD(t2, e2)
        RC_CHILD(t1, e1)
        RC_CHILD(t0, e0)
#undef RC_CHILD
        if (have) {
            S.cur = next;
        } else if (S.sp == 0) {
            if (!RC_LEAF_POSTPONE) return true;
            S.cur = kEmptyChild;  // nothing left but a leaf set aside
        } else {
            S.cur = stack_pop(S.reg, S.sp, stack, stride);
        }
    }
#if RC_LEAF_POSTPONE
    constexpr uint32_t kHold = ANY ? RC_LEAF_POSTPONE_ANY : RC_LEAF_POSTPONE;
    // Leaf postponing: a lane that reaches a leaf sets it aside (one slot)
    // and goes on with its stack; the wave tests the leaves set aside when at
    // least RC_LEAF_POSTPONE / 64 of its lanes hold one, or when no lane has a
    // node left to visit.  A lane whose next entry is a second leaf waits.
    if (S.cur != kEmptyChild && (S.cur & kLeafTag) && S.leaf == kEmptyChild) {
        S.leaf = S.cur;
        S.cur = S.sp == 0 ? kEmptyChild : stack_pop(S.reg, S.sp, stack, stride);
    }
    {
        const bool held = S.leaf != kEmptyChild;
        const uint64_t lanes = __ballot(1), visit = __ballot(!(S.cur & kLeafTag)), holding = __ballot(held);
        const bool run = visit == 0 || uint32_t(__popcll(holding)) * 64u >= uint32_t(__popcll(lanes)) * kHold;
        if (held && run) {
            const uint32_t first = S.leaf & ((1u << 26) - 1u), count = (S.leaf >> 26) & 31u;
            if (test_leaf<ANY, COUNT, DEC64>(tris, first, count, S.o, S.d, S.tmin, S.R.t, S.best, S.cull, S.R, n_tri))
                return true;
            S.leaf = kEmptyChild;
        }
    }
    return S.cur == kEmptyChild && S.leaf == kEmptyChild;
#else
    if (S.cur & kLeafTag) {
        const uint32_t first = S.cur & ((1u << 26) - 1u), count = (S.cur >> 26) & 31u;
        // an any-hit query's R.t holds tmax until the query is answered
        if (test_leaf<ANY, COUNT, DEC64>(tris, first, count, S.o, S.d, S.tmin, S.R.t, S.best, S.cull, S.R, n_tri))
            return true;
        if (S.sp == 0) return true;
        S.cur = stack_pop(S.reg, S.sp, stack, stride);
    }
    return false;
#endif  // RC_LEAF_POSTPONE
}

// A traversal parked mid-way (the extension kernel's step budget ran out):
// 32 words per ray, written and read as eight 16-byte vectors.  Words: 0 the
// queue entry, 1 the current entry, 2 the stack size, 3 the best hit's id,
// 4..6 its t, u, v, 7 the cull distance, 8..9 the best distance (a double
// with the FP64 decision, else a float in word 8), 10.. the register stack
// (top first), then the LDS rows in their order (only the first sp - kRegStack
// are meaningful); word 30 the best hit's w, word 31 the leaf set aside.
constexpr int kParkWords = 32;
constexpr int kParkW = 30;
static_assert(10 + kRegStack + kLdsStack <= kParkW, "a parked traversal does not fit its record");

// Word w of a parked record (compile-time w after unrolling).
template <bool DEC64>
__device__ __forceinline__ uint32_t park_word(const Trav<DEC64>& S, uint32_t entry, const uint32_t* stack, int stride,
                                              int w) {
    if (w == 0) return entry;
    if (w == 1) return S.cur;
    if (w == 2) return uint32_t(S.sp);
    if (w == 3) return S.R.id;
    if (w == 4) return __float_as_uint(S.R.t);
    if (w == 5) return __float_as_uint(S.R.u);
    if (w == 6) return __float_as_uint(S.R.v);
    if (w == 7) return __float_as_uint(S.cull);
    if (w == 8 || w == 9) {
        if constexpr (DEC64) {
            const unsigned long long b = __double_as_longlong(S.best);
            return w == 8 ? uint32_t(b) : uint32_t(b >> 32);
        } else {
            return w == 8 ? __float_as_uint(S.best) : 0u;
        }
    }
    if (w == kParkW) return __float_as_uint(S.R.w);
    if (w == kParkW + 1) return S.leaf;
    if (w < 10 + kRegStack) return S.reg[w - 10];
    const int j = w - 10 - kRegStack;  // LDS row
    return j < kLdsStack && j < S.sp - kRegStack ? stack[j * stride] : 0u;
}

template <bool DEC64>
__device__ __forceinline__ void trav_park(const Trav<DEC64>& S, uint32_t entry, const uint32_t* stack, int stride,
                                          uint32_t* __restrict__ rec) {
    uint4* r = reinterpret_cast<uint4*>(rec);
#pragma unroll
    for (int q = 0; q < kParkWords / 4; ++q)
        st_stream(r + q, make_uint4(park_word(S, entry, stack, stride, 4 * q),
                                    park_word(S, entry, stack, stride, 4 * q + 1),
                                    park_word(S, entry, stack, stride, 4 * q + 2),
                                    park_word(S, entry, stack, stride, 4 * q + 3)));
}

template <bool DEC64>
__device__ __forceinline__ void unpark_word(Trav<DEC64>& S, uint32_t* stack, int stride, int w, uint32_t v,
                                            uint32_t& best_lo) {
    if (w == 1) S.cur = v;
    else if (w == 2) S.sp = int(v);
    else if (w == 3) S.R.id = v;
    else if (w == 4) S.R.t = __uint_as_float(v);
    else if (w == 5) S.R.u = __uint_as_float(v);
    else if (w == 6) S.R.v = __uint_as_float(v);
    else if (w == 7) S.cull = __uint_as_float(v);
    else if (w == kParkW) S.R.w = __uint_as_float(v);
    else if (w == kParkW + 1) S.leaf = v;
    else if (w == 8) {
        best_lo = v;
        if constexpr (!DEC64) S.best = __uint_as_float(v);
    } else if (w == 9) {
        if constexpr (DEC64) S.best = __longlong_as_double((long long)(uint64_t(best_lo) | (uint64_t(v) << 32)));
    } else if (w >= 10 && w < 10 + kRegStack) {
        S.reg[w - 10] = v;
    } else if (w >= 10 + kRegStack) {
        const int j = w - 10 - kRegStack;
        if (j < kLdsStack && j < S.sp - kRegStack) stack[j * stride] = v;
    }
}

// Resumes a parked traversal: S must have been begun on the same ray
// (trav_begin), which restores the ray's own fields; this restores the rest
// and the lane's LDS rows (word 0, the queue entry, is the caller's).
template <bool DEC64>
__device__ __forceinline__ void trav_unpark(Trav<DEC64>& S, const uint32_t* __restrict__ rec, uint32_t* stack,
                                            int stride) {
    const uint4* r = reinterpret_cast<const uint4*>(rec);
    uint32_t best_lo = 0;
#pragma unroll
    for (int q = 0; q < kParkWords / 4; ++q) {
        const uint4 v = ld_stream(r + q);
        unpark_word(S, stack, stride, 4 * q, v.x, best_lo);
        unpark_word(S, stack, stride, 4 * q + 1, v.y, best_lo);
        unpark_word(S, stack, stride, 4 * q + 2, v.z, best_lo);
        unpark_word(S, stack, stride, 4 * q + 3, v.w, best_lo);
    }
}

template <bool ANY, bool COUNT, bool DEC64>
__device__ __forceinline__ TraceResult traverse_ifif(const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                     float3 o, float3 d, float tmin, float tmax, uint32_t* stack,
                                                     int stride, int max_depth, uint32_t& n_rec, uint32_t& n_tri,
                                                     const uint4* top = nullptr, uint32_t ntop = 0) {
    Trav<DEC64> S;
    trav_begin(S, o, d, tmin, tmax, top, ntop);
    while (!trav_step<ANY, COUNT, DEC64>(S, nodes, tris, stack, stride, max_depth, n_rec, n_tri)) {
    }
    return S.R;
}

// Closest hit (ANY = false) or any hit of one ray.  stack: the lane's memory
// part of the stack, entry k at stack[k * stride]; max_depth: entries the
// stack may hold in all (registers and memory).
template <bool ANY, bool COUNT, bool DEC64>
__device__ __forceinline__ TraceResult traverse(const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                float3 o, float3 d, float tmin, float tmax, uint32_t* stack,
                                                int stride, int max_depth, uint32_t& n_rec, uint32_t& n_tri,
                                                const uint4* top = nullptr, uint32_t ntop = 0) {
#if RC_QNODE
    return traverse_ifif<ANY, COUNT, DEC64>(nodes, tris, o, d, tmin, tmax, stack, stride, max_depth, n_rec, n_tri,
                                            top, ntop);
#else
    if (RC_TRAV_IFIF)
        return traverse_ifif<ANY, COUNT, DEC64>(nodes, tris, o, d, tmin, tmax, stack, stride, max_depth, n_rec,
                                                n_tri);
    return traverse_inline<ANY, COUNT, DEC64>(nodes, tris, o, d, tmin, tmax, stack, stride, max_depth, n_rec, n_tri);
#endif
}

#if RC_QNODE
// Closest hit of a wave of nearly coherent rays traversed as one packet (the
// camera bounce, whose waves each hold one pixel's samples).  The wave
// follows one stack and visits a child when any of its lanes enters it, so
// every lane visits a superset of the nodes its own traversal would; the
// exact-tie rule of test_leaf makes each answer independent of the order and
// of the extra visits, so the hits are those of traverse().  Children are
// taken nearest first by the first active lane's entry distances.  The
// stack entries are wave-uniform; each lane keeps its own copy in its LDS
// row (entry k at stack[k * stride]), up to max_depth entries, beyond which
// every lane of the wave reports an overflow.  Every active lane of the wave
// calls this (the loop and its exit are uniform).
template <bool COUNT>
__device__ __forceinline__ TraceResult traverse_packet(const QNode* __restrict__ nodes,
                                                       const TriRec* __restrict__ tris, float3 o, float3 d,
                                                       float tmin, float tmax, uint32_t* stack, int stride,
                                                       int max_depth, uint32_t& n_rec, uint32_t& n_tri) {
    const float INF = __builtin_inff();
    const float3 inv = safe_inv(d);
    tmin = fmaxf(tmin, 0.0f);
    float best = tmax;
    float cull = initial_cull<false>(tmax);
    TraceResult R{tmax, 0.0f, 0.0f, kNoHit, false, 1.0f};
    uint32_t cur = 0;  // root: interior record 0 (wave-uniform, as sp)
    int sp = 0;
    for (;;) {
        if (!(cur & kLeafTag)) {
            const uint4* p = reinterpret_cast<const uint4*>(nodes + cur);
            const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
            if (COUNT) n_rec += kWidth;
            const float ax = __uint_as_float((q0.w & 0xFFu) << 23) * inv.x;
            const float ay = __uint_as_float(((q0.w >> 8) & 0xFFu) << 23) * inv.y;
            const float az = __uint_as_float(((q0.w >> 16) & 0xFFu) << 23) * inv.z;
            const float bx = (__uint_as_float(q0.x) - o.x) * inv.x;
            const float by = (__uint_as_float(q0.y) - o.y) * inv.y;
            const float bz = (__uint_as_float(q0.z) - o.z) * inv.z;
            const uint32_t nx = inv.x < 0.0f ? q1.w : q1.x, fx = inv.x < 0.0f ? q1.x : q1.w;
            const uint32_t ny = inv.y < 0.0f ? q2.x : q1.y, fy = inv.y < 0.0f ? q1.y : q2.x;
            const uint32_t nz = inv.z < 0.0f ? q2.y : q1.z, fz = inv.z < 0.0f ? q1.z : q2.y;
            auto slab_q = [&](int k, uint32_t e) {
                const float tnx = fmaf(float((nx >> (8 * k)) & 0xFFu), ax, bx);
                const float tfx = fmaf(float((fx >> (8 * k)) & 0xFFu), ax, bx);
                const float tny = fmaf(float((ny >> (8 * k)) & 0xFFu), ay, by);
                const float tfy = fmaf(float((fy >> (8 * k)) & 0xFFu), ay, by);
                const float tnz = fmaf(float((nz >> (8 * k)) & 0xFFu), az, bz);
                const float tfz = fmaf(float((fz >> (8 * k)) & 0xFFu), az, bz);
                const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
                const float tf = fminf(fminf(tfx, tfy), fminf(tfz, cull));
                return tn <= tf * 1.000002f && e != kEmptyChild ? tn : INF;
            };
            const float t0 = slab_q(0, q3.x), t1 = slab_q(1, q3.y), t2 = slab_q(2, q3.z), t3 = slab_q(3, q3.w);
            // per child: entered by any lane (m), the first active lane's
            // distance (r) and the entry (e), all wave-uniform
            uint32_t m0 = __ballot(t0 < INF) != 0, m1 = __ballot(t1 < INF) != 0;
            uint32_t m2 = __ballot(t2 < INF) != 0, m3 = __ballot(t3 < INF) != 0;
            float r0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(t0)));
            float r1 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(t1)));
            float r2 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(t2)));
            float r3 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(t3)));
            // the entry of an entered child, else kEmptyChild (which sorts
            // with the misses)
            uint32_t e0 = m0 ? __builtin_amdgcn_readfirstlane(q3.x) : kEmptyChild;
            uint32_t e1 = m1 ? __builtin_amdgcn_readfirstlane(q3.y) : kEmptyChild;
            uint32_t e2 = m2 ? __builtin_amdgcn_readfirstlane(q3.z) : kEmptyChild;
            uint32_t e3 = m3 ? __builtin_amdgcn_readfirstlane(q3.w) : kEmptyChild;
            order2(r0, e0, r1, e1);
            order2(r2, e2, r3, e3);
            order2(r0, e0, r2, e2);
            order2(r1, e1, r3, e3);
            order2(r1, e1, r2, e2);
            // far to near: the nearest entered child is visited next, the
            // others pushed farthest first
            uint32_t next = kEmptyChild;
            bool overflow = false;
            auto take = [&](uint32_t e) {
                if (e == kEmptyChild) return;
                if (next != kEmptyChild) {
                    if (sp >= max_depth) {
                        overflow = true;
                        return;
                    }
                    stack[sp * stride] = next;
                    ++sp;
                }
                next = e;
            };
            take(e3);
            take(e2);
            take(e1);
            take(e0);
            if (overflow) {
                R.overflow = true;
                return R;
            }
            if (next != kEmptyChild) {
                cur = next;
                continue;
            }
        } else {
            // every active lane tests the leaf: a lane whose ray misses the
            // leaf's box finds no hit there it would not also find through
            // its own traversal
            test_leaf<false, COUNT, false>(tris, cur & 0x03FFFFFFu, (cur >> 26) & 0x1Fu, o, d, tmin, 0.0f, best, cull,
                                           R, n_tri);
        }
        if (sp == 0) return R;
        --sp;
        cur = __builtin_amdgcn_readfirstlane(stack[sp * stride]);
    }
}
#endif

}  // namespace rcd
