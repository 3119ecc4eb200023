// Wavefront path tracer and Phong raycast: the device counterpart of
// Tracer::run / generateMultiSamples (src/Raytracing/Tracers/Tracer.cpp:58-144),
// PathTracer::traceRecursive (PathTracer.cpp:25-75) and the Phong + point-light
// shadow part of Raytracer (Raytracer.cpp:312-407, 524-551).
//
// A render is a sequence of passes; a pass traces spp_per_pass samples of
// every pixel.  Slot s = j * npix + q holds sample j of the pass for the
// pixel at position q of the tiled pixel order (PixelOrder below).  Per
// bounce: extend (closest hit over the live queue) -> shade (emission, sky,
// sun next-event ray, cosine bounce, Russian roulette; survivors compacted
// into the next queue with one atomic per wave) -> shadow (any hit) -> apply.
// After the last bounce every pixel adds its spp_per_pass slot results to the
// film in a fixed order, with Kahan compensation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "rc_internal.hpp"

using namespace rcd;

namespace {

constexpr int kBlock = 256;
constexpr float kPi = 3.14159265358979323846f;
// Threads per block of the extension kernel (4 waves per queue reservation:
// 3% faster than 512 and 128 on the 2M-triangle terrain) and of the shadow
// kernel, unless a launch shape says otherwise (the batch autotuner).
#ifndef RC_EXT_BLOCK
#define RC_EXT_BLOCK 256
#endif
#ifndef RC_SHADOW_BLOCK
#define RC_SHADOW_BLOCK 128
#endif
constexpr uint32_t kExtBlockDefault = RC_EXT_BLOCK;
constexpr uint32_t kShadowBlockDefault = RC_SHADOW_BLOCK;
constexpr int kLdsRows = kLdsStack > 0 ? kLdsStack : 1;  // LDS stack entries per thread (at least one row)
// Minimum waves per SIMD the traversal kernels are compiled for, by triangle
// decision.  FP32: 8, which caps them at 64 VGPRs with no spills (59-62 used;
// 7 waves at 70 without the cap), 11% faster on the 2M-triangle terrain at
// 1080p x 64 spp (extension 59.5 ms against 66.9 per frame).  FP64: no
// minimum (88-92 VGPRs, 5 waves); at 8 it spills 28-39 VGPRs to scratch.
#ifndef RC_MIN_WAVES
#define RC_MIN_WAVES 8
#endif
// Park records per pass: 1 / RC_PARK_DIV of its paths (128 bytes each).
// Kernels per parked extension launch (the extension kernel and
// RC_PARK_PHASES - 1 resume kernels) and the iteration budget of the
// resume kernels that park again.
#ifndef RC_PARK_DIV
#define RC_PARK_DIV 4
#endif
#ifndef RC_PARK_PHASES
#define RC_PARK_PHASES 3
#endif
#ifndef RC_PARK_STEPS2
#define RC_PARK_STEPS2 20
#endif
static_assert(RC_PARK_PHASES >= 2, "a parking extension launch needs a resume kernel");
// Shadow rays (any hit) park after RC_SHADOW_PARK_STEPS iterations when
// parking is on (0: they never park), over RC_SHADOW_PARK_PHASES kernels.
#ifndef RC_SHADOW_PARK_STEPS
#define RC_SHADOW_PARK_STEPS 32
#endif
#ifndef RC_SHADOW_PARK_PHASES
#define RC_SHADOW_PARK_PHASES 3
#endif
static_assert(RC_SHADOW_PARK_PHASES >= 2, "a parking shadow launch needs a resume kernel");
#ifndef RC_MIN_WAVES_FP64
#define RC_MIN_WAVES_FP64 1
#endif
template <bool DEC64>
constexpr int kMinWaves = DEC64 ? RC_MIN_WAVES_FP64 : RC_MIN_WAVES;
// 1: each XCD takes a contiguous slice of the live queue entries (the grid is
// sized for the queue's capacity; only a prefix is live).  3% slower on the
// 2M-triangle terrain: the eight XCDs then sweep eight separate regions of the
// image, whose nodes compete for the shared Infinity Cache, while dealt
// round-robin they all sweep one band.  Off by default.
#ifndef RC_XCD_REMAP
#define RC_XCD_REMAP 0
#endif

template <int BLOCK>
__device__ __forceinline__ uint32_t trace_block(uint32_t n) {
    if (!RC_XCD_REMAP) return blockIdx.x;
    const uint32_t live = min((n + BLOCK - 1) / BLOCK, gridDim.x);  // blocks holding queue entries
    return blockIdx.x < live ? xcd_block(blockIdx.x, live) : blockIdx.x;
}

struct CamBasis {
    float3 pos, right, up, center;
    float ipw, iph, aspect;
};

// Camera::update / getRay basis in FP64 (EulerAngle.cpp:77-87, ONB.cpp:32-39,
// Camera.cpp:29-34,160,279-283), rounded to FP32 for the kernels.
CamBasis make_basis(const rc_camera& c, uint32_t w, uint32_t h) {
    const double d2r = M_PI / 180.0;
    double pitch = c.pitch, yaw = c.yaw;
    double fov = std::max(1.0, std::min(double(c.fov), 180.0));
    double f[3] = {-std::sin(yaw * d2r) * std::cos(pitch * d2r), std::sin(pitch * d2r),
                   -std::cos(yaw * d2r) * std::cos(pitch * d2r)};
    auto norm = [](double* v) {
        double inv = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) v[k] *= inv;
    };
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    norm(f);
    const double up0[3] = {0.0001, 1.0, 0.0001};  // Vector3::ALMOST_UP
    double r[3], u[3];
    cross(f, up0, r);
    norm(r);
    cross(r, f, u);
    norm(u);
    double dist = 0.5 / std::tan(fov / 2.0 * d2r);
    CamBasis b;
    b.pos = make_float3(c.position[0], c.position[1], c.position[2]);
    b.right = make_float3(float(r[0]), float(r[1]), float(r[2]));
    b.up = make_float3(float(u[0]), float(u[1]), float(u[2]));
    b.center = make_float3(float(c.position[0] + f[0] * dist), float(c.position[1] + f[1] * dist),
                           float(c.position[2] + f[2] * dist));
    b.ipw = float(w - 1);
    b.iph = float(h - 1);
    b.aspect = float(double(h) / double(w));
    return b;
}

__device__ __forceinline__ float3 cam_dir(const CamBasis& b, float px, float py) {
    float dx = (px / b.ipw) - 0.5f;
    float dy = (py / b.iph) - 0.5f;
    float3 p = b.center + (dx * b.right) + ((dy * b.aspect) * b.up);
    return normalize3(p - b.pos);
}

struct PassParams {
    CamBasis cam;
    PixelOrder order;
    uint32_t width, npix, spp_pass, sample0;
    uint32_t spp, sampler, jitter_n, filter;
    float radius, offset;
    float fa, fb;  // filter shape: gaussian alpha, beta; mitchell B, C
    uint32_t max_len, rr_start, estimator, frame;
    float3 sky, sun, sun_e;
    uint64_t seed;
    uint32_t has_sun;
    int stack_limit;  // on-chip traversal stack entries before the deep-stack kernel
};

// The pixel and sample a pass's queue slot traces.
__device__ __forceinline__ uint32_t slot_pixel(const PassParams& P, uint32_t slot) {
    return order_to_pixel(P.order, slot_index(slot, P.npix, P.spp_pass));
}
__device__ __forceinline__ uint32_t slot_sample(const PassParams& P, uint32_t slot) {
    return P.sample0 + slot_sample_in_pass(slot, P.npix, P.spp_pass);
}

// Filter::getWeightX of each FilterType (src/Rendering/Filters/*.cpp), in
// FP32; r is the filter's radius.
__device__ __forceinline__ float filter_1d(const PassParams& P, float x) {
    const float r = P.radius;
    switch (P.filter) {
    case RC_FILTER_BOX:  // BoxFilter.cpp:22-37
        return (x >= -r && x < r) ? 1.0f / (2.0f * r) : 0.0f;
    case RC_FILTER_TENT: {  // TentFilter.cpp:12-44
        float rinv = 1.0f / r;
        float s = fabsf(x * rinv);
        return (s < 1.0f ? 1.0f - s : 0.0f) * rinv;
    }
    case RC_FILTER_BELL: {  // BellFilter.cpp:12-22: the quadratic B-spline
        float s = fabsf(x);
        if (s < 0.5f) return 0.75f - s * s;
        if (s <= 1.5f) return 0.5f * (s - 1.5f) * (s - 1.5f);
        return 0.0f;
    }
    case RC_FILTER_GAUSSIAN:  // GaussianFilter.cpp:25-33: alpha exp(beta x^2)
        return P.fa * expf(x * x * P.fb);
    case RC_FILTER_MITCHELL: {  // MitchellFilter.cpp:12-23, Mitchell & Netravali's cubic
        const float B = P.fa, C = P.fb, s = fabsf(x), s2 = s * s, s3 = s2 * s;
        if (s <= 1.0f)
            return ((12.0f - 9.0f * B - 6.0f * C) * s3 + (-18.0f + 12.0f * B + 6.0f * C) * s2 + (6.0f - 2.0f * B)) *
                   (1.0f / 6.0f);
        if (s <= 2.0f)
            return ((-B - 6.0f * C) * s3 + (6.0f * B + 30.0f * C) * s2 + (-12.0f * B - 48.0f * C) * s +
                    (8.0f * B + 24.0f * C)) *
                   (1.0f / 6.0f);
        return 0.0f;
    }
    default: {  // LanczosSincFilter.cpp:12-30: sinc(s) sinc(s / r), sinc(x) = sin(pi x) / (pi x)
        float s = fabsf(x);
        if (s == 0.0f) return 1.0f;
        if (s > r) return 0.0f;
        const float q = s / r;
        return (sinpif(s) / (3.14159265358979f * s)) * (sinpif(q) / (3.14159265358979f * q));
    }
    }
}

// Tracer::generateMultiSamples (Tracer.cpp:115-144) for one slot: its camera
// ray (origin + tmin, direction + tmax, as rc_ray) and its filter weight.
__device__ __forceinline__ float camera_ray(const PassParams& P, uint32_t slot, float4& r0, float4& r1) {
    const uint32_t pixel = slot_pixel(P, slot), s = slot_sample(P, slot);
    float px = float(pixel % P.width), py = float(pixel / P.width);
    float w = 1.0f;
    if (!(P.sampler != RC_SAMPLER_RANDOM && P.spp == 1)) {
        float4 u = draw4(P.seed, pixel, s, P.frame, kCamDim);
        const uint32_t n = P.jitter_n, sy = P.sampler == RC_SAMPLER_RANDOM ? 0u : s / n, sx = s - sy * n;
        const uint64_t perm = P.sampler == RC_SAMPLER_CMJ ? pixel_permutation(P.seed, pixel, P.frame) : 0ull;
        const float2 pt = sample_point(P.sampler, sx, sy, n, perm, u.x, u.y);
        float ox = (pt.x - 0.5f) * 2.0f * P.radius;
        float oy = (pt.y - 0.5f) * 2.0f * P.radius;
        w = filter_1d(P, ox) * filter_1d(P, oy);
        px += ox;
        py += oy;
    }
    const float3 d = cam_dir(P.cam, px, py);
    r0 = make_float4(P.cam.pos.x, P.cam.pos.y, P.cam.pos.z, 0.0f);
    r1 = make_float4(d.x, d.y, d.z, __builtin_inff());
    return w;
}

// The pass's camera rays, path states, zeroed slot radiance and weights.
__global__ __launch_bounds__(kBlock) void k_camera(PassParams P, rc_ray* __restrict__ rays, float4* __restrict__ st,
                                                   float4* __restrict__ L, float* __restrict__ W,
                                                   DevCounters* __restrict__ ctr) {
    uint32_t n = P.npix * P.spp_pass;
    uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot == 0) {
        ctr->queue_count[0] = n;
        ctr->queue_count[1] = 0;
        ctr->queue_count[2] = 0;
    }
    if (slot >= n) return;
    float4 r0, r1;
    const float w = camera_ray(P, slot, r0, r1);
    float4* rp = reinterpret_cast<float4*>(rays + slot);
    rp[0] = r0;
    rp[1] = r1;
    st[slot] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(slot));
    st_stream(&L[slot], make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    W[slot] = w;
}

// 1: the camera bounce makes its camera rays in the extension kernel (no
// k_camera pass over the queues).
// RC_PACKET_CAMERA = 1: the camera bounce's waves (one pixel's samples each,
// with pixel-major slots) traverse as packets (traverse_packet); 0 (the
// default): each lane on its own, as the other bounces.  On the 2M-triangle
// terrain at 64 samples per pass the packets made the camera bounce 1.6 ms
// slower (12.5 against 14.1 ms): the union of a pixel's 64 traversals costs
// more visits than the per-lane traversal's leaf postponing and culling.
#ifndef RC_PACKET_CAMERA
#define RC_PACKET_CAMERA 0
#endif
#ifndef RC_FUSED_CAMERA
#define RC_FUSED_CAMERA 1
#endif
// Start of a pass whose camera bounce makes its own rays (RC_FUSED_CAMERA):
// the camera queue holds all n slots.
__global__ void k_pass_start(DevCounters* ctr, uint32_t n) {
    ctr->queue_count[0] = n;
    ctr->queue_count[1] = 0;
    ctr->queue_count[2] = 0;
}

// RC_PARK_KEEP > 0: a wave whose budget is spent keeps traversing while at
// least RC_PARK_KEEP of its lanes still are (up to twice the budget), so
// well-filled waves do not park; 0: every traversal parks at its budget.
#ifndef RC_PARK_KEEP
#define RC_PARK_KEEP 8
#endif
#ifndef RC_PARK_KEEP_ANY
#define RC_PARK_KEEP_ANY 0
#endif
// Whether a parking traversal (ANY: a shadow ray's) goes on after `steps`
// iterations of `budget`.
template <bool ANY>
__device__ __forceinline__ bool keep_going(bool done, uint32_t steps, uint32_t budget) {
    constexpr uint32_t keep = ANY ? RC_PARK_KEEP_ANY : RC_PARK_KEEP;
    if (keep == 0) return !done && steps < budget;
    const uint32_t busy = uint32_t(__popcll(__ballot(!done)));
    return !done && (steps < budget || (steps < 2 * budget && busy >= keep));
}

// Reserves `pred ? 1 : 0` entries of a queue: one atomic per wave, positions
// from the ballot prefix.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    uint64_t mask = __ballot(pred);
    uint64_t active = __ballot(1);
    uint32_t lane = __lane_id();
    uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t total = __popcll(mask);
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (int(lane) == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, leader, 64);
    return base + prefix;
}

// Reserves `pred ? 1 : 0` entries of a queue for a whole block with one
// atomic: per-wave ballots, the wave counts through LDS (scratch: BLOCK / 64
// + 1 words), positions in wave then lane order.  Every queue reservation is
// an atomic on one address; at one per 64-lane wave they serialized the
// extension kernel (the camera bounce took 37 ms instead of 16.5 at 132.7 M
// rays).  Every thread of the block calls it.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_append(bool pred, uint32_t* counter, uint32_t* scratch) {
    constexpr uint32_t kW = BLOCK / 64;
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(pred);
    if (lane == 0) scratch[wave] = uint32_t(__popcll(mask));
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kW; ++w) {
        const uint32_t c = scratch[w];
        before += w < wave ? c : 0u;
        total += c;
    }
    if (threadIdx.x == 0) scratch[kW] = total ? atomicAdd(counter, total) : 0u;
    __syncthreads();
    const uint32_t base = scratch[kW];
    __syncthreads();  // the scratch words may be reused by the next call
    return base + before + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// mapToHemisphere with distribution 1 (Sampler.cpp:50-64) in the ONB of a
// triangle without texcoords (Triangle.cpp:40-44: T = n x ALMOST_UP, B = T x n).
__device__ __forceinline__ float3 hemisphere(float3 n, float u1, float u2) {
    float3 T = normalize3(cross3(n, make_float3(0.0001f, 1.0f, 0.0001f)));
    float3 B = normalize3(cross3(T, n));
    float phi = 2.0f * kPi * u1;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float ct = sqrtf(1.0f - u2);
    float stt = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    return ((stt * cp) * T + (stt * sp) * B) + ct * n;
}

struct ShadeOut {
    bool alive, shadow;
    bool query;  // the path ends into the radiance cache: newo = position, newd = outgoing direction, rec = triangle
    bool train;  // a training vertex: tr holds the first four words of its record (emit_vertex writes it)
    float3 thr, newo, newd, contrib, so;
    uint32_t rec;  // the query's triangle, else kNoRecord
    float4 tr[4];
};
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;

// Training paths: a fixed-rate, frame-varying subset of the slots that run
// to the end instead of stopping into the cache.
__device__ __forceinline__ bool training_path(const PassParams& P, uint32_t slot, const CacheIO& CI) {
    const uint32_t stride = max(1u, CI.train_stride ? CI.train_stride : *CI.stride_dev);
    uint32_t pixel = slot_pixel(P, slot), sample = slot_sample(P, slot);
    uint32_t h = pixel * 2654435761u ^ (sample * 40503u + P.frame * 2246822519u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h % stride == 0u;
}

// One path vertex: PathTracer::traceRecursive (PathTracer.cpp:25-75) for the
// hit (t, tri) of ray (r0, r1), plus sky on a miss, sun next-event estimation
// and Russian roulette.  Radiance goes to L[slot]; the continuation and the
// shadow ray are returned for the caller to append.  CACHE false: the caller
// is the camera bounce, whose vertex is never the cache vertex.
template <bool CACHE = true>
__device__ __forceinline__ ShadeOut shade_vertex(const PassParams& P, uint32_t k, float4 r0, float4 r1, float4 s,
                                                 float t, int tri, const float4* __restrict__ shade,
                                                 const DevMaterial* __restrict__ mats, float4* __restrict__ L,
                                                 const CacheIO& CI) {
    ShadeOut O;
    O.alive = O.shadow = O.query = O.train = false;
    O.rec = kNoRecord;
    const bool more = (k + 1) < P.max_len;
    uint32_t slot = __float_as_uint(s.w);
    float3 thr = f3(s);
    if (tri < 0) {
        float4 l = ld_stream(&L[slot]);
        st_stream(&L[slot], make_float4(l.x + thr.x * P.sky.x, l.y + thr.y * P.sky.y, l.z + thr.z * P.sky.z, l.w));
        return O;
    }
    float4 sh = shade[tri];
    const DevMaterial& m = mats[__float_as_uint(sh.w)];
    float4 dif = m.diffuse;
    if (__float_as_uint(dif.w) & RC_MAT_EMISSIVE) {
        float4 e = m.emittance;
        float4 l = ld_stream(&L[slot]);
        st_stream(&L[slot], make_float4(l.x + thr.x * e.x, l.y + thr.y * e.y, l.z + thr.z * e.z, l.w));
        return O;
    }
    float3 nrm = f3(sh);
    float3 p = f3(r0) + (t * f3(r1));
    float3 rho = f3(dif);
    if (!more) return O;
    float4 u = draw4(P.seed, slot_pixel(P, slot), slot_sample(P, slot), P.frame, k);
    if (P.has_sun) {
        float cs = dot3(nrm, P.sun);
        if (cs > 0.0f) {
            float f = P.estimator == RC_EST_REFERENCE ? (2.0f / kPi) * cs * cs : (1.0f / kPi) * cs;
            O.contrib = mul3(thr, mul3(rho, P.sun_e)) * f;
            O.so = p + P.sun * P.offset;
            O.shadow = true;
        }
    }
    if (CACHE && CI.cache_vertex > 0 && k == CI.cache_vertex) {
        // The cache holds the indirect part of the radiance leaving this
        // vertex (the direct sun is the next-event ray above, for every path).
        const float3 wo = make_float3(-r1.x, -r1.y, -r1.z);
        if (!training_path(P, slot, CI)) {
            // the path ends here: its indirect radiance comes from the cache;
            // the caller appends the query (emit_vertex, one reservation per
            // block like the ray queues)
            O.query = true;
            O.newo = p;
            O.newd = wo;
            O.thr = thr;
            O.rec = uint32_t(tri);
            return O;
        }
        // training vertex: the radiance the path gathers after this vertex's
        // next-event ray, over the throughput arriving here, is the target;
        // emit_vertex reserves and writes the record (one reservation per
        // block), whose "before" value gets that ray's contribution added once
        // its occlusion is known (k_cache_nee_fix)
        O.train = true;
        O.tr[0] = make_float4(p.x, p.y, p.z, __uint_as_float(slot));
        O.tr[1] = make_float4(wo.x, wo.y, wo.z, thr.x);
        O.tr[2] = make_float4(nrm.x, nrm.y, nrm.z, thr.y);
        O.tr[3] = make_float4(rho.x, rho.y, rho.z, thr.z);
    }
    float3 w = hemisphere(nrm, u.x, u.y);
    if (P.estimator == RC_EST_REFERENCE) thr = mul3(thr, rho * (2.0f * fabsf(dot3(w, nrm))));
    else thr = mul3(thr, rho);
    O.alive = true;
    if (P.rr_start > 0 && k >= P.rr_start) {
        float q = fminf(1.0f, fmaxf(thr.x, fmaxf(thr.y, thr.z)));
        if (u.z >= q) O.alive = false;
        else thr = thr * (1.0f / q);
    }
    O.thr = thr;
    O.newo = p + w * P.offset;
    O.newd = w;
    return O;
}

// Appends the continuation and shadow ray of a shaded vertex (every lane of
// the wave calls this at the same point; with BLOCK > 0 every thread of the
// block, which reserves once per block, else once per wave).
template <int BLOCK, bool CACHE = true>
__device__ __forceinline__ void emit_vertex(const PassParams& P, uint32_t k, const ShadeOut& O, uint32_t slot,
                                            rc_ray* __restrict__ rays_out, float4* __restrict__ st_out,
                                            rc_ray* __restrict__ srays, float4* __restrict__ scontrib,
                                            DevCounters* __restrict__ ctr, uint32_t cout, const CacheIO& CI,
                                            const float4* __restrict__ shade, const DevMaterial* __restrict__ mats,
                                            const float4* __restrict__ L, uint32_t* append_scratch = nullptr) {
    uint32_t rec = kNoRecord;  // this vertex's training record
    if (CACHE && CI.cache_vertex > 0 && k == CI.cache_vertex) {  // uniform: paths end into the cache at this vertex
        const uint32_t r = BLOCK ? block_append<BLOCK>(O.train, &CI.counters[kCcRecords], append_scratch)
                                 : wave_append(O.train, &CI.counters[kCcRecords]);
        if (O.train && r < CI.cap) {
            const float4 lb = ld_stream(&L[slot]);
            float4* R = CI.records + 5 * size_t(r);
            R[0] = O.tr[0];
            R[1] = O.tr[1];
            R[2] = O.tr[2];
            R[3] = O.tr[3];
            R[4] = make_float4(lb.x, lb.y, lb.z, __uint_as_float(kNoRecord));
            rec = r;
        }
        const uint32_t qpos = BLOCK ? block_append<BLOCK>(O.query, &CI.counters[kCcQueries], append_scratch)
                                    : wave_append(O.query, &CI.counters[kCcQueries]);
        if (O.query && qpos < CI.cap) {
            const float4 sh = shade[O.rec];
            const float4 rho = mats[__float_as_uint(sh.w)].diffuse;
            float4* Q = CI.queries + 4 * size_t(qpos);
            st_stream(Q + 0, make_float4(O.newo.x, O.newo.y, O.newo.z, __uint_as_float(slot)));
            st_stream(Q + 1, make_float4(O.newd.x, O.newd.y, O.newd.z, O.thr.x));
            st_stream(Q + 2, make_float4(sh.x, sh.y, sh.z, O.thr.y));
            st_stream(Q + 3, make_float4(rho.x, rho.y, rho.z, O.thr.z));
        }
    }
    // the extension kernel reserves once per block, k_shade_listed once per wave
    uint32_t pos = BLOCK ? block_append<BLOCK>(O.alive, &ctr->queue_count[cout], append_scratch)
                         : wave_append(O.alive, &ctr->queue_count[cout]);
    if (O.alive) {
        float4* ro = reinterpret_cast<float4*>(rays_out + pos);
        st_stream(ro, make_float4(O.newo.x, O.newo.y, O.newo.z, 0.0f));
        st_stream(ro + 1, make_float4(O.newd.x, O.newd.y, O.newd.z, __builtin_inff()));
        st_stream(&st_out[pos], make_float4(O.thr.x, O.thr.y, O.thr.z, __uint_as_float(slot)));
    }
    uint32_t spos = BLOCK ? block_append<BLOCK>(O.shadow, &ctr->queue_count[2], append_scratch)
                          : wave_append(O.shadow, &ctr->queue_count[2]);
    if (O.shadow) {
        float4* ro = reinterpret_cast<float4*>(srays + spos);
        st_stream(ro, make_float4(O.so.x, O.so.y, O.so.z, 0.0f));
        st_stream(ro + 1, make_float4(P.sun.x, P.sun.y, P.sun.z, __builtin_inff()));
        st_stream(&scontrib[spos], make_float4(O.contrib.x, O.contrib.y, O.contrib.z, __uint_as_float(slot)));
        if (rec != kNoRecord) CI.records[5 * size_t(rec) + 4].w = __uint_as_float(spos);
    }
}

// Extend + shade: closest-hit traversal of queue entry i and, in the same
// thread, the shading of its hit.  A ray whose traversal stack outgrows LDS is
// listed for the deep-stack kernel and shaded afterwards by k_shade_listed.
// CAM: the camera bounce of a pass, whose entry i is camera slot i: the
// thread makes the ray itself (k_camera's work) instead of loading it, writes
// the slot's weight and zeroed radiance, and stores the ray and state only
// when it has to be listed for the deep-stack kernel (cam_rays, cam_st: the
// buffers `rays` and `st` name, written only in this mode).
// Blocks past the live entries return at once (the grid is sized for the
// queue's capacity; the test is uniform over the block, before any barrier).
// DEC64: the triangle decision mode; BLOCK: threads per block.
// Parking: an extension ray still traversing after `steps` loop iterations
// (the if-if loop's, one node visit and at most one leaf each) writes its
// traversal state to a park region and is continued by k_extend_resume,
// whose waves hold only such rays; a resume kernel may park its own long
// traversals again for the next one, the last runs them to the end.  A wave
// otherwise runs until its longest ray ends: in the diffuse bounces most rays
// leave for the sky after a few nodes while a few graze the terrain for many,
// and the waves' lanes idle.  Records beyond a region's `cap` are not
// written: those rays finish in place.
struct ParkIO {
    uint32_t* rec;   // kParkWords words per parked traversal
    uint32_t cap;    // records of the region
    uint32_t steps;  // budget of the kernel writing the region (0: no parking)
    uint32_t which;  // the region's counter: DevCounters::park_count[which]
};

template <bool COUNT, bool CAM, bool DEC64, int BLOCK, bool PARK>
__global__ __launch_bounds__(BLOCK, kMinWaves<DEC64>) void k_extend_shade(
    PassParams P, uint32_t k, const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
    const rc_ray* __restrict__ rays, const float4* __restrict__ st, const float4* __restrict__ shade,
    const DevMaterial* __restrict__ mats, rc_ray* __restrict__ rays_out, float4* __restrict__ st_out,
    rc_ray* __restrict__ srays, float4* __restrict__ scontrib, float4* __restrict__ L, uint32_t* __restrict__ ovf_list,
    DevCounters* __restrict__ ctr, uint32_t cin, uint32_t cout, CacheIO CI, float* __restrict__ W,
    rc_ray* __restrict__ cam_rays, float4* __restrict__ cam_st, ParkIO PK) {
    static_assert(!(PARK && CAM), "the camera bounce does not park");
    __shared__ uint32_t stack[kLdsRows * BLOCK];
    __shared__ uint32_t append_scratch[BLOCK / 64 + 1];
    __shared__ uint4 top[4 * kTopNodes];
    const uint32_t n = ctr->queue_count[cin];
    if (blockIdx.x * uint32_t(BLOCK) >= n) return;
    const uint32_t ntop = load_top(nodes, top);
    const uint32_t i = trace_block<BLOCK>(n) * BLOCK + threadIdx.x;
    if (i == 0) atomicAdd(&ctr->ext_rays, (unsigned long long)n);
    uint32_t n_rec = 0, n_tri = 0;
    ShadeOut O;
    O.alive = O.shadow = O.query = O.train = false;
    O.rec = kNoRecord;
    uint32_t slot = 0;
    if constexpr (PARK) {
        // the path state is read after the traversal, and the ray kept only
        // in S, so that fewer values live across the traversal loop
        Trav<DEC64> S;
        bool done = true;
        if (i < n) {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            const float4 r0 = ld_stream(rp), r1 = ld_stream(rp + 1);
            trav_begin(S, f3(r0), f3(r1), r0.w, r1.w, top, ntop);
            uint32_t steps = 0;
            do {
                done = trav_step<false, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, P.stack_limit,
                                                      n_rec, n_tri);
            } while (keep_going<false>(done, ++steps, PK.steps));
        }
        // every thread of the block reserves (one atomic per block)
        const uint32_t pos = block_append<BLOCK>(!done, &ctr->park_count[PK.which], append_scratch);
        bool parked = false;
        if (!done) {
            if (pos < PK.cap) {
                trav_park(S, i, stack + threadIdx.x, BLOCK, PK.rec + size_t(pos) * kParkWords);
                parked = true;
            } else {
                while (!trav_step<false, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, P.stack_limit,
                                                       n_rec, n_tri)) {
                }
            }
        }
        if (i < n && !parked) {
            if (S.R.overflow) {
                ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
            } else {
                const float4 s = ld_stream(&st[i]);
                slot = __float_as_uint(s.w);
                O = shade_vertex(P, k, make_float4(S.o.x, S.o.y, S.o.z, 0.0f), make_float4(S.d.x, S.d.y, S.d.z, 0.0f),
                                 s, S.R.t, S.R.id == kNoHit ? -1 : int(S.R.id), shade, mats, L, CI);
            }
        }
    } else if (i < n) {
        float4 r0, r1, s;
        if (CAM) {
            W[i] = camera_ray(P, i, r0, r1);
            s = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(i));
            st_stream(&L[i], make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        } else {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            r0 = ld_stream(rp);
            r1 = ld_stream(rp + 1);
            s = ld_stream(&st[i]);
        }
        slot = __float_as_uint(s.w);
        TraceResult R;
#if RC_QNODE && RC_PACKET_CAMERA
        if constexpr (CAM && !DEC64) {  // a wave holds one pixel's samples: one packet
            R = traverse_packet<COUNT>(nodes, tris, f3(r0), f3(r1), r0.w, r1.w, stack + threadIdx.x, BLOCK,
                                       min(P.stack_limit, kLdsRows), n_rec, n_tri);
        } else
#endif
        {
            R = traverse<false, COUNT, DEC64>(nodes, tris, f3(r0), f3(r1), r0.w, r1.w, stack + threadIdx.x, BLOCK,
                                              P.stack_limit, n_rec, n_tri, top, ntop);
        }
        if (R.overflow) {
            if (CAM) {  // the deep-stack kernel and k_shade_listed read the entry
                float4* rp = reinterpret_cast<float4*>(cam_rays + i);
                rp[0] = r0;
                rp[1] = r1;
                cam_st[i] = s;
            }
            ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
        } else {
            O = shade_vertex<!CAM>(P, k, r0, r1, s, R.t, R.id == kNoHit ? -1 : int(R.id), shade, mats, L, CI);
        }
    }
    emit_vertex<BLOCK, !CAM>(P, k, O, slot, rays_out, st_out, srays, scontrib, ctr, cout, CI, shade, mats, L,
                             append_scratch);
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

// The parked traversals of region `in`, continued and shaded as
// k_extend_shade would have (same rays, same steps, same hit), then their
// continuation and shadow rays appended; with PARK those still traversing
// after out.steps more iterations are parked again in region `out`.  One
// record per thread; the grid covers the region's capacity and blocks past
// the parked count return at once (uniform over the block, before any
// barrier).
template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
__global__ __launch_bounds__(BLOCK, kMinWaves<DEC64>) void k_extend_resume(
    PassParams P, uint32_t k, const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
    const rc_ray* __restrict__ rays, const float4* __restrict__ st, const float4* __restrict__ shade,
    const DevMaterial* __restrict__ mats, rc_ray* __restrict__ rays_out, float4* __restrict__ st_out,
    rc_ray* __restrict__ srays, float4* __restrict__ scontrib, float4* __restrict__ L, uint32_t* __restrict__ ovf_list,
    DevCounters* __restrict__ ctr, uint32_t cout, CacheIO CI, ParkIO in, ParkIO out) {
    __shared__ uint32_t stack[kLdsRows * BLOCK];
    __shared__ uint32_t append_scratch[BLOCK / 64 + 1];
    __shared__ uint4 top[4 * kTopNodes];
    const uint32_t n = min(ctr->park_count[in.which], in.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&ctr->park_total, (unsigned long long)n);
    if (blockIdx.x * uint32_t(BLOCK) >= n) return;
    const uint32_t ntop = load_top(nodes, top);
    const uint32_t q = blockIdx.x * uint32_t(BLOCK) + threadIdx.x;
    uint32_t n_rec = 0, n_tri = 0;
    ShadeOut O;
    O.alive = O.shadow = O.query = O.train = false;
    O.rec = kNoRecord;
    uint32_t slot = 0;
    Trav<DEC64> S;
    uint32_t i = 0;
    bool done = true;
    if (q < n) {
        const uint32_t* rec = in.rec + size_t(q) * kParkWords;
        i = rec[0];  // the queue entry (word 0)
        {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            const float4 r0 = ld_stream(rp), r1 = ld_stream(rp + 1);
            trav_begin(S, f3(r0), f3(r1), r0.w, r1.w, top, ntop);
        }
        trav_unpark(S, rec, stack + threadIdx.x, BLOCK);
        uint32_t steps = 0;
        do {
            done = trav_step<false, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, P.stack_limit, n_rec,
                                                  n_tri);
        } while (PARK ? keep_going<false>(done, ++steps, out.steps) : !done);
    }
    bool parked = false;
    if constexpr (PARK) {
        const uint32_t pos = block_append<BLOCK>(!done, &ctr->park_count[out.which], append_scratch);
        if (!done) {
            if (pos < out.cap) {
                trav_park(S, i, stack + threadIdx.x, BLOCK, out.rec + size_t(pos) * kParkWords);
                parked = true;
            } else {
                while (!trav_step<false, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, P.stack_limit,
                                                       n_rec, n_tri)) {
                }
            }
        }
    }
    if (q < n && !parked) {
        if (S.R.overflow) {
            ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
        } else {
            const float4 s = ld_stream(&st[i]);
            slot = __float_as_uint(s.w);
            O = shade_vertex(P, k, make_float4(S.o.x, S.o.y, S.o.z, 0.0f), make_float4(S.d.x, S.d.y, S.d.z, 0.0f), s,
                             S.R.t, S.R.id == kNoHit ? -1 : int(S.R.id), shade, mats, L, CI);
        }
    }
    emit_vertex<BLOCK>(P, k, O, slot, rays_out, st_out, srays, scontrib, ctr, cout, CI, shade, mats, L, append_scratch);
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

__global__ void k_zero_park(DevCounters* ctr, uint32_t which) { ctr->park_count[which] = 0; }

// Shading of the listed queue entries (rays the deep-stack kernel traced).
__global__ __launch_bounds__(kBlock) void k_shade_listed(PassParams P, uint32_t k, const rc_ray* __restrict__ rays,
                                                         const float4* __restrict__ st, const rc_hit* __restrict__ hits,
                                                         const float4* __restrict__ shade,
                                                         const DevMaterial* __restrict__ mats,
                                                         rc_ray* __restrict__ rays_out, float4* __restrict__ st_out,
                                                         rc_ray* __restrict__ srays, float4* __restrict__ scontrib,
                                                         float4* __restrict__ L, const uint32_t* __restrict__ list,
                                                         DevCounters* __restrict__ ctr, uint32_t cout, CacheIO CI) {
    const uint32_t n = ctr->ovf_count;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        uint32_t q = base + threadIdx.x;
        ShadeOut O;
        O.alive = O.shadow = O.query = O.train = false;
        O.rec = kNoRecord;
        uint32_t slot = 0;
        if (q < n) {
            uint32_t i = list[q];
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            float4 s = ld_stream(&st[i]);
            float4 h = reinterpret_cast<const float4*>(hits)[i];
            slot = __float_as_uint(s.w);
            O = shade_vertex(P, k, rp[0], rp[1], s, h.x, __float_as_int(h.w), shade, mats, L, CI);
        }
        emit_vertex<0>(P, k, O, slot, rays_out, st_out, srays, scontrib, ctr, cout, CI, shade, mats, L);
    }
}

// The result of a finished shadow ray: its occlusion flag (when kept) and,
// unoccluded, its contribution added to its slot.
__device__ __forceinline__ void shadow_result(uint32_t i, bool occluded, const float4* __restrict__ scontrib,
                                              float4* __restrict__ L, uint8_t* __restrict__ occ_out) {
    if (occ_out) occ_out[i] = occluded ? 1 : 0;
    if (!occluded) {
        float4 c = ld_stream(&scontrib[i]);
        uint32_t slot = __float_as_uint(c.w);
        float4 l = ld_stream(&L[slot]);
        st_stream(&L[slot], make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w));
    }
}

// Any hit of a shadow ray; an unoccluded ray adds its contribution to its
// slot.  Overflowing rays are listed for the deep-stack kernel and applied by
// k_shadow_apply_listed.  PARK: as the extension kernel, a ray still
// traversing after PK.steps iterations is parked for k_shadow_resume.
template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
__global__ __launch_bounds__(BLOCK, kMinWaves<DEC64>) void k_shadow(const NodeT* __restrict__ nodes,
                                                       const TriRec* __restrict__ tris, const rc_ray* __restrict__ srays,
                                                       const float4* __restrict__ scontrib, float4* __restrict__ L,
                                                       uint32_t* __restrict__ ovf_list, DevCounters* __restrict__ ctr,
                                                       uint8_t* __restrict__ occ_out, int stack_limit, ParkIO PK) {
    __shared__ uint32_t stack[kLdsRows * BLOCK];
    __shared__ uint4 top[4 * kTopNodes];
    const uint32_t n = ctr->queue_count[2];
    if (blockIdx.x * uint32_t(BLOCK) >= n) return;
    const uint32_t ntop = load_top(nodes, top);
    const uint32_t i = trace_block<BLOCK>(n) * BLOCK + threadIdx.x;
    if (i == 0) atomicAdd(&ctr->shadow_rays, (unsigned long long)n);
    uint32_t n_rec = 0, n_tri = 0;
    if constexpr (PARK) {
        __shared__ uint32_t append_scratch[BLOCK / 64 + 1];
        Trav<DEC64> S;
        bool done = true;
        if (i < n) {
            const float4* rp = reinterpret_cast<const float4*>(srays + i);
            const float4 r0 = ld_stream(rp), r1 = ld_stream(rp + 1);
            trav_begin(S, f3(r0), f3(r1), r0.w, r1.w, top, ntop);
            uint32_t steps = 0;
            do {
                done = trav_step<true, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, stack_limit, n_rec,
                                                     n_tri);
            } while (keep_going<true>(done, ++steps, PK.steps));
        }
        const uint32_t pos = block_append<BLOCK>(!done, &ctr->park_count[PK.which], append_scratch);
        bool parked = false;
        if (!done) {
            if (pos < PK.cap) {
                trav_park(S, i, stack + threadIdx.x, BLOCK, PK.rec + size_t(pos) * kParkWords);
                parked = true;
            } else {
                while (!trav_step<true, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, stack_limit, n_rec,
                                                      n_tri)) {
                }
            }
        }
        if (i < n && !parked) {
            if (S.R.overflow) ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
            else shadow_result(i, S.R.id != kNoHit, scontrib, L, occ_out);
        }
    } else if (i < n) {
        const float4* rp = reinterpret_cast<const float4*>(srays + i);
        float4 r0 = ld_stream(rp), r1 = ld_stream(rp + 1);
        TraceResult R = traverse<true, COUNT, DEC64>(nodes, tris, f3(r0), f3(r1), r0.w, r1.w, stack + threadIdx.x,
                                                     BLOCK, stack_limit, n_rec, n_tri, top, ntop);
        if (R.overflow) ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
        else shadow_result(i, R.id != kNoHit, scontrib, L, occ_out);
    }
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

// The parked shadow traversals of region `in`, continued (and with PARK
// parked again in region `out` after out.steps iterations) and applied.
template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
__global__ __launch_bounds__(BLOCK, kMinWaves<DEC64>) void k_shadow_resume(
    const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris, const rc_ray* __restrict__ srays,
    const float4* __restrict__ scontrib, float4* __restrict__ L, uint32_t* __restrict__ ovf_list,
    DevCounters* __restrict__ ctr, uint8_t* __restrict__ occ_out, int stack_limit, ParkIO in, ParkIO out) {
    __shared__ uint32_t stack[kLdsRows * BLOCK];
    __shared__ uint32_t append_scratch[BLOCK / 64 + 1];
    __shared__ uint4 top[4 * kTopNodes];
    const uint32_t n = min(ctr->park_count[in.which], in.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&ctr->park_total, (unsigned long long)n);
    if (blockIdx.x * uint32_t(BLOCK) >= n) return;
    const uint32_t ntop = load_top(nodes, top);
    const uint32_t q = blockIdx.x * uint32_t(BLOCK) + threadIdx.x;
    uint32_t n_rec = 0, n_tri = 0;
    Trav<DEC64> S;
    uint32_t i = 0;
    bool done = true;
    if (q < n) {
        const uint32_t* rec = in.rec + size_t(q) * kParkWords;
        i = rec[0];
        {
            const float4* rp = reinterpret_cast<const float4*>(srays + i);
            const float4 r0 = ld_stream(rp), r1 = ld_stream(rp + 1);
            trav_begin(S, f3(r0), f3(r1), r0.w, r1.w, top, ntop);
        }
        trav_unpark(S, rec, stack + threadIdx.x, BLOCK);
        uint32_t steps = 0;
        do {
            done = trav_step<true, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, stack_limit, n_rec, n_tri);
        } while (PARK ? keep_going<true>(done, ++steps, out.steps) : !done);
    }
    bool parked = false;
    if constexpr (PARK) {
        const uint32_t pos = block_append<BLOCK>(!done, &ctr->park_count[out.which], append_scratch);
        if (!done) {
            if (pos < out.cap) {
                trav_park(S, i, stack + threadIdx.x, BLOCK, out.rec + size_t(pos) * kParkWords);
                parked = true;
            } else {
                while (!trav_step<true, COUNT, DEC64>(S, nodes, tris, stack + threadIdx.x, BLOCK, stack_limit, n_rec,
                                                      n_tri)) {
                }
            }
        }
    }
    if (q < n && !parked) {
        if (S.R.overflow) ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
        else shadow_result(i, S.R.id != kNoHit, scontrib, L, occ_out);
    }
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

__global__ __launch_bounds__(kBlock) void k_shadow_apply_listed(const float4* __restrict__ scontrib,
                                                                const uint8_t* __restrict__ occ,
                                                                const uint32_t* __restrict__ list,
                                                                float4* __restrict__ L, DevCounters* __restrict__ ctr) {
    const uint32_t n = ctr->ovf_count;
    for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
        uint32_t i = list[q];
        if (occ[i]) continue;
        float4 c = ld_stream(&scontrib[i]);
        uint32_t slot = __float_as_uint(c.w);
        float4 l = ld_stream(&L[slot]);
        st_stream(&L[slot], make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w));
    }
}

// Training samples of a pass: query + target (L_final - L_before) / T for the
// recorded training vertices, appended to the frame's batch (capped); also
// adds the pass's query count to the frame total (kCcFrameQueries).
__global__ __launch_bounds__(kBlock) void k_cache_targets(const float4* __restrict__ records,
                                                          uint32_t* __restrict__ counters, uint32_t pass_cap,
                                                          const float4* __restrict__ L, float* __restrict__ batch_q,
                                                          float* __restrict__ batch_t, uint32_t batch_cap) {
    const uint32_t n = min(counters[kCcRecords], pass_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[kCcFrameQueries] += min(counters[kCcQueries], pass_cap);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float4* R = records + 5 * size_t(i);
        float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3], r4 = R[4];
        float4 lf = L[__float_as_uint(r0.w)];
        float T[3] = {r1.w, r2.w, r3.w};
        float d[3] = {lf.x - r4.x, lf.y - r4.y, lf.z - r4.z};
        uint32_t pos = atomicAdd(&counters[kCcSamples], 1u);
        if (pos >= batch_cap) continue;
        float* q = batch_q + size_t(pos) * 12;
        q[0] = r0.x; q[1] = r0.y; q[2] = r0.z;
        q[3] = r1.x; q[4] = r1.y; q[5] = r1.z;
        q[6] = r2.x; q[7] = r2.y; q[8] = r2.z;
        q[9] = r3.x; q[10] = r3.y; q[11] = r3.z;
        for (int c = 0; c < 3; ++c) batch_t[size_t(pos) * 3 + c] = T[c] > 1e-8f ? fmaxf(d[c], 0.0f) / T[c] : 0.0f;
    }
}

// Adds to each training record of the pass its vertex's next-event
// contribution when that shadow ray was unoccluded, so "before" covers the
// direct light at the cache vertex.
__global__ __launch_bounds__(kBlock) void k_cache_nee_fix(float4* __restrict__ records,
                                                          const uint32_t* __restrict__ counters, uint32_t pass_cap,
                                                          const float4* __restrict__ scontrib,
                                                          const uint8_t* __restrict__ occ) {
    const uint32_t n = min(counters[kCcRecords], pass_cap);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float4 r4 = records[5 * size_t(i) + 4];
        const uint32_t spos = __float_as_uint(r4.w);
        if (spos == kNoRecord || occ[spos]) continue;
        const float4 c = ld_stream(&scontrib[spos]);
        records[5 * size_t(i) + 4] = make_float4(r4.x + c.x, r4.y + c.y, r4.z + c.z, r4.w);
    }
}

__global__ void k_cache_reset(uint32_t* counters, uint32_t which) {
    counters[kCcQueries] = 0;
    counters[kCcRecords] = 0;
    if (which) counters[kCcSamples] = 0;
}

// Zeroes the queue the next shade writes, the shadow queue and the overflow list.
__global__ void k_prep(DevCounters* ctr, uint32_t cout) {
    ctr->queue_count[cout] = 0;
    ctr->queue_count[2] = 0;
    ctr->ovf_count = 0;
    ctr->park_count[0] = 0;
    ctr->park_count[1] = 0;
}

__global__ void k_reset_ovf(DevCounters* ctr) { ctr->ovf_count = 0; }
// Before a shadow pass: the overflow list and both park regions are free.
__global__ void k_reset_shadow(DevCounters* ctr) {
    ctr->ovf_count = 0;
    ctr->park_count[0] = 0;
    ctr->park_count[1] = 0;
}

__device__ __forceinline__ void kahan(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Film::addSample (Film.cpp:54-59) for the pass's samples of each pixel, in
// slot order, with FP32 Kahan compensation of the running sums.
__global__ __launch_bounds__(kBlock) void k_film_add(PixelOrder order, uint32_t npix, uint32_t spp_pass,
                                                     const float4* __restrict__ L, const float* __restrict__ W,
                                                     float4* __restrict__ sum, float4* __restrict__ comp) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    const uint32_t q = pixel_to_order(order, p);
    float4 s = sum[p], c = comp[p];
    for (uint32_t j = 0; j < spp_pass; ++j) {
        uint32_t slot = RC_PIXEL_MAJOR ? q * spp_pass + j : j * npix + q;
        float w = W[slot];
        float4 l = ld_stream(&L[slot]);
        kahan(s.x, c.x, w * l.x);
        kahan(s.y, c.y, w * l.y);
        kahan(s.z, c.z, w * l.z);
        kahan(s.w, c.w, w);
    }
    sum[p] = s;
    comp[p] = c;
}

// The same for pixel-major slots: a pixel's samples are adjacent, so a block
// of kFilmBlock pixels (consecutive in tile order) owns one contiguous run of
// slots, staged through LDS kFilmChunk samples at a time so the loads stay
// coalesced; each pixel still sums its samples in slot order.
constexpr uint32_t kFilmBlock = 256, kFilmChunk = 8;
__global__ __launch_bounds__(kFilmBlock) void k_film_add_pm(PixelOrder order, uint32_t npix, uint32_t spp_pass,
                                                            const float4* __restrict__ L,
                                                            const float* __restrict__ W, float4* __restrict__ sum,
                                                            float4* __restrict__ comp) {
    __shared__ float4 sl[kFilmBlock * kFilmChunk];
    __shared__ float sw[kFilmBlock * kFilmChunk];
    const uint32_t q0 = blockIdx.x * kFilmBlock, q = q0 + threadIdx.x;
    const uint32_t nq = min(kFilmBlock, npix - q0);
    const bool own = q < npix;
    uint32_t p = 0;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c = s;
    if (own) {
        p = order_to_pixel(order, q);
        s = sum[p];
        c = comp[p];
    }
    for (uint32_t j0 = 0; j0 < spp_pass; j0 += kFilmChunk) {
        const uint32_t ch = min(kFilmChunk, spp_pass - j0);
        for (uint32_t e = threadIdx.x; e < nq * ch; e += kFilmBlock) {
            const uint32_t i = e / ch, k = e - i * ch;
            const size_t slot = size_t(q0 + i) * spp_pass + j0 + k;
            sl[i * kFilmChunk + k] = ld_stream(&L[slot]);
            sw[i * kFilmChunk + k] = W[slot];
        }
        __syncthreads();
        if (own) {
            for (uint32_t k = 0; k < ch; ++k) {
                const float w = sw[threadIdx.x * kFilmChunk + k];
                const float4 l = sl[threadIdx.x * kFilmChunk + k];
                kahan(s.x, c.x, w * l.x);
                kahan(s.y, c.y, w * l.y);
                kahan(s.z, c.z, w * l.z);
                kahan(s.w, c.w, w);
            }
        }
        __syncthreads();
    }
    if (own) {
        sum[p] = s;
        comp[p] = c;
    }
}

// Film accumulation of one pass in the slot layout the kernels use.
void launch_film_add(hipStream_t st, PixelOrder order, uint32_t npix, uint32_t spp_pass, const float4* L,
                     const float* W, float4* sum, float4* comp) {
    if (RC_PIXEL_MAJOR)
        k_film_add_pm<<<(npix + kFilmBlock - 1) / kFilmBlock, kFilmBlock, 0, st>>>(order, npix, spp_pass, L, W, sum,
                                                                                    comp);
    else
        k_film_add<<<(npix + kBlock - 1) / kBlock, kBlock, 0, st>>>(order, npix, spp_pass, L, W, sum, comp);
}

// Launches of the extension and shadow kernels in every template variant:
// counting (default block only), camera bounce, decision mode, block size.
struct ExtendArgs {
    PassParams P;
    uint32_t k;
    const rc_scene* s;
    rc_ray* rin;
    float4* sin;
    rc_ray* rout;
    float4* sout;
    rc_ray* srays;
    float4* scontrib;
    float4* L;
    uint32_t* ovf;
    DevCounters* ctr;
    uint32_t cin, cout;
    CacheIO CI;
    float* W;
    ParkIO PK;  // PK.steps > 0: the non-camera bounces park long traversals
};

template <bool COUNT, bool CAM, bool DEC64, int BLOCK, bool PARK>
void extend_t(const ExtendArgs& a, uint32_t n, hipStream_t st) {
    k_extend_shade<COUNT, CAM, DEC64, BLOCK, PARK><<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(
        a.P, a.k, a.s->nodes, tri_array<DEC64>(a.s), a.rin, a.sin, a.s->shade, a.s->mats, a.rout, a.sout, a.srays, a.scontrib, a.L,
        a.ovf, a.ctr, a.cin, a.cout, a.CI, a.W, a.rin, a.sin, a.PK);
}

template <bool COUNT, bool CAM, bool DEC64>
void extend_b(const ExtendArgs& a, uint32_t n, hipStream_t st, uint32_t block) {
    if constexpr (!CAM) {
        if (a.PK.steps) {
            if constexpr (!COUNT) {
                if (block == 128) return extend_t<COUNT, CAM, DEC64, 128, true>(a, n, st);
            }
            return extend_t<COUNT, CAM, DEC64, 256, true>(a, n, st);
        }
    }
    if constexpr (!COUNT) {
        if (block == 128) return extend_t<COUNT, CAM, DEC64, 128, false>(a, n, st);
    }
    extend_t<COUNT, CAM, DEC64, 256, false>(a, n, st);
}

void launch_extend(const ExtendArgs& a, uint32_t n, hipStream_t st, bool count, bool cam, bool dec64, uint32_t block) {
#define RC_EXT(C, CM, D) extend_b<C, CM, D>(a, n, st, block)
    if (count) {
        if (cam) { if (dec64) RC_EXT(true, true, true); else RC_EXT(true, true, false); }
        else { if (dec64) RC_EXT(true, false, true); else RC_EXT(true, false, false); }
    } else {
        if (cam) { if (dec64) RC_EXT(false, true, true); else RC_EXT(false, true, false); }
        else { if (dec64) RC_EXT(false, false, true); else RC_EXT(false, false, false); }
    }
#undef RC_EXT
}

// Parked traversals: a grid over the input region's capacity (blocks past
// the parked count return at once).
template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
void resume_t(const ExtendArgs& a, const ParkIO& in, const ParkIO& out, hipStream_t st) {
    k_extend_resume<COUNT, DEC64, BLOCK, PARK><<<std::max(1u, (in.cap + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(
        a.P, a.k, a.s->nodes, tri_array<DEC64>(a.s), a.rin, a.sin, a.s->shade, a.s->mats, a.rout, a.sout, a.srays, a.scontrib, a.L,
        a.ovf, a.ctr, a.cout, a.CI, in, out);
}

template <bool COUNT, bool DEC64, bool PARK>
void resume_b(const ExtendArgs& a, const ParkIO& in, const ParkIO& out, hipStream_t st, uint32_t block) {
    if constexpr (!COUNT) {
        if (block == 128) return resume_t<COUNT, DEC64, 128, PARK>(a, in, out, st);
    }
    resume_t<COUNT, DEC64, 256, PARK>(a, in, out, st);
}

// The resume kernels of an extension launch that parked into region 0:
// `phases` - 1 of them, alternating between the two regions; all but the
// last park again after `steps2` iterations.
void launch_resumes(const ExtendArgs& a, const ParkIO regions[2], uint32_t phases, uint32_t steps2, hipStream_t st,
                    bool count, bool dec64, uint32_t block) {
    for (uint32_t ph = 1; ph < phases; ++ph) {
        ParkIO in = regions[(ph - 1) & 1u], out = regions[ph & 1u];
        const bool park = ph + 1 < phases;
        out.steps = park ? steps2 : 0u;
        if (park && ph > 1) k_zero_park<<<1, 1, 0, st>>>(a.ctr, out.which);  // its previous contents are consumed
#define RC_RES(C, D) (park ? resume_b<C, D, true>(a, in, out, st, block) : resume_b<C, D, false>(a, in, out, st, block))
        if (count) { if (dec64) RC_RES(true, true); else RC_RES(true, false); }
        else { if (dec64) RC_RES(false, true); else RC_RES(false, false); }
#undef RC_RES
    }
}

struct ShadowArgs {
    const rc_scene* s;
    const rc_ray* srays;
    const float4* scontrib;
    float4* L;
    uint32_t* ovf;
    DevCounters* ctr;
    uint8_t* occ_out;
    int stack_limit;
    ParkIO PK;  // PK.steps > 0: long traversals are parked
};

template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
void shadow_t(const ShadowArgs& a, uint32_t n, hipStream_t st) {
    k_shadow<COUNT, DEC64, BLOCK, PARK><<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(
        a.s->nodes, tri_array<DEC64>(a.s), a.srays, a.scontrib, a.L, a.ovf, a.ctr, a.occ_out, a.stack_limit, a.PK);
}

template <bool COUNT, bool DEC64, bool PARK>
void shadow_b(const ShadowArgs& a, uint32_t n, hipStream_t st, uint32_t block) {
    if constexpr (!COUNT) {
        if (block == 64) return shadow_t<COUNT, DEC64, 64, PARK>(a, n, st);
        if (block == 256) return shadow_t<COUNT, DEC64, 256, PARK>(a, n, st);
    }
    shadow_t<COUNT, DEC64, 128, PARK>(a, n, st);
}

void launch_shadow(const ShadowArgs& a, uint32_t n, hipStream_t st, bool count, bool dec64, uint32_t block) {
#define RC_SHD(C, D) (a.PK.steps ? shadow_b<C, D, true>(a, n, st, block) : shadow_b<C, D, false>(a, n, st, block))
    if (count) { if (dec64) RC_SHD(true, true); else RC_SHD(true, false); }
    else { if (dec64) RC_SHD(false, true); else RC_SHD(false, false); }
#undef RC_SHD
}

template <bool COUNT, bool DEC64, int BLOCK, bool PARK>
void shadow_resume_t(const ShadowArgs& a, const ParkIO& in, const ParkIO& out, hipStream_t st) {
    k_shadow_resume<COUNT, DEC64, BLOCK, PARK><<<std::max(1u, (in.cap + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(
        a.s->nodes, tri_array<DEC64>(a.s), a.srays, a.scontrib, a.L, a.ovf, a.ctr, a.occ_out, a.stack_limit, in, out);
}

template <bool COUNT, bool DEC64, bool PARK>
void shadow_resume_b(const ShadowArgs& a, const ParkIO& in, const ParkIO& out, hipStream_t st, uint32_t block) {
    if constexpr (!COUNT) {
        if (block == 64) return shadow_resume_t<COUNT, DEC64, 64, PARK>(a, in, out, st);
        if (block == 256) return shadow_resume_t<COUNT, DEC64, 256, PARK>(a, in, out, st);
    }
    shadow_resume_t<COUNT, DEC64, 128, PARK>(a, in, out, st);
}

// The resume kernels of a shadow launch that parked into region 0, as
// launch_resumes for the extension kernel.
void launch_shadow_resumes(const ShadowArgs& a, const ParkIO regions[2], uint32_t phases, uint32_t steps2,
                           hipStream_t st, bool count, bool dec64, uint32_t block) {
    for (uint32_t ph = 1; ph < phases; ++ph) {
        ParkIO in = regions[(ph - 1) & 1u], out = regions[ph & 1u];
        const bool park = ph + 1 < phases;
        out.steps = park ? steps2 : 0u;
        if (park && ph > 1) k_zero_park<<<1, 1, 0, st>>>(a.ctr, out.which);
#define RC_SRES(C, D) \
    (park ? shadow_resume_b<C, D, true>(a, in, out, st, block) : shadow_resume_b<C, D, false>(a, in, out, st, block))
        if (count) { if (dec64) RC_SRES(true, true); else RC_SRES(true, false); }
        else { if (dec64) RC_SRES(false, true); else RC_SRES(false, false); }
#undef RC_SRES
    }
}

uint32_t isqrt_exact(uint32_t v) {
    uint32_t r = uint32_t(std::sqrt(double(v)));
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

__global__ void k_zero_counters(DevCounters* ctr) {
    ctr->node_records = 0;
    ctr->tri_tests = 0;
    ctr->ext_rays = 0;
    ctr->shadow_rays = 0;
    ctr->ovf_total = 0;
    ctr->park_total = 0;
    ctr->ovf_count = 0;
    ctr->error = 0;
}

}  // namespace

void rc_zero_counters(DevCounters* ctr, hipStream_t st) { k_zero_counters<<<1, 1, 0, st>>>(ctr); }

int FrameBuffers::ensure(size_t npaths) {
    if (!counters) {
        RC_TRY_HIP(hipMalloc(&counters, sizeof(DevCounters)));
        own_counters = true;
        RC_TRY_HIP(hipMemset(counters, 0, sizeof(DevCounters)));
    }
    RC_TRY(ray_a.ensure(npaths * sizeof(rc_ray)));
    RC_TRY(ray_b.ensure(npaths * sizeof(rc_ray)));
    RC_TRY(st_a.ensure(npaths * sizeof(float4)));
    RC_TRY(st_b.ensure(npaths * sizeof(float4)));
    RC_TRY(hits.ensure(npaths * sizeof(rc_hit)));
    RC_TRY(srays.ensure(npaths * sizeof(rc_ray)));
    RC_TRY(scontrib.ensure(npaths * sizeof(float4)));
    RC_TRY(socc.ensure(npaths));
    RC_TRY(slotL.ensure(npaths * sizeof(float4)));
    RC_TRY(slotW.ensure(npaths * sizeof(float)));
    RC_TRY(ovf_list.ensure(npaths * sizeof(uint32_t)));
    RC_TRY(deep_stack.ensure(rc_deep_stack_bytes()));
    // parked traversals: 1 / RC_PARK_DIV of the pass's paths at most
    RC_TRY(park.ensure(std::max<size_t>(npaths / RC_PARK_DIV, 256) * kParkWords * sizeof(uint32_t)));
    return RC_OK;
}

void FrameBuffers::release() {
    DevBuf* b[] = {&ray_a, &ray_b, &st_a, &st_b, &hits, &srays, &scontrib, &socc, &slotL, &slotW, &ovf_list,
                   &deep_stack, &park};
    for (DevBuf* x : b) x->release();
    if (own_counters && counters) (void)hipFree(counters);
    if (own_counters) counters = nullptr;
    own_counters = false;
}

// The filter's radius and shape constants, in double as the reference's
// constructors compute them (GaussianFilter.cpp:15-23), rounded to float.
static void filter_shape(const rc_render_cfg& c, float& radius, float& a, float& b) {
    a = b = 0.0f;
    switch (c.filter) {
    case RC_FILTER_BELL: radius = 1.5f; break;
    case RC_FILTER_GAUSSIAN: {
        const double sd = c.filter_param[0];
        a = float(1.0 / (std::sqrt(2.0 * M_PI) * sd));
        b = float(-1.0 / (2.0 * sd * sd));
        radius = float((7.43384 * sd) / 2.0);
        break;
    }
    case RC_FILTER_MITCHELL:
        radius = 2.0f;
        a = c.filter_param[0];
        b = c.filter_param[1];
        break;
    default: radius = c.filter_radius; break;
    }
}

int rc_validate_cfg(const rc_render_cfg* cfg, const char* fn) {
    if (!cfg) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null config");
    if (cfg->width < 2 || cfg->height < 2 || cfg->width > 16384 || cfg->height > 16384)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": film size out of range");
    if (cfg->spp == 0 || cfg->spp > 65536) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": spp out of range");
    if (cfg->sampler > RC_SAMPLER_POISSON_DISC) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": unknown sampler");
    if (cfg->sampler != RC_SAMPLER_RANDOM) {
        uint32_t n = isqrt_exact(cfg->spp);
        if (n * n != cfg->spp)
            return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": this sampler needs spp = n*n");
    }
    if (cfg->max_path_length == 0 || cfg->max_path_length > 64)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": max_path_length out of range");
    if (cfg->filter > RC_FILTER_LANCZOS_SINC) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": unknown filter");
    if ((cfg->filter == RC_FILTER_BOX || cfg->filter == RC_FILTER_TENT) && !(cfg->filter_radius > 0.0f && cfg->filter_radius < 1e6f))
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": filter_radius must be positive");
    if (cfg->filter == RC_FILTER_LANCZOS_SINC &&
        !(cfg->filter_radius >= 1.0f && cfg->filter_radius <= 64.0f && cfg->filter_radius == std::floor(cfg->filter_radius)))
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": lanczos_sinc needs an integer filter_radius in [1, 64]");
    if (cfg->filter == RC_FILTER_GAUSSIAN && !(cfg->filter_param[0] > 0.0f && cfg->filter_param[0] < 1e6f))
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": gaussian needs a positive standard deviation");
    if (cfg->filter == RC_FILTER_MITCHELL && !(std::isfinite(cfg->filter_param[0]) && std::isfinite(cfg->filter_param[1])))
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": mitchell needs finite B and C");
    if (cfg->estimator > RC_EST_LAMBERT) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": bad estimator");
    return RC_OK;
}

size_t rc_pass_paths(const rc_render_cfg* cfg, uint32_t spp_pass) {
    spp_pass = std::max(1u, std::min(spp_pass, cfg->spp));
    return size_t(cfg->width) * cfg->height * spp_pass;
}

// The parameters of a frame's passes (the pass fields are set per pass).
static PassParams base_params(const rc_camera* cam, const rc_render_cfg* cfg) {
    PassParams P;
    std::memset(&P, 0, sizeof(P));
    P.cam = make_basis(*cam, cfg->width, cfg->height);
    P.order = make_order(cfg->width, cfg->height);
    P.width = cfg->width;
    P.npix = cfg->width * cfg->height;
    P.spp = cfg->spp;
    P.sampler = cfg->sampler;
    P.jitter_n = isqrt_exact(cfg->spp);
    P.filter = cfg->filter;
    filter_shape(*cfg, P.radius, P.fa, P.fb);
    P.offset = cfg->ray_start_offset;
    P.max_len = cfg->max_path_length;
    P.rr_start = cfg->rr_start;
    P.estimator = cfg->estimator;
    P.frame = cfg->frame;
    P.sky = make_float3(cfg->sky[0], cfg->sky[1], cfg->sky[2]);
    P.sun = make_float3(cfg->sun_dir[0], cfg->sun_dir[1], cfg->sun_dir[2]);
    P.sun_e = make_float3(cfg->sun_irradiance[0], cfg->sun_irradiance[1], cfg->sun_irradiance[2]);
    P.has_sun = (cfg->sun_irradiance[0] > 0.0f || cfg->sun_irradiance[1] > 0.0f || cfg->sun_irradiance[2] > 0.0f);
    P.seed = cfg->seed;
    return P;
}

int rc_launch_camera(rc_ctx* c, const rc_camera* cam, const rc_render_cfg* cfg, uint32_t sample0, uint32_t spp_pass,
                     rc_ray* rays, float4* st, float4* L, float* W) {
    PassParams P = base_params(cam, cfg);
    P.spp_pass = spp_pass;
    P.sample0 = sample0;
    const uint32_t n = P.npix * spp_pass;
    k_camera<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(P, rays, st, L, W, c->counters);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

int rc_launch_film_add(rc_ctx* c, rc_film* f, const rc_render_cfg* cfg, uint32_t spp_pass, const float4* L,
                       const float* W) {
    const uint32_t npix = cfg->width * cfg->height;
    launch_film_add(c->stream, make_order(cfg->width, cfg->height), npix, spp_pass, L, W, f->sum, f->comp);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

// One frame of a scene into a film: the passes of rc_render, enqueued on
// stream st with the buffers B.  Used by rc_render, the cached renderer and
// the batched graph.
int rc_enqueue_frame(rc_ctx* c, FrameBuffers& B, hipStream_t st, const rc_scene* s, rc_film* f,
                     const rc_camera* cam, const rc_render_cfg* cfg, const FrameShape& shape, rc_progress_fn progress,
                     void* user, const CacheIO* cache_io, rc_cache* cache) {
    CacheIO CI{};
    if (cache_io) CI = *cache_io;
    const uint32_t npix = cfg->width * cfg->height;
    const uint32_t spp_pass = std::max(1u, std::min(shape.spp_pass, cfg->spp));
    const size_t npaths = size_t(npix) * spp_pass;
    if (npaths > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_render: pass too large");
    if (B.slotW.bytes < npaths * sizeof(float) || B.ovf_list.bytes < npaths * sizeof(uint32_t) || !B.counters)
        return rc_host_fail(RC_ERR_INVALID, "rc_render: frame buffers are smaller than the pass");
    // park regions, sized for this pass (the resume kernels' grids cover a
    // region, and a frame far smaller than the buffers must not launch grids
    // for the largest): three quarters of the records for the extension
    // kernel's parks, a quarter for the resume kernels' (far fewer rays get
    // that far)
    const uint32_t park_recs = uint32_t(std::min<size_t>(B.park.bytes / (kParkWords * sizeof(uint32_t)),
                                                         std::max<size_t>(npaths / RC_PARK_DIV, 256)));
    if (c->tuning.park_steps && park_recs < 4)
        return rc_host_fail(RC_ERR_INVALID, "rc_render: no park records for the extension kernel");
    const uint32_t park_cap0 = park_recs / 4 * 3;
    const ParkIO regions[2] = {{B.park.as<uint32_t>(), park_cap0, 0u, 0u},
                               {B.park.as<uint32_t>() + size_t(park_cap0) * kParkWords, park_recs - park_cap0, 0u, 1u}};
    const bool timed = st == c->stream;  // profiling events live on the context stream
    PassParams P = base_params(cam, cfg);
    P.stack_limit = c->stack_depth();
    const uint32_t grid_small = uint32_t(c->cu_count) * 2u;  // list-driven cold-path kernels
    rc_ray* ra = B.ray_a.as<rc_ray>();
    rc_ray* rb = B.ray_b.as<rc_ray>();
    float4* sa = B.st_a.as<float4>();
    float4* sb = B.st_b.as<float4>();
    float4* L = B.slotL.as<float4>();
    float* W = B.slotW.as<float>();
    rc_ray* srays = B.srays.as<rc_ray>();
    float4* scontrib = B.scontrib.as<float4>();
    uint8_t* socc = B.socc.as<uint8_t>();
    uint32_t* ovf = B.ovf_list.as<uint32_t>();
    DevCounters* ctr = B.counters;
    TraceLaunch deep;
    deep.scene = s;
    deep.ovf_list = ovf;
    deep.deep_stack = B.deep_stack.as<uint32_t>();
    deep.ctr = ctr;
    deep.stream = st;
    deep.counting = c->counting;
    deep.dec64 = c->hit_fp64;
    for (uint32_t s0 = 0; s0 < cfg->spp; s0 += spp_pass) {
        if (progress) {
            RC_TRY_HIP(hipStreamSynchronize(st));
            uint64_t done = uint64_t(double(npix) * double(s0) / double(cfg->spp));
            if (progress(user, done)) return rc_host_fail(RC_ERR_INTERRUPTED, "rc_render: interrupted");
        }
        uint32_t this_pass = std::min(spp_pass, cfg->spp - s0);
        P.spp_pass = this_pass;
        P.sample0 = s0;
        const uint32_t n = npix * this_pass;
        const uint32_t grid_n = (n + kBlock - 1) / kBlock;
        hipEvent_t ev = nullptr;
        if (CI.cache_vertex) k_cache_reset<<<1, 1, 0, st>>>(CI.counters, 0);
        if (RC_FUSED_CAMERA) {
            k_pass_start<<<1, 1, 0, st>>>(ctr, n);
        } else {
            if (timed) RC_TRY(c->prof_begin(&ev));
            k_camera<<<grid_n, kBlock, 0, st>>>(P, ra, sa, L, W, ctr);
            if (timed) RC_TRY(c->prof_end(ev, PK_OTHER));
        }
        rc_ray* rin = ra;
        rc_ray* rout = rb;
        float4* sin_ = sa;
        float4* sout = sb;
        for (uint32_t k = 0; k < cfg->max_path_length; ++k) {
            uint32_t cin = k & 1u, cout = cin ^ 1u;
            k_prep<<<1, 1, 0, st>>>(ctr, cout);
            if (timed) RC_TRY(c->prof_begin(&ev));
            const bool camb = RC_FUSED_CAMERA && k == 0;
            ParkIO PK = regions[0];
            PK.steps = camb ? 0u : c->tuning.park_steps;
            ExtendArgs ea{P, k, s, rin, sin_, rout, sout, srays, scontrib, L, ovf, ctr, cin, cout, CI, W, PK};
            launch_extend(ea, n, st, c->counting, camb, c->hit_fp64, shape.ext_block);
            if (PK.steps)
                launch_resumes(ea, regions, RC_PARK_PHASES, RC_PARK_STEPS2, st, c->counting, c->hit_fp64,
                               shape.ext_block);
            if (timed) RC_TRY(c->prof_end(ev, PK_EXTEND));
            // cold path: rays whose stack outgrew the on-chip depth
            deep.any = false;
            deep.rays = rin;
            deep.hits = B.hits.as<rc_hit>();
            deep.occ = nullptr;
            RC_TRY(rc_launch_deep_with(deep));
            k_shade_listed<<<grid_small, kBlock, 0, st>>>(P, k, rin, sin_, B.hits.as<rc_hit>(), s->shade, s->mats, rout,
                                                          sout, srays, scontrib, L, ovf, ctr, cout, CI);
            if (P.has_sun && k + 1 < cfg->max_path_length) {
                // at the cache vertex the shadow results are kept for the training records
                const bool nee_fix = CI.cache_vertex && k == CI.cache_vertex;
                uint8_t* occ_out = nee_fix ? socc : nullptr;
                k_reset_shadow<<<1, 1, 0, st>>>(ctr);
                if (timed) RC_TRY(c->prof_begin(&ev));
                ParkIO SPK = regions[0];
                SPK.steps = c->tuning.park_steps ? uint32_t(RC_SHADOW_PARK_STEPS) : 0u;
                ShadowArgs sa_{s, srays, scontrib, L, ovf, ctr, occ_out, P.stack_limit, SPK};
                launch_shadow(sa_, n, st, c->counting, c->hit_fp64, shape.shadow_block);
                if (SPK.steps)
                    launch_shadow_resumes(sa_, regions, RC_SHADOW_PARK_PHASES, RC_PARK_STEPS2, st, c->counting,
                                          c->hit_fp64, shape.shadow_block);
                if (timed) RC_TRY(c->prof_end(ev, PK_SHADOW));
                deep.any = true;
                deep.rays = srays;
                deep.hits = nullptr;
                deep.occ = socc;
                RC_TRY(rc_launch_deep_with(deep));
                k_shadow_apply_listed<<<grid_small, kBlock, 0, st>>>(scontrib, socc, ovf, L, ctr);
                if (nee_fix)
                    k_cache_nee_fix<<<grid_small, kBlock, 0, st>>>(CI.records, CI.counters, CI.cap, scontrib, socc);
            }
            std::swap(rin, rout);
            std::swap(sin_, sout);
        }
        if (CI.cache_vertex) {
            // paths that stopped at the cache vertex: add thr * cache(x, w_o)
            if (timed) RC_TRY(c->prof_begin(&ev));
            RC_TRY(rc_cache_apply_device(c, st, cache, CI.queries, &CI.counters[kCcQueries], CI.cap, L));
            if (timed) RC_TRY(c->prof_end(ev, PK_CACHE));
            k_cache_targets<<<grid_small, kBlock, 0, st>>>(CI.records, CI.counters, CI.cap, L, CI.batch_q, CI.batch_t,
                                                          CI.batch_cap);
        }
        if (timed) RC_TRY(c->prof_begin(&ev));
        launch_film_add(st, P.order, npix, this_pass, L, W, f->sum, f->comp);
        if (timed) RC_TRY(c->prof_end(ev, PK_OTHER));
        RC_TRY_HIP(hipGetLastError());
    }
    return RC_OK;
}

namespace {
struct PointerProgress {
    volatile const int* interrupt;
    uint64_t* pixels;
};
int pointer_progress(void* user, uint64_t done) {
    auto* p = static_cast<PointerProgress*>(user);
    if (p->pixels) *p->pixels = done;
    return p->interrupt && *p->interrupt;
}
}  // namespace

extern "C" int rc_render(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                         volatile const int* interrupt, uint64_t* pixels_processed) {
    PointerProgress pp{interrupt, pixels_processed};
    bool poll = interrupt != nullptr || pixels_processed != nullptr;
    int rc = rc_render_progress(c, s, f, cam, cfg, poll ? pointer_progress : nullptr, &pp);
    if (rc == RC_OK && pixels_processed) *pixels_processed = uint64_t(cfg->width) * cfg->height;
    return rc;
}

namespace {
int check_frame_args(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                     const char* fn) {
    RC_TRY(rc_check_scene(c, s, fn));
    RC_TRY(rc_validate_cfg(cfg, fn));
    if (!f || f->ctx != c) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": bad film");
    if (!cam) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null camera");
    if (f->width != cfg->width || f->height != cfg->height)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": film size differs from the config");
    return RC_OK;
}

// One frame, synchronously, with the statistics of rc_get_stats.  With a
// cache, `train` optimizer steps on the frame's training samples are queued
// on the cache's own stream behind the frame (the next frame's traversal runs
// while they do; its cache inference waits for them).
int render_frame(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                 rc_progress_fn progress, void* user, rc_cache* cache, uint32_t train,
                 const FrameShape* forced = nullptr) {
    RC_TRY_HIP(hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    c->stats.extend_ms = c->stats.shadow_ms = c->stats.shade_ms = c->stats.other_ms = 0.0;
    c->stats.cache_ms = 0.0;
    c->stats.extend_launches = c->stats.shadow_launches = 0;
    c->stats.cache_queries = c->stats.train_samples = 0;
    rc_zero_counters(c->counters, c->stream);
    FrameShape shape;
    shape.spp_pass = pass_samples(c->tuning, cfg->width * cfg->height, cfg->spp, kPassPaths);
    if (forced) shape = *forced;
    RC_TRY(c->fb.ensure(rc_pass_paths(cfg, shape.spp_pass)));
    CacheIO io{};
    if (cache)
        RC_TRY(rc_cache_frame_io(c, c->stream, cache, cfg->cache_vertex, rc_pass_paths(cfg, shape.spp_pass),
                                 uint64_t(cfg->width) * cfg->height * cfg->spp, &io));
    int rc = rc_enqueue_frame(c, c->fb, c->stream, s, f, cam, cfg, shape, progress, user, cache ? &io : nullptr,
                              cache);
    uint32_t cache_counts[2] = {0, 0};
    if (rc == RC_OK && cache) rc = rc_cache_frame_end(c, c->stream, cache, train, cache_counts);
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    if (rc != RC_OK) return rc;
    DevCounters h;
    RC_TRY_HIP(hipMemcpy(&h, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost));
    if (h.error) return rc_host_fail(RC_ERR_HIP, "rc_render: traversal stack exceeded the deep-stack limit");
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    RC_TRY(c->prof_collect());
    c->stats.extension_rays = h.ext_rays;
    c->stats.shadow_rays = h.shadow_rays;
    c->stats.samples = uint64_t(cfg->width) * cfg->height * cfg->spp;
    c->stats.node_records = h.node_records;
    c->stats.tri_tests = h.tri_tests;
    c->stats.overflow_rays = h.ovf_total;
    c->stats.parked_rays = h.park_total;
    if (cache) {
        c->stats.cache_queries = cache_counts[0];
        c->stats.train_samples = train ? cache_counts[1] : 0;
    }
    if (progress) progress(user, uint64_t(cfg->width) * cfg->height);
    return RC_OK;
}
}  // namespace

int rc_render_shaped(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                     const FrameShape& shape) {
    RC_TRY(check_frame_args(c, s, f, cam, cfg, "rc_render"));
    return render_frame(c, s, f, cam, cfg, nullptr, nullptr, nullptr, 0, &shape);
}

extern "C" int rc_render_progress(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam,
                                  const rc_render_cfg* cfg, rc_progress_fn progress, void* user) {
    RC_TRY(check_frame_args(c, s, f, cam, cfg, "rc_render"));
    return render_frame(c, s, f, cam, cfg, progress, user, nullptr, 0);
}

extern "C" int rc_render_cached(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam,
                                const rc_render_cfg* cfg, rc_cache* cache, int train, float* loss) {
    RC_TRY(check_frame_args(c, s, f, cam, cfg, "rc_render_cached"));
    if (!cache || rc_cache_owner(cache) != c) return rc_host_fail(RC_ERR_INVALID, "rc_render_cached: bad cache");
    if (cfg->cache_vertex == 0 || cfg->cache_vertex + 1 >= cfg->max_path_length)
        return rc_host_fail(RC_ERR_INVALID,
                            "rc_render_cached: cache_vertex must satisfy 1 <= cache_vertex < max_path_length - 1");
    if (train < 0 || train > 16) return rc_host_fail(RC_ERR_INVALID, "rc_render_cached: train must be in [0, 16]");
    RC_TRY(render_frame(c, s, f, cam, cfg, nullptr, nullptr, cache, uint32_t(train)));
    if (loss) RC_TRY(rc_cache_last_loss(c, cache, loss));
    return RC_OK;
}

// ---------------------------------------------------------------- Phong raycast
namespace {

__global__ __launch_bounds__(kBlock) void k_cast_camera(CamBasis cam, uint32_t width, uint32_t npix, uint64_t seed,
                                                        rc_ray* __restrict__ rays) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float4 u = draw4(seed, p, 0, 0, kCamDim);
    float3 d = cam_dir(cam, float(p % width) + u.x, float(p / width) + u.y);
    float4* r = reinterpret_cast<float4*>(rays + p);
    r[0] = make_float4(cam.pos.x, cam.pos.y, cam.pos.z, 0.0f);
    r[1] = make_float4(d.x, d.y, d.z, __builtin_inff());
}

struct LightParams {
    float3 pos, color, ambient;
    float max_d, att, offset;
};

// Raytracer::calculateLightColor for one point light + calculatePhongShadingColor
// (Raytracer.cpp:312-407); the shadow ray of calculateShadowAmount (:524-551)
// starts rayStartOffset along the light direction and ends at the light.
__global__ __launch_bounds__(kBlock) void k_phong_prepare(uint32_t npix, LightParams Lp, const rc_ray* __restrict__ rays,
                                                          const rc_hit* __restrict__ hits,
                                                          const float4* __restrict__ shade,
                                                          const DevMaterial* __restrict__ mats,
                                                          rc_ray* __restrict__ srays, float4* __restrict__ direct,
                                                          float4* __restrict__ rgba, int32_t* __restrict__ ids,
                                                          float* __restrict__ hit_t) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float4 h = reinterpret_cast<const float4*>(hits)[p];
    int tri = __float_as_int(h.w);
    if (ids) ids[p] = tri;
    if (hit_t) hit_t[p] = h.x;
    float4* sr = reinterpret_cast<float4*>(srays + p);
    if (tri < 0) {
        sr[0] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);  // empty interval: never occluded
        sr[1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
        direct[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rgba[p] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        return;
    }
    const float4* rp = reinterpret_cast<const float4*>(rays + p);
    float3 o = f3(rp[0]), d = f3(rp[1]);
    float3 pos = o + (h.x * d);
    float4 sh = shade[tri];
    const DevMaterial& m = mats[__float_as_uint(sh.w)];
    float3 n = f3(sh), kd = f3(m.diffuse), ka = f3(m.ambient), ks = f3(m.specular);
    float shin = m.emittance.w;
    float3 col = mul3(Lp.ambient, ka);
    float3 tl = Lp.pos - pos;
    float dist = sqrtf(dot3(tl, tl));
    float3 l = normalize3(tl);
    float3 ph = make_float3(0.0f, 0.0f, 0.0f);
    float diff = dot3(l, n);
    if (diff > 0.0f) {
        ph = mul3(Lp.color * diff, kd);
        if (!(ks.x == 0.0f && ks.y == 0.0f && ks.z == 0.0f)) {
            float3 refl = normalize3((2.0f * diff * n) - l);
            float sa = dot3(refl, make_float3(-d.x, -d.y, -d.z));
            if (sa > 0.0f) ph = ph + mul3(Lp.color * powf(sa, shin), ks);
        }
    }
    float da = fminf(1.0f, dist / Lp.max_d);
    da = 1.0f - powf(da, Lp.att);
    float3 so = pos + l * Lp.offset;
    sr[0] = make_float4(so.x, so.y, so.z, 0.0f);
    sr[1] = make_float4(l.x, l.y, l.z, dist);
    direct[p] = make_float4(ph.x * da, ph.y * da, ph.z * da, 0.0f);
    rgba[p] = make_float4(col.x, col.y, col.z, 1.0f);
}

// Adds the unoccluded direct light.
__global__ __launch_bounds__(kBlock) void k_phong_finish(uint32_t npix, const uint8_t* __restrict__ occ,
                                                         const float4* __restrict__ direct, float4* __restrict__ rgba) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix || occ[p]) return;
    float4 c = rgba[p], dd = direct[p];
    rgba[p] = make_float4(c.x + dd.x, c.y + dd.y, c.z + dd.z, 1.0f);
}

}  // namespace

extern "C" int rc_raycast_phong(rc_ctx* c, const rc_scene* s, const rc_camera* cam, uint32_t w, uint32_t h,
                                uint64_t seed, float offset, const rc_point_light* light, int32_t* hit_ids, float* t,
                                float* rgba) {
    RC_TRY(rc_check_scene(c, s, "rc_raycast_phong"));
    if (!cam || !light) return rc_host_fail(RC_ERR_INVALID, "rc_raycast_phong: null camera or light");
    if (w < 2 || h < 2 || size_t(w) * h > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_raycast_phong: bad size");
    RC_TRY_HIP(hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    c->stats = rc_stats{};  // this call's statistics only
    const uint32_t npix = w * h;
    RC_TRY(c->fb.ensure(npix));
    CamBasis B = make_basis(*cam, w, h);
    LightParams Lp;
    Lp.pos = make_float3(light->position[0], light->position[1], light->position[2]);
    Lp.color = make_float3(light->color[0] * light->intensity, light->color[1] * light->intensity,
                           light->color[2] * light->intensity);
    Lp.ambient = make_float3(light->ambient_color[0] * light->ambient_intensity,
                             light->ambient_color[1] * light->ambient_intensity,
                             light->ambient_color[2] * light->ambient_intensity);
    Lp.max_d = light->max_distance;
    Lp.att = light->attenuation;
    Lp.offset = offset;
    const uint32_t grid = (npix + kBlock - 1) / kBlock;
    k_zero_counters<<<1, 1, 0, c->stream>>>(c->counters);
    hipEvent_t ev;
    RC_TRY(c->prof_begin(&ev));
    k_cast_camera<<<grid, kBlock, 0, c->stream>>>(B, w, npix, seed, c->fb.ray_a.as<rc_ray>());
    RC_TRY(c->prof_end(ev, PK_OTHER));
    FrameBuffers& FB = c->fb;
    RC_TRY(rc_launch_trace(c, s, false, FB.ray_a.as<rc_ray>(), nullptr, npix, FB.hits.as<rc_hit>(), nullptr));
    float4* rgba_d = FB.slotL.as<float4>();
    // the hit indices and distances go to two device arrays (the queue
    // buffers, free by now) as soon as the closest hits are known, and from
    // there to the caller's arrays on the copy stream while the shadow rays
    // are traced
    int32_t* ids_d = hit_ids ? reinterpret_cast<int32_t*>(FB.st_a.p) : nullptr;
    float* t_d = t ? reinterpret_cast<float*>(FB.st_b.p) : nullptr;
    RC_TRY(c->prof_begin(&ev));
    k_phong_prepare<<<grid, kBlock, 0, c->stream>>>(npix, Lp, FB.ray_a.as<rc_ray>(), FB.hits.as<rc_hit>(), s->shade,
                                                    s->mats, FB.srays.as<rc_ray>(), FB.scontrib.as<float4>(), rgba_d,
                                                    ids_d, t_d);
    RC_TRY(c->prof_end(ev, PK_SHADE));
    const bool side = hit_ids || t;
    if (side) {
        if (!c->copy_stream) RC_TRY_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        if (!c->copy_ev) RC_TRY_HIP(hipEventCreateWithFlags(&c->copy_ev, hipEventDisableTiming));
        RC_TRY_HIP(hipEventRecord(c->copy_ev, c->stream));
        RC_TRY_HIP(hipStreamWaitEvent(c->copy_stream, c->copy_ev, 0));
        if (hit_ids)
            RC_TRY_HIP(hipMemcpyAsync(hit_ids, ids_d, npix * sizeof(int32_t), hipMemcpyDeviceToHost, c->copy_stream));
        if (t) RC_TRY_HIP(hipMemcpyAsync(t, t_d, npix * sizeof(float), hipMemcpyDeviceToHost, c->copy_stream));
    }
    RC_TRY(rc_launch_trace(c, s, true, FB.srays.as<rc_ray>(), nullptr, npix, nullptr, FB.socc.as<uint8_t>()));
    RC_TRY(c->prof_begin(&ev));
    k_phong_finish<<<grid, kBlock, 0, c->stream>>>(npix, FB.socc.as<uint8_t>(), FB.scontrib.as<float4>(), rgba_d);
    RC_TRY(c->prof_end(ev, PK_OTHER));
    RC_TRY_HIP(hipGetLastError());
    if (rgba) RC_TRY_HIP(hipMemcpyAsync(rgba, rgba_d, npix * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    if (side) {
        // the ids and t buffers are the frame's queues: the next call on the
        // main stream must not overwrite them before the copies are done
        RC_TRY_HIP(hipEventRecord(c->copy_ev, c->copy_stream));
        RC_TRY_HIP(hipStreamWaitEvent(c->stream, c->copy_ev, 0));
        RC_TRY_HIP(hipStreamSynchronize(c->copy_stream));
    }
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    RC_TRY(c->prof_collect());
    c->stats.extension_rays = npix;
    c->stats.shadow_rays = npix;
    c->stats.samples = npix;
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RC_OK;
}
