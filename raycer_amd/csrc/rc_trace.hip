// Closest-hit and any-hit traversal kernels and the ray-query entry points
// (the device counterpart of FlatBVH::intersect + Triangle::intersect,
// src/Raytracing/Primitives/FlatBVH.cpp:22-75, Triangle.cpp:51-121).
//
// k_trace: persistent grid-stride loop, one ray per lane, the top of the
// traversal stack in registers and the rest in LDS laid out [depth][lane] so a
// wave's pushes and pops hit 64 distinct banks.  A ray whose stack would
// exceed the on-chip depth is queued for k_trace_deep, which walks it again
// with the memory part of its stack in global memory, as deep as the deepest
// BVH the upload accepts.  The node/triangle counting variant is a
// separate instantiation so the timed kernels carry no counters.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rc_internal.hpp"

using namespace rcd;

namespace {

#ifndef RC_TRACE_BLOCK
#define RC_TRACE_BLOCK 256
#endif
constexpr int kBlock = RC_TRACE_BLOCK;
constexpr int kDeepThreads = 2048;

__device__ __forceinline__ void store_result(bool any, const TraceResult& R, uint32_t i, rc_hit* hits, uint8_t* occ) {
    bool hit = R.id != kNoHit;
    if (any) {
        occ[i] = hit ? 1 : 0;
    } else {
        rc_hit h;
        h.t = hit ? R.t : __builtin_inff();
        h.u = hit ? bary_u(R) : 0.0f;
        h.v = hit ? bary_v(R) : 0.0f;
        h.tri = hit ? int32_t(R.id) : -1;
        reinterpret_cast<float4*>(hits)[i] = *reinterpret_cast<float4*>(&h);
    }
}

template <bool ANY, bool COUNT, bool DEC64>
__global__ __launch_bounds__(kBlock) void k_trace(const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                  const rc_ray* __restrict__ rays, const uint32_t* n_dev,
                                                  uint32_t n_host, rc_hit* __restrict__ hits, uint8_t* __restrict__ occ,
                                                  uint32_t* __restrict__ ovf_list, DevCounters* __restrict__ ctr,
                                                  int max_depth) {
    __shared__ uint32_t stack[(kLdsStack > 0 ? kLdsStack : 1) * kBlock];
    __shared__ uint4 top[4 * kTopNodes];
    const uint32_t ntop = load_top(nodes, top);
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    uint32_t n_rec = 0, n_tri = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        float4 r0 = rp[0], r1 = rp[1];
        TraceResult R = traverse<ANY, COUNT, DEC64>(nodes, tris, f3(r0), f3(r1), r0.w, r1.w, stack + threadIdx.x,
                                                    kBlock, max_depth, n_rec, n_tri, top, ntop);
        if (R.overflow) {
            ovf_list[atomicAdd(&ctr->ovf_count, 1u)] = i;
            continue;
        }
        store_result(ANY, R, i, hits, occ);
    }
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

// The rays listed in ovf_list (ctr->ovf_count of them), walked again with the
// memory part of each lane's stack in global memory (gstack: kDeepThreads x
// kDeepStack entries, lane-interleaved).
template <bool ANY, bool COUNT, bool DEC64>
__global__ __launch_bounds__(64) void k_trace_deep(const NodeT* __restrict__ nodes, const TriRec* __restrict__ tris,
                                                   const rc_ray* __restrict__ rays, rc_hit* __restrict__ hits,
                                                   uint8_t* __restrict__ occ, const uint32_t* __restrict__ ovf_list,
                                                   uint32_t* __restrict__ gstack, DevCounters* __restrict__ ctr) {
    const uint32_t n = ctr->ovf_count;
    const uint32_t gtid = blockIdx.x * 64 + threadIdx.x, nthreads = gridDim.x * 64;
    if (gtid == 0 && n) atomicAdd(&ctr->ovf_total, (unsigned long long)n);
    uint32_t n_rec = 0, n_tri = 0;
    for (uint32_t k = gtid; k < n; k += nthreads) {
        uint32_t i = ovf_list[k];
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        float4 r0 = rp[0], r1 = rp[1];
        TraceResult R = traverse<ANY, COUNT, DEC64>(nodes, tris, f3(r0), f3(r1), r0.w, r1.w, gstack + gtid, nthreads,
                                                    kRegStack + kDeepStack, n_rec, n_tri);
        if (R.overflow) {
            ctr->error = 1;  // unreachable: the upload rejects trees this deep
            continue;
        }
        store_result(ANY, R, i, hits, occ);
    }
    if (COUNT) {
        wave_add(&ctr->node_records, n_rec);
        wave_add(&ctr->tri_tests, n_tri);
    }
}

__global__ void k_reset_overflow(DevCounters* ctr) { ctr->ovf_count = 0; }

template <bool ANY, bool COUNT, bool DEC64>
void launch_pair(const TraceLaunch& L, uint32_t blocks) {
    k_trace<ANY, COUNT, DEC64><<<blocks, kBlock, 0, L.stream>>>(L.scene->nodes, tri_array<DEC64>(L.scene), L.rays, L.n_dev,
                                                                L.n_bound, L.hits, L.occ, L.ovf_list, L.ctr,
                                                                L.max_depth);
}

template <bool ANY, bool COUNT, bool DEC64>
void launch_deep(const TraceLaunch& L) {
    k_trace_deep<ANY, COUNT, DEC64><<<kDeepThreads / 64, 64, 0, L.stream>>>(L.scene->nodes, tri_array<DEC64>(L.scene), L.rays,
                                                                           L.hits, L.occ, L.ovf_list, L.deep_stack,
                                                                           L.ctr);
}

template <bool COUNT, bool DEC64>
void dispatch_trace(const TraceLaunch& L, uint32_t blocks) {
    if (L.any) launch_pair<true, COUNT, DEC64>(L, blocks);
    else launch_pair<false, COUNT, DEC64>(L, blocks);
}

template <bool COUNT, bool DEC64>
void dispatch_deep(const TraceLaunch& L) {
    if (L.any) launch_deep<true, COUNT, DEC64>(L);
    else launch_deep<false, COUNT, DEC64>(L);
}

}  // namespace

size_t rc_deep_stack_bytes() { return size_t(kDeepThreads) * kDeepStack * sizeof(uint32_t); }

int rc_launch_trace_with(const TraceLaunch& L) {
    if (L.n_bound == 0) return RC_OK;
    const uint32_t blocks = std::max(1u, std::min((L.n_bound + kBlock - 1) / kBlock, L.max_blocks));
    k_reset_overflow<<<1, 1, 0, L.stream>>>(L.ctr);
    if (L.counting) {
        if (L.dec64) dispatch_trace<true, true>(L, blocks);
        else dispatch_trace<true, false>(L, blocks);
    } else {
        if (L.dec64) dispatch_trace<false, true>(L, blocks);
        else dispatch_trace<false, false>(L, blocks);
    }
    return RC_OK;
}

int rc_launch_deep_with(const TraceLaunch& L) {
    if (L.counting) {
        if (L.dec64) dispatch_deep<true, true>(L);
        else dispatch_deep<true, false>(L);
    } else {
        if (L.dec64) dispatch_deep<false, true>(L);
        else dispatch_deep<false, false>(L);
    }
    return RC_OK;
}

namespace {
TraceLaunch context_launch(rc_ctx* c, const rc_scene* s, bool any, const rc_ray* rays, const uint32_t* n_dev,
                           uint32_t n_bound, rc_hit* hits, uint8_t* occ) {
    TraceLaunch L;
    L.scene = s;
    L.any = any;
    L.rays = rays;
    L.n_dev = n_dev;
    L.n_bound = n_bound;
    L.hits = hits;
    L.occ = occ;
    L.ovf_list = c->fb.ovf_list.as<uint32_t>();
    L.deep_stack = c->fb.deep_stack.as<uint32_t>();
    L.ctr = c->counters;
    L.stream = c->stream;
    L.max_depth = c->stack_depth();
    L.max_blocks = uint32_t(c->cu_count) * c->tuning.blocks_per_cu;
    L.counting = c->counting;
    L.dec64 = c->hit_fp64;
    return L;
}
}  // namespace

int rc_launch_trace(rc_ctx* c, const rc_scene* s, bool any, const rc_ray* rays, const uint32_t* n_dev,
                    uint32_t n_bound, rc_hit* hits, uint8_t* occ) {
    if (n_bound == 0) return RC_OK;
    RC_TRY(c->fb.ovf_list.ensure(size_t(n_bound) * sizeof(uint32_t)));
    RC_TRY(c->fb.deep_stack.ensure(rc_deep_stack_bytes()));
    // one lane per ray up to cu_count * blocks_per_cu blocks; beyond that the
    // kernel's grid-stride loop takes over
    const TraceLaunch L = context_launch(c, s, any, rays, n_dev, n_bound, hits, occ);
    hipEvent_t ev;
    RC_TRY(c->prof_begin(&ev));
    RC_TRY(rc_launch_trace_with(L));
    RC_TRY(c->prof_end(ev, any ? PK_SHADOW : PK_EXTEND));
    // rays whose stack outgrew the on-chip depth are walked again with the deep stack
    RC_TRY(rc_launch_deep_with(L));
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

// Deep-stack traversal of the rays listed in the overflow list (count in the
// device counters); hits (closest) or occluded flags (any) are written at the
// listed ray indices.
int rc_launch_deep(rc_ctx* c, const rc_scene* s, bool any, const rc_ray* rays, rc_hit* hits, uint8_t* occ) {
    RC_TRY(rc_launch_deep_with(context_launch(c, s, any, rays, nullptr, 0, hits, occ)));
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

namespace {

int reset_counters(rc_ctx* c) {
    RC_TRY_HIP(hipMemsetAsync(c->counters, 0, sizeof(DevCounters), c->stream));
    return RC_OK;
}

int read_counters(rc_ctx* c, DevCounters& h) {
    RC_TRY_HIP(hipMemcpyAsync(&h, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    if (h.error) return rc_host_fail(RC_ERR_HIP, "traversal stack exceeded the deep-stack limit");
    return RC_OK;
}

int trace_host(rc_ctx* c, const rc_scene* s, bool any, const rc_ray* rays, size_t n, rc_hit* hits, uint8_t* occ,
               const char* fn) {
    RC_TRY(rc_check_scene(c, s, fn));
    if (n == 0) return RC_OK;
    if (!rays || (!any && !hits) || (any && !occ)) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null array");
    if (n > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": too many rays");
    RC_TRY_HIP(hipSetDevice(c->device));
    RC_TRY(c->trace_rays.ensure(n * sizeof(rc_ray)));
    RC_TRY(any ? c->trace_occ.ensure(n) : c->trace_hits.ensure(n * sizeof(rc_hit)));
    RC_TRY_HIP(hipMemcpyAsync(c->trace_rays.p, rays, n * sizeof(rc_ray), hipMemcpyHostToDevice, c->stream));
    RC_TRY(reset_counters(c));
    RC_TRY(rc_launch_trace(c, s, any, c->trace_rays.as<rc_ray>(), nullptr, uint32_t(n), c->trace_hits.as<rc_hit>(),
                           c->trace_occ.as<uint8_t>()));
    if (any) RC_TRY_HIP(hipMemcpyAsync(occ, c->trace_occ.p, n, hipMemcpyDeviceToHost, c->stream));
    else RC_TRY_HIP(hipMemcpyAsync(hits, c->trace_hits.p, n * sizeof(rc_hit), hipMemcpyDeviceToHost, c->stream));
    DevCounters h;
    RC_TRY(read_counters(c, h));
    RC_TRY(c->prof_collect());
    c->stats.node_records = h.node_records;
    c->stats.tri_tests = h.tri_tests;
    c->stats.overflow_rays = h.ovf_total;
    if (any) c->stats.shadow_rays = n; else c->stats.extension_rays = n;
    return RC_OK;
}

}  // namespace

extern "C" int rc_trace_closest(rc_ctx* c, const rc_scene* s, const rc_ray* rays, size_t n, rc_hit* hits) {
    return trace_host(c, s, false, rays, n, hits, nullptr, "rc_trace_closest");
}

extern "C" int rc_trace_any(rc_ctx* c, const rc_scene* s, const rc_ray* rays, size_t n, uint8_t* occ) {
    return trace_host(c, s, true, rays, n, nullptr, occ, "rc_trace_any");
}

extern "C" int rc_trace_closest_device(rc_ctx* c, const rc_scene* s, const rc_ray* d_rays, size_t n, rc_hit* d_hits) {
    RC_TRY(rc_check_scene(c, s, "rc_trace_closest_device"));
    if (n > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_trace_closest_device: too many rays");
    if (n && (!d_rays || !d_hits)) return rc_host_fail(RC_ERR_INVALID, "rc_trace_closest_device: null array");
    return rc_launch_trace(c, s, false, d_rays, nullptr, uint32_t(n), d_hits, nullptr);
}

extern "C" int rc_trace_any_device(rc_ctx* c, const rc_scene* s, const rc_ray* d_rays, size_t n, uint8_t* d_occ) {
    RC_TRY(rc_check_scene(c, s, "rc_trace_any_device"));
    if (n > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_trace_any_device: too many rays");
    if (n && (!d_rays || !d_occ)) return rc_host_fail(RC_ERR_INVALID, "rc_trace_any_device: null array");
    return rc_launch_trace(c, s, true, d_rays, nullptr, uint32_t(n), nullptr, d_occ);
}
