// Host-side internals shared by the .hip translation units of libraycer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "host/rc_host_internal.hpp"
#include "raycer_hip.h"
#include "rc_device.hpp"

#define RC_TRY_HIP(expr)                                                                               \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return rc_host_fail(RC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define RC_TRY(expr)                 \
    do {                             \
        int r_ = (expr);             \
        if (r_ != RC_OK) return r_;  \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Device counters written by the kernels (one allocation per context).
struct DevCounters {
    unsigned long long node_records;
    unsigned long long tri_tests;
    unsigned long long ext_rays;
    unsigned long long shadow_rays;
    uint32_t ovf_count;
    uint32_t queue_count[4];   // wavefront queue sizes (ping, pong, shadow, scratch)
    uint32_t error;
};

enum ProfKind { PK_EXTEND = 0, PK_SHADOW = 1, PK_SHADE = 2, PK_OTHER = 3, PK_CACHE = 4 };

// Paths one pass of the path tracer may hold when the samples per pass are
// chosen automatically (about 13 GB of queues at 200 bytes per path); the
// Whitted tracer's per-level ray buffers are far larger per path.
constexpr size_t kPassPaths = size_t(1) << 26;
constexpr size_t kWhittedPassPaths = size_t(1) << 21;

// Samples per pass of a frame: the tuning's spp_per_pass, or with 0 (the
// default) as many as keep a pass within `budget` paths; at least one, at most
// the frame's.
inline uint32_t pass_samples(const rc_tuning& t, uint32_t npix, uint32_t spp, size_t budget) {
    const size_t sp = t.spp_per_pass ? t.spp_per_pass : budget / std::max<size_t>(npix, 1);
    return uint32_t(std::max<size_t>(1, std::min<size_t>(sp, spp)));
}

struct rc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string name;
    bool profiling = false;
    bool counting = false;
    rc_tuning tuning{256, 0, 64, 0};  // samples per pass chosen per frame
    rc_stats stats{};
    DevCounters* counters = nullptr;  // device
    DevBuf trace_rays, trace_hits, trace_occ, ovf_list, deep_stack;
    // wavefront state
    DevBuf ray_a, ray_b, st_a, st_b, hits, srays, scontrib, socc, slotL, slotW;
    // profiling events
    struct Pending {
        hipEvent_t a, b;
        int kind;
    };
    std::vector<hipEvent_t> ev_free;
    std::vector<Pending> pending;
    hipGraphExec_t batch_graph = nullptr;
    struct WhittedBufs* whitted = nullptr;  // the Whitted tracer's ray-tree levels

    int prof_begin(hipEvent_t* a);
    int prof_end(hipEvent_t a, int kind);
    int prof_collect();  // synchronises and accumulates pending timings
};

struct rc_scene {
    rc_ctx* ctx = nullptr;
    rcd::NodeT* nodes = nullptr;  // NodeRec records, or QNode nodes when RC_QNODE
    size_t nrec = 0;
    rcd::TriRec* tris = nullptr;
    size_t ntri = 0;
    float4* shade = nullptr;  // normal.xyz, material index bits
    rcd::DevMaterial* mats = nullptr;
    size_t nmat = 0;
    float bmin[3], bmax[3];
    uint32_t depth = 0;
    size_t bytes = 0;
};

struct rc_film {
    rc_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0;
    float4* sum = nullptr;   // w*rgb, w
    float4* comp = nullptr;  // Kahan compensation terms
    DevBuf linear, partial;
    double max_lum_average = 1.0;  // ReinhardToneMapper's maxLuminanceAverage (ReinhardToneMapper.cpp:12-15)
};

// Launches the closest-hit (ANY = false) or any-hit traversal over a device
// ray array.  n_dev (may be null) is a device count read by the kernel; n_bound
// bounds it and sizes the grid.
int rc_launch_trace(rc_ctx* ctx, const rc_scene* s, bool any, const rc_ray* rays, const uint32_t* n_dev,
                    uint32_t n_bound, rc_hit* hits, uint8_t* occ);
int rc_launch_deep(rc_ctx* ctx, const rc_scene* s, bool any, const rc_ray* rays, rc_hit* hits, uint8_t* occ);
// Camera rays of one pass of Tracer::generateMultiSamples (slot = j * npix + q
// for sample sample0 + j of the pixel at position q of the tiled order): ray,
// state (1, 1, 1, slot), L = 0 and the filter weight W.
int rc_launch_camera(rc_ctx* c, const rc_camera* cam, const rc_render_cfg* cfg, uint32_t sample0, uint32_t spp_pass,
                     rc_ray* rays, float4* st, float4* L, float* W);
// Film::addSample of a pass's slot results, per pixel in sample order.
int rc_launch_film_add(rc_ctx* c, rc_film* f, const rc_render_cfg* cfg, uint32_t spp_pass, const float4* L,
                       const float* W);
// Frees the Whitted tracer's buffers of a context.
void rc_whitted_release(rc_ctx* c);

// Radiance-cache plumbing of a frame (all null / 0 when the cache is off).
struct CacheIO {
    float4* queries;        // 4 float4 per query: (p, slot), (w_o, thr.r), (n, thr.g), (albedo, thr.b)
    float4* records;        // 5 float4 per training vertex: the query layout, then (L before, 0)
    uint32_t* counters;     // [0] queries, [1] records of the pass; [2] training samples, [3] queries of the frame
    float* batch_q;         // 12 floats per training sample
    float* batch_t;         // 3 floats per training sample
    uint32_t batch_cap;
    uint32_t cache_vertex;  // 0 = off
    uint32_t train_stride;  // one training path per train_stride slots
    uint32_t cap;           // record / query capacity of a pass
};

int rc_validate_cfg(const rc_render_cfg* cfg, const char* fn);
// Sizes every buffer a frame of this config needs (outside any capture).
int rc_prepare_frame(rc_ctx* c, const rc_render_cfg* cfg, uint32_t spp_pass);
// Enqueues the passes of one frame on the context stream; cache_io / cache are
// null for the plain renderer.
int rc_enqueue_frame(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                     uint32_t spp_pass, rc_progress_fn progress, void* user, const CacheIO* cache_io = nullptr,
                     rc_cache* cache = nullptr);
// Radiance cache, device side (rc_mlp.hip).
// L[slot] += thr * cache(query) for the *n_dev (<= n_bound) queries.
int rc_cache_apply_device(rc_ctx* c, rc_cache* k, const float4* queries, const uint32_t* n_dev, uint32_t n_bound,
                          float4* L);
// Sizes the cache's frame buffers for passes of pass_paths slots and a frame
// of frame_samples samples, fills io and enqueues the clearing of its counters.
int rc_cache_frame_io(rc_ctx* c, rc_cache* k, uint32_t cache_vertex, size_t pass_paths, uint64_t frame_samples,
                      CacheIO* io);
// After a frame: the frame's query and training-sample counts and `steps`
// optimizer steps over interleaved slices of the training samples it gathered.
int rc_cache_train_frame(rc_ctx* c, rc_cache* k, uint32_t steps, float* loss, uint32_t* samples, uint32_t* queries);
rc_ctx* rc_cache_owner(const rc_cache* k);
int rc_check_ctx(const rc_ctx* ctx, const char* fn);
int rc_check_scene(const rc_ctx* ctx, const rc_scene* s, const char* fn);
