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
    unsigned long long ovf_total;  // rays the deep-stack kernel walked, over every launch since the last reset
    unsigned long long park_total; // traversals parked for the resume kernel since the last reset
    uint32_t ovf_count;            // rays listed for the next deep-stack launch
    uint32_t queue_count[4];   // wavefront queue sizes (ping, pong, shadow, scratch)
    uint32_t park_count[2];    // traversals parked in the two park regions (ping-pong between phases)
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

// Device buffers of one frame's wavefront passes (path and shadow queues,
// slot radiance and weights, the cold path's overflow list and deep stack)
// and its counters.  The context owns one set (fb, whose counters are the
// context's); a batch owns one more per concurrent branch.
struct FrameBuffers {
    DevBuf ray_a, ray_b, st_a, st_b, hits, srays, scontrib, socc, slotL, slotW, ovf_list, deep_stack, park;
    DevCounters* counters = nullptr;
    bool own_counters = false;
    int ensure(size_t npaths);  // every buffer sized for a pass of npaths paths
    void release();
};

// Samples per pass and threads per block of one frame's passes.
struct FrameShape {
    uint32_t spp_pass = 1;
    uint32_t ext_block = 256;     // extension kernel: 128 or 256
    uint32_t shadow_block = 128;  // shadow kernel: 64, 128 or 256
};

struct rc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string name;
    bool profiling = false;
    bool counting = false;
    bool hit_fp64 = false;  // RC_HIT_FP64: FP32 filter, FP64 decision of triangle hits
    rc_tuning tuning{256, 0, 64, 0, RC_PARK_STEPS_DEFAULT};  // samples per pass chosen per frame, the whole on-chip stack

    // traversal stack entries held on chip before a ray goes to the deep-stack kernel
    int stack_depth() const {
        return tuning.stack_limit ? int(tuning.stack_limit) : rcd::kOnChipStack;
    }
    rc_stats stats{};
    DevCounters* counters = nullptr;  // device
    DevBuf trace_rays, trace_hits, trace_occ;
    FrameBuffers fb;  // wavefront state of rc_render and the ray queries' cold path (fb.counters == counters)
    // profiling events
    struct Pending {
        hipEvent_t a, b;
        int kind;
    };
    std::vector<hipEvent_t> ev_free;
    std::vector<Pending> pending;
    hipEvent_t train_ev0 = nullptr, train_ev1 = nullptr;  // the last cached frame's optimizer steps (timing)
    hipGraphExec_t batch_graph = nullptr;
    std::vector<uint8_t> batch_key;           // every value the cached batch graph's kernels were captured with
    std::vector<FrameBuffers*> branches;      // a batch's concurrent branches beyond the first (which uses fb)
    std::vector<hipStream_t> branch_streams;
    hipStream_t copy_stream = nullptr;  // device-to-host copies that overlap the main stream's kernels
    hipEvent_t copy_ev = nullptr;
    std::vector<hipEvent_t> branch_events;    // fork / join of the branches under capture
    struct WhittedBufs* whitted = nullptr;  // the Whitted tracer's ray-tree levels

    int prof_begin(hipEvent_t* a);
    int prof_end(hipEvent_t a, int kind);
    int prof_collect();  // synchronises and accumulates pending timings
};

struct rc_scene {
    rc_ctx* ctx = nullptr;
    rcd::NodeT* nodes = nullptr;  // NodeRec records, or QNode nodes when RC_QNODE
    size_t nrec = 0;
    rcd::TriRec* tris = nullptr;        // the FP32 decision's records (edges with RC_TRI_EDGES)
    rcd::TriRec* tris_exact = nullptr;  // v0, v1, v2 for the FP64 decision (tris itself without RC_TRI_EDGES)
    size_t ntri = 0;
    float4* shade = nullptr;  // normal.xyz, material index bits
    rcd::DevMaterial* mats = nullptr;
    size_t nmat = 0;
    float bmin[3], bmax[3];
    uint32_t depth = 0;
    size_t bytes = 0;
};

// The triangle records a traversal with the given decision reads.
template <bool DEC64>
inline const rcd::TriRec* tri_array(const rc_scene* s) {
    return DEC64 ? s->tris_exact : s->tris;
}

struct rc_film {
    rc_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0;
    float4* sum = nullptr;   // w*rgb, w
    float4* comp = nullptr;  // Kahan compensation terms
    DevBuf linear, partial;
    double max_lum_average = 1.0;  // ReinhardToneMapper's maxLuminanceAverage (ReinhardToneMapper.cpp:12-15)
};

// Launches the closest-hit (ANY = false) or any-hit traversal over a device
// ray array, then the deep-stack kernel over the rays whose stack outgrew the
// on-chip depth.  n_dev (may be null) is a device count read by the kernel;
// n_bound bounds it and sizes the grid.
int rc_launch_trace(rc_ctx* ctx, const rc_scene* s, bool any, const rc_ray* rays, const uint32_t* n_dev,
                    uint32_t n_bound, rc_hit* hits, uint8_t* occ);
int rc_launch_deep(rc_ctx* ctx, const rc_scene* s, bool any, const rc_ray* rays, rc_hit* hits, uint8_t* occ);

// One traversal launch with its scratch given explicitly: the overflow list
// (n_bound entries), the deep-stack kernel's stack (rc_deep_stack_bytes())
// and the counters.  rc_launch_trace fills it from the context's buffers.
struct TraceLaunch {
    const rc_scene* scene = nullptr;
    bool any = false;
    const rc_ray* rays = nullptr;
    const uint32_t* n_dev = nullptr;
    uint32_t n_bound = 0;
    rc_hit* hits = nullptr;
    uint8_t* occ = nullptr;
    uint32_t* ovf_list = nullptr;
    uint32_t* deep_stack = nullptr;
    DevCounters* ctr = nullptr;
    hipStream_t stream = nullptr;
    int max_depth = rcd::kOnChipStack;  // on-chip stack entries before the deep-stack kernel
    uint32_t max_blocks = 1;            // grid limit; the kernel's grid-stride loop covers the rest
    bool counting = false;
    bool dec64 = false;
};
size_t rc_deep_stack_bytes();
// Clears the overflow count and launches the traversal kernel.
int rc_launch_trace_with(const TraceLaunch& L);
// The deep-stack kernel over the rays the last traversal launch listed.
int rc_launch_deep_with(const TraceLaunch& L);
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
// Radiance-cache device counters (uint32 words).  Words that kernels hit at
// the same time sit one 128-byte line apart: every block's reservation
// atomics on the query and the record counter, and every wave's read of the
// training stride (sharing a line with those atomics made the cache-vertex
// bounce several times slower).
constexpr uint32_t kCcQueries = 0;        // queries of the pass
constexpr uint32_t kCcRecords = 32;       // training records of the pass
constexpr uint32_t kCcSamples = 64;       // training samples of the frame
constexpr uint32_t kCcFrameQueries = 65;  // cache queries of the frame (kCcSamples + 1: copied out together)
constexpr uint32_t kCcLastSamples = 66;   // training samples of the last finished frame
constexpr uint32_t kCcSteps = 67;         // Adam steps taken
constexpr uint32_t kCcStride = 96;        // adaptive training-path stride
constexpr uint32_t kCacheCounters = 128;

struct CacheIO {
    float4* queries;        // 4 float4 per query: (p, slot), (w_o, thr.r), (n, thr.g), (albedo, thr.b)
    float4* records;        // 5 float4 per training vertex: the query layout, then (L before, 0)
    uint32_t* counters;     // kCacheCounters words (kCc*)
    float* batch_q;         // 12 floats per training sample
    float* batch_t;         // 3 floats per training sample
    uint32_t batch_cap;
    uint32_t cache_vertex;  // 0 = off
    uint32_t train_stride;  // one training path per train_stride slots; 0: the adaptive stride *stride_dev
    const uint32_t* stride_dev;
    uint32_t cap;           // record / query capacity of a pass
};

int rc_validate_cfg(const rc_render_cfg* cfg, const char* fn);
// Paths of one pass of a frame (pixels x samples per pass, clamped to the spp).
size_t rc_pass_paths(const rc_render_cfg* cfg, uint32_t spp_pass);
// Enqueues the passes of one frame on stream st with the buffers B (sized
// beforehand for the frame's passes); cache_io / cache are null for the plain
// renderer.  Profiling events are recorded only when st is the context stream.
int rc_enqueue_frame(rc_ctx* c, FrameBuffers& B, hipStream_t st, const rc_scene* s, rc_film* f, const rc_camera* cam,
                     const rc_render_cfg* cfg, const FrameShape& shape, rc_progress_fn progress, void* user,
                     const CacheIO* cache_io = nullptr, rc_cache* cache = nullptr);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (current), once per device (rc_noalloc.hip).
int rc_noalloc_setup(int device);
// One frame with an explicit launch shape, synchronously (the autotuner).
int rc_render_shaped(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
                     const FrameShape& shape);
// Zeroes a frame's statistics counters (a kernel, so it may be captured).
void rc_zero_counters(DevCounters* ctr, hipStream_t st);
// Radiance cache, device side (rc_mlp.hip).
// L[slot] += thr * cache(query) for the *n_dev (<= n_bound) queries, on
// stream st once the optimizer steps in flight are done.
int rc_cache_apply_device(rc_ctx* c, hipStream_t st, rc_cache* k, const float4* queries, const uint32_t* n_dev,
                          uint32_t n_bound, float4* L);
// Sizes the cache's frame buffers for passes of pass_paths slots and a frame
// of frame_samples samples, fills io and enqueues the clearing of its
// per-frame counters on st.
int rc_cache_frame_io(rc_ctx* c, hipStream_t st, rc_cache* k, uint32_t cache_vertex, size_t pass_paths,
                      uint64_t frame_samples, CacheIO* io);
// After a frame's passes on st: freezes the frame's training samples, steers
// the adaptive stride and queues `steps` optimizer steps on the cache's own
// stream behind the frame; synchronises st and returns the frame's cache
// queries (counts[0]) and training samples (counts[1]).
int rc_cache_frame_end(rc_ctx* c, hipStream_t st, rc_cache* k, uint32_t steps, uint32_t counts[2]);
// Mean loss of the last frame's optimizer steps (waits for them).
int rc_cache_last_loss(rc_ctx* c, rc_cache* k, float* loss);
rc_ctx* rc_cache_owner(const rc_cache* k);
int rc_check_ctx(const rc_ctx* ctx, const char* fn);
int rc_check_scene(const rc_ctx* ctx, const rc_scene* s, const char* fn);
