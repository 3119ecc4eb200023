// Batched multi-scene rendering in one captured HIP graph.  No counterpart
// in the reference (its closest relative is NetworkRunner's static split of
// one job over servers, src/Runners/NetworkRunner.cpp:250-259).
//
// Every scene's frame (camera, bounces, shadow rays, film) is enqueued on the
// context stream under stream capture, the graph is instantiated once and
// launched once.  All device buffers are sized for the largest scene before
// capture starts, so no allocation happens inside the graph, and counters are
// cleared by kernels (no memset nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "rc_internal.hpp"

namespace {
__global__ void k_batch_zero(DevCounters* ctr) {
    ctr->node_records = 0;
    ctr->tri_tests = 0;
    ctr->ext_rays = 0;
    ctr->shadow_rays = 0;
    ctr->ovf_count = 0;
    ctr->error = 0;
}
}  // namespace

extern "C" int rc_render_batch(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                               const rc_render_cfg* cfgs, size_t n, const rc_tuning* tuning) {
    RC_TRY(rc_check_ctx(c, "rc_render_batch"));
    if (n == 0) return RC_OK;
    if (!scenes || !films || !cams || !cfgs) return rc_host_fail(RC_ERR_INVALID, "rc_render_batch: null array");
    rc_tuning t = tuning ? *tuning : c->tuning;
    if (t.spp_per_pass == 0) t.spp_per_pass = 1;
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_check_scene(c, scenes[i], "rc_render_batch"));
        RC_TRY(rc_validate_cfg(&cfgs[i], "rc_render_batch"));
        if (!films[i] || films[i]->ctx != c || films[i]->width != cfgs[i].width || films[i]->height != cfgs[i].height)
            return rc_host_fail(RC_ERR_INVALID, "rc_render_batch: film " + std::to_string(i) + " does not match its config");
    }
    RC_TRY_HIP(hipSetDevice(c->device));
    // size every scratch buffer for the largest frame before capturing
    for (size_t i = 0; i < n; ++i) RC_TRY(rc_prepare_frame(c, &cfgs[i], t.spp_per_pass));
    bool was_profiling = c->profiling;
    c->profiling = false;
    auto t0 = std::chrono::steady_clock::now();
    hipGraph_t graph = nullptr;
    RC_TRY_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    k_batch_zero<<<1, 1, 0, c->stream>>>(c->counters);
    int rc = RC_OK;
    for (size_t i = 0; i < n && rc == RC_OK; ++i)
        rc = rc_enqueue_frame(c, scenes[i], films[i], &cams[i], &cfgs[i], t.spp_per_pass, nullptr, nullptr);
    hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    c->profiling = was_profiling;
    if (rc != RC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return rc_host_fail(RC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return rc_host_fail(RC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    if (c->batch_graph) (void)hipGraphExecDestroy(c->batch_graph);
    c->batch_graph = exec;
    RC_TRY_HIP(hipGraphLaunch(exec, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    DevCounters h;
    RC_TRY_HIP(hipMemcpy(&h, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost));
    if (h.error) return rc_host_fail(RC_ERR_HIP, "rc_render_batch: traversal stack exceeded the deep-stack limit");
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->stats.extension_rays = h.ext_rays;
    c->stats.shadow_rays = h.shadow_rays;
    uint64_t samples = 0;
    for (size_t i = 0; i < n; ++i) samples += uint64_t(cfgs[i].width) * cfgs[i].height * cfgs[i].spp;
    c->stats.samples = samples;
    c->stats.overflow_rays = h.ovf_count;
    return RC_OK;
}
