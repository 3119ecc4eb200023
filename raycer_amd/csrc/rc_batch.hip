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
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

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

namespace {
int render_batch(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                 const rc_render_cfg* cfgs, size_t n, const uint32_t* spp_pass, uint32_t spp_all, const char* fn) {
    RC_TRY(rc_check_ctx(c, fn));
    if (n == 0) return RC_OK;
    if (!scenes || !films || !cams || !cfgs) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null array");
    std::vector<uint32_t> sp(n);
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_check_scene(c, scenes[i], fn));
        RC_TRY(rc_validate_cfg(&cfgs[i], fn));
        if (!films[i] || films[i]->ctx != c || films[i]->width != cfgs[i].width || films[i]->height != cfgs[i].height)
            return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": film " + std::to_string(i) + " does not match its config");
        if (spp_pass) {
            if (spp_pass[i] == 0 || spp_pass[i] > 64)
                return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": spp_per_pass must be in [1, 64]");
            sp[i] = spp_pass[i];
        } else {
            rc_tuning t{};
            t.spp_per_pass = spp_all;
            sp[i] = pass_samples(t, cfgs[i].width * cfgs[i].height, cfgs[i].spp, kPassPaths);
        }
    }
    RC_TRY_HIP(hipSetDevice(c->device));
    // size every scratch buffer for the largest frame before capturing
    for (size_t i = 0; i < n; ++i) RC_TRY(c->fb.ensure(rc_pass_paths(&cfgs[i], sp[i])));
    bool was_profiling = c->profiling;
    c->profiling = false;
    auto t0 = std::chrono::steady_clock::now();
    hipGraph_t graph = nullptr;
    RC_TRY_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    k_batch_zero<<<1, 1, 0, c->stream>>>(c->counters);
    int rc = RC_OK;
    for (size_t i = 0; i < n && rc == RC_OK; ++i) {
        FrameShape shape;
        shape.spp_pass = sp[i];
        rc = rc_enqueue_frame(c, c->fb, c->stream, scenes[i], films[i], &cams[i], &cfgs[i], shape, nullptr, nullptr);
    }
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
    if (h.error) return rc_host_fail(RC_ERR_HIP, std::string(fn) + ": traversal stack exceeded the deep-stack limit");
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->stats.extension_rays = h.ext_rays;
    c->stats.shadow_rays = h.shadow_rays;
    uint64_t samples = 0;
    for (size_t i = 0; i < n; ++i) samples += uint64_t(cfgs[i].width) * cfgs[i].height * cfgs[i].spp;
    c->stats.samples = samples;
    c->stats.overflow_rays = h.ovf_count;
    return RC_OK;
}
}  // namespace

extern "C" int rc_render_batch(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                               const rc_render_cfg* cfgs, size_t n, const rc_tuning* tuning) {
    RC_TRY(rc_check_ctx(c, "rc_render_batch"));
    const uint32_t sp = tuning ? tuning->spp_per_pass : c->tuning.spp_per_pass;
    return render_batch(c, scenes, films, cams, cfgs, n, nullptr, sp, "rc_render_batch");
}

extern "C" int rc_render_batch_plan(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films,
                                    const rc_camera* cams, const rc_render_cfg* cfgs, size_t n,
                                    const uint32_t* spp_per_pass) {
    if (!spp_per_pass && n) return rc_host_fail(RC_ERR_INVALID, "rc_render_batch_plan: null spp_per_pass");
    return render_batch(c, scenes, films, cams, cfgs, n, spp_per_pass, 1, "rc_render_batch_plan");
}

// ------------------------------------------------------------------ autotuner
// Chooses the samples per pass (the wavefront queue length is pixels x
// samples per pass) for each size class of a batch.  A size class is
// ceil(log2(pixels)); for each class present, one of its scenes is rendered
// in full with every candidate (1, 2, 4, ... 64, at most the frame's spp and
// at most 2^25 paths per pass) and the fastest wins.  Results are cached in a
// JSON object {"<device>/<class>": spp_per_pass, ...} when a path is given.
namespace {
constexpr uint64_t kMaxPassPaths = 1ull << 25;

uint32_t size_class(const rc_render_cfg& c) {
    uint64_t px = uint64_t(c.width) * c.height, k = 0;
    while ((1ull << k) < px) ++k;
    return uint32_t(k);
}

std::map<std::string, uint32_t> read_tuning(const std::string& path) {
    std::map<std::string, uint32_t> m;
    std::ifstream f(path);
    if (!f.good()) return m;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    size_t i = 0;
    while ((i = s.find('"', i)) != std::string::npos) {
        size_t e = s.find('"', i + 1);
        if (e == std::string::npos) break;
        std::string key = s.substr(i + 1, e - i - 1);
        size_t colon = s.find(':', e);
        if (colon == std::string::npos) break;
        char* end = nullptr;
        unsigned long v = std::strtoul(s.c_str() + colon + 1, &end, 10);
        if (end != s.c_str() + colon + 1 && v >= 1 && v <= 64) m[key] = uint32_t(v);
        i = end ? size_t(end - s.c_str()) : e + 1;
    }
    return m;
}

int write_tuning(const std::string& path, const std::map<std::string, uint32_t>& m) {
    std::ofstream f(path, std::ios::trunc);
    if (!f.good()) return rc_host_fail(RC_ERR_INVALID, "rc_autotune_batch: cannot write " + path);
    f << "{";
    bool first = true;
    for (auto& kv : m) {
        f << (first ? "\n  \"" : ",\n  \"") << kv.first << "\": " << kv.second;
        first = false;
    }
    f << "\n}\n";
    return RC_OK;
}
}  // namespace

extern "C" int rc_autotune_batch(rc_ctx* c, const rc_scene* const* scenes, const rc_camera* cams,
                                 const rc_render_cfg* cfgs, size_t n, const char* cache_path,
                                 uint32_t* spp_per_pass) {
    RC_TRY(rc_check_ctx(c, "rc_autotune_batch"));
    if (n && (!scenes || !cams || !cfgs || !spp_per_pass))
        return rc_host_fail(RC_ERR_INVALID, "rc_autotune_batch: null array");
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_check_scene(c, scenes[i], "rc_autotune_batch"));
        RC_TRY(rc_validate_cfg(&cfgs[i], "rc_autotune_batch"));
    }
    std::map<std::string, uint32_t> table;
    if (cache_path && *cache_path) table = read_tuning(cache_path);
    bool changed = false;
    std::map<uint32_t, uint32_t> chosen;  // class -> spp per pass
    const rc_tuning saved = c->tuning;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t cls = size_class(cfgs[i]);
        if (chosen.count(cls)) continue;
        const std::string key = c->name + "/" + std::to_string(cls);
        auto it = table.find(key);
        if (it != table.end()) {
            chosen[cls] = it->second;
            continue;
        }
        // tune on this scene, the first of its class
        rc_film* film = nullptr;
        RC_TRY(rc_film_create(c, cfgs[i].width, cfgs[i].height, &film));
        double best_ms = 1e300;
        uint32_t best = 1;
        int rc = RC_OK;
        const uint64_t px = uint64_t(cfgs[i].width) * cfgs[i].height;
        for (uint32_t sp = 1; sp <= 64 && rc == RC_OK; sp *= 2) {
            if (sp > 1 && (sp > cfgs[i].spp || px * sp > kMaxPassPaths)) break;
            c->tuning.spp_per_pass = sp;
            double ms = 1e300;
            for (int rep = 0; rep < 2 && rc == RC_OK; ++rep) {  // the first run also sizes the buffers
                auto t0 = std::chrono::steady_clock::now();
                rc = rc_render_progress(c, scenes[i], film, &cams[i], &cfgs[i], nullptr, nullptr);
                ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            if (rc == RC_OK && ms < best_ms * 0.98) {  // a larger queue must win by 2%
                best_ms = ms;
                best = sp;
            }
        }
        c->tuning = saved;
        rc_film_free(film);
        if (rc != RC_OK) return rc;
        chosen[cls] = best;
        table[key] = best;
        changed = true;
    }
    for (size_t i = 0; i < n; ++i) spp_per_pass[i] = chosen[size_class(cfgs[i])];
    if (changed && cache_path && *cache_path) RC_TRY(write_tuning(cache_path, table));
    return RC_OK;
}
