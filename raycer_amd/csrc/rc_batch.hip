// Batched multi-scene rendering in one captured HIP graph, and its
// autotuner.  No counterpart in the reference (its closest relative is
// NetworkRunner's static split of one job over servers,
// src/Runners/NetworkRunner.cpp:250-259, and CLTracer's untuned local size,
// src/OpenCL/CLTracer.cpp:288-297).
//
// Every scene's frame (bounces, shadow rays, film) is enqueued under stream
// capture.  Frames are dealt to up to eight branches, longest first to the
// least loaded; each branch is its own stream with its own wavefront buffers
// and counters, forked from and joined back to the context stream, so small
// frames overlap each other's tails.  All buffers are sized before capture
// (no allocation inside the graph) and counters are cleared by kernels (no
// memset nodes).  The executable graph is kept with a key of every value its
// kernels were captured with: the same batch relaunches it, a changed batch is
// captured again and the executable updated in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "rc_internal.hpp"

namespace {

constexpr uint32_t kMaxBranches = 8;

template <class T>
void put(std::vector<uint8_t>& k, const T& v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

void put_buffers(std::vector<uint8_t>& k, const FrameBuffers& B) {
    const DevBuf* b[] = {&B.ray_a, &B.ray_b, &B.st_a, &B.st_b, &B.hits, &B.srays, &B.scontrib, &B.socc, &B.slotL,
                         &B.slotW, &B.ovf_list, &B.deep_stack, &B.park};
    for (const DevBuf* x : b) put(k, x->p);
    put(k, B.counters);
}

// Estimated cost of a frame: its samples times the depth of its tree.
double frame_cost(const rc_scene* s, const rc_render_cfg& c) {
    return double(c.width) * c.height * c.spp * std::log2(double(s->ntri) + 2.0);
}

int check_plan(const rc_batch_plan& p, const char* fn, rc_batch_plan& out) {
    out = p;
    if (out.ext_block == 0) out.ext_block = 256;
    if (out.shadow_block == 0) out.shadow_block = 128;
    if (out.spp_per_pass == 0 || out.spp_per_pass > 64)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": spp_per_pass must be in [1, 64]");
    if (out.ext_block != 128 && out.ext_block != 256)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": ext_block must be 128 or 256");
    if (out.shadow_block != 64 && out.shadow_block != 128 && out.shadow_block != 256)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": shadow_block must be 64, 128 or 256");
    out.pad = 0;
    return RC_OK;
}

int ensure_branches(rc_ctx* c, uint32_t branches) {
    while (c->branches.size() + 1 < branches) {
        c->branches.push_back(new FrameBuffers());
        hipStream_t st = nullptr;
        RC_TRY_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        c->branch_streams.push_back(st);
    }
    while (c->branch_events.size() < 2 * kMaxBranches) {
        hipEvent_t e = nullptr;
        RC_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->branch_events.push_back(e);
    }
    return RC_OK;
}

FrameBuffers& branch_buffers(rc_ctx* c, uint32_t b) { return b == 0 ? c->fb : *c->branches[b - 1]; }
hipStream_t branch_stream(rc_ctx* c, uint32_t b) { return b == 0 ? c->stream : c->branch_streams[b - 1]; }

__global__ void k_batch_zero(DevCounters* ctr) {
    ctr->node_records = 0;
    ctr->tri_tests = 0;
    ctr->ext_rays = 0;
    ctr->shadow_rays = 0;
    ctr->ovf_total = 0;
    ctr->park_total = 0;
    ctr->ovf_count = 0;
    ctr->error = 0;
}

// Captures the batch's frames over `branches` branches (dealt by `assign`).
int capture(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
            const rc_render_cfg* cfgs, const std::vector<rc_batch_plan>& plans,
            const std::vector<std::vector<size_t>>& assign, hipGraph_t* graph) {
    const uint32_t nb = uint32_t(assign.size());
    RC_TRY_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = RC_OK;
    for (uint32_t b = 0; b < nb; ++b) k_batch_zero<<<1, 1, 0, c->stream>>>(branch_buffers(c, b).counters);
    if (nb > 1) {
        if (hipEventRecord(c->branch_events[0], c->stream) != hipSuccess) rc = RC_ERR_HIP;
        for (uint32_t b = 1; b < nb && rc == RC_OK; ++b)
            if (hipStreamWaitEvent(branch_stream(c, b), c->branch_events[0], 0) != hipSuccess) rc = RC_ERR_HIP;
    }
    for (uint32_t b = 0; b < nb && rc == RC_OK; ++b) {
        for (size_t i : assign[b]) {
            FrameShape shape;
            shape.spp_pass = plans[i].spp_per_pass;
            shape.ext_block = plans[i].ext_block;
            shape.shadow_block = plans[i].shadow_block;
            rc = rc_enqueue_frame(c, branch_buffers(c, b), branch_stream(c, b), scenes[i], films[i], &cams[i],
                                  &cfgs[i], shape, nullptr, nullptr);
            if (rc != RC_OK) break;
        }
    }
    if (nb > 1 && rc == RC_OK) {
        for (uint32_t b = 1; b < nb && rc == RC_OK; ++b) {
            if (hipEventRecord(c->branch_events[b], branch_stream(c, b)) != hipSuccess ||
                hipStreamWaitEvent(c->stream, c->branch_events[b], 0) != hipSuccess)
                rc = RC_ERR_HIP;
        }
    }
    hipError_t ec = hipStreamEndCapture(c->stream, graph);
    if (rc != RC_OK) {
        if (*graph) (void)hipGraphDestroy(*graph);
        *graph = nullptr;
        return rc == RC_ERR_HIP ? rc_host_fail(RC_ERR_HIP, "batch capture: fork or join failed") : rc;
    }
    if (ec != hipSuccess) return rc_host_fail(RC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    return RC_OK;
}

int render_batch(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                 const rc_render_cfg* cfgs, size_t n, const rc_batch_plan* plan_in, uint32_t branches,
                 const char* fn) {
    RC_TRY(rc_check_ctx(c, fn));
    if (n == 0) return RC_OK;
    if (!scenes || !films || !cams || !cfgs || !plan_in)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null array");
    if (branches == 0 || branches > kMaxBranches)
        return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": branches must be in [1, 8]");
    std::vector<rc_batch_plan> plans(n);
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_check_scene(c, scenes[i], fn));
        RC_TRY(rc_validate_cfg(&cfgs[i], fn));
        if (!films[i] || films[i]->ctx != c || films[i]->width != cfgs[i].width || films[i]->height != cfgs[i].height)
            return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": film " + std::to_string(i) +
                                                    " does not match its config");
        RC_TRY(check_plan(plan_in[i], fn, plans[i]));
    }
    branches = uint32_t(std::min<size_t>(branches, n));
    RC_TRY_HIP(hipSetDevice(c->device));
    RC_TRY(ensure_branches(c, branches));
    // longest processing time first: each frame to the least loaded branch
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t(0));
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return frame_cost(scenes[a], cfgs[a]) > frame_cost(scenes[b], cfgs[b]);
    });
    std::vector<std::vector<size_t>> assign(branches);
    std::vector<double> load(branches, 0.0);
    for (size_t i : order) {
        const uint32_t b = uint32_t(std::min_element(load.begin(), load.end()) - load.begin());
        assign[b].push_back(i);
        load[b] += frame_cost(scenes[i], cfgs[i]);
    }
    // every branch's buffers sized for its largest pass, before any capture
    for (uint32_t b = 0; b < branches; ++b)
        for (size_t i : assign[b]) RC_TRY(branch_buffers(c, b).ensure(rc_pass_paths(&cfgs[i], plans[i].spp_per_pass)));
    // the key: every value the captured kernels would embed
    std::vector<uint8_t> key;
    put(key, n);
    put(key, branches);
    put(key, c->counting);
    put(key, c->hit_fp64);
    put(key, c->stack_depth());
    put(key, c->tuning.park_steps);
    for (uint32_t b = 0; b < branches; ++b) {
        put_buffers(key, branch_buffers(c, b));
        put(key, assign[b].size());
        for (size_t i : assign[b]) put(key, i);
    }
    for (size_t i = 0; i < n; ++i) {
        put(key, scenes[i]->nodes);
        put(key, scenes[i]->tris);
        put(key, scenes[i]->shade);
        put(key, scenes[i]->mats);
        put(key, films[i]->sum);
        put(key, films[i]->comp);
        put(key, cams[i]);
        put(key, cfgs[i]);
        put(key, plans[i]);
    }
    const bool was_profiling = c->profiling;
    c->stats.graph_ms = 0.0;
    c->stats.graph_relaunched = 0;
    if (c->batch_graph && key == c->batch_key) {
        c->stats.graph_relaunched = 1;
    } else {
        c->profiling = false;
        auto tg = std::chrono::steady_clock::now();
        hipGraph_t graph = nullptr;
        int rc = capture(c, scenes, films, cams, cfgs, plans, assign, &graph);
        c->profiling = was_profiling;
        if (rc != RC_OK) return rc;
        bool updated = false;
        if (c->batch_graph) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res;
            updated = hipGraphExecUpdate(c->batch_graph, graph, &err_node, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(c->batch_graph);
                c->batch_graph = nullptr;
            }
        }
        if (!updated) {
            hipError_t ei = hipGraphInstantiate(&c->batch_graph, graph, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(graph);
                c->batch_graph = nullptr;
                c->batch_key.clear();
                return rc_host_fail(RC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            }
        }
        (void)hipGraphDestroy(graph);
        c->batch_key = key;
        c->stats.graph_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
    }
    auto t0 = std::chrono::steady_clock::now();
    RC_TRY_HIP(hipGraphLaunch(c->batch_graph, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    c->stats.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    uint64_t ext = 0, shadow = 0, ovf = 0, parked = 0;
    for (uint32_t b = 0; b < branches; ++b) {
        DevCounters h;
        RC_TRY_HIP(hipMemcpy(&h, branch_buffers(c, b).counters, sizeof(DevCounters), hipMemcpyDeviceToHost));
        if (h.error) return rc_host_fail(RC_ERR_HIP, std::string(fn) + ": traversal stack exceeded the deep-stack limit");
        ext += h.ext_rays;
        shadow += h.shadow_rays;
        ovf += h.ovf_total;
        parked += h.park_total;
    }
    c->stats.extension_rays = ext;
    c->stats.shadow_rays = shadow;
    c->stats.overflow_rays = ovf;
    c->stats.parked_rays = parked;
    uint64_t samples = 0;
    for (size_t i = 0; i < n; ++i) samples += uint64_t(cfgs[i].width) * cfgs[i].height * cfgs[i].spp;
    c->stats.samples = samples;
    return RC_OK;
}

std::vector<rc_batch_plan> spp_plans(size_t n, const uint32_t* spp) {
    std::vector<rc_batch_plan> p(n);
    for (size_t i = 0; i < n; ++i) p[i] = rc_batch_plan{spp[i], 256, 128, 0};
    return p;
}

}  // namespace

extern "C" int rc_render_batch(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                               const rc_render_cfg* cfgs, size_t n, const rc_tuning* tuning) {
    RC_TRY(rc_check_ctx(c, "rc_render_batch"));
    if (n && !cfgs) return rc_host_fail(RC_ERR_INVALID, "rc_render_batch: null array");
    rc_tuning t = c->tuning;
    if (tuning) t.spp_per_pass = tuning->spp_per_pass;
    std::vector<uint32_t> sp(n);
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_validate_cfg(&cfgs[i], "rc_render_batch"));
        sp[i] = pass_samples(t, cfgs[i].width * cfgs[i].height, cfgs[i].spp, kPassPaths);
    }
    const std::vector<rc_batch_plan> p = spp_plans(n, sp.data());
    return render_batch(c, scenes, films, cams, cfgs, n, p.data(), 1, "rc_render_batch");
}

extern "C" int rc_render_batch_plan(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films,
                                    const rc_camera* cams, const rc_render_cfg* cfgs, size_t n,
                                    const uint32_t* spp_per_pass) {
    if (!spp_per_pass && n) return rc_host_fail(RC_ERR_INVALID, "rc_render_batch_plan: null spp_per_pass");
    const std::vector<rc_batch_plan> p = spp_plans(n, spp_per_pass);
    return render_batch(c, scenes, films, cams, cfgs, n, p.data(), 1, "rc_render_batch_plan");
}

extern "C" int rc_render_batch_tuned(rc_ctx* c, const rc_scene* const* scenes, rc_film* const* films,
                                     const rc_camera* cams, const rc_render_cfg* cfgs, size_t n,
                                     const rc_batch_plan* plans, uint32_t branches) {
    return render_batch(c, scenes, films, cams, cfgs, n, plans, branches, "rc_render_batch_tuned");
}

// ------------------------------------------------------------------ autotuner
// Per size class (ceil(log2(pixels))) the first scene of the class is
// rendered in full: with 1, 2, 4, ... 64 samples per pass (at most the frame's
// spp and 2^25 paths per pass), then at the best of those with every block
// shape.  A candidate must beat the incumbent by 2% (larger queues and
// non-default shapes have to earn their place).  Then the whole batch is
// timed as relaunches of its graph with 1, 2, 4 and 8 branches.
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
        if (end != s.c_str() + colon + 1 && v >= 1 && v <= 4096) m[key] = uint32_t(v);
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

// Wall time of one full frame with the given shape (the second of two runs:
// the first also sizes the buffers).
int time_frame(rc_ctx* c, const rc_scene* s, rc_film* f, const rc_camera* cam, const rc_render_cfg* cfg,
               const FrameShape& shape, double* ms) {
    for (int rep = 0; rep < 2; ++rep) {
        auto t0 = std::chrono::steady_clock::now();
        RC_TRY(rc_render_shaped(c, s, f, cam, cfg, shape));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return RC_OK;
}

int tune(rc_ctx* c, const rc_scene* const* scenes, const rc_camera* cams, const rc_render_cfg* cfgs, size_t n,
         const char* cache_path, bool shapes, rc_batch_plan* plans, uint32_t* branches, const char* fn) {
    RC_TRY(rc_check_ctx(c, fn));
    if (n && (!scenes || !cams || !cfgs || !plans)) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null array");
    for (size_t i = 0; i < n; ++i) {
        RC_TRY(rc_check_scene(c, scenes[i], fn));
        RC_TRY(rc_validate_cfg(&cfgs[i], fn));
    }
    std::map<std::string, uint32_t> table;
    if (cache_path && *cache_path) table = read_tuning(cache_path);
    bool changed = false;
    std::map<uint32_t, rc_batch_plan> chosen;  // class -> plan
    for (size_t i = 0; i < n; ++i) {
        const uint32_t cls = size_class(cfgs[i]);
        if (chosen.count(cls)) continue;
        const std::string key = c->name + "/" + std::to_string(cls);
        rc_batch_plan p{0, 256, 128, 0};
        auto it = table.find(key + "/spp");
        if (it == table.end()) it = table.find(key);  // the samples-per-pass-only form
        const bool have_spp = it != table.end();
        if (have_spp) p.spp_per_pass = it->second;
        const bool have_shape = table.count(key + "/ext") && table.count(key + "/shadow");
        if (have_shape) {
            p.ext_block = table[key + "/ext"];
            p.shadow_block = table[key + "/shadow"];
        }
        if (have_spp && (have_shape || !shapes)) {
            chosen[cls] = p;
            continue;
        }
        rc_film* film = nullptr;
        RC_TRY(rc_film_create(c, cfgs[i].width, cfgs[i].height, &film));
        int rc = RC_OK;
        const uint64_t px = uint64_t(cfgs[i].width) * cfgs[i].height;
        if (!have_spp) {
            double best_ms = 1e300;
            for (uint32_t sp = 1; sp <= 64 && rc == RC_OK; sp *= 2) {
                if (sp > 1 && (sp > cfgs[i].spp || px * sp > kMaxPassPaths)) break;
                FrameShape sh;
                sh.spp_pass = sp;
                double ms = 1e300;
                rc = time_frame(c, scenes[i], film, &cams[i], &cfgs[i], sh, &ms);
                if (rc == RC_OK && ms < best_ms * 0.98) {
                    best_ms = ms;
                    p.spp_per_pass = sp;
                }
            }
        }
        if (shapes && !have_shape && rc == RC_OK) {
            FrameShape sh;
            sh.spp_pass = p.spp_per_pass;
            double best_ms = 1e300;
            rc = time_frame(c, scenes[i], film, &cams[i], &cfgs[i], sh, &best_ms);  // the default shape
            for (uint32_t eb : {256u, 128u})
                for (uint32_t sb : {128u, 64u, 256u}) {
                    if (rc != RC_OK || (eb == 256 && sb == 128)) continue;
                    sh.ext_block = eb;
                    sh.shadow_block = sb;
                    double ms = 1e300;
                    rc = time_frame(c, scenes[i], film, &cams[i], &cfgs[i], sh, &ms);
                    if (rc == RC_OK && ms < best_ms * 0.98) {
                        best_ms = ms;
                        p.ext_block = eb;
                        p.shadow_block = sb;
                    }
                }
        }
        rc_film_free(film);
        if (rc != RC_OK) return rc;
        chosen[cls] = p;
        table[key + "/spp"] = p.spp_per_pass;
        if (shapes) {
            table[key + "/ext"] = p.ext_block;
            table[key + "/shadow"] = p.shadow_block;
        }
        changed = true;
    }
    for (size_t i = 0; i < n; ++i) plans[i] = chosen[size_class(cfgs[i])];
    if (branches) {
        *branches = 1;
        const std::string bkey = c->name + "/branches/" + std::to_string(n);
        auto it = table.find(bkey);
        if (it != table.end()) {
            *branches = std::max(1u, std::min(it->second, kMaxBranches));
        } else if (n > 1) {
            std::vector<rc_film*> films(n, nullptr);
            int rc = RC_OK;
            for (size_t i = 0; i < n && rc == RC_OK; ++i) rc = rc_film_create(c, cfgs[i].width, cfgs[i].height, &films[i]);
            double best_ms = 1e300;
            for (uint32_t b = 1; b <= kMaxBranches && b <= n && rc == RC_OK; b *= 2) {
                double ms = 1e300;
                for (int rep = 0; rep < 2 && rc == RC_OK; ++rep) {  // the first captures, the second relaunches
                    rc = render_batch(c, scenes, films.data(), cams, cfgs, n, plans, b, fn);
                    ms = c->stats.render_ms;
                }
                if (rc == RC_OK && ms < best_ms * 0.98) {
                    best_ms = ms;
                    *branches = b;
                }
            }
            for (rc_film* f : films)
                if (f) rc_film_free(f);
            if (rc != RC_OK) return rc;
            table[bkey] = *branches;
            changed = true;
        }
    }
    if (changed && cache_path && *cache_path) RC_TRY(write_tuning(cache_path, table));
    return RC_OK;
}
}  // namespace

extern "C" int rc_autotune_batch(rc_ctx* c, const rc_scene* const* scenes, const rc_camera* cams,
                                 const rc_render_cfg* cfgs, size_t n, const char* cache_path,
                                 uint32_t* spp_per_pass) {
    if (n && !spp_per_pass) return rc_host_fail(RC_ERR_INVALID, "rc_autotune_batch: null array");
    std::vector<rc_batch_plan> plans(n);
    RC_TRY(tune(c, scenes, cams, cfgs, n, cache_path, false, plans.data(), nullptr, "rc_autotune_batch"));
    for (size_t i = 0; i < n; ++i) spp_per_pass[i] = plans[i].spp_per_pass;
    return RC_OK;
}

extern "C" int rc_autotune_batch_plan(rc_ctx* c, const rc_scene* const* scenes, const rc_camera* cams,
                                      const rc_render_cfg* cfgs, size_t n, const char* cache_path,
                                      rc_batch_plan* plans, uint32_t* branches) {
    if (n && (!plans || !branches)) return rc_host_fail(RC_ERR_INVALID, "rc_autotune_batch_plan: null array");
    return tune(c, scenes, cams, cfgs, n, cache_path, true, plans, branches, "rc_autotune_batch_plan");
}
