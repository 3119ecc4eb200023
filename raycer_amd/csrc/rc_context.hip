// Context, error reporting, scene upload (FlatBVH repack) and device memory
// entry points.  Counterpart of CLManager + CLTracer::initializeBuffers +
// CLScene::readSceneFull (src/OpenCL/CLManager.cpp, CLTracer.cpp:173-238,
// CLScene.cpp:72-263).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>

#include "rc_internal.hpp"

// 1: the FlatBVH is collapsed into 4-wide nodes by the SAH-optimal dynamic
// program in rc_scene_upload (1% faster on the 2M-triangle terrain, 18% fewer
// nodes); 0: greedily, by opening the largest child.
#ifndef RC_COLLAPSE
#define RC_COLLAPSE 1
#endif
#ifndef RC_COLLAPSE_TRI
#define RC_COLLAPSE_TRI 2.0  // cost of a triangle test relative to a wide-node visit
#endif
#ifndef RC_COLLAPSE_LEAF
#define RC_COLLAPSE_LEAF 8   // most triangles the collapse gathers into one leaf
#endif

using namespace rcd;


int DevBuf::ensure(size_t need) {
    if (need <= bytes) return RC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t alloc = std::max<size_t>(need, 256);
    hipError_t e = hipMalloc(&p, alloc);
    if (e != hipSuccess) {
        p = nullptr;
        return rc_host_fail(RC_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(alloc) + "): " + hipGetErrorString(e));
    }
    bytes = alloc;
    return RC_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int rc_check_ctx(const rc_ctx* ctx, const char* fn) {
    if (!ctx) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null context");
    return RC_OK;
}

int rc_check_scene(const rc_ctx* ctx, const rc_scene* s, const char* fn) {
    RC_TRY(rc_check_ctx(ctx, fn));
    if (!s) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": null scene");
    if (s->ctx != ctx) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": scene belongs to another context");
    return RC_OK;
}

int rc_ctx::prof_begin(hipEvent_t* a) {
    *a = nullptr;
    if (!profiling) return RC_OK;
    if (ev_free.empty()) {
        hipEvent_t e;
        RC_TRY_HIP(hipEventCreate(&e));
        ev_free.push_back(e);
    }
    *a = ev_free.back();
    ev_free.pop_back();
    RC_TRY_HIP(hipEventRecord(*a, stream));
    return RC_OK;
}

int rc_ctx::prof_end(hipEvent_t a, int kind) {
    if (!profiling || !a) return RC_OK;
    if (ev_free.empty()) {
        hipEvent_t e;
        RC_TRY_HIP(hipEventCreate(&e));
        ev_free.push_back(e);
    }
    hipEvent_t b = ev_free.back();
    ev_free.pop_back();
    RC_TRY_HIP(hipEventRecord(b, stream));
    pending.push_back({a, b, kind});
    return RC_OK;
}

int rc_ctx::prof_collect() {
    RC_TRY_HIP(hipStreamSynchronize(stream));
    for (const Pending& p : pending) {
        float ms = 0.0f;
        RC_TRY_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        switch (p.kind) {
            case PK_EXTEND: stats.extend_ms += ms; stats.extend_launches++; break;
            case PK_SHADOW: stats.shadow_ms += ms; stats.shadow_launches++; break;
            case PK_SHADE: stats.shade_ms += ms; break;
            case PK_CACHE: stats.cache_ms += ms; break;
            default: stats.other_ms += ms; break;
        }
        ev_free.push_back(p.a);
        ev_free.push_back(p.b);
    }
    pending.clear();
    return RC_OK;
}

extern "C" int rc_create(int device, rc_ctx** out) {
    if (!out) return rc_host_fail(RC_ERR_INVALID, "rc_create: null out");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return rc_host_fail(RC_ERR_NO_DEVICE, std::string("rc_create: no HIP device (") + hipGetErrorString(e) + ")");
    if (device < 0 || device >= n) return rc_host_fail(RC_ERR_INVALID, "rc_create: device index out of range");
    hipDeviceProp_t prop;
    RC_TRY_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return rc_host_fail(RC_ERR_NO_DEVICE, std::string("rc_create: device is ") + prop.gcnArchName +
                                                  ", this library is built for gfx950 only");
    RC_TRY_HIP(hipSetDevice(device));
    rc_ctx* c = new (std::nothrow) rc_ctx();
    if (!c) return rc_host_fail(RC_ERR_ALLOC, "rc_create: out of memory");
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->name = prop.name[0] ? prop.name : prop.gcnArchName;
    hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
        delete c;
        return rc_host_fail(RC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(se));
    }
    hipError_t me = hipMalloc(&c->counters, sizeof(DevCounters));
    if (me != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc_host_fail(RC_ERR_ALLOC, std::string("hipMalloc(counters): ") + hipGetErrorString(me));
    }
    (void)hipMemset(c->counters, 0, sizeof(DevCounters));
    c->fb.counters = c->counters;  // not owned by the frame buffers
    int se2 = rc_noalloc_setup(device);  // this device's copy of the allocation-free scratch
    if (se2 != RC_OK) {
        (void)hipFree(c->counters);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return se2;
    }
    *out = c;
    return RC_OK;
}

extern "C" int rc_destroy(rc_ctx* c) {
    if (!c) return RC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->batch_graph) (void)hipGraphExecDestroy(c->batch_graph);
    rc_whitted_release(c);
    DevBuf* bufs[] = {&c->trace_rays, &c->trace_hits, &c->trace_occ};
    for (DevBuf* b : bufs) b->release();
    c->fb.release();
    for (FrameBuffers* b : c->branches) {
        b->release();
        delete b;
    }
    c->branches.clear();
    for (hipStream_t b : c->branch_streams) (void)hipStreamDestroy(b);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->copy_ev) (void)hipEventDestroy(c->copy_ev);
    c->branch_streams.clear();
    for (hipEvent_t e : c->branch_events) (void)hipEventDestroy(e);
    c->branch_events.clear();
    for (auto& p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : c->ev_free) (void)hipEventDestroy(e);
    if (c->counters) (void)hipFree(c->counters);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RC_OK;
}

extern "C" int rc_device_name(rc_ctx* c, char* buf, size_t len, int* cu) {
    RC_TRY(rc_check_ctx(c, "rc_device_name"));
    if (buf && len) {
        size_t n = std::min(len - 1, c->name.size());
        std::memcpy(buf, c->name.data(), n);
        buf[n] = 0;
    }
    if (cu) *cu = c->cu_count;
    return RC_OK;
}

extern "C" int rc_set_profiling(rc_ctx* c, int enable) {
    RC_TRY(rc_check_ctx(c, "rc_set_profiling"));
    c->profiling = enable != 0;
    return RC_OK;
}

extern "C" int rc_set_counting(rc_ctx* c, int enable) {
    RC_TRY(rc_check_ctx(c, "rc_set_counting"));
    c->counting = enable != 0;
    return RC_OK;
}

extern "C" int rc_set_hit_decision(rc_ctx* c, int mode) {
    RC_TRY(rc_check_ctx(c, "rc_set_hit_decision"));
    if (mode != RC_HIT_FP32 && mode != RC_HIT_FP64)
        return rc_host_fail(RC_ERR_INVALID, "rc_set_hit_decision: mode must be RC_HIT_FP32 or RC_HIT_FP64");
    c->hit_fp64 = mode == RC_HIT_FP64;
    return RC_OK;
}

extern "C" int rc_set_tuning(rc_ctx* c, const rc_tuning* t) {
    RC_TRY(rc_check_ctx(c, "rc_set_tuning"));
    if (!t) return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: null tuning");
    if (t->block != 64 && t->block != 128 && t->block != 256)
        return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: block must be 64, 128 or 256");
    if (t->spp_per_pass > 64)
        return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: spp_per_pass must be 0 (automatic) or in [1, 64]");
    if (t->blocks_per_cu == 0 || t->blocks_per_cu > 64)
        return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: blocks_per_cu must be in [1, 64]");
    if (t->stack_limit > uint32_t(rcd::kOnChipStack))
        return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: stack_limit must be 0 (all) or in [1, " +
                                                std::to_string(rcd::kOnChipStack) + "]");
    if (t->park_steps > 1024)
        return rc_host_fail(RC_ERR_INVALID, "rc_set_tuning: park_steps must be 0 (never park) or in [1, 1024]");
    c->tuning = *t;
    return RC_OK;
}

extern "C" int rc_get_stats(rc_ctx* c, rc_stats* out) {
    RC_TRY(rc_check_ctx(c, "rc_get_stats"));
    if (!out) return rc_host_fail(RC_ERR_INVALID, "rc_get_stats: null out");
    RC_TRY(c->prof_collect());
    if (c->train_ev1) {  // the last cached frame's optimizer steps, timed on the cache's stream
        RC_TRY_HIP(hipEventSynchronize(c->train_ev1));
        float ms = 0.0f;
        RC_TRY_HIP(hipEventElapsedTime(&ms, c->train_ev0, c->train_ev1));
        c->stats.train_ms = ms;
    }
    *out = c->stats;
    return RC_OK;
}

extern "C" int rc_synchronize(rc_ctx* c) {
    RC_TRY(rc_check_ctx(c, "rc_synchronize"));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

// ---------------------------------------------------------------- scene upload
namespace {

float round_down(double x) {
    float f = float(x);
    if (double(f) > x) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double x) {
    float f = float(x);
    if (double(f) < x) f = std::nextafter(f, INFINITY);
    return f;
}

// One wide node (four child records) as a QNode.  Per axis the scale is the
// smallest power of two 2^e with 253 * 2^e >= the node's extent and 2^e >= 8
// ulps of its largest coordinate; children are quantized one unit outward on
// each side, so the decoded box origin + q * 2^e contains the exact one with a
// unit to spare, which covers the traversal's rounding of the decode.
void quantize_node(const NodeRec* r, QNode& q) {
    std::memset(&q, 0, sizeof(q));
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < kWidth; ++k) {
        if (r[k].count == kEmptyChild) continue;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], double(r[k].lo[a]));
            hi[a] = std::max(hi[a], double(r[k].hi[a]));
        }
    }
    int ebias[3];
    double scale[3];
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a])) lo[a] = hi[a] = 0.0;  // no child: any box
        const double ext = hi[a] - lo[a], mag = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
        int e = -126;
        if (ext > 0.0) e = std::max(e, int(std::ceil(std::log2(ext / 253.0))));
        if (mag > 0.0) e = std::max(e, int(std::floor(std::log2(mag))) - 23 + 3);
        while (std::ldexp(253.0, e) < ext) ++e;  // guard log2 rounding
        e = std::min(e, 127);
        ebias[a] = e + 127;
        scale[a] = std::ldexp(1.0, e);
        q.origin[a] = float(lo[a]);  // lo is a float (a child's bound)
    }
    q.exps = uint32_t(ebias[0]) | (uint32_t(ebias[1]) << 8) | (uint32_t(ebias[2]) << 16);
    for (int k = 0; k < kWidth; ++k) {
        const NodeRec& c = r[k];
        uint32_t meta;
        if (c.count == kEmptyChild) {
            meta = 0xFFu;
        } else {
            meta = c.count;  // 0 interior, else 1..31 triangles
            for (int a = 0; a < 3; ++a) {
                double ql = std::floor((double(c.lo[a]) - double(q.origin[a])) / scale[a]) - 1.0;
                double qh = std::ceil((double(c.hi[a]) - double(q.origin[a])) / scale[a]) + 1.0;
                ql = std::max(0.0, std::min(255.0, ql));
                qh = std::max(0.0, std::min(255.0, qh));
                q.qlo[a] |= uint32_t(ql) << (8 * k);
                q.qhi[a] |= uint32_t(qh) << (8 * k);
            }
        }
        q.meta |= meta << (8 * k);
        // the traversal's stack entry: an interior child's node index, a leaf
        // tagged with its count and first triangle, or kEmptyChild
        if (c.count == kEmptyChild)
            q.link[k] = kEmptyChild;
        else if (c.count == 0)
            q.link[k] = c.link / kWidth;
        else
            q.link[k] = kLeafTag | (c.count << 26) | c.link;
    }
}

}  // namespace

extern "C" int rc_scene_upload(rc_ctx* c, const float* v, size_t ntri, const uint32_t* tri_mat,
                               const rc_material* mats, size_t nmat, const rc_flat_node* nodes, size_t nnodes,
                               const uint64_t* ordered, rc_scene** out) {
    RC_TRY(rc_check_ctx(c, "rc_scene_upload"));
    if (!out || !mats || !nodes || nnodes == 0 || nmat == 0 || (ntri != 0 && (!v || !tri_mat || !ordered)))
        return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: null or empty input");
    // A scene without triangles has the single empty leaf FlatBVH::build makes
    // over no primitives (FlatBVH.cpp:112-139); it uploads as one wide node of
    // empty slots, so every ray misses.
    const bool empty = ntri == 0;
    if (empty && (nnodes != 1 || nodes[0].right_offset != 0 || nodes[0].primitive_count != 0))
        return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: a scene without triangles needs the single empty leaf");
    if (ntri > kMaxTriangles)
        return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: more than 2^26 triangles");
    *out = nullptr;
    // Validate the FlatBVH: a permutation of the triangles, leaves inside the
    // ordered range, right offsets pointing forward inside the array.
    {
        std::vector<uint8_t> seen(ntri, 0);
        for (size_t i = 0; i < ntri; ++i) {
            if (ordered[i] >= ntri || seen[ordered[i]])
                return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: orderedPrimitiveIds is not a permutation");
            seen[ordered[i]] = 1;
        }
        for (size_t i = 0; i < ntri; ++i)
            if (tri_mat[i] >= nmat) return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: material index out of range");
        for (size_t i = 0; i < (empty ? 0 : nnodes); ++i) {
            const rc_flat_node& n = nodes[i];
            if (n.right_offset == 0) {
                if (n.primitive_count == 0 || n.start_offset + n.primitive_count > ntri)
                    return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: leaf range outside the triangles");
            } else if (n.right_offset < 2 || i + uint64_t(n.right_offset) >= nnodes) {
                return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: bad rightOffset");
            }
        }
    }
    // Collapse the binary FlatBVH into 4-wide nodes: starting from an interior
    // node's two children, the interior child with the largest surface area is
    // replaced by its own two children until there are four.  Each wide node is
    // stored as kWidth consecutive child records in preorder; empty slots carry
    // count = kEmptyChild and are never entered.
    auto area = [&](size_t i) {
        const rc_flat_node& n = nodes[i];
        double ex = n.aabb_max[0] - n.aabb_min[0], ey = n.aabb_max[1] - n.aabb_min[1],
               ez = n.aabb_max[2] - n.aabb_min[2];
        return ex * ey + ey * ez + ez * ex;
    };
    auto is_leaf = [&](size_t i) { return nodes[i].right_offset == 0; };
#if RC_COLLAPSE
    // SAH-optimal collapse (dynamic programming over the binary tree):
    // cost[n][i-1] is the least expected cost of the subtree at n given as at
    // most i child slots, a wide node costing area * kNodeCost and a leaf
    // area * triangles * kTriCost.  An interior node whose leaves hold a
    // contiguous triangle range of at most kCollapseLeaf triangles may become
    // one leaf.  Children follow their parent in a FlatBVH, so one reverse
    // pass sees them first.
    constexpr double kNodeCost = 1.0, kTriCost = RC_COLLAPSE_TRI;
    constexpr uint64_t kCollapseLeaf = RC_COLLAPSE_LEAF;
    static_assert(kCollapseLeaf < kMaxLeafRecord, "collapsed leaves must fit a leaf record");
    std::vector<std::array<double, kWidth>> cost(empty ? 0 : nnodes);
    // pick[n][0]: 0 the subtree is one leaf, else it is a wide node whose
    // slots are the left child's best pick[n][0] and the right child's best
    // kWidth - pick[n][0]; pick[n][i - 1], i >= 2: 0 the best at most i slots
    // are the best at most i - 1, else the left child's share of them
    std::vector<std::array<uint8_t, kWidth>> pick(empty ? 0 : nnodes);
    std::vector<uint32_t> first(empty ? 0 : nnodes), last(empty ? 0 : nnodes);  // leaves' range; first > last: not contiguous
    for (size_t n = empty ? 0 : nnodes; n-- > 0;) {
        const rc_flat_node& x = nodes[n];
        const double a = area(n);
        auto& C = cost[n];
        auto& P = pick[n];
        if (is_leaf(n)) {
            first[n] = uint32_t(x.start_offset);  // validated: inside 2^26 triangles
            last[n] = uint32_t(x.start_offset + x.primitive_count);
            C.fill(a * double(x.primitive_count) * kTriCost);
            P.fill(0);
            continue;
        }
        const size_t L = n + 1, R = n + size_t(x.right_offset);
        const bool contiguous = first[L] <= last[L] && first[R] <= last[R] && last[L] == first[R];
        first[n] = contiguous ? first[L] : 1;
        last[n] = contiguous ? last[R] : 0;
        double dist[kWidth + 1];
        uint8_t share[kWidth + 1];
        for (int j = 2; j <= kWidth; ++j) {
            dist[j] = INFINITY;
            share[j] = 1;
            for (int k = 1; k < j; ++k) {
                const double c2 = cost[L][k - 1] + cost[R][j - k - 1];
                if (c2 < dist[j]) { dist[j] = c2; share[j] = uint8_t(k); }
            }
        }
        C[0] = a * kNodeCost + dist[kWidth];
        P[0] = share[kWidth];
        if (contiguous && last[n] - first[n] <= kCollapseLeaf) {
            const double lc = a * double(last[n] - first[n]) * kTriCost;
            if (lc <= C[0]) { C[0] = lc; P[0] = 0; }
        }
        for (int i = 2; i <= kWidth; ++i) {
            if (dist[i] < C[i - 2]) {
                C[i - 1] = dist[i];
                P[i - 1] = share[i];
            } else {
                C[i - 1] = C[i - 2];
                P[i - 1] = 0;
            }
        }
    }
    // a leaf slot: a FlatBVH leaf, or an interior node the collapse made one
    auto slot_leaf = [&](size_t i) { return is_leaf(i) || pick[i][0] == 0; };
    // the slots of the wide node at interior node n
    auto slots = [&](size_t n, size_t* out) -> int {
        int cnt = 0;
        struct Item { size_t node; int i; };
        Item st[2 * kWidth];
        int sp = 0;
        st[sp++] = {n + size_t(nodes[n].right_offset), kWidth - int(pick[n][0])};
        st[sp++] = {n + 1, int(pick[n][0])};
        while (sp > 0) {
            Item it = st[--sp];
            int i = it.i;
            while (i > 1 && pick[it.node][i - 1] == 0) --i;
            if (i == 1 || is_leaf(it.node)) {
                out[cnt++] = it.node;
                continue;
            }
            const int k = pick[it.node][i - 1];
            st[sp++] = {it.node + size_t(nodes[it.node].right_offset), i - k};
            st[sp++] = {it.node + 1, k};
        }
        return cnt;
    };
#endif
    std::vector<NodeRec> recs;
    recs.reserve(2 * nnodes / 3 + kWidth);
    // explicit stack of (binary node, record slot that links to its wide node);
    // a leaf of kMaxLeafRecord or more triangles becomes an interior record
    // whose wide node splits the range (split = true, range start / count)
    struct Pending { size_t node; size_t slot; bool split; uint64_t start, count; };
    std::vector<Pending> todo;
    auto fill = [&](size_t at, size_t child, uint32_t link) {
        NodeRec& r = recs[at];
        const rc_flat_node& n = nodes[child];
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = round_down(n.aabb_min[k]);
            r.hi[k] = round_up(n.aabb_max[k]);
        }
#if RC_COLLAPSE
        if (!is_leaf(child) && pick[child][0] == 0) {
            r.link = uint32_t(first[child]);
            r.count = uint32_t(last[child] - first[child]);
            return;
        }
#endif
        if (is_leaf(child) && n.primitive_count < kMaxLeafRecord) {
            r.link = uint32_t(n.start_offset);
            r.count = uint32_t(n.primitive_count);
        } else {
            r.link = link;
            r.count = 0;
            if (is_leaf(child)) todo.push_back({child, at, true, n.start_offset, n.primitive_count});
        }
    };
    // the wide node of a split leaf: up to four parts of the range, each a
    // leaf record or (still too long) another split; every part keeps the
    // leaf's box
    auto emit_split = [&](size_t leaf, uint64_t start, uint64_t count) -> uint32_t {
        uint32_t base = uint32_t(recs.size());
        recs.resize(recs.size() + kWidth);
        const uint64_t part = (count + kWidth - 1) / kWidth;
        for (int k = 0; k < kWidth; ++k) {
            NodeRec& r = recs[base + k];
            const uint64_t s0 = start + k * part, c0 = s0 < start + count ? std::min(part, start + count - s0) : 0;
            if (c0 == 0) {
                for (int q = 0; q < 3; ++q) { r.lo[q] = 0.0f; r.hi[q] = 0.0f; }
                r.link = 0;
                r.count = kEmptyChild;
                continue;
            }
            for (int q = 0; q < 3; ++q) {
                r.lo[q] = round_down(nodes[leaf].aabb_min[q]);
                r.hi[q] = round_up(nodes[leaf].aabb_max[q]);
            }
            if (c0 < kMaxLeafRecord) {
                r.link = uint32_t(s0);
                r.count = uint32_t(c0);
            } else {
                r.link = 0;
                r.count = 0;
                todo.push_back({leaf, base + size_t(k), true, s0, c0});
            }
        }
        return base;
    };
    auto emit = [&](size_t n) -> uint32_t {  // returns the first record of n's wide node
        size_t kids[kWidth];
        int nk = 0;
        if (is_leaf(n)) {
            kids[nk++] = n;  // a single-leaf tree: one leaf child
        } else if (RC_COLLAPSE) {
#if RC_COLLAPSE
            // only the root can be a node the collapse made a leaf: one leaf child
            if (pick[n][0] == 0)
                kids[nk++] = n;
            else
                nk = slots(n, kids);
#endif
        } else {
            kids[nk++] = n + 1;
            kids[nk++] = n + size_t(nodes[n].right_offset);
            while (nk < kWidth) {
                int best = -1;
                double ba = -1.0;
                for (int k = 0; k < nk; ++k)
                    if (!is_leaf(kids[k]) && area(kids[k]) > ba) { ba = area(kids[k]); best = k; }
                if (best < 0) break;
                size_t c = kids[best];
                kids[best] = c + 1;
                kids[nk++] = c + size_t(nodes[c].right_offset);
            }
        }
        uint32_t base = uint32_t(recs.size());
        recs.resize(recs.size() + kWidth);
        for (int k = 0; k < kWidth; ++k) {
            NodeRec& r = recs[base + k];
            if (k >= nk) {
                for (int q = 0; q < 3; ++q) { r.lo[q] = 0.0f; r.hi[q] = 0.0f; }
                r.link = 0;
                r.count = kEmptyChild;
                continue;
            }
            fill(base + k, kids[k], 0);
#if RC_COLLAPSE
            if (!slot_leaf(kids[k])) todo.push_back({kids[k], base + size_t(k), false, 0, 0});
#else
            if (!is_leaf(kids[k])) todo.push_back({kids[k], base + size_t(k), false, 0, 0});
#endif
        }
        return base;
    };
    if (empty) {
        recs.resize(kWidth);
        for (NodeRec& r : recs) {
            for (int q = 0; q < 3; ++q) { r.lo[q] = 0.0f; r.hi[q] = 0.0f; }
            r.link = 0;
            r.count = kEmptyChild;
        }
    } else {
        emit(0);
    }
    while (!todo.empty()) {
        Pending p = todo.back();
        todo.pop_back();
        uint32_t b = p.split ? emit_split(p.node, p.start, p.count) : emit(p.node);
        recs[p.slot].link = b;
    }
    size_t nrec = recs.size();
    if (nrec >= 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: too many nodes");
    // depth of the wide tree; links must point forward so traversal terminates
    uint32_t depth = 0;
    {
        std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 1u}};
        while (!st.empty()) {
            auto [r, d] = st.back();
            st.pop_back();
            depth = std::max(depth, d);
            for (int k = 0; k < kWidth; ++k) {
                const NodeRec& x = recs[r + k];
                if (x.count == 0) {
                    if (x.link <= r || x.link + kWidth > nrec || x.link % kWidth != 0)
                        return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: node link does not point forward");
                    st.push_back({x.link, d + 1});
                } else if (x.count != kEmptyChild &&
                           (x.count >= kMaxLeafRecord || uint64_t(x.link) + x.count > ntri)) {
                    return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: leaf record outside the triangles");
                }
            }
        }
    }
    // a traversal stack holds at most (kWidth - 1) entries per level
    if (depth * (kWidth - 1) >= uint32_t(kDeepStack))
        return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: BVH too deep for the traversal stack");
#if RC_NODE_ORDER
    // Each wide node's interior children side by side (one block per node),
    // the blocks in depth-first order: siblings a ray visits one after the
    // other share 128-byte lines, and a subtree stays in one region.
    {
        const size_t nn = nrec / kWidth;
        std::vector<uint32_t> order, st;
        order.reserve(nn);
        if (nn) {
            order.push_back(0);
            st.push_back(0);
        }
        while (!st.empty()) {
            const uint32_t n = st.back();
            st.pop_back();
            uint32_t kids[kWidth];
            int nk = 0;
            for (int k = 0; k < kWidth; ++k)
                if (recs[size_t(n) * kWidth + k].count == 0) kids[nk++] = recs[size_t(n) * kWidth + k].link / kWidth;
            for (int k = 0; k < nk; ++k) order.push_back(kids[k]);
            for (int k = nk - 1; k >= 0; --k) st.push_back(kids[k]);
        }
        if (order.size() != nn) return rc_host_fail(RC_ERR_INVALID, "rc_scene_upload: wide nodes not a tree");
        std::vector<uint32_t> to(nn);
        for (uint32_t t = 0; t < nn; ++t) to[order[t]] = t;
        std::vector<NodeRec> moved(nrec);
        for (size_t n = 0; n < nn; ++n)
            for (int k = 0; k < kWidth; ++k) {
                NodeRec r = recs[n * kWidth + k];
                if (r.count == 0) r.link = to[r.link / kWidth] * kWidth;
                moved[size_t(to[n]) * kWidth + k] = r;
            }
        recs.swap(moved);
    }
#endif
    // The top of the tree (breadth first, at most kTopNodes wide nodes) moves
    // to the front of the node array, so the traversal kernels can keep nodes
    // [0, ntop) in LDS; the rest keep their order.  Links still point forward.
    uint32_t ntop = 0;
    {
        const size_t nn = nrec / kWidth;
        std::vector<uint32_t> top;  // breadth-first from the root
        if (nn) top.push_back(0);
        for (size_t q = 0; q < top.size() && top.size() < kTopNodes; ++q) {
            const uint32_t r = top[q] * kWidth;
            for (int k = 0; k < kWidth && top.size() < kTopNodes; ++k)
                if (recs[r + k].count == 0) top.push_back(recs[r + k].link / kWidth);
        }
        ntop = uint32_t(top.size());
        std::vector<uint32_t> to(nn, 0xffffffffu);
        for (uint32_t t = 0; t < ntop; ++t) to[top[t]] = t;
        uint32_t next = ntop;
        for (size_t n = 0; n < nn; ++n)
            if (to[n] == 0xffffffffu) to[n] = next++;
        std::vector<NodeRec> moved(nrec);
        for (size_t n = 0; n < nn; ++n)
            for (int k = 0; k < kWidth; ++k) {
                NodeRec r = recs[n * kWidth + k];
                if (r.count == 0) r.link = to[r.link / kWidth] * kWidth;
                moved[size_t(to[n]) * kWidth + k] = r;
            }
        recs.swap(moved);
    }
#if RC_QNODE
    std::vector<QNode> qn(nrec / kWidth);
    for (size_t n = 0; n < qn.size(); ++n) quantize_node(&recs[n * kWidth], qn[n]);
    if (!qn.empty()) qn[0].pad = RC_LDS_TOP ? ntop : 0u;  // the root carries the size of the LDS top
    const void* node_src = qn.data();
    const size_t b_nodes = qn.size() * sizeof(QNode);
#else
    const void* node_src = recs.data();
    const size_t b_nodes = nrec * sizeof(NodeRec);
#endif
    std::vector<TriRec> tris(ntri), tri_edges(RC_TRI_EDGES ? ntri : 0);
    std::vector<float4> shade(ntri);
    std::vector<DevMaterial> dm(nmat);
    for (size_t i = 0; i < nmat; ++i) {
        const rc_material& m = mats[i];
        uint32_t fl = m.flags;
        float flf;
        std::memcpy(&flf, &fl, 4);
        dm[i].diffuse = make_float4(m.diffuse[0], m.diffuse[1], m.diffuse[2], flf);
        dm[i].emittance = make_float4(m.emittance[0], m.emittance[1], m.emittance[2], m.shininess);
        dm[i].ambient = make_float4(m.ambient[0], m.ambient[1], m.ambient[2], 0.0f);
        dm[i].specular = make_float4(m.specular[0], m.specular[1], m.specular[2], 0.0f);
    }
    float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < ntri; ++i) {
        const float* p = v + 9 * i;
        double e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = double(p[3 + k]) - double(p[k]);
            e2[k] = double(p[6 + k]) - double(p[k]);
            for (int q = 0; q < 3; ++q) {
                bmin[k] = std::min(bmin[k], p[3 * q + k]);
                bmax[k] = std::max(bmax[k], p[3 * q + k]);
            }
        }
        // Triangle::initialize normal = normalize(e1 x e2) in FP64 (Triangle.cpp:27)
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double inv = 1.0 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        uint32_t mi = tri_mat[i];
        float mf;
        std::memcpy(&mf, &mi, 4);
        shade[i] = make_float4(float(n[0] * inv), float(n[1] * inv), float(n[2] * inv), mf);
    }
    for (size_t j = 0; j < ntri; ++j) {
        uint32_t id = uint32_t(ordered[j]);
        const float* p = v + 9 * size_t(id);
        uint32_t fl = mats[tri_mat[id]].flags;
        float idf, flf;
        std::memcpy(&idf, &id, 4);
        std::memcpy(&flf, &fl, 4);
        tris[j].v0 = make_float4(p[0], p[1], p[2], idf);
        tris[j].v1 = make_float4(p[3], p[4], p[5], flf);
        tris[j].v2 = make_float4(p[6], p[7], p[8], 0.0f);
        if (RC_TRI_EDGES) {
            tri_edges[j] = tris[j];
            // the FP32 differences the kernel would otherwise form (IEEE, round to nearest)
            tri_edges[j].v1 = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], flf);
            tri_edges[j].v2 = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.0f);
        }
    }
    rc_scene* s = new (std::nothrow) rc_scene();
    if (!s) return rc_host_fail(RC_ERR_ALLOC, "rc_scene_upload: out of memory");
    s->ctx = c;
    s->nrec = nrec;
    s->ntri = ntri;
    s->nmat = nmat;
    s->depth = depth;
    if (empty)
        for (int k = 0; k < 3; ++k) bmin[k] = bmax[k] = 0.0f;
    std::memcpy(s->bmin, bmin, sizeof(bmin));
    std::memcpy(s->bmax, bmax, sizeof(bmax));
    // no zero-byte allocations: an empty scene still holds one (never read)
    // triangle record
    const size_t nalloc = std::max<size_t>(ntri, 1);
    tris.resize(nalloc);
    if (RC_TRI_EDGES) tri_edges.resize(nalloc);
    shade.resize(nalloc);
    size_t b_tris = nalloc * sizeof(TriRec), b_shade = nalloc * sizeof(float4), b_mats = nmat * sizeof(DevMaterial);
    s->bytes = b_nodes + (RC_TRI_EDGES ? 2 : 1) * b_tris + b_shade + b_mats;
    auto cleanup = [&](int code, const std::string& msg) {
        if (s->nodes) (void)hipFree(s->nodes);
        if (s->tris_exact && s->tris_exact != s->tris) (void)hipFree(s->tris_exact);
        if (s->tris) (void)hipFree(s->tris);
        if (s->shade) (void)hipFree(s->shade);
        if (s->mats) (void)hipFree(s->mats);
        delete s;
        return rc_host_fail(code, msg);
    };
    if (hipMalloc(&s->nodes, b_nodes) != hipSuccess || hipMalloc(&s->tris, b_tris) != hipSuccess ||
        (RC_TRI_EDGES && hipMalloc(&s->tris_exact, b_tris) != hipSuccess) ||
        hipMalloc(&s->shade, b_shade) != hipSuccess || hipMalloc(&s->mats, b_mats) != hipSuccess)
        return cleanup(RC_ERR_ALLOC, "rc_scene_upload: device allocation failed");
    if (!RC_TRI_EDGES) s->tris_exact = s->tris;
    if (hipMemcpy(s->nodes, node_src, b_nodes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->tris, (RC_TRI_EDGES ? tri_edges : tris).data(), b_tris, hipMemcpyHostToDevice) != hipSuccess ||
        (RC_TRI_EDGES && hipMemcpy(s->tris_exact, tris.data(), b_tris, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(s->shade, shade.data(), b_shade, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->mats, dm.data(), b_mats, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(RC_ERR_HIP, "rc_scene_upload: copy to device failed");
    *out = s;
    return RC_OK;
}

extern "C" int rc_scene_free(rc_scene* s) {
    if (!s) return RC_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(s->nodes);
    if (s->tris_exact != s->tris) (void)hipFree(s->tris_exact);
    (void)hipFree(s->tris);
    (void)hipFree(s->shade);
    (void)hipFree(s->mats);
    delete s;
    return RC_OK;
}

extern "C" int rc_scene_info(const rc_scene* s, size_t* ntri, size_t* nrec, size_t* bytes) {
    if (!s) return rc_host_fail(RC_ERR_INVALID, "rc_scene_info: null scene");
    if (ntri) *ntri = s->ntri;
    if (nrec) *nrec = s->nrec;
    if (bytes) *bytes = s->bytes;
    return RC_OK;
}

// ---------------------------------------------------------------- device memory
extern "C" int rc_device_alloc(rc_ctx* c, size_t bytes, void** p) {
    RC_TRY(rc_check_ctx(c, "rc_device_alloc"));
    if (!p || bytes == 0) return rc_host_fail(RC_ERR_INVALID, "rc_device_alloc: bad argument");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return rc_host_fail(RC_ERR_ALLOC, std::string("rc_device_alloc: ") + hipGetErrorString(e));
    return RC_OK;
}
extern "C" int rc_device_free(rc_ctx* c, void* p) {
    RC_TRY(rc_check_ctx(c, "rc_device_free"));
    if (p) {
        RC_TRY_HIP(hipStreamSynchronize(c->stream));
        RC_TRY_HIP(hipFree(p));
    }
    return RC_OK;
}
extern "C" int rc_host_alloc(rc_ctx* c, size_t bytes, void** p) {
    RC_TRY(rc_check_ctx(c, "rc_host_alloc"));
    if (!p || bytes == 0) return rc_host_fail(RC_ERR_INVALID, "rc_host_alloc: bad argument");
    RC_TRY_HIP(hipSetDevice(c->device));
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return rc_host_fail(RC_ERR_ALLOC, std::string("rc_host_alloc: ") + hipGetErrorString(e));
    return RC_OK;
}
extern "C" int rc_host_free(rc_ctx* c, void* p) {
    RC_TRY(rc_check_ctx(c, "rc_host_free"));
    if (p) {
        RC_TRY_HIP(hipStreamSynchronize(c->stream));
        if (c->copy_stream) RC_TRY_HIP(hipStreamSynchronize(c->copy_stream));
        RC_TRY_HIP(hipHostFree(p));
    }
    return RC_OK;
}
extern "C" int rc_memcpy_h2d(rc_ctx* c, void* dst, const void* src, size_t bytes) {
    RC_TRY(rc_check_ctx(c, "rc_memcpy_h2d"));
    RC_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}
extern "C" int rc_memcpy_d2h(rc_ctx* c, void* dst, const void* src, size_t bytes) {
    RC_TRY(rc_check_ctx(c, "rc_memcpy_d2h"));
    RC_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

// ---------------------------------------------------------------- bandwidth probes
namespace {
// Streaming copy, four 16-byte loads in flight per thread before their
// stores; NT: nontemporal loads and stores (the data is touched once).
template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * 256;
    size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f* s4 = reinterpret_cast<const v4f*>(src);
    v4f* d4 = reinterpret_cast<v4f*>(dst);
    auto ld = [&](size_t k) { return NT ? __builtin_nontemporal_load(s4 + k) : s4[k]; };
    auto st = [&](size_t k, v4f v) {
        if (NT) __builtin_nontemporal_store(v, d4 + k);
        else d4[k] = v;
    };
    for (; i + 3 * stride < n; i += 4 * stride) {
        const v4f a = ld(i), b = ld(i + stride), c = ld(i + 2 * stride), d = ld(i + 3 * stride);
        st(i, a);
        st(i + stride, b);
        st(i + 2 * stride, c);
        st(i + 3 * stride, d);
    }
    for (; i < n; i += stride) st(i, ld(i));
}

// Reads n float4 `iters` times (each thread the same addresses every time, so
// a working set that fits a cache level stays in it) and folds them into one
// value per thread, which is stored only if it is impossible (keeps the loads).
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ src, size_t n, int iters,
                                              float* __restrict__ sink) {
    const size_t stride = size_t(gridDim.x) * 256;
    float acc = 0.0f;
    for (int it = 0; it < iters; ++it) {
        size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
        for (; i + 3 * stride < n; i += 4 * stride) {
            const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
            acc += (a.x + b.y) + (c.z + d.w);
        }
        for (; i < n; i += stride) acc += src[i].x;
    }
    if (acc == -1.0f) sink[blockIdx.x * 256 + threadIdx.x] = acc;
}

// Reads one 64-byte record per thread at a distinct pseudo-random record of
// a 2^log2m-record buffer (thread t reads record t * 0x9E3779B1 mod 2^log2m,
// a bijection of the record indices, so no record is read twice): the access
// pattern of a traversal's node fetches, with no reuse for a cache to serve.
__global__ __launch_bounds__(256) void k_gather(const float4* __restrict__ src, uint32_t records, uint32_t log2m,
                                                float* __restrict__ sink) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= records) return;
    const uint64_t idx = (uint64_t(t) * 0x9E3779B1ull) & ((uint64_t(1) << log2m) - 1u);
    const float4* r = src + 4 * idx;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3];
    const float acc = (a.x + b.y) + (c.z + d.w);
    if (acc == -1.0f) sink[t] = acc;
}

__global__ void k_fill(float4* p, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
        p[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

// Times `iters` launches of f(grid) with HIP events on the context stream (ms).
template <class F>
int time_launches(rc_ctx* c, int iters, F f, float* ms) {
    hipEvent_t e0, e1;
    RC_TRY_HIP(hipEventCreate(&e0));
    RC_TRY_HIP(hipEventCreate(&e1));
    f();  // warm-up
    RC_TRY_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) f();
    RC_TRY_HIP(hipEventRecord(e1, c->stream));
    RC_TRY_HIP(hipEventSynchronize(e1));
    RC_TRY_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}
}  // namespace

extern "C" int rc_copy_bandwidth(rc_ctx* c, size_t bytes, int iters, double* gbps) {
    RC_TRY(rc_check_ctx(c, "rc_copy_bandwidth"));
    if (!gbps || bytes < 4096 || iters <= 0) return rc_host_fail(RC_ERR_INVALID, "rc_copy_bandwidth: bad argument");
    RC_TRY_HIP(hipSetDevice(c->device));
    size_t n = bytes / sizeof(float4);
    DevBuf a, b;
    RC_TRY(a.ensure(n * sizeof(float4)));
    int rc = b.ensure(n * sizeof(float4));
    if (rc != RC_OK) { a.release(); return rc; }
    double best = 0.0;
    // the best of four grid sizes, plain and nontemporal: the achievable rate,
    // not one launch shape's
    for (int nt = 0; nt < 2; ++nt)
        for (unsigned per_cu : {4u, 8u, 16u, 32u}) {
            if (rc != RC_OK) break;
            const unsigned grid =
                unsigned(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, size_t(c->cu_count) * per_cu)));
            float ms = 0.0f;
            rc = time_launches(
                c, iters,
                [&]() {
                    if (nt) k_copy<true><<<grid, 256, 0, c->stream>>>(a.as<float4>(), b.as<float4>(), n);
                    else k_copy<false><<<grid, 256, 0, c->stream>>>(a.as<float4>(), b.as<float4>(), n);
                },
                &ms);
            if (rc == RC_OK) best = std::max(best, 2.0 * double(n * sizeof(float4)) * iters / (double(ms) * 1e-3) / 1e9);
        }
    a.release();
    b.release();
    if (rc != RC_OK) return rc;
    *gbps = best;
    return RC_OK;
}

extern "C" int rc_read_bandwidth(rc_ctx* c, size_t bytes, int iters, double* gbps) {
    RC_TRY(rc_check_ctx(c, "rc_read_bandwidth"));
    if (!gbps || bytes < 4096 || iters <= 0 || iters > 4096)
        return rc_host_fail(RC_ERR_INVALID, "rc_read_bandwidth: bad argument");
    RC_TRY_HIP(hipSetDevice(c->device));
    const size_t n = bytes / sizeof(float4);
    DevBuf a, sink;
    RC_TRY(a.ensure(n * sizeof(float4)));
    const unsigned grid_max = unsigned(c->cu_count) * 32u;
    int rc = sink.ensure(size_t(grid_max) * 256 * sizeof(float));
    if (rc != RC_OK) { a.release(); return rc; }
    k_fill<<<grid_max, 256, 0, c->stream>>>(a.as<float4>(), n);
    double best = 0.0;
    for (unsigned per_cu : {8u, 16u, 32u}) {
        if (rc != RC_OK) break;
        const unsigned grid = unsigned(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, size_t(c->cu_count) * per_cu)));
        float ms = 0.0f;
        // one launch reads the set `iters` times, so launch overhead does not count
        rc = time_launches(c, 3, [&]() { k_read<<<grid, 256, 0, c->stream>>>(a.as<float4>(), n, iters, sink.as<float>()); },
                           &ms);
        if (rc == RC_OK) best = std::max(best, double(n * sizeof(float4)) * iters * 3 / (double(ms) * 1e-3) / 1e9);
    }
    a.release();
    sink.release();
    if (rc != RC_OK) return rc;
    *gbps = best;
    return RC_OK;
}

extern "C" int rc_gather_bandwidth(rc_ctx* c, size_t bytes, size_t records, int iters, double* gbps) {
    RC_TRY(rc_check_ctx(c, "rc_gather_bandwidth"));
    if (!gbps || bytes < 4096 || records == 0 || records > (size_t(1) << 31) || iters <= 0 || iters > 4096)
        return rc_host_fail(RC_ERR_INVALID, "rc_gather_bandwidth: bad argument");
    uint32_t log2m = 0;
    while ((size_t(64) << (log2m + 1)) <= bytes) ++log2m;  // the largest power-of-two record count that fits
    if ((size_t(1) << log2m) < records)
        return rc_host_fail(RC_ERR_INVALID, "rc_gather_bandwidth: more records than the buffer holds");
    RC_TRY_HIP(hipSetDevice(c->device));
    const size_t n = (size_t(64) << log2m) / sizeof(float4);
    DevBuf a, sink;
    RC_TRY(a.ensure(n * sizeof(float4)));
    int rc = sink.ensure(records * sizeof(float));
    if (rc != RC_OK) { a.release(); return rc; }
    k_fill<<<unsigned(c->cu_count) * 32u, 256, 0, c->stream>>>(a.as<float4>(), n);
    float ms = 0.0f;
    const unsigned grid = unsigned((records + 255) / 256);
    rc = time_launches(
        c, iters, [&]() { k_gather<<<grid, 256, 0, c->stream>>>(a.as<float4>(), uint32_t(records), log2m, sink.as<float>()); },
        &ms);
    a.release();
    sink.release();
    if (rc != RC_OK) return rc;
    *gbps = double(records) * 64.0 * iters / (double(ms) * 1e-3) / 1e9;
    return RC_OK;
}
