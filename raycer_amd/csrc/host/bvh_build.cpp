// Host FP64 SAH BVH builder: the algorithm of raycer's FlatBVH::build
// (src/Raytracing/Primitives/FlatBVH.cpp:89-204) and its SAH split selection
// (calculateSAHSplit / calculateSAHScore / calculateMedianPoint, FlatBVH.cpp:241-342),
// producing the same preorder node array and orderedPrimitiveIds.
//
// The reference builds with an explicit stack on one thread and sorts the
// centroids of every node to find medians.  Here subtrees are built
// recursively; the two children of a large node are built on separate threads
// into private node arrays that are concatenated in preorder (rightOffset is
// relative, so concatenation needs no fix-up), and the median comes from
// nth_element, which yields the same values as the sorted array.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "raycer_host.h"
#include "rc_host_internal.hpp"

namespace {

struct Prim {
    double c[3];   // AABB::getCenter of the triangle box
    double lo[3];
    double hi[3];
    uint64_t id;
};

struct Box {
    double lo[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double hi[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                    std::numeric_limits<double>::lowest()};
    void add(const Prim& p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p.lo[k]);
            hi[k] = std::max(hi[k], p.hi[k]);
        }
    }
    double area() const {
        double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        return 2.0 * (ex * ey + ez * ey + ex * ez);
    }
    double center(int a) const { return (lo[a] + hi[a]) * 0.5; }
};

struct Builder {
    std::vector<Prim> prims;
    uint64_t max_leaf = 5;
    unsigned spawn_depth = 0;
    static constexpr uint64_t kParallelMin = 20000;

    double score(uint64_t s, uint64_t e, int axis, double split, double node_area) const {
        Box L, R;
        uint64_t nl = 0, nr = 0;
        for (uint64_t i = s; i < e; ++i) {
            if (prims[i].c[axis] <= split) { L.add(prims[i]); ++nl; }
            else { R.add(prims[i]); ++nr; }
        }
        double sc = 0.0;
        if (nl > 0) sc += (L.area() / node_area) * double(nl);
        if (nr > 0) sc += (R.area() / node_area) * double(nr);
        return sc;
    }

    double median(uint64_t s, uint64_t e, int axis, std::vector<double>& tmp) const {
        uint64_t n = e - s;
        tmp.resize(n);
        for (uint64_t i = 0; i < n; ++i) tmp[i] = prims[s + i].c[axis];
        uint64_t k = n / 2;
        std::nth_element(tmp.begin(), tmp.begin() + k, tmp.end());
        double hi = tmp[k];
        if (n % 2 == 0) {
            double lo = *std::max_element(tmp.begin(), tmp.begin() + k);
            return (lo + hi) / 2.0;
        }
        return hi;
    }

    // Appends the subtree over prims[s, e) to out in preorder; returns its depth.
    uint32_t build(uint64_t s, uint64_t e, std::vector<rc_flat_node>& out, unsigned depth) {
        Box box;
        for (uint64_t i = s; i < e; ++i) box.add(prims[i]);
        rc_flat_node n;
        for (int k = 0; k < 3; ++k) { n.aabb_min[k] = box.lo[k]; n.aabb_max[k] = box.hi[k]; }
        n.start_offset = s;
        n.primitive_count = e - s;
        n.right_offset = 0;
        size_t self = out.size();
        out.push_back(n);
        if (e - s <= max_leaf) return 1;

        // 6 candidates in the reference order: for each axis the box centre,
        // then the centroid median; strict '<' keeps the first best.  A node box
        // of zero area makes every score NaN: the reference then leaves
        // axis/split uninitialised, here it is axis 0 at the box centre.
        std::vector<double> tmp;
        double node_area = box.area();
        double best = std::numeric_limits<double>::max();
        int axis = 0;
        double split = box.center(0);
        for (int a = 0; a <= 2; ++a) {
            double c = box.center(a);
            double sc = score(s, e, a, c, node_area);
            if (sc < best) { best = sc; axis = a; split = c; }
            double m = median(s, e, a, tmp);
            sc = score(s, e, a, m, node_area);
            if (sc < best) { best = sc; axis = a; split = m; }
        }
        // In-place partition with the reference's swap sequence (FlatBVH.cpp:166-173).
        uint64_t mid = s;
        for (uint64_t i = s; i < e; ++i) {
            if (prims[i].c[axis] <= split) {
                std::swap(prims[i], prims[mid]);
                ++mid;
            }
        }
        if (mid == s || mid == e) mid = s + (e - s) / 2;  // FlatBVH.cpp:175-177

        uint32_t dl, dr;
        if (depth < spawn_depth && (e - s) >= kParallelMin) {
            std::vector<rc_flat_node> left, right;
            left.reserve(2 * (mid - s) / max_leaf + 8);
            right.reserve(2 * (e - mid) / max_leaf + 8);
            std::thread th([&] { dl = build(s, mid, left, depth + 1); });
            dr = build(mid, e, right, depth + 1);
            th.join();
            out[self].right_offset = int64_t(1 + left.size());
            out.insert(out.end(), left.begin(), left.end());
            out.insert(out.end(), right.begin(), right.end());
        } else {
            dl = build(s, mid, out, depth + 1);
            out[self].right_offset = int64_t(out.size() - self);
            dr = build(mid, e, out, depth + 1);
        }
        return 1 + std::max(dl, dr);
    }
};

// Binned SAH builder (RCH_BVH_BINNED): no counterpart in the reference.  Per
// node, centroids are binned into 32 slabs along each axis with a non-zero
// centroid extent; the split plane between slabs with the lowest surface area
// cost C_t * A(node) + C_i * (A(L) N(L) + A(R) N(R)) wins, and a node of at
// most max_leaf primitives becomes a leaf when C_i * N * A(node) is no worse
// (C_t = 1, C_i = tri_cost).
// Output format, preorder and threading are those of Builder.
struct Binned {
    static constexpr int kBins = 32;
    std::vector<Prim> prims;
    uint64_t max_leaf = 5;
    double tri_cost = 1.0;  // C_i, with C_t = 1
    unsigned spawn_depth = 0;
    static constexpr uint64_t kParallelMin = 20000;

    uint32_t build(uint64_t s, uint64_t e, std::vector<rc_flat_node>& out, unsigned depth) {
        Box box, cb;
        for (uint64_t i = s; i < e; ++i) {
            box.add(prims[i]);
            for (int k = 0; k < 3; ++k) {
                cb.lo[k] = std::min(cb.lo[k], prims[i].c[k]);
                cb.hi[k] = std::max(cb.hi[k], prims[i].c[k]);
            }
        }
        rc_flat_node n;
        for (int k = 0; k < 3; ++k) { n.aabb_min[k] = box.lo[k]; n.aabb_max[k] = box.hi[k]; }
        n.start_offset = s;
        n.primitive_count = e - s;
        n.right_offset = 0;
        size_t self = out.size();
        out.push_back(n);
        const uint64_t cnt = e - s;
        if (cnt <= 1) return 1;

        double best = std::numeric_limits<double>::infinity();
        int axis = -1, split_bin = 0;
        for (int a = 0; a < 3; ++a) {
            double ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0)) continue;
            double scale = kBins / ext;
            Box bb[kBins];
            uint64_t bn[kBins] = {};
            for (uint64_t i = s; i < e; ++i) {
                int k = std::min(kBins - 1, int((prims[i].c[a] - cb.lo[a]) * scale));
                bb[k].add(prims[i]);
                ++bn[k];
            }
            double rarea[kBins];
            uint64_t rcount[kBins];
            Box acc;
            uint64_t c = 0;
            for (int k = kBins - 1; k > 0; --k) {
                if (bn[k]) {
                    for (int d = 0; d < 3; ++d) {
                        acc.lo[d] = std::min(acc.lo[d], bb[k].lo[d]);
                        acc.hi[d] = std::max(acc.hi[d], bb[k].hi[d]);
                    }
                }
                c += bn[k];
                rarea[k] = c ? acc.area() : 0.0;
                rcount[k] = c;
            }
            Box lacc;
            uint64_t lc = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                if (bn[k]) {
                    for (int d = 0; d < 3; ++d) {
                        lacc.lo[d] = std::min(lacc.lo[d], bb[k].lo[d]);
                        lacc.hi[d] = std::max(lacc.hi[d], bb[k].hi[d]);
                    }
                }
                lc += bn[k];
                if (lc == 0 || rcount[k + 1] == 0) continue;
                double cost = lacc.area() * double(lc) + rarea[k + 1] * double(rcount[k + 1]);
                if (cost < best) { best = cost; axis = a; split_bin = k; }
            }
        }
        const double area = box.area();
        const double split_cost = area + tri_cost * best;
        if (cnt <= max_leaf && (axis < 0 || tri_cost * double(cnt) * area <= split_cost)) return 1;

        uint64_t mid;
        if (axis < 0) {
            mid = s + cnt / 2;  // coincident centroids: halve the list
        } else {
            const double lo = cb.lo[axis], scale = kBins / (cb.hi[axis] - cb.lo[axis]);
            auto it = std::partition(prims.begin() + s, prims.begin() + e, [&](const Prim& p) {
                return std::min(kBins - 1, int((p.c[axis] - lo) * scale)) <= split_bin;
            });
            mid = uint64_t(it - prims.begin());
            if (mid == s || mid == e) mid = s + cnt / 2;
        }
        uint32_t dl, dr;
        if (depth < spawn_depth && cnt >= kParallelMin) {
            std::vector<rc_flat_node> left, right;
            std::thread th([&] { dl = build(s, mid, left, depth + 1); });
            dr = build(mid, e, right, depth + 1);
            th.join();
            out[self].right_offset = int64_t(1 + left.size());
            out.insert(out.end(), left.begin(), left.end());
            out.insert(out.end(), right.begin(), right.end());
        } else {
            dl = build(s, mid, out, depth + 1);
            out[self].right_offset = int64_t(out.size() - self);
            dr = build(mid, e, out, depth + 1);
        }
        return 1 + std::max(dl, dr);
    }
};

}  // namespace

struct rch_bvh {
    std::vector<rc_flat_node> nodes;
    std::vector<uint64_t> ordered;
    double build_ms = 0.0;
    uint32_t depth = 0;
};

extern "C" int rch_bvh_build(const float* v, size_t ntri, const rch_bvh_build_info* info, rch_bvh** out) {
    // no triangles: one leaf with AABB()'s empty box and no primitives, as
    // FlatBVH::build makes over an empty list (FlatBVH.cpp:112-139)
    if (!out || (!v && ntri != 0)) return rc_host_fail(RC_ERR_INVALID, "rch_bvh_build: null argument");
    if (ntri > 0x7fffffffull) return rc_host_fail(RC_ERR_INVALID, "rch_bvh_build: more than 2^31-1 triangles");
    auto t0 = std::chrono::steady_clock::now();
    rch_bvh* b = new (std::nothrow) rch_bvh();
    if (!b) return rc_host_fail(RC_ERR_ALLOC, "rch_bvh_build: out of memory");
    const uint32_t method = info ? info->method : RCH_BVH_REFERENCE;
    if (method != RCH_BVH_REFERENCE && method != RCH_BVH_BINNED) {
        delete b;
        return rc_host_fail(RC_ERR_INVALID, "rch_bvh_build: unknown method");
    }
    const uint64_t max_leaf = info && info->max_leaf_size ? info->max_leaf_size : 5;
    unsigned threads = info && info->threads ? info->threads : std::max(1u, std::thread::hardware_concurrency());
    unsigned spawn_depth = 0;
    while ((1u << spawn_depth) < threads * 2 && spawn_depth < 12) ++spawn_depth;
    if (threads == 1) spawn_depth = 0;
    std::vector<Prim> prims(ntri);
    for (size_t i = 0; i < ntri; ++i) {
        const float* p = v + 9 * i;
        Prim& q = prims[i];
        for (int k = 0; k < 3; ++k) {
            double a = p[k], bb = p[3 + k], c = p[6 + k];
            q.lo[k] = std::min(a, std::min(bb, c));
            q.hi[k] = std::max(a, std::max(bb, c));
            q.c[k] = (q.lo[k] + q.hi[k]) * 0.5;
        }
        q.id = i;
    }
    b->nodes.reserve(2 * ntri / max_leaf + 16);
    const std::vector<Prim>* done;
    Builder B;
    Binned N;
    if (method == RCH_BVH_BINNED) {
        N.prims.swap(prims);
        N.max_leaf = max_leaf;
        N.tri_cost = info && info->tri_cost > 0.0f ? double(info->tri_cost) : 1.0;
        N.spawn_depth = spawn_depth;
        b->depth = N.build(0, ntri, b->nodes, 0);
        done = &N.prims;
    } else {
        B.prims.swap(prims);
        B.max_leaf = max_leaf;
        B.spawn_depth = spawn_depth;
        b->depth = B.build(0, ntri, b->nodes, 0);
        done = &B.prims;
    }
    b->ordered.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) b->ordered[i] = (*done)[i].id;
    b->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = b;
    return RC_OK;
}

extern "C" size_t rch_bvh_node_count(const rch_bvh* b) { return b ? b->nodes.size() : 0; }
extern "C" int rch_bvh_get(const rch_bvh* b, rc_flat_node* nodes, uint64_t* ordered) {
    if (!b) return rc_host_fail(RC_ERR_INVALID, "rch_bvh_get: null bvh");
    if (nodes) std::memcpy(nodes, b->nodes.data(), b->nodes.size() * sizeof(rc_flat_node));
    if (ordered) std::memcpy(ordered, b->ordered.data(), b->ordered.size() * sizeof(uint64_t));
    return RC_OK;
}
extern "C" double rch_bvh_build_ms(const rch_bvh* b) { return b ? b->build_ms : 0.0; }
extern "C" uint32_t rch_bvh_depth(const rch_bvh* b) { return b ? b->depth : 0; }
extern "C" int rch_bvh_free(rch_bvh* b) {
    delete b;
    return RC_OK;
}

// ------------------------------------------------------------------ save / load
// A built FlatBVH on disk: the counterpart of the reference's serialized
// FlatBVH (flatNodes + orderedPrimitiveIds, src/Raytracing/Primitives/FlatBVH.h:99-105)
// and its restore on load, which skips the build (FlatBVH.cpp:206-214,
// Scene.cpp:487-488).  Binary, little-endian: magic "RCBVH001", node count,
// triangle count, FNV-1a hash of the triangles' vertex bytes, the nodes, the
// ordered ids.  Loading checks the hash against the triangles it is given.
namespace {
constexpr char kMagic[8] = {'R', 'C', 'B', 'V', 'H', '0', '0', '1'};

uint64_t fnv1a(const float* v, size_t ntri) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* p = reinterpret_cast<const unsigned char*>(v);
    for (size_t i = 0; i < ntri * 9 * sizeof(float); ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct FileHeader {
    char magic[8];
    uint64_t nnodes, ntri, hash;
};
}  // namespace

extern "C" int rch_bvh_save(const char* path, const float* vertices, size_t ntri, const rc_flat_node* nodes,
                            size_t nnodes, const uint64_t* ordered) {
    if (!path || !nodes || nnodes == 0 || (ntri != 0 && (!vertices || !ordered)))
        return rc_host_fail(RC_ERR_INVALID, "rch_bvh_save: null or empty argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_save: cannot write ") + path);
    FileHeader h;
    std::memcpy(h.magic, kMagic, 8);
    h.nnodes = nnodes;
    h.ntri = ntri;
    h.hash = fnv1a(vertices, ntri);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(nodes, sizeof(rc_flat_node), nnodes, f) == nnodes &&
              std::fwrite(ordered, sizeof(uint64_t), ntri, f) == ntri;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RC_OK : rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_save: write failed for ") + path);
}

namespace {
int read_header(FILE* f, const char* path, FileHeader& h) {
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0)
        return rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_load: not a saved BVH: ") + path);
    return RC_OK;
}
}  // namespace

extern "C" int rch_bvh_file_counts(const char* path, size_t* nnodes, size_t* ntri) {
    if (!path || !nnodes || !ntri) return rc_host_fail(RC_ERR_INVALID, "rch_bvh_file_counts: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_file_counts: cannot open ") + path);
    FileHeader h;
    int rc = read_header(f, path, h);
    std::fclose(f);
    if (rc != RC_OK) return rc;
    *nnodes = size_t(h.nnodes);
    *ntri = size_t(h.ntri);
    return RC_OK;
}

extern "C" int rch_bvh_load(const char* path, const float* vertices, size_t ntri, rc_flat_node* nodes,
                            uint64_t* ordered) {
    if (!path || !nodes || (ntri != 0 && (!vertices || !ordered)))
        return rc_host_fail(RC_ERR_INVALID, "rch_bvh_load: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_load: cannot open ") + path);
    FileHeader h;
    int rc = read_header(f, path, h);
    if (rc == RC_OK && (h.ntri != ntri || h.hash != fnv1a(vertices, ntri)))
        rc = rc_host_fail(RC_ERR_INVALID, "rch_bvh_load: the saved BVH was built over other triangles");
    if (rc == RC_OK && (std::fread(nodes, sizeof(rc_flat_node), h.nnodes, f) != h.nnodes ||
                        std::fread(ordered, sizeof(uint64_t), ntri, f) != ntri))
        rc = rc_host_fail(RC_ERR_INVALID, std::string("rch_bvh_load: truncated file ") + path);
    std::fclose(f);
    return rc;
}
