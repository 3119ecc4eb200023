// FP64 restatement of raycer's FlatBVH build and traversal.  TEST INFRASTRUCTURE ONLY.
#include <omp.h>

#include <algorithm>
#include <vector>

#include "core.hpp"
#include "oracle.h"

using namespace orc;

struct or_bvh {
    std::vector<or_flat_node> nodes;
    std::vector<uint64_t> ordered;  // orderedPrimitiveIds
};

namespace {

struct Prim {
    uint64_t id;
    Box box;
    V3 center;
};

struct BuildEntry {
    uint64_t start, end;
    int64_t parent;
};

// FlatBVH::calculateSAHScore (FlatBVH.cpp:288-321)
double sah_score(const std::vector<Prim>& p, int axis, double split, const Box& node, const BuildEntry& e) {
    Box L, R;
    uint64_t nl = 0, nr = 0;
    for (uint64_t i = e.start; i < e.end; ++i) {
        if (p[i].center.get(axis) <= split) { L.expand(p[i].box); ++nl; }
        else { R.expand(p[i].box); ++nr; }
    }
    double s = 0.0;
    if (nl > 0) s += (L.area() / node.area()) * double(nl);
    if (nr > 0) s += (R.area() / node.area()) * double(nr);
    return s;
}

// FlatBVH::calculateMedianPoint (FlatBVH.cpp:323-342): sorted centroids, mean of
// the two middle values for an even count, the middle value otherwise.
double median_point(const std::vector<Prim>& p, int axis, const BuildEntry& e) {
    std::vector<double> c;
    c.reserve(e.end - e.start);
    for (uint64_t i = e.start; i < e.end; ++i) c.push_back(p[i].center.get(axis));
    std::sort(c.begin(), c.end());
    uint64_t n = c.size();
    if (n % 2 == 0) return (c[n / 2 - 1] + c[n / 2]) / 2.0;
    return c[n / 2];
}

// FlatBVH::calculateSAHSplit (FlatBVH.cpp:241-286) with regularSAHSplits = 0:
// per axis the node-box centre, then the centroid median; strict '<' keeps the
// first best.  If every score is NaN (a node box of zero area) the reference
// leaves axis/split uninitialised; the oracle takes axis 0 at the box centre.
void sah_split(const std::vector<Prim>& p, const Box& node, const BuildEntry& e, int& axis, double& split) {
    double best = std::numeric_limits<double>::max();
    axis = 0;
    split = node.center().x;
    for (int a = 0; a <= 2; ++a) {
        double s = node.center().get(a);
        double sc = sah_score(p, a, s, node, e);
        if (sc < best) { axis = a; split = s; best = sc; }
        s = median_point(p, a, e);
        sc = sah_score(p, a, s, node, e);
        if (sc < best) { axis = a; split = s; best = sc; }
    }
}

}  // namespace

extern "C" int or_bvh_build(const float* vertices, size_t ntri, uint64_t max_leaf, or_bvh** out) {
    if (!out || (ntri && !vertices)) return 1;  // no triangles: one empty leaf, as the reference builds
    std::vector<Prim> prims(ntri);
    for (uint64_t i = 0; i < ntri; ++i) {
        Tri t = load_tri(vertices, i);
        prims[i].id = i;
        prims[i].box = tri_box(t);
        prims[i].center = prims[i].box.center();
    }
    auto* b = new or_bvh();
    // FlatBVH::build (FlatBVH.cpp:89-204): explicit stack, preorder emission,
    // the parent's rightOffset fixed up when its second child is emitted.
    const int64_t ROOT = -4, UNVISITED = -3, VISITED_TWICE = -1;
    std::vector<BuildEntry> stack;
    stack.push_back({0, ntri, ROOT});
    uint64_t node_count = 0;
    while (!stack.empty()) {
        BuildEntry e = stack.back();
        stack.pop_back();
        ++node_count;
        or_flat_node n;
        Box box;
        for (uint64_t i = e.start; i < e.end; ++i) box.expand(prims[i].box);
        n.right_offset = UNVISITED;
        n.start_offset = e.start;
        n.primitive_count = e.end - e.start;
        for (int k = 0; k < 3; ++k) { n.aabb_min[k] = box.mn.get(k); n.aabb_max[k] = box.mx.get(k); }
        if (n.primitive_count <= max_leaf) n.right_offset = 0;
        b->nodes.push_back(n);
        if (e.parent != ROOT) {
            or_flat_node& par = b->nodes[uint64_t(e.parent)];
            par.right_offset++;
            if (par.right_offset == VISITED_TWICE) par.right_offset = int64_t(node_count) - 1 - e.parent;
        }
        if (n.right_offset == 0) continue;
        int axis;
        double split;
        sah_split(prims, box, e, axis, split);
        uint64_t middle = e.start;
        for (uint64_t i = e.start; i < e.end; ++i) {
            if (prims[i].center.get(axis) <= split) {
                std::swap(prims[i], prims[middle]);
                ++middle;
            }
        }
        if (middle == e.start || middle == e.end) middle = e.start + (e.end - e.start) / 2;
        stack.push_back({middle, e.end, int64_t(node_count) - 1});
        stack.push_back({e.start, middle, int64_t(node_count) - 1});
    }
    b->ordered.resize(ntri);
    for (uint64_t i = 0; i < ntri; ++i) b->ordered[i] = prims[i].id;
    *out = b;
    return 0;
}

extern "C" size_t or_bvh_node_count(const or_bvh* b) { return b->nodes.size(); }
extern "C" void or_bvh_nodes(const or_bvh* b, or_flat_node* out) {
    std::memcpy(out, b->nodes.data(), b->nodes.size() * sizeof(or_flat_node));
}
extern "C" void or_bvh_ordered_ids(const or_bvh* b, uint64_t* out) {
    std::memcpy(out, b->ordered.data(), b->ordered.size() * sizeof(uint64_t));
}
extern "C" void or_bvh_free(or_bvh* b) { delete b; }

extern "C" int or_intersect_triangle(const double v[9], const double org[3], const double dir[3],
                                     double min_d, double max_d, double* t, double* u, double* vv) {
    Tri tr{V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), V3(v[6], v[7], v[8])};
    double det;
    return mt_intersect(tr, V3(org[0], org[1], org[2]), V3(dir[0], dir[1], dir[2]), min_d, max_d, *t, *u, *vv, det) ? 1 : 0;
}

namespace orc {

inline RayD load_ray(const float* r) {
    RayD x;
    x.o = V3(r[0], r[1], r[2]);
    x.tmin = r[3];
    x.d = V3(r[4], r[5], r[6]);
    x.tmax = r[7];
    // Ray::precalculate: inverseDirection = 1 / direction (Ray.cpp:10-13)
    x.inv = V3(1.0 / x.d.x, 1.0 / x.d.y, 1.0 / x.d.z);
    return x;
}

// FlatBVH::intersect (FlatBVH.cpp:22-75) + Triangle::intersect: unordered
// traversal, right child pushed first, left popped first, boxes tested against
// t >= 0 only; a hit replaces the current one unless t > current distance, so
// the last tested triangle wins an exact tie.  any=true is fastOcclusion.
bool trace_ref(const or_bvh* b, const float* vertices, const RayD& r, bool any, double& best_t, int64_t& best_id,
               const uint32_t* tri_flags, uint64_t* nodes_visited, uint64_t* tris_tested) {
    best_t = std::numeric_limits<double>::max();
    best_id = -1;
    uint64_t stack[128];
    uint64_t sp = 0;
    stack[sp++] = 0;
    bool found = false;
    while (sp > 0) {
        uint64_t idx = stack[--sp];
        const or_flat_node& n = b->nodes[idx];
        if (nodes_visited) ++*nodes_visited;
        if (n.right_offset == 0) {
            for (uint64_t i = 0; i < n.primitive_count; ++i) {
                uint64_t id = b->ordered[n.start_offset + i];
                if (any && tri_flags && (tri_flags[id] & 2u)) continue;  // nonShadowing
                if (tris_tested) ++*tris_tested;
                Tri tr = load_tri(vertices, id);
                double t, u, v, det;
                if (!mt_intersect(tr, r.o, r.d, r.tmin, r.tmax, t, u, v, det)) continue;
                if (t > best_t) continue;
                best_t = t;
                best_id = int64_t(id);
                if (any) return true;
                found = true;
            }
        } else {
            uint64_t ri = idx + uint64_t(n.right_offset);
            if (box_hit_ref(b->nodes[ri].aabb_min, b->nodes[ri].aabb_max, r.o, r.inv)) stack[sp++] = ri;
            if (box_hit_ref(b->nodes[idx + 1].aabb_min, b->nodes[idx + 1].aabb_max, r.o, r.inv)) stack[sp++] = idx + 1;
        }
    }
    return found;
}

// Ordered (near child first), t-culled FP64 traversal.  The strict closest hit
// uses the canonical tie rule (lower original id wins an exact tie); a second
// pass with widened barycentric bounds lists the rays whose answer is not
// robust to FP32 rounding.
void trace_ordered(const or_bvh* b, const float* vertices, const RayD& r, double rel_tie, double band, Diag& D) {
    struct Cand { double t; int64_t id; };
    std::vector<Cand> cands;
    uint64_t stack[256];
    int sp = 0;
    const or_flat_node& root = b->nodes[0];
    double cull = r.tmax;
    if (box_entry(root.aabb_min, root.aabb_max, r.o, r.inv, r.tmax) == std::numeric_limits<double>::infinity()) return;
    stack[sp++] = 0;
    while (sp > 0) {
        uint64_t idx = stack[--sp];
        const or_flat_node& n = b->nodes[idx];
        if (n.right_offset == 0) {
            for (uint64_t i = 0; i < n.primitive_count; ++i) {
                uint64_t id = b->ordered[n.start_offset + i];
                Tri tr = load_tri(vertices, id);
                double t, u, v, det;
                if (mt_intersect(tr, r.o, r.d, r.tmin, r.tmax, t, u, v, det)) {
                    if (t < D.t || (t == D.t && int64_t(id) < D.id)) {
                        D.t = t; D.id = int64_t(id); D.u = u; D.v = v; D.det = det;
                        V3 e1 = tr.v1 - tr.v0, e2 = tr.v2 - tr.v0;
                        D.cond = std::abs(det) / (length(e1) * length(e2) * length(r.d));
                    }
                }
                if (band > 0.0 && mt_intersect_band(tr, r.o, r.d, r.tmin, r.tmax, band, t, u, v, det))
                    cands.push_back({t, int64_t(id)});
            }
            double c = std::isinf(D.t) ? r.tmax : D.t * (1.0 + 4.0 * rel_tie);
            if (c < cull) cull = c;
        } else {
            uint64_t li = idx + 1, ri = idx + uint64_t(n.right_offset);
            double tl = box_entry(b->nodes[li].aabb_min, b->nodes[li].aabb_max, r.o, r.inv, cull);
            double tr = box_entry(b->nodes[ri].aabb_min, b->nodes[ri].aabb_max, r.o, r.inv, cull);
            const double inf = std::numeric_limits<double>::infinity();
            bool hl = tl != inf && tl <= cull, hr = tr != inf && tr <= cull;
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = ri; stack[sp++] = li; }
                else { stack[sp++] = li; stack[sp++] = ri; }
            } else if (hl) stack[sp++] = li;
            else if (hr) stack[sp++] = ri;
            if (sp > 250) sp = 250;  // unreachable for trees of depth < 125
        }
    }
    if (band <= 0.0) return;
    if (D.id < 0) {
        if (!cands.empty()) D.flags |= 1u;
        return;
    }
    double margin = std::min(D.u, std::min(D.v, 1.0 - D.u - D.v));
    if (margin < band) D.flags |= 2u;
    if (D.cond < 1e-5) D.flags |= 4u;
    for (const Cand& c : cands) {
        if (c.id == D.id) continue;
        if (c.t <= D.t * (1.0 + rel_tie)) D.flags |= 1u;
    }
}

}  // namespace orc

extern "C" void or_trace_closest_ref(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                                     double* t_out, int64_t* id_out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1000)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayD r = load_ray(rays + 8 * i);
        double t;
        int64_t id;
        if (trace_ref(b, vertices, r, false, t, id, nullptr, nullptr)) { t_out[i] = t; id_out[i] = id; }
        else { t_out[i] = std::numeric_limits<double>::infinity(); id_out[i] = -1; }
    }
}

extern "C" void or_trace_closest_diag(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                                      double rel_tie, double band, double* t_out, int64_t* id_out,
                                      uint32_t* flags_out, double* cond_out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayD r = load_ray(rays + 8 * i);
        Diag D;
        trace_ordered(b, vertices, r, rel_tie, band, D);
        t_out[i] = D.t;
        id_out[i] = D.id;
        if (flags_out) flags_out[i] = D.flags;
        if (cond_out) cond_out[i] = D.id >= 0 ? D.cond : 1.0;
    }
}

extern "C" void or_trace_closest_brute(const float* vertices, size_t ntri, const float* rays, size_t n,
                                       double* t_out, int64_t* id_out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayD r = load_ray(rays + 8 * i);
        double best = std::numeric_limits<double>::infinity();
        int64_t bid = -1;
        for (uint64_t k = 0; k < ntri; ++k) {
            Tri tr = load_tri(vertices, k);
            double t, u, v, det;
            if (!mt_intersect(tr, r.o, r.d, r.tmin, r.tmax, t, u, v, det)) continue;
            if (t < best || (t == best && int64_t(k) < bid)) { best = t; bid = int64_t(k); }
        }
        t_out[i] = best;
        id_out[i] = bid;
    }
}

extern "C" void or_trace_any(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                             double band, uint8_t* occ_out, uint8_t* flags_out, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayD r = load_ray(rays + 8 * i);
        double t;
        int64_t id;
        bool occ = trace_ref(b, vertices, r, true, t, id, nullptr, nullptr);
        occ_out[i] = occ ? 1 : 0;
        if (flags_out) {
            // A decision is fragile when a widened test flips it: some triangle
            // is hit only inside the band, or the only hits lie near tmax.
            uint8_t f = 0;
            if (band > 0.0) {
                Diag D;
                RayD r2 = r;
                r2.tmax = r.tmax * (1.0 + 1e-5);
                trace_ordered(b, vertices, r2, 1e-6, band, D);
                bool widened = D.flags & 1u;
                if (occ && ((D.flags & 2u) || D.t > r.tmax * (1.0 - 1e-5))) f = 1;
                if (!occ && (widened || D.id >= 0)) f = 1;
            }
            flags_out[i] = f;
        }
    }
}

extern "C" size_t or_bvh_ordered_count(const or_bvh* b) { return b->ordered.size(); }
