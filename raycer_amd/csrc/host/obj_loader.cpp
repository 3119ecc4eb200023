// Wavefront OBJ + MTL (and PLY, below) loader for triangle scenes: the behaviour of raycer's
// ModelLoader::readObjFile (src/Utils/ObjModelLoader.cpp:332-482) with
// processMaterialFile (:35-230) and processFace (:232-330), reduced to what the
// triangle path tracer uses.
//
//   v x y z         vertex, transformed by translate * rotateXYZ * scale
//   f a b c ...     polygon of 3+ vertex references in any of the forms
//                   i, i/t, i//n, i/t/n; negative indices count back from the
//                   last vertex; triangulated as a fan (0, k-1, k)
//   mtllib file     materials, relative to the OBJ file's directory
//   usemtl name     material of the following faces (unknown: the default)
//   newmtl / Ka / Kd / Ks / Ke / Ns in the MTL file; Ke != 0 makes it emissive
//
// Numbers are read as StringUtils::parseDouble reads them
// (src/Utils/StringUtils.cpp:70-100): optional '-', integer digits, optional
// '.' and fraction digits accumulated with a running factor of 0.1, nothing
// else; so the vertices are the reference's to the last bit.  vn / vt are
// parsed for validation only (the tracer uses geometric normals); texture
// maps, groups and instances are ignored.  A face that references a missing
// vertex is skipped, as the reference skips it with a warning.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "raycer_host.h"
#include "rc_host_internal.hpp"

namespace {

// StringUtils::readUntilSpace: skips ' ', '\t', '\r', '\n', then reads to the next one.
bool next_token(const std::string& s, size_t& i, std::string& out) {
    auto space = [](char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; };
    while (i < s.size() && space(s[i])) ++i;
    if (i >= s.size()) return false;
    size_t e = i;
    while (e < s.size() && !space(s[e])) ++e;
    out.assign(s, i, e - i);
    i = e;
    return true;
}

double parse_number(const std::string& t) {
    const char* p = t.c_str();
    double sign = 1.0, acc = 0.0;
    if (*p == '-') {
        sign = -1.0;
        ++p;
    }
    while (*p >= '0' && *p <= '9') acc = acc * 10 + (*p++ - '0');
    if (*p == '.') {
        double k = 0.1;
        ++p;
        while (*p >= '0' && *p <= '9') {
            acc += (*p++ - '0') * k;
            k *= 0.1;
        }
    }
    return sign * acc;
}

double read_number(const std::string& line, size_t& i) {
    std::string t;
    next_token(line, i, t);  // a missing number reads as 0, as in the reference
    return parse_number(t);
}

struct M4 {
    double m[4][4];
    static M4 identity() {
        M4 r{};
        for (int k = 0; k < 4; ++k) r.m[k][k] = 1.0;
        return r;
    }
    M4 operator*(const M4& b) const {
        M4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) r.m[i][j] += m[i][k] * b.m[k][j];
        return r;
    }
};

// Matrix4x4::rotateX/Y/Z (src/Math/Matrix4x4.cpp:388-437), angles in degrees.
M4 rotation(int axis, double deg) {
    M4 r = M4::identity();
    double s = std::sin(deg * (M_PI / 180.0)), c = std::cos(deg * (M_PI / 180.0));
    int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
    r.m[a][a] = c;
    r.m[b][b] = c;
    if (axis == 1) {  // rotateY: m[0][2] = s, m[2][0] = -s
        r.m[a][b] = s;
        r.m[b][a] = -s;
    } else {
        r.m[a][b] = -s;
        r.m[b][a] = s;
    }
    return r;
}

M4 model_transform(const rch_obj_info& info) {
    M4 s = M4::identity(), t = M4::identity();
    for (int k = 0; k < 3; ++k) {
        s.m[k][k] = info.scale[k];
        t.m[k][3] = info.translate[k];
    }
    // translation * rotateXYZ(pitch, yaw, roll) * scaling
    return t * (rotation(0, info.rotate[0]) * rotation(1, info.rotate[1]) * rotation(2, info.rotate[2])) * s;
}

struct Loader {
    rch_obj_info info;
    std::vector<rc_material> materials;  // [0] = the base material
    std::map<std::string, uint32_t> names;
    std::vector<double> verts;           // transformed, xyz
    size_t normals = 0, texcoords = 0;
    std::vector<float> tris;             // 9 per triangle
    std::vector<uint32_t> tri_mat;
    M4 xf = M4::identity();

    int load_mtl(const std::string& path) {
        std::ifstream f(path);
        if (!f.good()) return rc_host_fail(RC_ERR_INVALID, "rch_obj_load: cannot open the MTL file " + path);
        std::string line, t;
        bool pending = false;
        rc_material cur = info.base;
        auto color = [&](const std::string& l, size_t& i, float* c) {
            for (int k = 0; k < 3; ++k) c[k] = float(read_number(l, i));
        };
        while (std::getline(f, line)) {
            size_t i = 0;
            if (!next_token(line, i, t)) continue;
            if (t == "newmtl") {
                if (pending) materials.push_back(cur);
                pending = true;
                cur = info.base;
                std::string name;
                next_token(line, i, name);
                names[name] = uint32_t(materials.size());
            } else if (t == "Ka") {
                color(line, i, cur.ambient);
            } else if (t == "Kd") {
                color(line, i, cur.diffuse);
            } else if (t == "Ks") {
                color(line, i, cur.specular);
            } else if (t == "Ke") {
                color(line, i, cur.emittance);
                if (cur.emittance[0] != 0.0f || cur.emittance[1] != 0.0f || cur.emittance[2] != 0.0f)
                    cur.flags |= RC_MAT_EMISSIVE;
            } else if (t == "Ns") {
                cur.shininess = float(read_number(line, i));
            }
        }
        if (pending) materials.push_back(cur);
        return RC_OK;
    }

    // processFace: resolves the references of one polygon, then fans it.
    void face(const std::string& line, size_t i, uint32_t mat) {
        std::vector<size_t> idx;
        std::string t;
        size_t probe = i;
        if (!next_token(line, probe, t)) return;
        size_t slashes = size_t(std::count(t.begin(), t.end(), '/'));
        bool dbl = t.find("//") != std::string::npos;
        bool has_t = slashes > 0 && !dbl, has_n = slashes > 1;
        auto resolve = [](long long k, size_t n, size_t& out) {
            long long r = k < 0 ? (long long)n + k : k - 1;
            if (r < 0 || r >= (long long)n) return false;
            out = size_t(r);
            return true;
        };
        while (next_token(line, i, t)) {
            std::replace(t.begin(), t.end(), '/', ' ');
            size_t j = 0;
            std::string part;
            next_token(t, j, part);
            size_t v;
            if (!resolve(std::strtoll(part.c_str(), nullptr, 10), verts.size() / 3, v)) return;
            idx.push_back(v);
            size_t dummy;
            if (has_t) {
                next_token(t, j, part);
                if (!resolve(std::strtoll(part.c_str(), nullptr, 10), texcoords, dummy)) return;
            }
            if (has_n) {
                next_token(t, j, part);
                if (!resolve(std::strtoll(part.c_str(), nullptr, 10), normals, dummy)) return;
            }
        }
        if (idx.size() < 3) return;
        for (size_t k = 2; k < idx.size(); ++k) {
            const size_t c[3] = {idx[0], idx[k - 1], idx[k]};
            for (size_t a : c)
                for (int d = 0; d < 3; ++d) tris.push_back(float(verts[3 * a + d]));
            tri_mat.push_back(mat);
        }
    }

    int load(const char* path) {
        FILE* f = std::fopen(path, "r");
        if (!f) return rc_host_fail(RC_ERR_INVALID, std::string("rch_obj_load: cannot open ") + path);
        std::string dir(path);
        size_t slash = dir.find_last_of('/');
        dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);
        xf = model_transform(info);
        materials.assign(1, info.base);
        uint32_t cur_mat = 0;
        char buf[1024];
        std::string line, t;
        int rc = RC_OK;
        while (std::fgets(buf, sizeof(buf), f)) {
            line.assign(buf);
            size_t i = 0;
            if (!next_token(line, i, t)) continue;
            if (t == "mtllib" && !info.ignore_materials) {
                std::string file;
                next_token(line, i, file);
                rc = load_mtl(dir + file);
                if (rc != RC_OK) break;
            } else if (t == "usemtl" && !info.ignore_materials) {
                std::string name;
                next_token(line, i, name);
                auto it = names.find(name);
                cur_mat = it == names.end() ? 0u : it->second;
            } else if (t == "v") {
                double p[3];
                for (int k = 0; k < 3; ++k) p[k] = read_number(line, i);
                for (int r = 0; r < 3; ++r)
                    verts.push_back(xf.m[r][0] * p[0] + xf.m[r][1] * p[1] + xf.m[r][2] * p[2] + xf.m[r][3]);
            } else if (t == "vn") {
                ++normals;
            } else if (t == "vt") {
                ++texcoords;
            } else if (t == "f") {
                face(line, i, cur_mat);
            }
        }
        std::fclose(f);
        return rc;
    }
};

// ------------------------------------------------------------------ PLY
// ModelLoader::readPlyFile (src/Utils/PlyModelLoader.cpp): header with
// element / property lines (scalar types char..double and lists), ascii or
// binary bodies (binary read in the host's little-endian order, as the
// reference reads it), vertex x y z, face vertex_indices fanned like OBJ
// faces; every triangle gets material 0.
enum class PlyT { I8, U8, I16, U16, I32, U32, F32, F64, LIST, NONE };

PlyT ply_type(const std::string& t) {
    if (t == "char" || t == "int8") return PlyT::I8;
    if (t == "uchar" || t == "uint8") return PlyT::U8;
    if (t == "short" || t == "int16") return PlyT::I16;
    if (t == "ushort" || t == "uint16") return PlyT::U16;
    if (t == "int" || t == "int32") return PlyT::I32;
    if (t == "uint" || t == "uint32") return PlyT::U32;
    if (t == "float" || t == "float32") return PlyT::F32;
    if (t == "double" || t == "float64") return PlyT::F64;
    if (t == "list") return PlyT::LIST;
    return PlyT::NONE;
}

struct PlyProp {
    PlyT type = PlyT::NONE, len = PlyT::NONE, item = PlyT::NONE;
    std::string name;
};
struct PlyElem {
    std::string name;
    uint64_t count = 0;
    std::vector<PlyProp> props;
};

bool read_binary(std::ifstream& f, PlyT t, double& out) {
    auto rd = [&](auto v) {
        f.read(reinterpret_cast<char*>(&v), sizeof(v));
        out = double(v);
    };
    switch (t) {
        case PlyT::I8: rd(int8_t()); break;
        case PlyT::U8: rd(uint8_t()); break;
        case PlyT::I16: rd(int16_t()); break;
        case PlyT::U16: rd(uint16_t()); break;
        case PlyT::I32: rd(int32_t()); break;
        case PlyT::U32: rd(uint32_t()); break;
        case PlyT::F32: rd(float()); break;
        case PlyT::F64: rd(double()); break;
        default: return false;
    }
    return bool(f);
}

int load_ply(const char* path, const M4& xf, std::vector<float>& tris, std::vector<uint32_t>& tri_mat) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) return rc_host_fail(RC_ERR_INVALID, std::string("rch_ply_load: cannot open ") + path);
    auto bad = [&](const char* why) { return rc_host_fail(RC_ERR_INVALID, std::string("rch_ply_load: ") + why); };
    std::string line, t;
    if (!std::getline(f, line) || line.find("ply") != 0) return bad("not a PLY file");
    bool ascii = true, ended = false;
    std::vector<PlyElem> elems;
    while (std::getline(f, line)) {
        size_t i = 0;
        if (!next_token(line, i, t)) continue;
        if (t == "format") {
            next_token(line, i, t);
            if (t == "ascii") ascii = true;
            else if (t == "binary_little_endian" || t == "binary_big_endian") ascii = false;
            else return bad("unknown PLY format");
            next_token(line, i, t);
            if (t != "1.0") return bad("unknown PLY version");
        } else if (t == "element") {
            PlyElem e;
            next_token(line, i, e.name);
            next_token(line, i, t);
            e.count = std::strtoull(t.c_str(), nullptr, 10);
            elems.push_back(e);
        } else if (t == "property") {
            if (elems.empty()) return bad("property before element");
            PlyProp p;
            next_token(line, i, t);
            p.type = ply_type(t);
            if (p.type == PlyT::NONE) return bad("unknown PLY type");
            if (p.type == PlyT::LIST) {
                next_token(line, i, t);
                p.len = ply_type(t);
                next_token(line, i, t);
                p.item = ply_type(t);
                if (p.len == PlyT::NONE || p.len == PlyT::LIST || p.item == PlyT::NONE || p.item == PlyT::LIST)
                    return bad("unknown PLY type");
            }
            next_token(line, i, p.name);
            elems.back().props.push_back(p);
        } else if (t == "end_header") {
            ended = true;
            break;
        }
    }
    if (!ended) return bad("invalid PLY file");
    bool has_v = false, has_f = false;
    for (auto& e : elems) {
        auto has = [&](const char* n) {
            for (auto& p : e.props)
                if (p.name == n) return true;
            return false;
        };
        if (e.name == "vertex") {
            has_v = true;
            if (!has("x") || !has("y") || !has("z")) return bad("vertex element without x, y, z");
        }
        if (e.name == "face") {
            has_f = true;
            if (!has("vertex_indices")) return bad("face element without vertex_indices");
        }
    }
    if (!has_v || !has_f) return bad("no vertex or face element");
    std::vector<double> verts;
    std::vector<std::vector<uint64_t>> faces;
    for (auto& e : elems) {
        for (uint64_t r = 0; r < e.count; ++r) {
            size_t i = 0;
            if (ascii && !std::getline(f, line)) return bad("truncated PLY body");
            double xyz[3] = {0.0, 0.0, 0.0};
            std::vector<uint64_t> idx;
            for (auto& p : e.props) {
                if (p.type == PlyT::LIST) {
                    double n = 0.0;
                    if (ascii) {
                        next_token(line, i, t);
                        n = double(std::strtoull(t.c_str(), nullptr, 10));
                    } else if (!read_binary(f, p.len, n)) {
                        return bad("truncated PLY body");
                    }
                    const bool keep = e.name == "face" && p.name == "vertex_indices";
                    for (uint64_t k = 0; k < uint64_t(n); ++k) {
                        double v = 0.0;
                        if (ascii) {
                            next_token(line, i, t);
                            v = double(std::strtoull(t.c_str(), nullptr, 10));
                        } else if (!read_binary(f, p.item, v)) {
                            return bad("truncated PLY body");
                        }
                        if (keep) idx.push_back(uint64_t(v));
                    }
                    continue;
                }
                double v = 0.0;
                if (ascii) {
                    next_token(line, i, t);
                    v = parse_number(t);
                } else if (!read_binary(f, p.type, v)) {
                    return bad("truncated PLY body");
                }
                if (e.name == "vertex") {
                    if (p.name == "x") xyz[0] = v;
                    else if (p.name == "y") xyz[1] = v;
                    else if (p.name == "z") xyz[2] = v;
                }
            }
            if (e.name == "vertex")
                for (int row = 0; row < 3; ++row)
                    verts.push_back(xf.m[row][0] * xyz[0] + xf.m[row][1] * xyz[1] + xf.m[row][2] * xyz[2] + xf.m[row][3]);
            if (e.name == "face") faces.push_back(idx);
        }
    }
    const uint64_t nv = verts.size() / 3;
    for (auto& fc : faces) {
        if (fc.size() < 3) continue;
        for (uint64_t a : fc)
            if (a >= nv) return bad("face index out of range");
        for (size_t k = 2; k < fc.size(); ++k) {
            const uint64_t c[3] = {fc[0], fc[k - 1], fc[k]};
            for (uint64_t a : c)
                for (int d = 0; d < 3; ++d) tris.push_back(float(verts[3 * a + d]));
            tri_mat.push_back(0);
        }
    }
    return RC_OK;
}

}  // namespace

struct rch_obj {
    std::vector<float> tris;
    std::vector<uint32_t> tri_mat;
    std::vector<rc_material> materials;
};

extern "C" void rch_obj_default_info(rch_obj_info* info) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    for (int k = 0; k < 3; ++k) info->scale[k] = 1.0;
    info->base.shininess = 2.0f;  // Material::specularShininess default
}

extern "C" int rch_obj_load(const char* path, const rch_obj_info* info, rch_obj** out) {
    if (!path || !out) return rc_host_fail(RC_ERR_INVALID, "rch_obj_load: null argument");
    Loader L;
    if (info) L.info = *info;
    else rch_obj_default_info(&L.info);
    int rc = L.load(path);
    if (rc != RC_OK) return rc;
    if (L.tri_mat.empty()) return rc_host_fail(RC_ERR_INVALID, std::string("rch_obj_load: no triangles in ") + path);
    rch_obj* o = new (std::nothrow) rch_obj();
    if (!o) return rc_host_fail(RC_ERR_ALLOC, "rch_obj_load: out of memory");
    o->tris.swap(L.tris);
    o->tri_mat.swap(L.tri_mat);
    o->materials.swap(L.materials);
    *out = o;
    return RC_OK;
}

extern "C" int rch_ply_load(const char* path, const rch_obj_info* info, rch_obj** out) {
    if (!path || !out) return rc_host_fail(RC_ERR_INVALID, "rch_ply_load: null argument");
    rch_obj_info in;
    if (info) in = *info;
    else rch_obj_default_info(&in);
    std::vector<float> tris;
    std::vector<uint32_t> tri_mat;
    int rc = load_ply(path, model_transform(in), tris, tri_mat);
    if (rc != RC_OK) return rc;
    if (tri_mat.empty()) return rc_host_fail(RC_ERR_INVALID, std::string("rch_ply_load: no triangles in ") + path);
    rch_obj* o = new (std::nothrow) rch_obj();
    if (!o) return rc_host_fail(RC_ERR_ALLOC, "rch_ply_load: out of memory");
    o->tris.swap(tris);
    o->tri_mat.swap(tri_mat);
    o->materials.assign(1, in.base);
    *out = o;
    return RC_OK;
}

extern "C" size_t rch_obj_triangle_count(const rch_obj* o) { return o ? o->tri_mat.size() : 0; }
extern "C" size_t rch_obj_material_count(const rch_obj* o) { return o ? o->materials.size() : 0; }

extern "C" int rch_obj_get(const rch_obj* o, float* vertices, uint32_t* tri_material, rc_material* materials) {
    if (!o) return rc_host_fail(RC_ERR_INVALID, "rch_obj_get: null model");
    if (vertices) std::memcpy(vertices, o->tris.data(), o->tris.size() * sizeof(float));
    if (tri_material) std::memcpy(tri_material, o->tri_mat.data(), o->tri_mat.size() * sizeof(uint32_t));
    if (materials) std::memcpy(materials, o->materials.data(), o->materials.size() * sizeof(rc_material));
    return RC_OK;
}

extern "C" int rch_obj_free(rch_obj* o) {
    delete o;
    return RC_OK;
}
