// Scene files: Scene::loadFromFile / saveToFile (src/Raytracing/Scene.cpp:67-183)
// for cereal's XML and JSON archives of Scene::serialize (Scene.h:330-345).
//
// Both text forms are read into one small tree (an element or object member is
// a named node; a sequence's entries are nodes named value0, value1, ... as
// cereal's XML writes them) and fields are looked up by name, so member order
// does not matter on input.  Output writes every member in the order of the
// reference's serialize functions, with the reference's defaults for what this
// library does not model (textures, other primitives, camera lens settings).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <memory>
#include <string>
#include <vector>

#include "host/rc_host_internal.hpp"
#include "raycer_host.h"

namespace {

// A Node& returned by add() or get() stays valid only until its parent
// receives another child.
struct Node {
    std::string name, text;
    bool seq = false;  // a cereal sequence (size="dynamic" / JSON array)
    std::vector<Node> kids;
    const Node* get(const char* n) const {
        for (const Node& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
    Node& add(const std::string& n, const std::string& t = std::string()) {
        kids.push_back(Node{n, t, false, {}});
        return kids.back();
    }
};

struct ParseError {
    std::string msg;
};

// Nesting the readers accept: scene files nest a few levels; a hostile file
// ("[[[[..." or "<a><a><a>...") is rejected instead of exhausting the stack.
constexpr int kMaxNesting = 256;
struct NestGuard {
    int& depth;
    explicit NestGuard(int& d) : depth(d) {
        if (++depth > kMaxNesting) {
            --depth;
            throw ParseError{"nesting deeper than " + std::to_string(kMaxNesting) + " levels"};
        }
    }
    ~NestGuard() { --depth; }
};

// ---------------------------------------------------------------- XML
struct XmlParser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool starts(const char* s) const { return size_t(e - p) >= std::strlen(s) && std::strncmp(p, s, std::strlen(s)) == 0; }
    void skip_misc() {
        for (;;) {
            ws();
            if (starts("<?")) {
                const char* q = std::strstr(p, "?>");
                if (!q || q >= e) throw ParseError{"unterminated <?"};
                p = q + 2;
            } else if (starts("<!--")) {
                const char* q = std::strstr(p, "-->");
                if (!q || q >= e) throw ParseError{"unterminated comment"};
                p = q + 3;
            } else {
                return;
            }
        }
    }
    static std::string unescape(const std::string& s) {
        std::string o;
        for (size_t i = 0; i < s.size(); ++i) {
            if (s[i] != '&') {
                o += s[i];
                continue;
            }
            size_t j = s.find(';', i);
            std::string ent = j == std::string::npos ? std::string() : s.substr(i + 1, j - i - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else throw ParseError{"unknown entity &" + ent + ";"};
            i = j;
        }
        return o;
    }
    Node element() {
        NestGuard guard(depth);
        skip_misc();
        if (p >= e || *p != '<') throw ParseError{"expected an element"};
        ++p;
        Node n;
        while (p < e && !std::strchr(" \t\r\n/>", *p)) n.name += *p++;
        // attributes: only size="dynamic" matters
        for (;;) {
            ws();
            if (p >= e) throw ParseError{"unterminated tag <" + n.name};
            if (*p == '/' || *p == '>') break;
            std::string an, av;
            while (p < e && *p != '=' && *p != ' ' && *p != '>') an += *p++;
            ws();
            if (p < e && *p == '=') {
                ++p;
                ws();
                if (p >= e) throw ParseError{"unterminated attribute"};
                char q = *p++;
                while (p < e && *p != q) av += *p++;
                if (p >= e) throw ParseError{"unterminated attribute"};
                ++p;
            } else if (an.empty()) {
                throw ParseError{"bad tag <" + n.name};
            }
            if (an == "size" && av == "dynamic") n.seq = true;
        }
        if (*p == '/') {
            if (p + 1 >= e || p[1] != '>') throw ParseError{"bad tag <" + n.name};
            p += 2;  // "/>"
            return n;
        }
        ++p;  // '>'
        std::string text;
        for (;;) {
            const char* t = p;
            while (p < e && *p != '<') ++p;
            text.append(t, p);
            if (p >= e) throw ParseError{"unterminated element <" + n.name + ">"};
            if (starts("<!--")) {
                skip_misc();
                continue;
            }
            if (starts("</")) {
                p += 2;
                std::string close;
                while (p < e && *p != '>') close += *p++;
                if (p >= e) throw ParseError{"unterminated </" + close};
                ++p;
                while (!close.empty() && std::isspace((unsigned char)close.back())) close.pop_back();
                if (close != n.name) throw ParseError{"</" + close + "> closes <" + n.name + ">"};
                break;
            }
            n.kids.push_back(element());
        }
        if (n.kids.empty()) {
            size_t a = text.find_first_not_of(" \t\r\n"), b = text.find_last_not_of(" \t\r\n");
            n.text = a == std::string::npos ? std::string() : unescape(text.substr(a, b - a + 1));
        }
        return n;
    }
    Node parse() {
        skip_misc();
        Node root = element();  // <cereal>
        skip_misc();
        return root;
    }
};

// ---------------------------------------------------------------- JSON
struct JsonParser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    std::string str() {
        if (p >= e || *p != '"') throw ParseError{"expected a string"};
        ++p;
        std::string o;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) break;
                switch (*p) {
                case 'n': o += '\n'; break;
                case 't': o += '\t'; break;
                case 'r': o += '\r'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'u': {
                    if (e - p < 5) throw ParseError{"bad \\u escape"};
                    unsigned v = unsigned(std::strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16));
                    if (v < 0x80) o += char(v);
                    else if (v < 0x800) { o += char(0xC0 | (v >> 6)); o += char(0x80 | (v & 0x3F)); }
                    else { o += char(0xE0 | (v >> 12)); o += char(0x80 | ((v >> 6) & 0x3F)); o += char(0x80 | (v & 0x3F)); }
                    p += 4;
                    break;
                }
                default: o += *p; break;
                }
                ++p;
            } else {
                o += *p++;
            }
        }
        if (p >= e) throw ParseError{"unterminated string"};
        ++p;
        return o;
    }
    void value(Node& n) {
        NestGuard guard(depth);
        ws();
        if (p >= e) throw ParseError{"unexpected end"};
        if (*p == '{') {
            ++p;
            ws();
            if (p < e && *p == '}') { ++p; return; }
            for (;;) {
                ws();
                std::string k = str();
                ws();
                if (p >= e || *p != ':') throw ParseError{"expected ':'"};
                ++p;
                Node& c = n.add(k);
                value(c);
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; return; }
                throw ParseError{"expected ',' or '}'"};
            }
        } else if (*p == '[') {
            ++p;
            n.seq = true;
            ws();
            if (p < e && *p == ']') { ++p; return; }
            for (size_t i = 0;; ++i) {
                Node& c = n.add("value" + std::to_string(i));
                value(c);
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; return; }
                throw ParseError{"expected ',' or ']'"};
            }
        } else if (*p == '"') {
            n.text = str();
        } else {
            const char* t = p;
            while (p < e && !std::strchr(",}] \t\r\n", *p)) ++p;
            n.text.assign(t, p);
        }
    }
    Node parse() {
        Node root;
        value(root);
        return root;
    }
};

// ---------------------------------------------------------------- field readers
double num(const Node* n, const char* f, double dflt) {
    const Node* c = n ? n->get(f) : nullptr;
    if (!c || c->text.empty()) return dflt;
    char* end = nullptr;
    double v = std::strtod(c->text.c_str(), &end);
    if (end == c->text.c_str()) throw ParseError{std::string("field ") + f + " is not a number: " + c->text};
    return v;
}
bool flag(const Node* n, const char* f, bool dflt) {
    const Node* c = n ? n->get(f) : nullptr;
    if (!c) return dflt;
    return c->text == "true" || c->text == "1";
}
void vec3(const Node* n, const char* f, double d0, double d1, double d2, double out[3], const char* a = "x",
          const char* b = "y", const char* cc = "z") {
    const Node* c = n ? n->get(f) : nullptr;
    out[0] = num(c, a, d0);
    out[1] = num(c, b, d1);
    out[2] = num(c, cc, d2);
}
void color3(const Node* n, const char* f, double d, float out[3]) {
    double v[3];
    vec3(n, f, d, d, d, v, "r", "g", "b");
    for (int k = 0; k < 3; ++k) out[k] = float(v[k]);
}

// reference SamplerType / ToneMapperType order <-> this library's enums
const uint32_t kSamplerFromRef[6] = {RC_SAMPLER_CENTER, RC_SAMPLER_RANDOM, RC_SAMPLER_REGULAR,
                                     RC_SAMPLER_JITTERED, RC_SAMPLER_CMJ, RC_SAMPLER_POISSON_DISC};
uint32_t sampler_from_ref(double v) {
    uint32_t i = uint32_t(v);
    if (i > 5) throw ParseError{"unknown sampler type " + std::to_string(i)};
    return kSamplerFromRef[i];
}
uint32_t sampler_to_ref(uint32_t s) {
    for (uint32_t i = 0; i < 6; ++i)
        if (kSamplerFromRef[i] == s) return i;
    return 4;
}
const uint32_t kToneFromRef[4] = {RC_TONEMAP_PASSTHROUGH, RC_TONEMAP_LINEAR, RC_TONEMAP_SIMPLE, RC_TONEMAP_REINHARD};
uint32_t tone_to_ref(uint32_t t) {
    for (uint32_t i = 0; i < 4; ++i)
        if (kToneFromRef[i] == t) return i;
    return 1;
}

void read_material(const Node* m, rc_material& a, rc_ray_material& r) {
    a = rc_material{};
    color3(m, "diffuseReflectance", 0.0, a.diffuse);
    color3(m, "emittance", 0.0, a.emittance);
    color3(m, "ambientReflectance", 0.0, a.ambient);
    color3(m, "specularReflectance", 0.0, a.specular);
    a.shininess = float(num(m, "specularShininess", 2.0));
    a.flags = (flag(m, "isEmissive", false) ? RC_MAT_EMISSIVE : 0u) | (flag(m, "nonShadowing", false) ? RC_MAT_NON_SHADOWING : 0u);
    r = rc_ray_material{};
    r.skip_lighting = flag(m, "skipLighting", false);
    r.fresnel_reflection = flag(m, "fresnelReflection", false);
    r.refractive_index = float(num(m, "refractiveIndex", 1.0));
    r.ray_reflectance = float(num(m, "rayReflectance", 0.0));
    r.refl_sampler = sampler_from_ref(num(m, "rayReflectanceGlossinessSamplerType", 4));
    r.refl_n = uint32_t(num(m, "rayReflectanceGlossinessSampleCountSqrt", 0));
    r.refl_glossiness = float(num(m, "rayReflectanceGlossiness", 1000.0));
    r.ray_transmittance = float(num(m, "rayTransmittance", 0.0));
    r.trans_sampler = sampler_from_ref(num(m, "rayTransmittanceGlossinessSamplerType", 4));
    r.trans_n = uint32_t(num(m, "rayTransmittanceGlossinessSampleCountSqrt", 0));
    r.trans_glossiness = float(num(m, "rayTransmittanceGlossiness", 1000.0));
    r.enable_attenuation = flag(m, "enableRayTransmissionAttenuation", false);
    r.attenuation_factor = float(num(m, "rayTransmissionAttenuationFactor", 1.0));
    color3(m, "rayTransmissionAttenuationColor", 0.0, r.attenuation_color);
}

// Light's colour and intensity; nested under "light" once per base class.
void read_light_base(const Node* l, rc_light& out) {
    color3(l, "color", 1.0, out.color);
    out.intensity = float(num(l, "intensity", 1.0));
}
void read_point(const Node* pnode, rc_light& out) {
    read_light_base(pnode ? pnode->get("light") : nullptr, out);
    double v[3];
    vec3(pnode, "position", 0, 0, 0, v);
    for (int k = 0; k < 3; ++k) out.position[k] = float(v[k]);
    out.max_distance = float(num(pnode, "maxDistance", 1000.0));
    out.attenuation = float(num(pnode, "attenuation", 1.0));
    out.area_light = flag(pnode, "enableAreaLight", false);
    out.area_sampler = sampler_from_ref(num(pnode, "areaLightSamplerType", 4));
    out.area_n = uint32_t(num(pnode, "areaLightSampleCountSqrt", 0));
    out.area_radius = float(num(pnode, "areaLightRadius", 1.0));
}

template <class T>
T* copy_out(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    if (!p) throw std::bad_alloc();
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

rch_scene_file_data* build(const Node& root) {
    // cereal's root holds the scene as its first member ("scene" when saved with a name)
    const Node* sc = root.get("scene");
    if (!sc && !root.kids.empty()) sc = &root.kids.front();
    if (!sc) throw ParseError{"no scene in the file"};
    auto* d = static_cast<rch_scene_file_data*>(std::calloc(1, sizeof(rch_scene_file_data)));
    if (!d) throw std::bad_alloc();
    try {
        const Node* g = sc->get("general");
        d->tracer_type = uint32_t(num(g, "tracerType", 0));
        rc_render_cfg& c = d->cfg;
        const uint32_t n = uint32_t(num(g, "multiSampleCountSqrt", 1));
        c.spp = n * n;
        c.sampler = sampler_from_ref(num(g, "multiSamplerType", 4));
        c.filter = uint32_t(num(g, "multiSamplerFilterType", 4));
        if (c.filter > RC_FILTER_LANCZOS_SINC) throw ParseError{"unknown filter type"};
        // the reference filters' own sizes (BoxFilter.h, TentFilter.h, GaussianFilter.h, MitchellFilter.h, LanczosSincFilter.h)
        const float radius[6] = {0.5f, 1.0f, 1.5f, 0.0f, 2.0f, 2.0f};
        c.filter_radius = radius[c.filter];
        if (c.filter == RC_FILTER_GAUSSIAN) c.filter_param[0] = 1.0f;
        if (c.filter == RC_FILTER_MITCHELL) c.filter_param[0] = c.filter_param[1] = float(1.0 / 3.0);
        c.max_path_length = uint32_t(num(g, "maxPathLength", 3));
        c.ray_start_offset = float(num(g, "rayStartOffset", 0.00001));
        rc_ray_settings& r = d->ray;
        color3(g, "backgroundColor", 0.0, r.background);
        r.max_ray_iterations = uint32_t(num(g, "maxRayIterations", 3));
        r.visualize_depth = flag(g, "visualizeDepth", false);
        r.visualize_depth_max = float(num(g, "visualizeDepthMaxDistance", 25.0));

        const Node* cam = sc->get("camera");
        double v[3];
        vec3(cam, "position", 0, 0, 0, v);
        for (int k = 0; k < 3; ++k) d->camera.position[k] = float(v[k]);
        vec3(cam, "orientation", 0, 0, 0, v, "pitch", "yaw", "roll");
        d->camera.pitch = float(v[0]);
        d->camera.yaw = float(v[1]);
        d->camera.roll = float(v[2]);
        d->camera.fov = float(num(cam, "fov", 75.0));
        if (num(cam, "projectionType", 0) != 0) throw ParseError{"only the perspective camera is supported"};

        const Node* tm = sc->get("toneMapper");
        const uint32_t tref = uint32_t(num(tm, "type", 1));
        if (tref > 3) throw ParseError{"unknown tone mapper type"};
        d->tonemap.type = kToneFromRef[tref];
        d->tonemap.apply_gamma = flag(tm, "applyGamma", true);
        d->tonemap.clamp = flag(tm, "shouldClamp", true);
        d->tonemap.gamma = float(num(tm, "gamma", 2.2));
        d->tonemap.exposure = float(num(tm, "exposure", 0.0));
        d->tonemap.key = float(num(tm, "key", 0.18));
        d->tonemap.enable_averaging = flag(tm, "enableAveraging", false);
        d->tonemap.averaging_alpha = float(num(tm, "averagingAlpha", 0.1));
        d->tonemap.fast_pow = 1;

        const Node* fog = sc->get("simpleFog");
        r.fog_enabled = flag(fog, "enabled", false);
        color3(fog, "color", 1.0, r.fog_color);
        r.fog_distance = float(num(fog, "distance", 100.0));
        r.fog_steepness = float(num(fog, "steepness", 1.0));
        r.fog_height_dispersion = flag(fog, "heightDispersion", false);
        r.fog_height = float(num(fog, "height", 100.0));
        r.fog_height_steepness = float(num(fog, "heightSteepness", 1.0));

        // materials, by Material::id
        std::vector<rc_material> mats;
        std::vector<rc_ray_material> rmats;
        std::vector<uint64_t> ids;
        std::map<uint64_t, uint32_t> by_id;
        if (const Node* ms = sc->get("materials"))
            for (const Node& m : ms->kids) {
                rc_material a;
                rc_ray_material b;
                read_material(&m, a, b);
                uint64_t id = uint64_t(num(&m, "id", 0));
                by_id[id] = uint32_t(mats.size());
                mats.push_back(a);
                rmats.push_back(b);
                ids.push_back(id);
            }

        // lights: the ambient light into the settings, then directional, point, spot
        std::vector<rc_light> lights;
        if (const Node* ls = sc->get("lights")) {
            const Node* amb = ls->get("ambientLight");
            const Node* al = amb ? amb->get("light") : nullptr;
            color3(al, "color", 1.0, r.ambient_color);
            r.ambient_intensity = float(num(al, "intensity", 1.0));
            r.ao_enabled = flag(amb, "enableAmbientOcclusion", false);
            r.ao_sampler = sampler_from_ref(num(amb, "ambientOcclusionSamplerType", 4));
            r.ao_n = uint32_t(num(amb, "ambientOcclusionSampleCountSqrt", 0));
            r.ao_max_distance = float(num(amb, "ambientOcclusionMaxSampleDistance", 1.0));
            r.ao_distribution = float(num(amb, "ambientOcclusionSampleDistribution", 1.0));
            if (const Node* dl = ls->get("directionalLights"))
                for (const Node& x : dl->kids) {
                    rc_light L{};
                    L.type = RC_LIGHT_DIRECTIONAL;
                    read_light_base(x.get("light"), L);
                    vec3(&x, "direction", 0, -1, 0, v);
                    for (int k = 0; k < 3; ++k) L.direction[k] = float(v[k]);
                    lights.push_back(L);
                }
            if (const Node* pl = ls->get("pointLights"))
                for (const Node& x : pl->kids) {
                    rc_light L{};
                    L.type = RC_LIGHT_POINT;
                    read_point(&x, L);
                    lights.push_back(L);
                }
            if (const Node* sl = ls->get("spotLights"))
                for (const Node& x : sl->kids) {
                    rc_light L{};
                    L.type = RC_LIGHT_SPOT;
                    read_point(x.get("light"), L);  // SpotLight's PointLight base
                    vec3(&x, "direction", 0, -1, 0, v);
                    for (int k = 0; k < 3; ++k) L.direction[k] = float(v[k]);
                    L.angle = float(num(&x, "angle", 45.0));
                    L.side_attenuation = float(num(&x, "sideAttenuation", 1.0));
                    lights.push_back(L);
                }
        } else {
            r.ambient_color[0] = r.ambient_color[1] = r.ambient_color[2] = 1.0f;
            r.ambient_intensity = 1.0f;
        }

        // models
        std::vector<rch_scene_model> models;
        if (const Node* ms = sc->get("models"))
            for (const Node& x : ms->kids) {
                rch_scene_model m{};
                const Node* pth = x.get("modelFilePath");
                std::string path = pth ? pth->text : std::string();
                if (path.size() >= sizeof(m.path)) throw ParseError{"model path too long"};
                std::memcpy(m.path, path.c_str(), path.size() + 1);
                rch_obj_default_info(&m.info);
                vec3(&x, "scale", 1, 1, 1, m.info.scale);
                vec3(&x, "rotate", 0, 0, 0, m.info.rotate, "pitch", "yaw", "roll");
                vec3(&x, "translate", 0, 0, 0, m.info.translate);
                if (const Node* bm = x.get("baseMaterial")) {
                    rc_ray_material unused;
                    read_material(bm, m.info.base, unused);
                }
                m.info.ignore_materials = flag(&x, "ignoreMaterials", false);
                m.invisible_triangles = flag(&x, "invisibleTriangles", false);
                models.push_back(m);
            }

        // visible triangles (Scene.cpp:423-470 keeps invisible ones out of the root BVH)
        std::vector<float> verts;
        std::vector<uint32_t> tri_mat;
        std::vector<uint64_t> tri_ids;
        std::map<uint64_t, uint32_t> tri_by_id;
        size_t skipped = 0;
        if (const Node* pr = sc->get("primitives")) {
            for (const Node& grp : pr->kids)
                if (grp.name != "triangles") skipped += grp.kids.size();
            if (const Node* ts = pr->get("triangles"))
                for (const Node& t : ts->kids) {
                    const Node* base = t.get("primitive");
                    if (flag(base, "invisible", false)) {
                        ++skipped;
                        continue;
                    }
                    const Node* vs = t.get("vertices");
                    if (!vs || vs->kids.size() != 3) throw ParseError{"a triangle without three vertices"};
                    for (const Node& q : vs->kids) {
                        verts.push_back(float(num(&q, "x", 0)));
                        verts.push_back(float(num(&q, "y", 0)));
                        verts.push_back(float(num(&q, "z", 0)));
                    }
                    uint64_t mid = uint64_t(num(base, "materialId", 0));
                    auto it = by_id.find(mid);
                    if (it == by_id.end())
                        throw ParseError{"a triangle uses material id " + std::to_string(mid) + ", which the file does not define"};
                    tri_by_id[uint64_t(num(base, "id", 0))] = uint32_t(tri_mat.size());
                    tri_mat.push_back(it->second);
                    tri_ids.push_back(uint64_t(num(base, "id", 0)));
                }
        }

        // a built root BVH over exactly these triangles
        std::vector<rc_flat_node> nodes;
        std::vector<uint64_t> ordered;
        const Node* rb = sc->get("rootBVH");
        const Node* bvh = rb ? rb->get("bvh") : nullptr;
        if (bvh && flag(bvh, "hasBeenBuilt", false) && bvh->get("flatNodes") && bvh->get("orderedPrimitiveIds")) {
            bool ok = true;
            for (const Node& x : bvh->get("orderedPrimitiveIds")->kids) {
                auto it = tri_by_id.find(uint64_t(std::strtoull(x.text.c_str(), nullptr, 10)));
                if (it == tri_by_id.end()) {
                    ok = false;
                    break;
                }
                ordered.push_back(it->second);
            }
            ok = ok && ordered.size() == tri_mat.size();
            if (ok)
                for (const Node& x : bvh->get("flatNodes")->kids) {
                    rc_flat_node fn{};
                    const Node* ab = x.get("aabb");
                    vec3(ab, "min", 0, 0, 0, fn.aabb_min);
                    vec3(ab, "max", 0, 0, 0, fn.aabb_max);
                    fn.right_offset = int64_t(num(&x, "rightOffset", 0));
                    fn.start_offset = uint64_t(num(&x, "startOffset", 0));
                    fn.primitive_count = uint64_t(num(&x, "primitiveCount", 0));
                    if (fn.right_offset == 0 && fn.start_offset + fn.primitive_count > ordered.size()) {
                        ok = false;
                        break;
                    }
                    nodes.push_back(fn);
                }
            if (!ok) {
                nodes.clear();
                ordered.clear();
            }
        }

        d->nmat = mats.size();
        d->materials = copy_out(mats);
        d->ray_materials = copy_out(rmats);
        d->material_ids = copy_out(ids);
        d->ntri = tri_mat.size();
        d->vertices = copy_out(verts);
        d->tri_material = copy_out(tri_mat);
        d->tri_ids = copy_out(tri_ids);
        d->nlights = lights.size();
        d->lights = copy_out(lights);
        d->nmodels = models.size();
        d->models = copy_out(models);
        d->nnodes = nodes.size();
        d->nodes = copy_out(nodes);
        d->ordered = copy_out(ordered);
        d->skipped_primitives = skipped;
        return d;
    } catch (...) {
        rch_scene_file_free(d);
        throw;
    }
}

// ---------------------------------------------------------------- writer
std::string fmt(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}
std::string fmt_u(uint64_t v) { return std::to_string(v); }
const char* fmt_b(bool v) { return v ? "true" : "false"; }

void put_vec(Node& n, const char* name, double x, double y, double z) {
    Node& c = n.add(name);
    c.add("x", fmt(x));
    c.add("y", fmt(y));
    c.add("z", fmt(z));
}
void put_color(Node& n, const char* name, const float* c3) {
    Node& c = n.add(name);
    c.add("r", fmt(c3[0]));
    c.add("g", fmt(c3[1]));
    c.add("b", fmt(c3[2]));
    c.add("a", "1");
}
void put_euler(Node& n, const char* name, double p, double y, double r) {
    Node& c = n.add(name);
    c.add("pitch", fmt(p));
    c.add("yaw", fmt(y));
    c.add("roll", fmt(r));
}
Node& put_seq(Node& n, const char* name) {
    Node& c = n.add(name);
    c.seq = true;
    return c;
}

// Material::serialize (Material.h:80-120)
void put_material(Node& n, const rc_material& a, const rc_ray_material& r, uint64_t id) {
    n.add("id", fmt_u(id));
    n.add("skipLighting", fmt_b(r.skip_lighting));
    n.add("nonShadowing", fmt_b(a.flags & RC_MAT_NON_SHADOWING));
    n.add("normalInterpolation", "false");
    n.add("invertNormal", "false");
    n.add("enableCSG", "false");
    n.add("isEmissive", fmt_b(a.flags & RC_MAT_EMISSIVE));
    Node& ts = n.add("texcoordScale");
    ts.add("x", "1");
    ts.add("y", "1");
    put_color(n, "ambientReflectance", a.ambient);
    n.add("ambientMapTextureId", "0");
    put_color(n, "diffuseReflectance", a.diffuse);
    n.add("diffuseMapTextureId", "0");
    put_color(n, "specularReflectance", a.specular);
    n.add("specularShininess", fmt(a.shininess));
    n.add("specularMapTextureId", "0");
    put_color(n, "emittance", a.emittance);
    n.add("emittanceMapTextureId", "0");
    n.add("fresnelReflection", fmt_b(r.fresnel_reflection));
    n.add("refractiveIndex", fmt(r.refractive_index));
    n.add("rayReflectance", fmt(r.ray_reflectance));
    n.add("rayReflectanceMapTextureId", "0");
    n.add("rayReflectanceGlossinessSamplerType", fmt_u(sampler_to_ref(r.refl_sampler)));
    n.add("rayReflectanceGlossinessSampleCountSqrt", fmt_u(r.refl_n));
    n.add("rayReflectanceGlossiness", fmt(r.refl_glossiness));
    n.add("rayTransmittance", fmt(r.ray_transmittance));
    n.add("rayTransmittanceMapTextureId", "0");
    n.add("rayTransmittanceGlossinessSamplerType", fmt_u(sampler_to_ref(r.trans_sampler)));
    n.add("rayTransmittanceGlossinessSampleCountSqrt", fmt_u(r.trans_n));
    n.add("rayTransmittanceGlossiness", fmt(r.trans_glossiness));
    n.add("enableRayTransmissionAttenuation", fmt_b(r.enable_attenuation));
    n.add("rayTransmissionAttenuationFactor", fmt(r.attenuation_factor));
    put_color(n, "rayTransmissionAttenuationColor", r.attenuation_color);
    n.add("normalMapTextureId", "0");
    n.add("maskMapTextureId", "0");
    n.add("heightMapTextureId", "0");
}

void put_light_base(Node& n, const rc_light& L) {
    Node& l = n.add("light");
    put_color(l, "color", L.color);
    l.add("intensity", fmt(L.intensity));
}
void put_point(Node& n, const rc_light& L) {
    put_light_base(n, L);
    put_vec(n, "position", L.position[0], L.position[1], L.position[2]);
    n.add("maxDistance", fmt(L.max_distance));
    n.add("attenuation", fmt(L.attenuation));
    n.add("enableAreaLight", fmt_b(L.area_light));
    n.add("areaLightSamplerType", fmt_u(sampler_to_ref(L.area_sampler)));
    n.add("areaLightSampleCountSqrt", fmt_u(L.area_n));
    n.add("areaLightRadius", fmt(L.area_radius));
}

Node scene_tree(const rch_scene_file_data& d) {
    Node root{"cereal", "", false, {}};
    Node& s = root.add("scene");
    const rc_render_cfg& c = d.cfg;
    const rc_ray_settings& r = d.ray;
    const float black[3] = {0, 0, 0}, white[3] = {1, 1, 1};
    // Scene::General (Scene.h:89-131)
    Node& g = s.add("general");
    g.add("tracerType", fmt_u(d.tracer_type));
    g.add("maxRayIterations", fmt_u(r.max_ray_iterations));
    g.add("maxPathLength", fmt_u(c.max_path_length));
    g.add("pathSampleCount", "1");
    g.add("rayStartOffset", fmt(c.ray_start_offset));
    put_color(g, "backgroundColor", r.background);
    put_color(g, "offLensColor", black);
    g.add("multiSamplerType", fmt_u(sampler_to_ref(c.sampler)));
    g.add("multiSamplerFilterType", fmt_u(c.filter));
    uint32_t n = 1;
    while ((n + 1) * (n + 1) <= c.spp) ++n;
    g.add("multiSampleCountSqrt", fmt_u(n));
    g.add("timeSamplerType", "3");
    g.add("timeSampleCount", "1");
    g.add("cameraSamplerType", "4");
    g.add("cameraSampleCountSqrt", "1");
    g.add("visualizeDepth", fmt_b(r.visualize_depth));
    g.add("visualizeDepthMaxDistance", fmt(r.visualize_depth_max));
    g.add("enableNormalMapping", "true");
    // Camera::serialize (Camera.h:88-103)
    Node& cam = s.add("camera");
    put_vec(cam, "position", d.camera.position[0], d.camera.position[1], d.camera.position[2]);
    put_euler(cam, "orientation", d.camera.pitch, d.camera.yaw, d.camera.roll);
    cam.add("projectionType", "0");
    cam.add("fov", fmt(d.camera.fov));
    cam.add("orthoSize", "10");
    cam.add("fishEyeAngle", "180");
    cam.add("apertureSize", "0.10000000000000001");
    cam.add("focalDistance", "10");
    cam.add("isTimeVariant", "false");
    put_vec(cam, "translateInTime", 0, 0, 0);
    put_euler(cam, "rotateInTime", 0, 0, 0);
    // Scene::ToneMapper
    Node& tm = s.add("toneMapper");
    tm.add("type", fmt_u(tone_to_ref(d.tonemap.type)));
    tm.add("applyGamma", fmt_b(d.tonemap.apply_gamma));
    tm.add("shouldClamp", fmt_b(d.tonemap.clamp));
    tm.add("gamma", fmt(d.tonemap.gamma));
    tm.add("exposure", fmt(d.tonemap.exposure));
    tm.add("key", fmt(d.tonemap.key));
    tm.add("enableAveraging", fmt_b(d.tonemap.enable_averaging));
    tm.add("averagingAlpha", fmt(d.tonemap.averaging_alpha));
    // Scene::SimpleFog, Scene::VolumetricFog
    Node& f = s.add("simpleFog");
    f.add("enabled", fmt_b(r.fog_enabled));
    put_color(f, "color", r.fog_color);
    f.add("distance", fmt(r.fog_distance));
    f.add("steepness", fmt(r.fog_steepness));
    f.add("heightDispersion", fmt_b(r.fog_height_dispersion));
    f.add("height", fmt(r.fog_height));
    f.add("heightSteepness", fmt(r.fog_height_steepness));
    Node& vf = s.add("volumetricFog");
    vf.add("enabled", "false");
    put_color(vf, "color", white);
    vf.add("density", "1");
    vf.add("steps", "0");
    // Scene::RootBVH with FlatBVH::serialize (FlatBVH.h:96-105)
    Node& rb = s.add("rootBVH");
    rb.add("enabled", "true");
    Node& bi = rb.add("buildInfo");
    bi.add("maxLeafSize", "5");
    bi.add("useSAH", "true");
    bi.add("regularSAHSplits", "0");
    bi.add("axisSelection", "0");
    bi.add("axisSplit", "1");
    Node& bvh = rb.add("bvh");
    Node& bp = bvh.add("primitive");
    bp.add("id", "0");
    bp.add("materialId", "0");
    bp.add("invisible", "false");
    bvh.add("hasBeenBuilt", fmt_b(d.nnodes > 0));
    Node& fn = put_seq(bvh, "flatNodes");
    for (size_t i = 0; i < d.nnodes; ++i) {
        const rc_flat_node& x = d.nodes[i];
        Node& e = fn.add("value" + std::to_string(i));
        Node& ab = e.add("aabb");
        put_vec(ab, "min", x.aabb_min[0], x.aabb_min[1], x.aabb_min[2]);
        put_vec(ab, "max", x.aabb_max[0], x.aabb_max[1], x.aabb_max[2]);
        e.add("rightOffset", std::to_string(x.right_offset));
        e.add("startOffset", fmt_u(x.start_offset));
        e.add("primitiveCount", fmt_u(x.primitive_count));
    }
    Node& op = put_seq(bvh, "orderedPrimitiveIds");
    for (size_t i = 0; i < (d.nnodes ? d.ntri : 0); ++i)
        op.add("value" + std::to_string(i), fmt_u(d.tri_ids ? d.tri_ids[d.ordered[i]] : d.ordered[i] + 1));
    // Scene::BoundingBoxes, Scene::Textures
    Node& bb = s.add("boundingBoxes");
    bb.add("enabled", "false");
    bb.add("useDefaultMaterial", "true");
    {
        rc_material m{};
        m.shininess = 2.0f;
        rc_ray_material rm{};
        rm.refractive_index = 1.0f;
        rm.refl_sampler = rm.trans_sampler = RC_SAMPLER_CMJ;
        rm.refl_glossiness = rm.trans_glossiness = 1000.0f;
        rm.attenuation_factor = 1.0f;
        put_material(bb.add("material"), m, rm, 0);
    }
    Node& tx = s.add("textures");
    for (const char* t : {"colorTextures", "colorGradientTextures", "checkerTextures", "imageTextures",
                          "perlinNoiseTextures", "valueNoiseTextures", "cellNoiseTextures", "marbleTextures",
                          "woodTextures", "fireTextures", "atmosphereTextures", "voronoiTextures"})
        put_seq(tx, t);
    Node& ms = put_seq(s, "materials");
    for (size_t i = 0; i < d.nmat; ++i)
        put_material(ms.add("value" + std::to_string(i)), d.materials[i], d.ray_materials[i],
                     d.material_ids ? d.material_ids[i] : uint64_t(i + 1));
    // Scene::Lights (Scene.h:240-255, Lights.h)
    Node& ls = s.add("lights");
    Node& amb = ls.add("ambientLight");
    Node& al = amb.add("light");
    put_color(al, "color", r.ambient_color);
    al.add("intensity", fmt(r.ambient_intensity));
    amb.add("enableAmbientOcclusion", fmt_b(r.ao_enabled));
    amb.add("ambientOcclusionSamplerType", fmt_u(sampler_to_ref(r.ao_sampler)));
    amb.add("ambientOcclusionSampleCountSqrt", fmt_u(r.ao_n));
    amb.add("ambientOcclusionMaxSampleDistance", fmt(r.ao_max_distance));
    amb.add("ambientOcclusionSampleDistribution", fmt(r.ao_distribution));
    // one list at a time: a Node reference stays valid only until its parent grows
    for (uint32_t type : {uint32_t(RC_LIGHT_DIRECTIONAL), uint32_t(RC_LIGHT_POINT), uint32_t(RC_LIGHT_SPOT)}) {
        Node& seq = put_seq(ls, type == RC_LIGHT_DIRECTIONAL ? "directionalLights"
                                : type == RC_LIGHT_POINT     ? "pointLights"
                                                             : "spotLights");
        for (size_t i = 0; i < d.nlights; ++i) {
            const rc_light& L = d.lights[i];
            if (L.type != type) continue;
            Node& e = seq.add("value" + std::to_string(seq.kids.size()));
            if (type == RC_LIGHT_DIRECTIONAL) {
                put_light_base(e, L);
                put_vec(e, "direction", L.direction[0], L.direction[1], L.direction[2]);
            } else if (type == RC_LIGHT_POINT) {
                put_point(e, L);
            } else {
                put_point(e.add("light"), L);
                put_vec(e, "direction", L.direction[0], L.direction[1], L.direction[2]);
                e.add("angle", fmt(L.angle));
                e.add("sideAttenuation", fmt(L.side_attenuation));
            }
        }
    }
    // ModelLoaderInfo::serialize (ModelLoader.h:66-84)
    Node& mo = put_seq(s, "models");
    for (size_t i = 0; i < d.nmodels; ++i) {
        const rch_scene_model& m = d.models[i];
        Node& e = mo.add("value" + std::to_string(i));
        e.add("modelFilePath", m.path);
        put_vec(e, "scale", m.info.scale[0], m.info.scale[1], m.info.scale[2]);
        put_euler(e, "rotate", m.info.rotate[0], m.info.rotate[1], m.info.rotate[2]);
        put_vec(e, "translate", m.info.translate[0], m.info.translate[1], m.info.translate[2]);
        rc_ray_material rm{};
        rm.refractive_index = 1.0f;
        rm.refl_sampler = rm.trans_sampler = RC_SAMPLER_CMJ;
        rm.refl_glossiness = rm.trans_glossiness = 1000.0f;
        rm.attenuation_factor = 1.0f;
        put_material(e.add("baseMaterial"), m.info.base, rm, 0);
        e.add("invisibleTriangles", fmt_b(m.invisible_triangles));
        e.add("ignoreMaterials", fmt_b(m.info.ignore_materials));
        for (const char* k : {"enableGroups", "enableGroupsInstances", "enableCombinedGroup", "enableCombinedGroupInstance"})
            e.add(k, "false");
        for (const char* k : {"defaultMaterialId", "idStartOffset", "combinedGroupId", "combinedGroupInstanceId"})
            e.add(k, "0");
    }
    // Scene::Primitives: triangles only (Triangle.h:44-51, Primitive.h:45-51)
    Node& pr = s.add("primitives");
    Node& tr = put_seq(pr, "triangles");
    for (size_t i = 0; i < d.ntri; ++i) {
        Node& e = tr.add("value" + std::to_string(i));
        Node& b = e.add("primitive");
        b.add("id", fmt_u(d.tri_ids ? d.tri_ids[i] : uint64_t(i + 1)));
        b.add("materialId", fmt_u(d.material_ids ? d.material_ids[d.tri_material[i]] : uint64_t(d.tri_material[i] + 1)));
        b.add("invisible", "false");
        const float* v = d.vertices + 9 * i;
        double e1[3] = {double(v[3]) - v[0], double(v[4]) - v[1], double(v[5]) - v[2]};
        double e2[3] = {double(v[6]) - v[0], double(v[7]) - v[1], double(v[8]) - v[2]};
        double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
        double nl = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (nl > 0) { nx /= nl; ny /= nl; nz /= nl; }
        {
            Node& vs = put_seq(e, "vertices");
            for (int k = 0; k < 3; ++k)
                put_vec(vs, ("value" + std::to_string(k)).c_str(), v[3 * k], v[3 * k + 1], v[3 * k + 2]);
        }
        {
            Node& ns = put_seq(e, "normals");
            for (int k = 0; k < 3; ++k) put_vec(ns, ("value" + std::to_string(k)).c_str(), nx, ny, nz);
        }
        {
            Node& tc = put_seq(e, "texcoords");
            for (int k = 0; k < 3; ++k) {
                Node& t = tc.add("value" + std::to_string(k));
                t.add("x", "0");
                t.add("y", "0");
            }
        }
    }
    for (const char* k : {"planes", "spheres", "boxes", "cylinders", "toruses", "blinnBlobs", "csgs",
                          "primitiveGroups", "instances"})
        put_seq(pr, k);
    return root;
}

std::string xml_escape(const std::string& s) {
    std::string o;
    for (char ch : s) {
        switch (ch) {
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        case '&': o += "&amp;"; break;
        case '"': o += "&quot;"; break;
        default: o += ch; break;
        }
    }
    return o;
}
void write_xml(std::ostream& os, const Node& n, int depth) {
    std::string ind(size_t(depth), '\t');
    os << ind << '<' << n.name << (n.seq ? " size=\"dynamic\"" : "");
    if (n.kids.empty() && n.text.empty()) {
        os << "/>\n";
        return;
    }
    if (n.kids.empty()) {
        os << '>' << xml_escape(n.text) << "</" << n.name << ">\n";
        return;
    }
    os << ">\n";
    for (const Node& k : n.kids) write_xml(os, k, depth + 1);
    os << ind << "</" << n.name << ">\n";
}
std::string json_escape(const std::string& s) {
    std::string o = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') o += '\\';
        if (ch == '\n') { o += "\\n"; continue; }
        o += ch;
    }
    return o + "\"";
}
bool json_number(const std::string& t) {
    if (t.empty()) return false;
    if (t == "true" || t == "false") return true;
    char* end = nullptr;
    std::strtod(t.c_str(), &end);
    return *end == '\0';
}
void write_json(std::ostream& os, const Node& n, int depth, const std::string& leaf_hint) {
    std::string ind(size_t(depth) * 4, ' ');
    if (n.seq) {
        if (n.kids.empty()) { os << "[]"; return; }
        os << "[\n";
        for (size_t i = 0; i < n.kids.size(); ++i) {
            os << ind << "    ";
            write_json(os, n.kids[i], depth + 1, leaf_hint);
            os << (i + 1 < n.kids.size() ? ",\n" : "\n");
        }
        os << ind << ']';
    } else if (n.kids.empty()) {
        // path names stay strings; numbers and booleans are bare
        if (leaf_hint == "modelFilePath" || !json_number(n.text)) os << json_escape(n.text);
        else os << n.text;
    } else {
        os << "{\n";
        for (size_t i = 0; i < n.kids.size(); ++i) {
            os << ind << "    " << json_escape(n.kids[i].name) << ": ";
            write_json(os, n.kids[i], depth + 1, n.kids[i].name);
            os << (i + 1 < n.kids.size() ? ",\n" : "\n");
        }
        os << ind << '}';
    }
}

// ---------------------------------------------------------------- cereal binary
// cereal's BinaryOutputArchive (the ".bin" scene files of Scene::saveToFile,
// Scene.cpp:144-146, read back at :88-90) writes the values of
// Scene::serialize in member order with no names: arithmetic values as their
// little-endian bytes (bool one byte, enum classes as their int), std::string
// and std::vector as a uint64 count then the elements, C arrays as their
// elements alone.  The schema below follows the serialize functions it cites;
// a binary file is read into the same name tree the XML and JSON readers
// build (and written from the tree the XML and JSON writers use), so one
// interpreter serves all three forms.  Lists of textures and of non-triangle
// primitives have no schema here: a file holding any is rejected by name.
enum class BK { U64, I64, I32, BOOL, F64, STR, STRUCT, VEC, ARR, NONE };
struct BType {
    BK kind;
    std::vector<std::pair<std::string, const BType*>> fields;  // STRUCT
    const BType* elem = nullptr;                               // VEC, ARR
    size_t n = 0;                                              // ARR
    std::string what;                                          // NONE: what the list holds
    size_t min_bytes() const {
        switch (kind) {
        case BK::U64: case BK::I64: case BK::F64: case BK::STR: case BK::VEC: return 8;
        case BK::I32: return 4;
        case BK::BOOL: return 1;
        case BK::ARR: return n * elem->min_bytes();
        case BK::STRUCT: {
            size_t b = 0;
            for (const auto& f : fields) b += f.second->min_bytes();
            return b;
        }
        default: return 1;
        }
    }
};

const BType* leaf(BK k) {
    static const BType u64{BK::U64}, i64{BK::I64}, i32{BK::I32}, b{BK::BOOL}, f64{BK::F64}, str{BK::STR};
    switch (k) {
    case BK::U64: return &u64;
    case BK::I64: return &i64;
    case BK::I32: return &i32;
    case BK::BOOL: return &b;
    case BK::F64: return &f64;
    default: return &str;
    }
}

struct Schema {
    std::vector<std::unique_ptr<BType>> pool;
    const BType* st(std::vector<std::pair<std::string, const BType*>> f) {
        pool.push_back(std::unique_ptr<BType>(new BType{BK::STRUCT, std::move(f)}));
        return pool.back().get();
    }
    const BType* vec(const BType* e) {
        pool.push_back(std::unique_ptr<BType>(new BType{BK::VEC, {}, e}));
        return pool.back().get();
    }
    const BType* arr(const BType* e, size_t n) {
        pool.push_back(std::unique_ptr<BType>(new BType{BK::ARR, {}, e, n}));
        return pool.back().get();
    }
    const BType* none(const std::string& what) {
        pool.push_back(std::unique_ptr<BType>(new BType{BK::NONE, {}, nullptr, 0, what}));
        return vec(pool.back().get());
    }
    const BType* scene = nullptr;

    Schema() {
        const BType *U = leaf(BK::U64), *I = leaf(BK::I64), *E = leaf(BK::I32), *B = leaf(BK::BOOL),
                    *D = leaf(BK::F64), *S = leaf(BK::STR);
        const BType* color = st({{"r", D}, {"g", D}, {"b", D}, {"a", D}});            // Color.h:78-84
        const BType* v3 = st({{"x", D}, {"y", D}, {"z", D}});                         // Vector3.h:74-79
        const BType* v2 = st({{"x", D}, {"y", D}});                                   // Vector2.h:63-67
        const BType* euler = st({{"pitch", D}, {"yaw", D}, {"roll", D}});             // EulerAngle.h:55-60
        const BType* general = st({{"tracerType", E}, {"maxRayIterations", U}, {"maxPathLength", U},  // Scene.h:89-131
                                   {"pathSampleCount", U}, {"rayStartOffset", D}, {"backgroundColor", color},
                                   {"offLensColor", color}, {"multiSamplerType", E}, {"multiSamplerFilterType", E},
                                   {"multiSampleCountSqrt", U}, {"timeSamplerType", E}, {"timeSampleCount", U},
                                   {"cameraSamplerType", E}, {"cameraSampleCountSqrt", U}, {"visualizeDepth", B},
                                   {"visualizeDepthMaxDistance", D}, {"enableNormalMapping", B}});
        const BType* camera = st({{"position", v3}, {"orientation", euler}, {"projectionType", E}, {"fov", D},  // Camera.h:88-103
                                  {"orthoSize", D}, {"fishEyeAngle", D}, {"apertureSize", D}, {"focalDistance", D},
                                  {"isTimeVariant", B}, {"translateInTime", v3}, {"rotateInTime", euler}});
        const BType* tone = st({{"type", E}, {"applyGamma", B}, {"shouldClamp", B}, {"gamma", D},  // Scene.h:134-160
                                {"exposure", D}, {"key", D}, {"enableAveraging", B}, {"averagingAlpha", D}});
        const BType* sfog = st({{"enabled", B}, {"color", color}, {"distance", D}, {"steepness", D},  // Scene.h:162-183
                                {"heightDispersion", B}, {"height", D}, {"heightSteepness", D}});
        const BType* vfog = st({{"enabled", B}, {"color", color}, {"density", D}, {"steps", U}});  // Scene.h:185-200
        const BType* build = st({{"maxLeafSize", U}, {"useSAH", B}, {"regularSAHSplits", U},  // FlatBVH.h:19-36
                                 {"axisSelection", E}, {"axisSplit", E}});
        const BType* aabb = st({{"min", v3}, {"max", v3}});                            // AABB.h:44-48
        const BType* fnode = st({{"aabb", aabb}, {"rightOffset", I}, {"startOffset", U},  // FlatBVH.h:38-53
                                 {"primitiveCount", U}});
        const BType* prim = st({{"id", U}, {"materialId", U}, {"invisible", B}});       // Primitive.h:46-51
        const BType* bvh = st({{"primitive", prim}, {"hasBeenBuilt", B}, {"flatNodes", vec(fnode)},  // FlatBVH.h:99-105
                               {"orderedPrimitiveIds", vec(U)}});
        const BType* root_bvh = st({{"enabled", B}, {"buildInfo", build}, {"bvh", bvh}});  // Scene.h:202-216
        const BType* material = st({{"id", U}, {"skipLighting", B}, {"nonShadowing", B},  // Material.h:80-120
            {"normalInterpolation", B}, {"invertNormal", B}, {"enableCSG", B}, {"isEmissive", B},
            {"texcoordScale", v2}, {"ambientReflectance", color}, {"ambientMapTextureId", U},
            {"diffuseReflectance", color}, {"diffuseMapTextureId", U}, {"specularReflectance", color},
            {"specularShininess", D}, {"specularMapTextureId", U}, {"emittance", color},
            {"emittanceMapTextureId", U}, {"fresnelReflection", B}, {"refractiveIndex", D}, {"rayReflectance", D},
            {"rayReflectanceMapTextureId", U}, {"rayReflectanceGlossinessSamplerType", E},
            {"rayReflectanceGlossinessSampleCountSqrt", U}, {"rayReflectanceGlossiness", D},
            {"rayTransmittance", D}, {"rayTransmittanceMapTextureId", U},
            {"rayTransmittanceGlossinessSamplerType", E}, {"rayTransmittanceGlossinessSampleCountSqrt", U},
            {"rayTransmittanceGlossiness", D}, {"enableRayTransmissionAttenuation", B},
            {"rayTransmissionAttenuationFactor", D}, {"rayTransmissionAttenuationColor", color},
            {"normalMapTextureId", U}, {"maskMapTextureId", U}, {"heightMapTextureId", U}});
        const BType* boxes = st({{"enabled", B}, {"useDefaultMaterial", B}, {"material", material}});  // Scene.h:218-232
        const BType* textures = st({{"colorTextures", none("color textures")},  // Scene.h:234-266
            {"colorGradientTextures", none("color gradient textures")}, {"checkerTextures", none("checker textures")},
            {"imageTextures", none("image textures")}, {"perlinNoiseTextures", none("Perlin noise textures")},
            {"valueNoiseTextures", none("value noise textures")}, {"cellNoiseTextures", none("cell noise textures")},
            {"marbleTextures", none("marble textures")}, {"woodTextures", none("wood textures")},
            {"fireTextures", none("fire textures")}, {"atmosphereTextures", none("atmosphere textures")},
            {"voronoiTextures", none("Voronoi textures")}});
        const BType* light = st({{"color", color}, {"intensity", D}});                 // Lights.h:14-24
        const BType* ambient = st({{"light", light}, {"enableAmbientOcclusion", B},    // Lights.h:26-44
                                   {"ambientOcclusionSamplerType", E}, {"ambientOcclusionSampleCountSqrt", U},
                                   {"ambientOcclusionMaxSampleDistance", D},
                                   {"ambientOcclusionSampleDistribution", D}});
        const BType* dir = st({{"light", light}, {"direction", v3}});                   // Lights.h:46-57
        const BType* point = st({{"light", light}, {"position", v3}, {"maxDistance", D}, {"attenuation", D},  // Lights.h:59-81
                                 {"enableAreaLight", B}, {"areaLightSamplerType", E}, {"areaLightSampleCountSqrt", U},
                                 {"areaLightRadius", D}});
        const BType* spot = st({{"light", point}, {"direction", v3}, {"angle", D}, {"sideAttenuation", D}});  // Lights.h:83-97
        const BType* lights = st({{"ambientLight", ambient}, {"directionalLights", vec(dir)},  // Scene.h:272-287
                                  {"pointLights", vec(point)}, {"spotLights", vec(spot)}});
        const BType* model = st({{"modelFilePath", S}, {"scale", v3}, {"rotate", euler}, {"translate", v3},  // ModelLoader.h:66-84
            {"baseMaterial", material}, {"invisibleTriangles", B}, {"ignoreMaterials", B}, {"enableGroups", B},
            {"enableGroupsInstances", B}, {"enableCombinedGroup", B}, {"enableCombinedGroupInstance", B},
            {"defaultMaterialId", U}, {"idStartOffset", U}, {"combinedGroupId", U}, {"combinedGroupInstanceId", U}});
        const BType* tri = st({{"primitive", prim}, {"vertices", arr(v3, 3)}, {"normals", arr(v3, 3)},  // Triangle.h:45-51
                               {"texcoords", arr(v2, 3)}});
        const BType* prims = st({{"triangles", vec(tri)}, {"planes", none("planes")}, {"spheres", none("spheres")},  // Scene.h:291-320
            {"boxes", none("boxes")}, {"cylinders", none("cylinders")}, {"toruses", none("toruses")},
            {"blinnBlobs", none("blinn blobs")}, {"csgs", none("CSG primitives")},
            {"primitiveGroups", none("primitive groups")}, {"instances", none("instances")}});
        scene = st({{"general", general}, {"camera", camera}, {"toneMapper", tone}, {"simpleFog", sfog},  // Scene.h:331-348
                    {"volumetricFog", vfog}, {"rootBVH", root_bvh}, {"boundingBoxes", boxes},
                    {"textures", textures}, {"materials", vec(material)}, {"lights", lights},
                    {"models", vec(model)}, {"primitives", prims}});
    }
};

const Schema& schema() {
    static const Schema s;
    return s;
}

struct BinReader {
    const uint8_t* p;
    const uint8_t* e;
    template <class T>
    T get() {
        if (size_t(e - p) < sizeof(T)) throw ParseError{"binary scene file ends early"};
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    void read(const BType& t, Node& n) {
        switch (t.kind) {
        case BK::U64: n.text = std::to_string(get<uint64_t>()); break;
        case BK::I64: n.text = std::to_string(get<int64_t>()); break;
        case BK::I32: n.text = std::to_string(get<int32_t>()); break;
        case BK::BOOL: {
            const uint8_t b = get<uint8_t>();
            if (b > 1) throw ParseError{"binary scene file: a bool holds " + std::to_string(b)};
            n.text = b ? "true" : "false";
            break;
        }
        case BK::F64: n.text = fmt(get<double>()); break;
        case BK::STR: {
            const uint64_t len = get<uint64_t>();
            if (len > uint64_t(e - p)) throw ParseError{"binary scene file: a string runs past the end"};
            n.text.assign(reinterpret_cast<const char*>(p), size_t(len));
            p += len;
            break;
        }
        case BK::STRUCT:
            for (const auto& f : t.fields) {
                n.kids.push_back(Node{f.first, std::string(), false, {}});
                read(*f.second, n.kids.back());
            }
            break;
        case BK::VEC: {
            const uint64_t count = get<uint64_t>();
            n.seq = true;
            if (count == 0) break;
            if (t.elem->kind == BK::NONE)
                throw ParseError{"binary scene files with " + t.elem->what + " are not supported"};
            if (count > uint64_t(e - p) / std::max<size_t>(1, t.elem->min_bytes()))
                throw ParseError{"binary scene file: a list of " + std::to_string(count) + " runs past the end"};
            n.kids.reserve(size_t(count));
            for (uint64_t i = 0; i < count; ++i) {
                n.kids.push_back(Node{"value" + std::to_string(i), std::string(), false, {}});
                read(*t.elem, n.kids.back());
            }
            break;
        }
        case BK::ARR:
            for (size_t i = 0; i < t.n; ++i) {
                n.kids.push_back(Node{"value" + std::to_string(i), std::string(), false, {}});
                read(*t.elem, n.kids.back());
            }
            break;
        default: throw ParseError{"binary scene file: no schema"};
        }
    }
};

Node parse_binary(const uint8_t* p, size_t len) {
    BinReader R{p, p + len};
    Node root{"cereal", std::string(), false, {}};
    root.kids.push_back(Node{"scene", std::string(), false, {}});
    R.read(*schema().scene, root.kids.back());
    if (R.p != R.e) throw ParseError{"binary scene file: " + std::to_string(R.e - R.p) + " bytes after the scene"};
    return root;
}

template <class T>
void put_bytes(std::string& out, T v) {
    out.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

// Writes the tree node n (null: the member is absent, its zero value is
// written) as type t.
void write_binary(const BType& t, const Node* n, std::string& out) {
    const std::string empty;
    const std::string& tx = n ? n->text : empty;
    switch (t.kind) {
    case BK::U64: put_bytes<uint64_t>(out, std::strtoull(tx.c_str(), nullptr, 10)); break;
    case BK::I64: put_bytes<int64_t>(out, std::strtoll(tx.c_str(), nullptr, 10)); break;
    case BK::I32: put_bytes<int32_t>(out, int32_t(std::strtol(tx.c_str(), nullptr, 10))); break;
    case BK::BOOL: put_bytes<uint8_t>(out, tx == "true" || tx == "1" ? 1 : 0); break;
    case BK::F64: put_bytes<double>(out, std::strtod(tx.c_str(), nullptr)); break;
    case BK::STR:
        put_bytes<uint64_t>(out, tx.size());
        out += tx;
        break;
    case BK::STRUCT:
        for (const auto& f : t.fields) write_binary(*f.second, n ? n->get(f.first.c_str()) : nullptr, out);
        break;
    case BK::VEC: {
        const size_t count = n ? n->kids.size() : 0;
        if (count && t.elem->kind == BK::NONE)
            throw ParseError{"binary scene files with " + t.elem->what + " are not supported"};
        put_bytes<uint64_t>(out, count);
        for (size_t i = 0; i < count; ++i) write_binary(*t.elem, &n->kids[i], out);
        break;
    }
    case BK::ARR:
        for (size_t i = 0; i < t.n; ++i) write_binary(*t.elem, n && i < n->kids.size() ? &n->kids[i] : nullptr, out);
        break;
    default: throw ParseError{"binary scene file: no schema"};
    }
}

bool ends_with(const std::string& s, const char* suf) {
    size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

extern "C" int rch_scene_file_parse(const char* text, size_t len, int format, rch_scene_file_data** out) {
    if ((!text && len) || !out) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_parse: null argument");
    if (format < 0 || format > RCH_SCENE_BINARY) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_parse: unknown format");
    try {
        // a NUL-terminated copy: the parsers may look one byte past the end
        const std::string buf(text ? text : "", len);
        Node root;
        if (format == RCH_SCENE_BINARY) {
            root = parse_binary(reinterpret_cast<const uint8_t*>(buf.data()), buf.size());
        } else if (format == RCH_SCENE_JSON) {
            JsonParser P{buf.c_str(), buf.c_str() + buf.size()};
            root = P.parse();
        } else {
            XmlParser P{buf.c_str(), buf.c_str() + buf.size()};
            root = P.parse();
        }
        *out = build(root);
        return RC_OK;
    } catch (const ParseError& e) {
        return rc_host_fail(RC_ERR_INVALID, "scene file: " + e.msg);
    } catch (const std::bad_alloc&) {
        return rc_host_fail(RC_ERR_ALLOC, "scene file: out of memory");
    }
}

extern "C" int rch_scene_file_load(const char* path, rch_scene_file_data** out) {
    if (!path || !out) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_load: null argument");
    const std::string p(path);
    const int format = ends_with(p, ".json") ? RCH_SCENE_JSON : ends_with(p, ".xml") ? RCH_SCENE_XML
                       : ends_with(p, ".bin") ? RCH_SCENE_BINARY : -1;
    if (format < 0)
        return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_load: unsupported scene file format (.xml, .json or .bin)");
    std::ifstream f(p, std::ios::binary);
    if (!f.good()) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_load: could not open " + p);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    return rch_scene_file_parse(text.data(), text.size(), format, out);
}

extern "C" int rch_scene_file_save(const char* path, const rch_scene_file_data* d) {
    if (!path || !d) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: null argument");
    if ((d->nmat && (!d->materials || !d->ray_materials)) || (d->ntri && (!d->vertices || !d->tri_material)) ||
        (d->nlights && !d->lights) || (d->nmodels && !d->models) || (d->nnodes && (!d->nodes || !d->ordered)))
        return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: a count without its array");
    for (size_t i = 0; i < d->ntri; ++i)
        if (d->tri_material[i] >= d->nmat) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: bad material index");
    for (size_t i = 0; d->nnodes && i < d->ntri; ++i)
        if (d->ordered[i] >= d->ntri) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: bad ordered index");
    const std::string p(path);
    const bool json = ends_with(p, ".json"), bin = ends_with(p, ".bin");
    if (!json && !bin && !ends_with(p, ".xml"))
        return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: unsupported scene file format (.xml, .json or .bin)");
    try {
        Node root = scene_tree(*d);
        std::string bytes;
        if (bin) write_binary(*schema().scene, &root.kids.front(), bytes);
        std::ofstream f(p, std::ios::binary | std::ios::trunc);
        if (!f.good()) return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: could not open " + p);
        if (bin) {
            f.write(bytes.data(), std::streamsize(bytes.size()));
        } else if (json) {
            f << "{\n    \"scene\": ";
            write_json(f, root.kids.front(), 1, "scene");
            f << "\n}\n";
        } else {
            f << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
            write_xml(f, root, 0);
        }
        return f.good() ? RC_OK : rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: write failed");
    } catch (const ParseError& e) {
        return rc_host_fail(RC_ERR_INVALID, "rch_scene_file_save: " + e.msg);
    } catch (const std::bad_alloc&) {
        return rc_host_fail(RC_ERR_ALLOC, "rch_scene_file_save: out of memory");
    }
}

extern "C" void rch_scene_file_free(rch_scene_file_data* d) {
    if (!d) return;
    std::free(d->materials);
    std::free(d->ray_materials);
    std::free(d->material_ids);
    std::free(d->vertices);
    std::free(d->tri_material);
    std::free(d->tri_ids);
    std::free(d->lights);
    std::free(d->models);
    std::free(d->nodes);
    std::free(d->ordered);
    std::free(d);
}
