// Drives the host half of libraycer (OBJ/PLY loaders, the cereal XML, JSON
// and binary scene reader and writer, the BVH builders and BVH files, the
// scene generators) and the FP64 oracle through their C interfaces, for a
// build with AddressSanitizer and UndefinedBehaviorSanitizer (any finding
// aborts).  The scene reader also gets every truncation of the reference's
// sample.xml and of its JSON and binary forms, seeded byte mutations of all
// three, and hostile nesting; all must end in an error return or a scene,
// never in a sanitizer report.
//
//   host_sanitize <tests/golden directory> <scratch directory>
// prints "host_sanitize: N checks, 0 failures" on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "oracle.h"
#include "raycer_host.h"

namespace {

int g_checks = 0, g_failures = 0;

void check(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        ++g_failures;
        std::printf("FAIL: %s\n", what.c_str());
    }
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void load_models(const std::string& dir) {
    const char* objs[] = {"cube1.obj", "cube2.obj", "cube3.obj", "cube4.obj", "cube5.obj", "cornellbox.obj"};
    for (const char* name : objs) {
        rch_obj* m = nullptr;
        const int rc = rch_obj_load((dir + "/meshes/" + name).c_str(), nullptr, &m);
        check(rc == RC_OK && m, std::string("obj ") + name);
        if (!m) continue;
        const size_t nt = rch_obj_triangle_count(m), nm = rch_obj_material_count(m);
        std::vector<float> v(nt * 9);
        std::vector<uint32_t> tm(nt);
        std::vector<rc_material> mats(nm);
        check(rch_obj_get(m, v.data(), tm.data(), mats.data()) == RC_OK, std::string("obj get ") + name);
        for (uint32_t t : tm) check(t < nm, "obj material index");
        rch_obj_free(m);
    }
    const char* plys[] = {"cube1.ply", "cube2.ply", "cube3.ply"};
    for (const char* name : plys) {
        rch_obj* m = nullptr;
        const int rc = rch_ply_load((dir + "/meshes/" + name).c_str(), nullptr, &m);
        check(rc == RC_OK && m, std::string("ply ") + name);
        if (m) rch_obj_free(m);
    }
    rch_obj* m = nullptr;
    check(rch_obj_load((dir + "/meshes/missing.obj").c_str(), nullptr, &m) != RC_OK, "missing obj fails");
}

// Parses text; a scene or an error, and the scene's arrays are walked.
void parse_any(const std::string& text, int json) {
    rch_scene_file_data* sf = nullptr;
    const int rc = rch_scene_file_parse(text.data(), text.size(), json, &sf);
    if (rc == RC_OK && sf) {
        double acc = 0.0;
        for (size_t i = 0; i < sf->ntri * 9; ++i) acc += sf->vertices[i];
        for (size_t i = 0; i < sf->ntri; ++i) acc += sf->tri_material[i] < sf->nmat ? 0.0 : 1e300;
        for (size_t i = 0; i < sf->nnodes; ++i) acc += sf->nodes[i].aabb_min[0];
        check(std::isfinite(acc) || sf->ntri == 0, "parsed scene arrays readable");
        rch_scene_file_free(sf);
    }
}

void scene_files(const std::string& dir, const std::string& tmp) {
    const std::string xml = read_file(dir + "/scenes/sample.xml");
    check(!xml.empty(), "sample.xml read");
    rch_scene_file_data* sf = nullptr;
    check(rch_scene_file_load((dir + "/scenes/sample.xml").c_str(), &sf) == RC_OK && sf, "sample.xml loads");
    if (!sf) return;
    const std::string xml2 = tmp + "/round.xml", json2 = tmp + "/round.json", bin2 = tmp + "/round.bin";
    check(rch_scene_file_save(xml2.c_str(), sf) == RC_OK, "save xml");
    check(rch_scene_file_save(json2.c_str(), sf) == RC_OK, "save json");
    check(rch_scene_file_save(bin2.c_str(), sf) == RC_OK, "save binary");
    rch_scene_file_free(sf);
    for (const std::string& p : {xml2, json2, bin2}) {
        rch_scene_file_data* r = nullptr;
        check(rch_scene_file_load(p.c_str(), &r) == RC_OK && r, "round trip " + p);
        if (r) rch_scene_file_free(r);
    }
    const std::string json = read_file(json2), bin = read_file(bin2);
    auto form = [&](const std::string* t) { return t == &json ? RCH_SCENE_JSON : t == &bin ? RCH_SCENE_BINARY : RCH_SCENE_XML; };
    // every truncation of the three forms (bounded to a few thousand cut points)
    for (const std::string* t : {&xml, &json, &bin}) {
        const int js = form(t);
        const size_t step = std::max<size_t>(1, t->size() / 3000);
        for (size_t cut = 0; cut < t->size(); cut += step) parse_any(t->substr(0, cut), js);
    }
    // seeded byte mutations: flips, insertions of structure characters, deletions
    std::mt19937 rng(12345);
    const char specials[] = "<>/\"{}[],:&;= \n\\";
    for (const std::string* t : {&xml, &json, &bin}) {
        const int js = form(t);
        for (int k = 0; k < 1500; ++k) {
            std::string m = *t;
            const int edits = 1 + int(rng() % 4);
            for (int e = 0; e < edits && !m.empty(); ++e) {
                const size_t pos = rng() % m.size();
                switch (rng() % 3) {
                case 0: m[pos] = char(rng() % 256); break;
                case 1: m.insert(pos, 1, specials[rng() % (sizeof(specials) - 1)]); break;
                default: m.erase(pos, 1 + rng() % 16); break;
                }
            }
            parse_any(m, js);
        }
    }
    // hostile nesting is rejected, not recursed into
    for (int js = 0; js < 2; ++js) {
        std::string deep;
        for (int i = 0; i < 100000; ++i) deep += js ? "[" : "<a>";
        rch_scene_file_data* r = nullptr;
        check(rch_scene_file_parse(deep.data(), deep.size(), js, &r) == RC_ERR_INVALID, "deep nesting rejected");
        char msg[256];
        rc_last_error(msg, sizeof(msg));
        check(std::strstr(msg, "nesting") != nullptr, std::string("nesting message: ") + msg);
    }
    parse_any("", 0);
    parse_any("", 1);
    parse_any("{\"value0\": ", 1);
    parse_any("<cereal><value0 size=\"dynamic\">", 0);
}

void bvh_files(const std::string& tmp) {
    for (uint32_t kind : {uint32_t(RCH_SCENE_BOX), uint32_t(RCH_SCENE_SOUP), uint32_t(RCH_SCENE_TERRAIN)}) {
        rch_scene_params p{kind, 3, 3, 33, kind == RCH_SCENE_SOUP ? 5000u : 0u};
        rch_scene_desc d;
        check(rch_scene_describe(&p, &d) == RC_OK, "describe");
        std::vector<float> v(d.ntri * 9);
        std::vector<uint32_t> tm(d.ntri);
        std::vector<rc_material> mats(d.nmat);
        check(rch_scene_generate(&p, v.data(), tm.data(), mats.data()) == RC_OK, "generate");
        for (uint32_t method : {uint32_t(RCH_BVH_REFERENCE), uint32_t(RCH_BVH_BINNED)}) {
            rch_bvh_build_info info{5, 2, method, 0.0f, 0};
            rch_bvh* b = nullptr;
            check(rch_bvh_build(v.data(), d.ntri, &info, &b) == RC_OK && b, "bvh build");
            if (!b) continue;
            const size_t nn = rch_bvh_node_count(b);
            std::vector<rc_flat_node> nodes(nn);
            std::vector<uint64_t> ord(d.ntri);
            check(rch_bvh_get(b, nodes.data(), ord.data()) == RC_OK, "bvh get");
            rch_bvh_free(b);
            const std::string path = tmp + "/tree.bvh";
            check(rch_bvh_save(path.c_str(), v.data(), d.ntri, nodes.data(), nn, ord.data()) == RC_OK, "bvh save");
            size_t n2 = 0, t2 = 0;
            check(rch_bvh_file_counts(path.c_str(), &n2, &t2) == RC_OK && n2 == nn && t2 == d.ntri, "bvh counts");
            std::vector<rc_flat_node> nodes2(nn);
            std::vector<uint64_t> ord2(d.ntri);
            check(rch_bvh_load(path.c_str(), v.data(), d.ntri, nodes2.data(), ord2.data()) == RC_OK, "bvh load");
            check(std::memcmp(nodes.data(), nodes2.data(), nn * sizeof(rc_flat_node)) == 0 && ord == ord2,
                  "bvh round trip");
            // a file cut short is rejected
            const std::string bytes = read_file(path);
            std::ofstream(path, std::ios::binary | std::ios::trunc).write(bytes.data(), std::streamsize(bytes.size() / 2));
            check(rch_bvh_load(path.c_str(), v.data(), d.ntri, nodes2.data(), ord2.data()) != RC_OK, "short bvh file");
        }
    }
}

void oracle_paths() {
    rch_scene_params p{RCH_SCENE_BOX, 0, 0, 0, 0};
    rch_scene_desc d;
    rch_scene_describe(&p, &d);
    std::vector<float> v(d.ntri * 9);
    std::vector<uint32_t> tm(d.ntri);
    std::vector<rc_material> mats(d.nmat);
    rch_scene_generate(&p, v.data(), tm.data(), mats.data());
    or_bvh* b = nullptr;
    check(or_bvh_build(v.data(), d.ntri, 5, &b) == 0 && b, "oracle bvh");
    if (!b) return;
    const size_t n = 512;
    std::vector<float> rays(n * 8);
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    for (size_t i = 0; i < n; ++i) {
        float* r = &rays[8 * i];
        r[0] = 0.5f + 0.3f * U(rng); r[1] = 0.5f + 0.3f * U(rng); r[2] = 0.5f + 0.3f * U(rng); r[3] = 0.0f;
        float dx = U(rng), dy = U(rng), dz = U(rng), l = std::sqrt(dx * dx + dy * dy + dz * dz) + 1e-6f;
        r[4] = dx / l; r[5] = dy / l; r[6] = dz / l; r[7] = INFINITY;
    }
    std::vector<double> t(n), cond(n);
    std::vector<int64_t> id(n);
    std::vector<uint32_t> flags(n);
    or_trace_closest_diag(b, v.data(), rays.data(), n, 1e-6, 2e-5, t.data(), id.data(), flags.data(), cond.data(), 2);
    std::vector<uint8_t> occ(n), fr(n);
    or_trace_any(b, v.data(), rays.data(), n, 2e-5, occ.data(), fr.data(), 2);
    std::vector<double> mat_arr(d.nmat * OR_MAT_STRIDE, 0.0);
    for (size_t i = 0; i < d.nmat; ++i) {
        double* m = &mat_arr[i * OR_MAT_STRIDE];
        for (int k = 0; k < 3; ++k) {
            m[k] = mats[i].diffuse[k];
            m[3 + k] = mats[i].emittance[k];
        }
        m[13] = mats[i].flags;
    }
    const double cam[7] = {d.camera.position[0], d.camera.position[1], d.camera.position[2], d.camera.pitch,
                           d.camera.yaw, d.camera.roll, d.camera.fov};
    or_render_cfg c{};
    c.width = 8; c.height = 8; c.spp = 4; c.sampler = 0; c.max_path_length = 6; c.rr_start = 3; c.filter = 1;
    c.estimator = 0; c.traversal = 1; c.filter_radius = 1.0; c.ray_start_offset = 1e-5; c.seed = 1;
    c.cache_vertex = 2;
    std::vector<double> st(64 * 14), sp(64 * 7);  // or_render: 14 doubles per pixel
    std::vector<uint32_t> ids(64);
    for (uint32_t i = 0; i < 64; ++i) ids[i] = i;
    check(or_render(b, v.data(), d.ntri, tm.data(), mat_arr.data(), d.nmat, cam, &c, nullptr, 0, st.data(), 2,
                    nullptr) == 0, "oracle render");
    check(or_render_split_tail(b, v.data(), d.ntri, tm.data(), mat_arr.data(), cam, &c, ids.data(), 64, sp.data(), 2) == 0,
          "oracle split tail");
    or_bvh_free(b);
    const size_t np = or_mlp_param_count(81, 64, 4, 3);
    std::vector<double> prm(np), g(np), x(16 * 81), tg(16 * 3), y(16 * 3);
    std::normal_distribution<double> N(0.0, 0.1);
    for (double& q : prm) q = N(rng);
    for (double& q : x) q = N(rng);
    for (double& q : tg) q = std::abs(N(rng));
    or_mlp_forward(prm.data(), 81, 64, 4, 3, x.data(), 16, y.data());
    const double loss = or_mlp_loss_grad(prm.data(), 81, 64, 4, 3, x.data(), tg.data(), 16, 1, 0.01, g.data());
    check(std::isfinite(loss), "mlp loss");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: host_sanitize <golden dir> <scratch dir>\n");
        return 2;
    }
    const std::string golden = argv[1], tmp = argv[2];
    load_models(golden);
    scene_files(golden, tmp);
    bvh_files(tmp);
    oracle_paths();
    std::printf("host_sanitize: %d checks, %d failures\n", g_checks, g_failures);
    return g_failures ? 1 : 0;
}
