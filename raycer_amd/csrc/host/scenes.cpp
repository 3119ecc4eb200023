// Seeded scene generators for the benchmark workloads (the counterpart of
// raycer's Scene::createTestScene, src/TestScenes/TestScene*.cpp).  Every
// choice the workload descriptions leave open is fixed here:
//
// BOX (256x256, 16 spp): unit box open towards the camera (walls at y=0, y=1,
//   z=1, x=0, x=1; the open face is z=0 because the camera sits at z=-1.4 and
//   triangles are double-sided).  The camera basis right = forward x ALMOST_UP
//   is -x for a camera looking along +z, so the wall on the image's left is
//   x=1: it gets the red albedo (0.65,0.05,0.05), x=0 the green one.  Inner
//   boxes: 0.3x0.6x0.3 rotated +15 deg about y (Matrix4x4::rotateY convention),
//   bottom-face centre (0.3,0,0.35); 0.3^3 rotated -18 deg at (0.65,0,0.6).
//   Light: 0.25x0.25 quad at y=0.999 centred on x=z=0.5, facing down,
//   emittance (17,12,4).  Wall normals face the interior.
// SPHERE_GROUND: icosphere of level L (20*4^L faces) radius 0.5 centred at
//   (0,0.65,0), each vertex moved along its direction by 0.15*fBm (improved
//   Perlin noise, 6 octaves, lacunarity 2, gain 0.5, at 2*dir, seeded
//   permutation); then every face is split at the midpoint of its longest edge
//   (midpoint of the displaced endpoints, so neighbours stay watertight):
//   20*4^7*2 = 655,360 triangles at L=7.  Ground: G x G quads over
//   [-1.5,1.5]^2 with per-vertex heights U[-0.02,0.02]: 524,288 at G=512.
// TERRAIN: n x n vertices over [0,1]^2, height 0.2 * ridged fBm (8 octaves,
//   frequency 3, offset 1, gain 2, normalised by the octave weights),
//   2 (n-1)^2 triangles; grey albedo U[0.2,0.8] per 64x64-quad tile; sky 0.5;
//   sun towards normalize(0.4,0.7,-0.6) with irradiance (3.0,2.8,2.5).
// SOUP: triangle count floor(exp(U[ln 1e3, ln 5e5])) unless given; centres
//   U[0,1]^3; each vertex = centre + s * (random unit vector) with
//   s = exp(U[ln 1e-3, ln 5e-2]); 4 albedos U[0.3,0.9]^3; resolution drawn
//   from {128,256,512,1024}^2; 32 random samples, depth 5.
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "raycer_host.h"
#include "rc_host_internal.hpp"

namespace {

// splitmix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA'14)
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
};

// Improved noise (K. Perlin, "Improving noise", SIGGRAPH 2002) over a
// permutation shuffled by the seed.
struct Perlin {
    int p[512];
    explicit Perlin(uint64_t seed) {
        int q[256];
        for (int i = 0; i < 256; ++i) q[i] = i;
        Rng r(seed * 0x2545F4914F6CDD1Dull + 7);
        for (int i = 255; i > 0; --i) {
            int j = int(r.next() % uint64_t(i + 1));
            std::swap(q[i], q[j]);
        }
        for (int i = 0; i < 512; ++i) p[i] = q[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
    static double lerp(double t, double a, double b) { return a + t * (b - a); }
    static double grad(int h, double x, double y, double z) {
        h &= 15;
        double u = h < 8 ? x : y;
        double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
        return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
    }
    double noise(double x, double y, double z) const {
        double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int X = int(fx) & 255, Y = int(fy) & 255, Z = int(fz) & 255;
        x -= fx; y -= fy; z -= fz;
        double u = fade(x), v = fade(y), w = fade(z);
        int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
        int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
        return lerp(w,
                    lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                         lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
    }
    double fbm(double x, double y, double z, int octaves) const {
        double sum = 0.0, amp = 1.0, f = 1.0;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * noise(x * f, y * f, z * f);
            amp *= 0.5;
            f *= 2.0;
        }
        return sum;
    }
    // Ridged multifractal (Musgrave), normalised by the sum of octave weights.
    double ridged(double x, double y, double z, int octaves) const {
        double sum = 0.0, amp = 1.0, f = 1.0, weight = 1.0, norm = 0.0;
        for (int o = 0; o < octaves; ++o) {
            double s = 1.0 - std::abs(noise(x * f, y * f, z * f));
            s *= s;
            s *= weight;
            weight = std::min(1.0, std::max(0.0, s * 2.0));
            sum += s * amp;
            norm += amp;
            amp *= 0.5;
            f *= 2.0;
        }
        return sum / norm;
    }
};

struct V {
    double x, y, z;
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double len(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V unit(V a) { return a * (1.0 / len(a)); }

struct Mesh {
    std::vector<float> v;
    std::vector<uint32_t> m;
    void tri(V a, V b, V c, uint32_t mat) {
        const V p[3] = {a, b, c};
        for (const V& q : p) {
            v.push_back(float(q.x));
            v.push_back(float(q.y));
            v.push_back(float(q.z));
        }
        m.push_back(mat);
    }
    void quad(V a, V b, V c, V d, uint32_t mat) {  // normal (b-a)x(c-a)
        tri(a, b, c, mat);
        tri(a, c, d, mat);
    }
};

rc_material lambert(float r, float g, float b) {
    rc_material m;
    std::memset(&m, 0, sizeof(m));
    m.diffuse[0] = m.ambient[0] = r;
    m.diffuse[1] = m.ambient[1] = g;
    m.diffuse[2] = m.ambient[2] = b;
    m.shininess = 2.0f;
    return m;
}

void default_cfg(rc_render_cfg& c) {
    std::memset(&c, 0, sizeof(c));
    c.spp = 1;
    c.sampler = RC_SAMPLER_JITTERED;
    c.max_path_length = 3;
    c.filter = RC_FILTER_BOX;
    c.filter_radius = 0.5f;
    c.estimator = RC_EST_REFERENCE;
    c.ray_start_offset = 1e-5f;  // Scene::General::rayStartOffset default (Scene.h:95)
}

// Box with bottom-face centre at `pos`, rotated by `deg` about y
// (Matrix4x4::rotateY: x' = c x + s z, z' = -s x + c z), outward normals.
void inner_box(Mesh& M, V size, double deg, V pos, uint32_t mat) {
    double r = deg * (M_PI / 180.0), c = std::cos(r), s = std::sin(r);
    auto P = [&](double x, double y, double z) {
        V l{x * size.x * 0.5, y * size.y, z * size.z * 0.5};
        return V{c * l.x + s * l.z + pos.x, l.y + pos.y, -s * l.x + c * l.z + pos.z};
    };
    V a = P(-1, 0, -1), b = P(1, 0, -1), cc = P(1, 0, 1), d = P(-1, 0, 1);
    V e = P(-1, 1, -1), f = P(1, 1, -1), g = P(1, 1, 1), h = P(-1, 1, 1);
    M.quad(a, b, cc, d, mat);  // bottom (-y)
    M.quad(e, h, g, f, mat);   // top (+y)
    M.quad(a, e, f, b, mat);   // -z
    M.quad(d, cc, g, h, mat);  // +z
    M.quad(a, d, h, e, mat);   // -x
    M.quad(b, f, g, cc, mat);  // +x
}

void gen_box(Mesh& M, std::vector<rc_material>& mats, rch_scene_desc& D) {
    mats = {lambert(0.73f, 0.73f, 0.73f), lambert(0.65f, 0.05f, 0.05f), lambert(0.12f, 0.45f, 0.15f),
            lambert(0.0f, 0.0f, 0.0f)};
    mats[3].emittance[0] = 17.0f;
    mats[3].emittance[1] = 12.0f;
    mats[3].emittance[2] = 4.0f;
    mats[3].flags = RC_MAT_EMISSIVE;
    M.quad({0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}, 0);  // floor, +y
    M.quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}, 0);  // ceiling, -y
    M.quad({0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, 0);  // back wall, -z
    M.quad({1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, 1);  // x=1 (image left), -x
    M.quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}, 2);  // x=0 (image right), +x
    inner_box(M, {0.3, 0.6, 0.3}, 15.0, {0.3, 0.0, 0.35}, 0);
    inner_box(M, {0.3, 0.3, 0.3}, -18.0, {0.65, 0.0, 0.6}, 0);
    M.quad({0.375, 0.999, 0.375}, {0.625, 0.999, 0.375}, {0.625, 0.999, 0.625}, {0.375, 0.999, 0.625}, 3);
    D.camera = rc_camera{{0.5f, 0.5f, -1.4f}, 0.0f, 180.0f, 0.0f, 40.0f};
    default_cfg(D.cfg);
    D.cfg.width = D.cfg.height = 256;
    D.cfg.spp = 16;
    D.cfg.max_path_length = 8;
    D.cfg.rr_start = 3;
    D.cfg.filter = RC_FILTER_BOX;
    D.cfg.filter_radius = 0.5f;
    D.cfg.seed = 0;
}

void gen_sphere_ground(const rch_scene_params& p, Mesh& M, std::vector<rc_material>& mats, rch_scene_desc& D,
                       bool emit) {
    uint32_t level = p.level ? p.level : 7, G = p.grid ? p.grid : 512;
    rc_material sphere = lambert(0.8f, 0.55f, 0.35f);
    sphere.specular[0] = sphere.specular[1] = sphere.specular[2] = 0.3f;
    sphere.shininess = 32.0f;
    mats = {sphere, lambert(0.5f, 0.5f, 0.5f)};
    D.camera = rc_camera{{0.0f, 0.9f, -2.6f}, -12.0f, 180.0f, 0.0f, 55.0f};
    default_cfg(D.cfg);
    D.cfg.width = 1920;
    D.cfg.height = 1080;
    D.cfg.spp = 1;
    D.cfg.sampler = RC_SAMPLER_RANDOM;
    D.cfg.seed = p.seed;
    D.light = rc_point_light{{-2.0f, 3.0f, -2.5f}, {1.0f, 1.0f, 1.0f}, 1.0f, 1000.0f, 1.0f, {1.0f, 1.0f, 1.0f}, 0.1f};
    D.ntri = size_t(20) * (size_t(1) << (2 * level)) * 2 + size_t(2) * G * G;
    if (!emit) return;

    // icosahedron
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<V> vs = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (V& q : vs) q = unit(q);
    std::vector<uint32_t> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                               4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3,
                               6, 8, 3, 8,  9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            vs.push_back(unit((vs[a] + vs[b]) * 0.5));
            uint32_t id = uint32_t(vs.size() - 1);
            mid.emplace(key, id);
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add, add + 12);
        }
        f.swap(nf);
    }
    Perlin noise(p.seed);
    const V centre{0.0, 0.65, 0.0};
    std::vector<V> disp(vs.size());
    for (size_t i = 0; i < vs.size(); ++i) {
        V u = vs[i];
        double h = 0.15 * noise.fbm(2.0 * u.x + 11.5, 2.0 * u.y + 3.25, 2.0 * u.z + 7.75, 6);
        disp[i] = centre + u * (0.5 + h);
    }
    for (size_t i = 0; i < f.size(); i += 3) {
        V a = disp[f[i]], b = disp[f[i + 1]], c = disp[f[i + 2]];
        double lab = len(b - a), lbc = len(c - b), lca = len(a - c);
        if (lab >= lbc && lab >= lca) {
            V m = (a + b) * 0.5;
            M.tri(a, m, c, 0);
            M.tri(m, b, c, 0);
        } else if (lbc >= lca) {
            V m = (b + c) * 0.5;
            M.tri(a, b, m, 0);
            M.tri(a, m, c, 0);
        } else {
            V m = (c + a) * 0.5;
            M.tri(a, b, m, 0);
            M.tri(m, b, c, 0);
        }
    }
    Rng r(uint64_t(p.seed) * 0x9E3779B97F4A7C15ull + 0x51ull);
    std::vector<double> h(size_t(G + 1) * (G + 1));
    for (double& q : h) q = r.range(-0.02, 0.02);
    auto gp = [&](uint32_t i, uint32_t j) {
        return V{-1.5 + 3.0 * double(i) / G, h[size_t(j) * (G + 1) + i], -1.5 + 3.0 * double(j) / G};
    };
    for (uint32_t j = 0; j < G; ++j)
        for (uint32_t i = 0; i < G; ++i) M.quad(gp(i, j), gp(i, j + 1), gp(i + 1, j + 1), gp(i + 1, j), 1);
}

void gen_terrain(const rch_scene_params& p, Mesh& M, std::vector<rc_material>& mats, rch_scene_desc& D, bool emit) {
    uint32_t n = p.grid ? p.grid : 1024;
    uint32_t tiles = (n - 1 + 63) / 64;
    D.nmat = size_t(tiles) * tiles;
    D.ntri = size_t(2) * (n - 1) * (n - 1);
    D.camera = rc_camera{{0.5f, 0.42f, -0.3f}, -28.0f, 180.0f, 0.0f, 65.0f};
    default_cfg(D.cfg);
    D.cfg.width = 1920;
    D.cfg.height = 1080;
    D.cfg.spp = 64;
    D.cfg.filter = RC_FILTER_TENT;
    D.cfg.filter_radius = 1.0f;
    D.cfg.max_path_length = 6;
    D.cfg.rr_start = 3;
    D.cfg.sky[0] = D.cfg.sky[1] = D.cfg.sky[2] = 0.5f;
    V sun = unit({0.4, 0.7, -0.6});
    D.cfg.sun_dir[0] = float(sun.x);
    D.cfg.sun_dir[1] = float(sun.y);
    D.cfg.sun_dir[2] = float(sun.z);
    D.cfg.sun_irradiance[0] = 3.0f;
    D.cfg.sun_irradiance[1] = 2.8f;
    D.cfg.sun_irradiance[2] = 2.5f;
    D.cfg.seed = p.seed;
    if (!emit) return;
    Rng r(uint64_t(p.seed) * 0x9E3779B97F4A7C15ull + 0x77ull);
    mats.resize(D.nmat);
    for (auto& m : mats) {
        float a = float(r.range(0.2, 0.8));
        m = lambert(a, a, a);
    }
    Perlin noise(p.seed);
    std::vector<double> h(size_t(n) * n);
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t i = 0; i < n; ++i) {
            double x = double(i) / (n - 1), z = double(j) / (n - 1);
            h[size_t(j) * n + i] = 0.2 * noise.ridged(3.0 * x + 0.37, 0.5, 3.0 * z + 0.81, 8);
        }
    auto gp = [&](uint32_t i, uint32_t j) { return V{double(i) / (n - 1), h[size_t(j) * n + i], double(j) / (n - 1)}; };
    for (uint32_t j = 0; j + 1 < n; ++j)
        for (uint32_t i = 0; i + 1 < n; ++i) {
            uint32_t mat = (j / 64) * tiles + (i / 64);
            V a = gp(i, j), b = gp(i, j + 1), c = gp(i + 1, j + 1), d = gp(i + 1, j);
            M.tri(a, b, c, mat);
            M.tri(a, c, d, mat);
        }
}

void gen_soup(const rch_scene_params& p, Mesh& M, std::vector<rc_material>& mats, rch_scene_desc& D, bool emit) {
    Rng r(uint64_t(p.seed) * 0xD1B54A32D192ED03ull + 0x1234ull);
    uint64_t ntri = p.tri_count ? p.tri_count : uint64_t(std::floor(std::exp(r.range(std::log(1e3), std::log(5e5)))));
    static const uint32_t res[4] = {128, 256, 512, 1024};
    uint32_t w = res[r.next() % 4];
    D.ntri = ntri;
    D.nmat = 4;
    D.camera = rc_camera{{0.5f, 0.5f, -1.0f}, 0.0f, 180.0f, 0.0f, 60.0f};
    default_cfg(D.cfg);
    D.cfg.width = D.cfg.height = w;
    D.cfg.spp = 32;
    D.cfg.sampler = RC_SAMPLER_RANDOM;
    D.cfg.max_path_length = 5;
    D.cfg.rr_start = 3;
    D.cfg.sky[0] = D.cfg.sky[1] = D.cfg.sky[2] = 0.5f;
    V sun = unit({0.3, 0.8, -0.5});
    D.cfg.sun_dir[0] = float(sun.x);
    D.cfg.sun_dir[1] = float(sun.y);
    D.cfg.sun_dir[2] = float(sun.z);
    D.cfg.sun_irradiance[0] = D.cfg.sun_irradiance[1] = D.cfg.sun_irradiance[2] = 2.0f;
    D.cfg.seed = p.seed;
    if (!emit) return;
    mats.resize(4);
    for (auto& m : mats) m = lambert(float(r.range(0.3, 0.9)), float(r.range(0.3, 0.9)), float(r.range(0.3, 0.9)));
    auto rand_unit = [&]() {
        double z = r.range(-1.0, 1.0), phi = r.range(0.0, 2.0 * M_PI);
        double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        return V{s * std::cos(phi), s * std::sin(phi), z};
    };
    M.v.reserve(ntri * 9);
    M.m.reserve(ntri);
    for (uint64_t i = 0; i < ntri; ++i) {
        V c{r.uniform(), r.uniform(), r.uniform()};
        double s = std::exp(r.range(std::log(1e-3), std::log(5e-2)));
        V a = c + rand_unit() * s, b = c + rand_unit() * s, d = c + rand_unit() * s;
        M.tri(a, b, d, uint32_t(r.next() % 4));
    }
}

int run(const rch_scene_params* p, rch_scene_desc& D, Mesh* M, std::vector<rc_material>* mats) {
    std::memset(&D, 0, sizeof(D));
    Mesh tmpM;
    std::vector<rc_material> tmpMats;
    bool emit = M != nullptr;
    if (!M) { M = &tmpM; mats = &tmpMats; }
    switch (p->kind) {
        case RCH_SCENE_BOX:
            gen_box(*M, *mats, D);
            D.ntri = M->m.size();
            D.nmat = mats->size();
            break;
        case RCH_SCENE_SPHERE_GROUND:
            if (p->level > 9 || p->grid > 4096) return rc_host_fail(RC_ERR_INVALID, "sphere/ground size out of range");
            gen_sphere_ground(*p, *M, *mats, D, emit);
            D.nmat = 2;
            break;
        case RCH_SCENE_TERRAIN:
            if (p->grid == 1 || p->grid > 8192) return rc_host_fail(RC_ERR_INVALID, "terrain grid out of range");
            gen_terrain(*p, *M, *mats, D, emit);
            break;
        case RCH_SCENE_SOUP:
            if (p->tri_count > 50000000ull) return rc_host_fail(RC_ERR_INVALID, "soup too large");
            gen_soup(*p, *M, *mats, D, emit);
            break;
        default:
            return rc_host_fail(RC_ERR_INVALID, "unknown scene kind");
    }
    if (emit && M->m.size() != D.ntri) return rc_host_fail(RC_ERR_INVALID, "scene generator size mismatch");
    // bounds
    for (int k = 0; k < 3; ++k) { D.bounds_min[k] = 1e30f; D.bounds_max[k] = -1e30f; }
    for (size_t i = 0; i < M->v.size(); i += 3)
        for (int k = 0; k < 3; ++k) {
            D.bounds_min[k] = std::min(D.bounds_min[k], M->v[i + k]);
            D.bounds_max[k] = std::max(D.bounds_max[k], M->v[i + k]);
        }
    return RC_OK;
}

}  // namespace

extern "C" int rch_scene_describe(const rch_scene_params* p, rch_scene_desc* d) {
    if (!p || !d) return rc_host_fail(RC_ERR_INVALID, "rch_scene_describe: null argument");
    // Box and soup sizes/bounds need the geometry; the large scenes are sized analytically.
    if (p->kind == RCH_SCENE_BOX) {
        Mesh M;
        std::vector<rc_material> m;
        return run(p, *d, &M, &m);
    }
    return run(p, *d, nullptr, nullptr);
}

extern "C" int rch_scene_generate(const rch_scene_params* p, float* vertices, uint32_t* tri_material,
                                  rc_material* materials) {
    if (!p || !vertices || !tri_material || !materials)
        return rc_host_fail(RC_ERR_INVALID, "rch_scene_generate: null argument");
    Mesh M;
    std::vector<rc_material> mats;
    rch_scene_desc D;
    int rc = run(p, D, &M, &mats);
    if (rc != RC_OK) return rc;
    std::memcpy(vertices, M.v.data(), M.v.size() * sizeof(float));
    std::memcpy(tri_material, M.m.data(), M.m.size() * sizeof(uint32_t));
    std::memcpy(materials, mats.data(), mats.size() * sizeof(rc_material));
    return RC_OK;
}
