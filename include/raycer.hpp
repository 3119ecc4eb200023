// raycer.hpp — C++ interface of the MI355X path tracer, modelled on raycer's
// own classes so a raycer caller can switch renderers:
//   Scene, Scene::General/ToneMapper/RootBVH   (src/Raytracing/Scene.h:89-217)
//   FlatBVH, BVHBuildInfo, FlatBVHNode          (src/Raytracing/Primitives/FlatBVH.h:19-106)
//   Film                                        (src/Rendering/Film.h:23-52)
//   TracerState                                 (src/Raytracing/Tracers/TracerState.h:13-26)
//   HipTracer with CLTracer's lifecycle         (src/OpenCL/CLTracer.h:23-90)
// Everything calls the C ABI of libraycer (raycer_hip.h, raycer_host.h);
// failures throw RaycerError (std::runtime_error), like CLManager::checkError.
#pragma once

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "raycer_hip.h"
#include "raycer_host.h"

namespace raycer {

class RaycerError : public std::runtime_error {
public:
    RaycerError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

inline void check(int rc) {
    if (rc != RC_OK) {
        char buf[2048];
        rc_last_error(buf, sizeof(buf));
        throw RaycerError(rc, std::string("raycer error ") + std::to_string(rc) + ": " + buf);
    }
}

enum class TracerType { RAY, PATH };
// The reference's enums, in its order (Filter.h:12, Sampler.h:23, ToneMapper.h:14).
enum class FilterType { BOX, TENT, BELL, GAUSSIAN, MITCHELL, LANCZOS_SINC };
enum class SamplerType { CENTER, RANDOM, REGULAR, JITTERED, CMJ, POISSON_DISC };
enum class ToneMapperType { PASSTHROUGH, LINEAR, SIMPLE, REINHARD };

inline uint32_t to_rc(FilterType f) { return uint32_t(f); }  // RC_FILTER_* share the order
inline uint32_t to_rc(SamplerType s) {
    switch (s) {
    case SamplerType::CENTER: return RC_SAMPLER_CENTER;
    case SamplerType::RANDOM: return RC_SAMPLER_RANDOM;
    case SamplerType::REGULAR: return RC_SAMPLER_REGULAR;
    case SamplerType::JITTERED: return RC_SAMPLER_JITTERED;
    case SamplerType::CMJ: return RC_SAMPLER_CMJ;
    default: return RC_SAMPLER_POISSON_DISC;
    }
}
inline SamplerType sampler_of(uint32_t s) {
    switch (s) {
    case RC_SAMPLER_CENTER: return SamplerType::CENTER;
    case RC_SAMPLER_RANDOM: return SamplerType::RANDOM;
    case RC_SAMPLER_REGULAR: return SamplerType::REGULAR;
    case RC_SAMPLER_CMJ: return SamplerType::CMJ;
    case RC_SAMPLER_POISSON_DISC: return SamplerType::POISSON_DISC;
    default: return SamplerType::JITTERED;
    }
}
inline uint32_t to_rc(ToneMapperType t) {
    switch (t) {
    case ToneMapperType::PASSTHROUGH: return RC_TONEMAP_PASSTHROUGH;
    case ToneMapperType::LINEAR: return RC_TONEMAP_LINEAR;
    case ToneMapperType::SIMPLE: return RC_TONEMAP_SIMPLE;
    default: return RC_TONEMAP_REINHARD;
    }
}

struct Color {
    double r = 0.0, g = 0.0, b = 0.0;
};

struct Vector3 {
    double x = 0.0, y = 0.0, z = 0.0;
};

struct EulerAngle {
    double pitch = 0.0, yaw = 0.0, roll = 0.0;
};

// Material (Material.h:20-66) without the texture maps.
struct Material {
    Color diffuseReflectance;
    Color emittance;
    Color ambientReflectance;
    Color specularReflectance;
    double specularShininess = 2.0;
    bool isEmissive = false;
    bool nonShadowing = false;
    bool skipLighting = false;
    bool fresnelReflection = false;
    double refractiveIndex = 1.0;
    double rayReflectance = 0.0;
    SamplerType rayReflectanceGlossinessSamplerType = SamplerType::CMJ;
    uint64_t rayReflectanceGlossinessSampleCountSqrt = 0;
    double rayReflectanceGlossiness = 1000.0;
    double rayTransmittance = 0.0;
    SamplerType rayTransmittanceGlossinessSamplerType = SamplerType::CMJ;
    uint64_t rayTransmittanceGlossinessSampleCountSqrt = 0;
    double rayTransmittanceGlossiness = 1000.0;
    bool enableRayTransmissionAttenuation = false;
    double rayTransmissionAttenuationFactor = 1.0;
    Color rayTransmissionAttenuationColor;
};

// Lights.h: the ambient light (with ambient occlusion) and the directional,
// point (optionally an area light) and spot lights of the ray tracer.
struct Light {
    Color color{1.0, 1.0, 1.0};
    double intensity = 1.0;
};
struct AmbientLight : Light {
    bool enableAmbientOcclusion = false;
    SamplerType ambientOcclusionSamplerType = SamplerType::CMJ;
    uint64_t ambientOcclusionSampleCountSqrt = 0;
    double ambientOcclusionMaxSampleDistance = 1.0;
    double ambientOcclusionSampleDistribution = 1.0;
};
struct DirectionalLight : Light {
    Vector3 direction;
};
struct PointLight : Light {
    Vector3 position;
    double maxDistance = 1000.0;
    double attenuation = 1.0;
    bool enableAreaLight = false;
    SamplerType areaLightSamplerType = SamplerType::CMJ;
    uint64_t areaLightSampleCountSqrt = 0;
    double areaLightRadius = 1.0;
};
struct SpotLight : PointLight {
    Vector3 direction;
    double angle = 45.0;
    double sideAttenuation = 1.0;
};

// Material <-> the C ABI's rc_material (shading) and rc_ray_material (ray tracing).
inline rc_material to_rc_material(const Material& m) {
    rc_material r{};
    const Color* src[4] = {&m.diffuseReflectance, &m.emittance, &m.ambientReflectance, &m.specularReflectance};
    float* dst[4] = {r.diffuse, r.emittance, r.ambient, r.specular};
    for (int k = 0; k < 4; ++k) {
        dst[k][0] = float(src[k]->r);
        dst[k][1] = float(src[k]->g);
        dst[k][2] = float(src[k]->b);
    }
    r.shininess = float(m.specularShininess);
    r.flags = (m.isEmissive ? RC_MAT_EMISSIVE : 0u) | (m.nonShadowing ? RC_MAT_NON_SHADOWING : 0u);
    return r;
}
inline rc_ray_material to_rc_ray_material(const Material& m) {
    rc_ray_material x{};
    x.ray_reflectance = float(m.rayReflectance);
    x.ray_transmittance = float(m.rayTransmittance);
    x.refractive_index = float(m.refractiveIndex);
    x.fresnel_reflection = m.fresnelReflection ? 1u : 0u;
    x.skip_lighting = m.skipLighting ? 1u : 0u;
    x.refl_sampler = to_rc(m.rayReflectanceGlossinessSamplerType);
    x.refl_n = uint32_t(m.rayReflectanceGlossinessSampleCountSqrt);
    x.refl_glossiness = float(m.rayReflectanceGlossiness);
    x.trans_sampler = to_rc(m.rayTransmittanceGlossinessSamplerType);
    x.trans_n = uint32_t(m.rayTransmittanceGlossinessSampleCountSqrt);
    x.trans_glossiness = float(m.rayTransmittanceGlossiness);
    x.enable_attenuation = m.enableRayTransmissionAttenuation ? 1u : 0u;
    x.attenuation_factor = float(m.rayTransmissionAttenuationFactor);
    x.attenuation_color[0] = float(m.rayTransmissionAttenuationColor.r);
    x.attenuation_color[1] = float(m.rayTransmissionAttenuationColor.g);
    x.attenuation_color[2] = float(m.rayTransmissionAttenuationColor.b);
    return x;
}
inline Material from_rc_material(const rc_material& a, const rc_ray_material& r) {
    Material m;
    m.diffuseReflectance = Color{a.diffuse[0], a.diffuse[1], a.diffuse[2]};
    m.emittance = Color{a.emittance[0], a.emittance[1], a.emittance[2]};
    m.ambientReflectance = Color{a.ambient[0], a.ambient[1], a.ambient[2]};
    m.specularReflectance = Color{a.specular[0], a.specular[1], a.specular[2]};
    m.specularShininess = a.shininess;
    m.isEmissive = a.flags & RC_MAT_EMISSIVE;
    m.nonShadowing = a.flags & RC_MAT_NON_SHADOWING;
    m.skipLighting = r.skip_lighting != 0;
    m.fresnelReflection = r.fresnel_reflection != 0;
    m.refractiveIndex = r.refractive_index;
    m.rayReflectance = r.ray_reflectance;
    m.rayReflectanceGlossinessSamplerType = sampler_of(r.refl_sampler);
    m.rayReflectanceGlossinessSampleCountSqrt = r.refl_n;
    m.rayReflectanceGlossiness = r.refl_glossiness;
    m.rayTransmittance = r.ray_transmittance;
    m.rayTransmittanceGlossinessSamplerType = sampler_of(r.trans_sampler);
    m.rayTransmittanceGlossinessSampleCountSqrt = r.trans_n;
    m.rayTransmittanceGlossiness = r.trans_glossiness;
    m.enableRayTransmissionAttenuation = r.enable_attenuation != 0;
    m.rayTransmissionAttenuationFactor = r.attenuation_factor;
    m.rayTransmissionAttenuationColor = Color{r.attenuation_color[0], r.attenuation_color[1], r.attenuation_color[2]};
    return m;
}

struct Triangle {
    Vector3 vertices[3];
    uint64_t materialId = 0;  // index into Scene::materials
};

// BVHBuildInfo (FlatBVH.h:19-36): only the SAH path is provided; binnedSAH
// selects this library's 32-bin SAH instead of the reference's 6-candidate one.
struct BVHBuildInfo {
    uint64_t maxLeafSize = 5;
    bool useSAH = true;
    uint32_t threads = 0;  // 0 = all hardware threads
    bool binnedSAH = false;
};

using FlatBVHNode = rc_flat_node;

class FlatBVH {
public:
    // FlatBVH::build (FlatBVH.cpp:89-204) over the triangles' boxes.
    void build(const std::vector<float>& vertices, const BVHBuildInfo& info) {
        if (!info.useSAH) throw RaycerError(RC_ERR_INVALID, "only the SAH build is provided");
        rch_bvh_build_info bi{info.maxLeafSize, info.threads,
                              uint32_t(info.binnedSAH ? RCH_BVH_BINNED : RCH_BVH_REFERENCE), 0.0f, 0u};
        rch_bvh* h = nullptr;
        check(rch_bvh_build(vertices.data(), vertices.size() / 9, &bi, &h));
        flatNodes.resize(rch_bvh_node_count(h));
        orderedPrimitiveIds.resize(vertices.size() / 9);
        int rc = rch_bvh_get(h, flatNodes.data(), orderedPrimitiveIds.data());
        buildMilliseconds = rch_bvh_build_ms(h);
        rch_bvh_free(h);
        check(rc);
        hasBeenBuilt = true;
    }

    // Serialization of the built tree (FlatBVH.h:99-105) and its restore on
    // load (FlatBVH.cpp:206-214); restore() refuses a file saved for other
    // triangles.
    void save(const std::string& path, const std::vector<float>& vertices) const {
        check(rch_bvh_save(path.c_str(), vertices.data(), vertices.size() / 9, flatNodes.data(), flatNodes.size(),
                           orderedPrimitiveIds.data()));
    }
    void restore(const std::string& path, const std::vector<float>& vertices) {
        size_t nn = 0, nt = 0;
        check(rch_bvh_file_counts(path.c_str(), &nn, &nt));
        if (nt != vertices.size() / 9) throw RaycerError(RC_ERR_INVALID, "saved BVH was built over other triangles");
        flatNodes.resize(nn);
        orderedPrimitiveIds.resize(nt);
        check(rch_bvh_load(path.c_str(), vertices.data(), nt, flatNodes.data(), orderedPrimitiveIds.data()));
        hasBeenBuilt = true;
    }

    bool hasBeenBuilt = false;
    std::vector<FlatBVHNode> flatNodes;
    std::vector<uint64_t> orderedPrimitiveIds;
    double buildMilliseconds = 0.0;
};

struct Camera {
    Vector3 position;
    EulerAngle orientation;  // pitch, yaw, roll in degrees
    double fov = 75.0;       // horizontal, degrees
};

class Scene {
public:
    // Scene::General (Scene.h:89-131) subset plus the estimator terms this
    // renderer adds; defaults are the reference's.
    struct General {
        TracerType tracerType = TracerType::PATH;
        uint64_t maxPathLength = 3;
        uint64_t multiSampleCountSqrt = 1;
        SamplerType multiSamplerType = SamplerType::JITTERED;
        uint64_t randomSampleCount = 0;  // RANDOM sampler: samples per pixel
        FilterType multiSamplerFilterType = FilterType::BOX;
        double filterRadius = 0.5;              // box, tent; lanczos_sinc (integer)
        double gaussianStandardDeviation = 1.0; // GaussianFilter's default
        double mitchellB = 1.0 / 3.0, mitchellC = 1.0 / 3.0;  // MitchellFilter's defaults
        double rayStartOffset = 0.00001;
        uint64_t russianRouletteStart = 0;  // path vertex where roulette starts; 0 = off
        bool lambertEstimator = false;      // false: the reference's 2 rho |cos| weight
        Color skyRadiance;                  // the reference returns black on a miss
        Vector3 sunDirection;               // towards the sun
        Color sunIrradiance;
        uint64_t seed = 0;
        // ray tracer (Scene.h:92-105)
        uint64_t maxRayIterations = 3;
        Color backgroundColor;
        bool visualizeDepth = false;
        double visualizeDepthMaxDistance = 25.0;
    } general;

    struct Lights {
        AmbientLight ambientLight;
        std::vector<DirectionalLight> directionalLights;
        std::vector<PointLight> pointLights;
        std::vector<SpotLight> spotLights;
    } lights;

    // Scene::SimpleFog (Scene.h:162-183)
    struct SimpleFog {
        bool enabled = false;
        Color color{1.0, 1.0, 1.0};
        double distance = 100.0;
        double steepness = 1.0;
        bool heightDispersion = false;
        double height = 100.0;
        double heightSteepness = 1.0;
    } simpleFog;

    Camera camera;

    // Scene::ToneMapper (Scene.h:137-160), with the reference's defaults.
    struct ToneMapper {
        ToneMapperType type = ToneMapperType::LINEAR;
        bool applyGamma = true;
        bool shouldClamp = true;
        double gamma = 2.2;
        double exposure = 0.0;
        double key = 0.18;
        bool enableAveraging = false;
        double averagingAlpha = 0.1;
        bool fastPow = true;  // MathUtils::fastPow, as the reference
    } toneMapper;

    struct RootBVH {
        bool enabled = true;
        BVHBuildInfo buildInfo;
        FlatBVH bvh;
    } rootBVH;

    struct Primitives {
        std::vector<Triangle> triangles;
    } primitives;

    std::vector<Material> materials;

    // ModelLoaderInfo (src/Utils/ModelLoader.h:27-46): models listed here are
    // read by initialize(), OBJ or PLY by file extension.
    struct ModelLoaderInfo {
        std::string modelFilePath;
        Vector3 scale{1.0, 1.0, 1.0};
        EulerAngle rotate{0.0, 0.0, 0.0};
        Vector3 translate{0.0, 0.0, 0.0};
        bool ignoreMaterials = false;
    };
    std::vector<ModelLoaderInfo> models;

    // Scene::initialize (Scene.cpp:203-500): loads the models, flattens the
    // triangles and builds the root BVH unless one was restored.
    void initialize() {
        for (const ModelLoaderInfo& mi : models) loadModel(mi);
        models.clear();
        vertices_.resize(primitives.triangles.size() * 9);
        materialIds_.resize(primitives.triangles.size());
        for (size_t i = 0; i < primitives.triangles.size(); ++i) {
            const Triangle& t = primitives.triangles[i];
            if (t.materialId >= materials.size()) throw RaycerError(RC_ERR_INVALID, "a triangle has a non-existent material id");
            for (int v = 0; v < 3; ++v) {
                vertices_[9 * i + 3 * v + 0] = float(t.vertices[v].x);
                vertices_[9 * i + 3 * v + 1] = float(t.vertices[v].y);
                vertices_[9 * i + 3 * v + 2] = float(t.vertices[v].z);
            }
            materialIds_[i] = uint32_t(t.materialId);
        }
        if (!rootBVH.bvh.hasBeenBuilt) rootBVH.bvh.build(vertices_, rootBVH.buildInfo);
    }

    // Scene::createTestScene counterpart: the seeded benchmark scenes.
    static Scene createTestScene(uint32_t kind, uint32_t seed = 0, uint32_t level = 0, uint32_t grid = 0,
                                 uint64_t triCount = 0) {
        rch_scene_params p{kind, seed, level, grid, triCount};
        rch_scene_desc d;
        check(rch_scene_describe(&p, &d));
        std::vector<float> v(d.ntri * 9);
        std::vector<uint32_t> m(d.ntri);
        std::vector<rc_material> mats(d.nmat);
        check(rch_scene_generate(&p, v.data(), m.data(), mats.data()));
        Scene s;
        s.primitives.triangles.resize(d.ntri);
        for (size_t i = 0; i < d.ntri; ++i) {
            for (int k = 0; k < 3; ++k)
                s.primitives.triangles[i].vertices[k] = Vector3{v[9 * i + 3 * k], v[9 * i + 3 * k + 1], v[9 * i + 3 * k + 2]};
            s.primitives.triangles[i].materialId = m[i];
        }
        for (const rc_material& rm : mats) {
            Material mt;
            mt.diffuseReflectance = Color{rm.diffuse[0], rm.diffuse[1], rm.diffuse[2]};
            mt.emittance = Color{rm.emittance[0], rm.emittance[1], rm.emittance[2]};
            mt.ambientReflectance = Color{rm.ambient[0], rm.ambient[1], rm.ambient[2]};
            mt.specularReflectance = Color{rm.specular[0], rm.specular[1], rm.specular[2]};
            mt.specularShininess = rm.shininess;
            mt.isEmissive = rm.flags & RC_MAT_EMISSIVE;
            mt.nonShadowing = rm.flags & RC_MAT_NON_SHADOWING;
            s.materials.push_back(mt);
        }
        const rc_render_cfg& c = d.cfg;
        s.camera.position = Vector3{d.camera.position[0], d.camera.position[1], d.camera.position[2]};
        s.camera.orientation = EulerAngle{d.camera.pitch, d.camera.yaw, d.camera.roll};
        s.camera.fov = d.camera.fov;
        s.general.maxPathLength = c.max_path_length;
        s.general.multiSamplerType = sampler_of(c.sampler);
        if (c.sampler != RC_SAMPLER_RANDOM) {
            uint64_t n = 1;
            while ((n + 1) * (n + 1) <= c.spp) ++n;
            s.general.multiSampleCountSqrt = n;
        } else {
            s.general.randomSampleCount = c.spp;
        }
        s.general.multiSamplerFilterType = FilterType(c.filter);
        s.general.filterRadius = c.filter_radius;
        if (c.filter == RC_FILTER_GAUSSIAN) s.general.gaussianStandardDeviation = c.filter_param[0];
        if (c.filter == RC_FILTER_MITCHELL) {
            s.general.mitchellB = c.filter_param[0];
            s.general.mitchellC = c.filter_param[1];
        }
        s.toneMapper.type = ToneMapperType::REINHARD;  // the benchmark scenes are shown with Reinhard
        s.general.rayStartOffset = c.ray_start_offset;
        s.general.russianRouletteStart = c.rr_start;
        s.general.lambertEstimator = c.estimator == RC_EST_LAMBERT;
        s.general.skyRadiance = Color{c.sky[0], c.sky[1], c.sky[2]};
        s.general.sunDirection = Vector3{c.sun_dir[0], c.sun_dir[1], c.sun_dir[2]};
        s.general.sunIrradiance = Color{c.sun_irradiance[0], c.sun_irradiance[1], c.sun_irradiance[2]};
        s.general.seed = c.seed;
        const rc_point_light& l = d.light;
        if (l.max_distance > 0.0f) {  // the scenes lit for the ray tracer carry a point light
            PointLight pl;
            pl.position = Vector3{l.position[0], l.position[1], l.position[2]};
            pl.color = Color{l.color[0], l.color[1], l.color[2]};
            pl.intensity = l.intensity;
            pl.maxDistance = l.max_distance;
            pl.attenuation = l.attenuation;
            s.lights.pointLights.push_back(pl);
        }
        s.lights.ambientLight.color = Color{l.ambient_color[0], l.ambient_color[1], l.ambient_color[2]};
        s.lights.ambientLight.intensity = l.ambient_intensity;
        s.defaultWidth = c.width;
        s.defaultHeight = c.height;
        return s;
    }

    // TestScene3 (src/TestScenes/TestScene3.cpp:11-47): the Cornell box model,
    // camera at (0, 1, 3.5) looking down -z, ambient 0.01 and one point light,
    // ray tracer.
    static Scene createTestScene3(const std::string& cornellBoxObjPath) {
        Scene s;
        s.general.tracerType = TracerType::RAY;
        s.camera.position = Vector3{0.0, 1.0, 3.5};
        s.camera.orientation = EulerAngle{0.0, 0.0, 0.0};
        ModelLoaderInfo mi;
        mi.modelFilePath = cornellBoxObjPath;
        s.models.push_back(mi);
        s.lights.ambientLight.color = Color{1.0, 1.0, 1.0};
        s.lights.ambientLight.intensity = 0.01;
        PointLight pl;
        pl.color = Color{1.0, 0.71, 0.24};
        pl.intensity = 1.0;
        pl.position = Vector3{0.0, 1.9, 0.0};
        pl.maxDistance = 3.0;
        pl.attenuation = 1.0;
        s.lights.pointLights.push_back(pl);
        s.defaultWidth = 1280;
        s.defaultHeight = 800;
        return s;
    }

    // Scene::loadFromFile / saveToFile (Scene.cpp:67-183): cereal XML, JSON or
    // binary by the file's extension (.xml, .json, .bin).  Triangles,
    // materials, lights, settings, the model list and a built root BVH; other
    // primitive types are skipped (XML, JSON) or rejected (binary).
    static Scene loadFromFile(const std::string& path) {
        rch_scene_file_data* d = nullptr;
        check(rch_scene_file_load(path.c_str(), &d));
        Scene s;
        try {
            s.readFileData(*d);
        } catch (...) {
            rch_scene_file_free(d);
            throw;
        }
        rch_scene_file_free(d);
        return s;
    }
    inline void saveToFile(const std::string& path) const;  // defined after HipTracer

    const std::vector<float>& flatVertices() const { return vertices_; }
    const std::vector<uint32_t>& materialIds() const { return materialIds_; }
    uint32_t defaultWidth = 0, defaultHeight = 0;

private:
    void readFileData(const rch_scene_file_data& d) {
        auto col = [](const float* c) { return Color{c[0], c[1], c[2]}; };
        auto vec = [](const float* v) { return Vector3{v[0], v[1], v[2]}; };
        const rc_render_cfg& c = d.cfg;
        const rc_ray_settings& r = d.ray;
        general.tracerType = d.tracer_type == 0 ? TracerType::RAY : TracerType::PATH;
        general.maxRayIterations = r.max_ray_iterations;
        general.maxPathLength = c.max_path_length;
        general.rayStartOffset = c.ray_start_offset;
        general.backgroundColor = col(r.background);
        general.multiSamplerType = sampler_of(c.sampler);
        uint64_t n = 1;
        while ((n + 1) * (n + 1) <= c.spp) ++n;
        general.multiSampleCountSqrt = n;
        general.multiSamplerFilterType = FilterType(c.filter);
        general.filterRadius = c.filter_radius;
        if (c.filter == RC_FILTER_GAUSSIAN) general.gaussianStandardDeviation = c.filter_param[0];
        if (c.filter == RC_FILTER_MITCHELL) {
            general.mitchellB = c.filter_param[0];
            general.mitchellC = c.filter_param[1];
        }
        general.visualizeDepth = r.visualize_depth != 0;
        general.visualizeDepthMaxDistance = r.visualize_depth_max;
        camera.position = vec(d.camera.position);
        camera.orientation = EulerAngle{d.camera.pitch, d.camera.yaw, d.camera.roll};
        camera.fov = d.camera.fov;
        const ToneMapperType tt[4] = {ToneMapperType::REINHARD, ToneMapperType::LINEAR, ToneMapperType::SIMPLE,
                                      ToneMapperType::PASSTHROUGH};
        toneMapper.type = tt[d.tonemap.type & 3u];
        toneMapper.applyGamma = d.tonemap.apply_gamma != 0;
        toneMapper.shouldClamp = d.tonemap.clamp != 0;
        toneMapper.gamma = d.tonemap.gamma;
        toneMapper.exposure = d.tonemap.exposure;
        toneMapper.key = d.tonemap.key;
        toneMapper.enableAveraging = d.tonemap.enable_averaging != 0;
        toneMapper.averagingAlpha = d.tonemap.averaging_alpha;
        simpleFog.enabled = r.fog_enabled != 0;
        simpleFog.color = col(r.fog_color);
        simpleFog.distance = r.fog_distance;
        simpleFog.steepness = r.fog_steepness;
        simpleFog.heightDispersion = r.fog_height_dispersion != 0;
        simpleFog.height = r.fog_height;
        simpleFog.heightSteepness = r.fog_height_steepness;
        AmbientLight& a = lights.ambientLight;
        a.color = col(r.ambient_color);
        a.intensity = r.ambient_intensity;
        a.enableAmbientOcclusion = r.ao_enabled != 0;
        a.ambientOcclusionSamplerType = sampler_of(r.ao_sampler);
        a.ambientOcclusionSampleCountSqrt = r.ao_n;
        a.ambientOcclusionMaxSampleDistance = r.ao_max_distance;
        a.ambientOcclusionSampleDistribution = r.ao_distribution;
        for (size_t i = 0; i < d.nlights; ++i) {
            const rc_light& L = d.lights[i];
            if (L.type == RC_LIGHT_DIRECTIONAL) {
                DirectionalLight x;
                x.color = col(L.color);
                x.intensity = L.intensity;
                x.direction = vec(L.direction);
                lights.directionalLights.push_back(x);
                continue;
            }
            SpotLight x;
            x.color = col(L.color);
            x.intensity = L.intensity;
            x.position = vec(L.position);
            x.maxDistance = L.max_distance;
            x.attenuation = L.attenuation;
            x.enableAreaLight = L.area_light != 0;
            x.areaLightSamplerType = sampler_of(L.area_sampler);
            x.areaLightSampleCountSqrt = L.area_n;
            x.areaLightRadius = L.area_radius;
            if (L.type == RC_LIGHT_POINT) {
                lights.pointLights.push_back(x);
            } else {
                x.direction = vec(L.direction);
                x.angle = L.angle;
                x.sideAttenuation = L.side_attenuation;
                lights.spotLights.push_back(x);
            }
        }
        for (size_t i = 0; i < d.nmat; ++i) materials.push_back(from_rc_material(d.materials[i], d.ray_materials[i]));
        primitives.triangles.resize(d.ntri);
        for (size_t i = 0; i < d.ntri; ++i) {
            for (int k = 0; k < 3; ++k) primitives.triangles[i].vertices[k] = vec(d.vertices + 9 * i + 3 * k);
            primitives.triangles[i].materialId = d.tri_material[i];
        }
        for (size_t i = 0; i < d.nmodels; ++i) {
            const rch_scene_model& m = d.models[i];
            ModelLoaderInfo mi;
            mi.modelFilePath = m.path;
            mi.scale = Vector3{m.info.scale[0], m.info.scale[1], m.info.scale[2]};
            mi.rotate = EulerAngle{m.info.rotate[0], m.info.rotate[1], m.info.rotate[2]};
            mi.translate = Vector3{m.info.translate[0], m.info.translate[1], m.info.translate[2]};
            mi.ignoreMaterials = m.info.ignore_materials != 0;
            models.push_back(mi);
        }
        if (d.nnodes) {
            rootBVH.bvh.flatNodes.assign(d.nodes, d.nodes + d.nnodes);
            rootBVH.bvh.orderedPrimitiveIds.assign(d.ordered, d.ordered + d.ntri);
            rootBVH.bvh.hasBeenBuilt = true;
        }
    }

    void loadModel(const ModelLoaderInfo& mi) {
        rch_obj_info info;
        rch_obj_default_info(&info);
        info.scale[0] = mi.scale.x; info.scale[1] = mi.scale.y; info.scale[2] = mi.scale.z;
        info.rotate[0] = mi.rotate.pitch; info.rotate[1] = mi.rotate.yaw; info.rotate[2] = mi.rotate.roll;
        info.translate[0] = mi.translate.x; info.translate[1] = mi.translate.y; info.translate[2] = mi.translate.z;
        info.ignore_materials = mi.ignoreMaterials ? 1u : 0u;
        const std::string& f = mi.modelFilePath;
        const bool ply = f.size() >= 4 && (f.compare(f.size() - 4, 4, ".ply") == 0 || f.compare(f.size() - 4, 4, ".PLY") == 0);
        rch_obj* h = nullptr;
        check(ply ? rch_ply_load(f.c_str(), &info, &h) : rch_obj_load(f.c_str(), &info, &h));
        const size_t nt = rch_obj_triangle_count(h), nm = rch_obj_material_count(h);
        std::vector<float> v(nt * 9);
        std::vector<uint32_t> m(nt);
        std::vector<rc_material> mats(nm);
        int rc = rch_obj_get(h, v.data(), m.data(), mats.data());
        rch_obj_free(h);
        check(rc);
        const uint64_t base = materials.size();
        for (const rc_material& rm : mats) {
            Material mt;
            mt.diffuseReflectance = Color{rm.diffuse[0], rm.diffuse[1], rm.diffuse[2]};
            mt.emittance = Color{rm.emittance[0], rm.emittance[1], rm.emittance[2]};
            mt.ambientReflectance = Color{rm.ambient[0], rm.ambient[1], rm.ambient[2]};
            mt.specularReflectance = Color{rm.specular[0], rm.specular[1], rm.specular[2]};
            mt.specularShininess = rm.shininess;
            mt.isEmissive = rm.flags & RC_MAT_EMISSIVE;
            materials.push_back(mt);
        }
        for (size_t i = 0; i < nt; ++i) {
            Triangle t;
            for (int k = 0; k < 3; ++k) t.vertices[k] = Vector3{v[9 * i + 3 * k], v[9 * i + 3 * k + 1], v[9 * i + 3 * k + 2]};
            t.materialId = base + m[i];
            primitives.triangles.push_back(t);
        }
    }

    std::vector<float> vertices_;
    std::vector<uint32_t> materialIds_;
};

// RGBA float image, row 0 at the bottom (src/Rendering/Image.h:10-14).
struct Image {
    uint64_t width = 0, height = 0;
    std::vector<float> pixels;  // width * height * 4

    // Portable float map: stored bottom row first, like this image.
    bool savePFM(const std::string& path) const {
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) return false;
        std::fprintf(f, "PF\n%llu %llu\n-1.0\n", (unsigned long long)width, (unsigned long long)height);
        std::vector<float> rgb(width * 3);
        for (uint64_t y = 0; y < height; ++y) {
            for (uint64_t x = 0; x < width; ++x)
                for (int c = 0; c < 3; ++c) rgb[3 * x + c] = pixels[4 * (y * width + x) + c];
            std::fwrite(rgb.data(), sizeof(float), rgb.size(), f);
        }
        return std::fclose(f) == 0;
    }

    // 8-bit binary PPM, top row first (Image::save flips rows, Image.cpp:118-124).
    bool savePPM(const std::string& path) const {
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) return false;
        std::fprintf(f, "P6\n%llu %llu\n255\n", (unsigned long long)width, (unsigned long long)height);
        std::vector<unsigned char> row(width * 3);
        for (uint64_t yy = 0; yy < height; ++yy) {
            uint64_t y = height - 1 - yy;
            for (uint64_t x = 0; x < width; ++x)
                for (int c = 0; c < 3; ++c) {
                    float v = pixels[4 * (y * width + x) + c];
                    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                    row[3 * x + c] = (unsigned char)(v * 255.0f + 0.5f);
                }
            std::fwrite(row.data(), 1, row.size(), f);
        }
        return std::fclose(f) == 0;
    }
};

// Film (Film.h:23-52): the host side holds the images; the accumulators live
// on the device inside HipTracer, as CLTracer keeps them in its own buffers.
class Film {
public:
    void resize(uint64_t w, uint64_t h) {
        width_ = w;
        height_ = h;
    }
    void setToneMappedImage(const Image& img) { toneMapped_ = img; }
    void setLinearImage(const Image& img) { linear_ = img; }
    const Image& getToneMappedImage() const { return toneMapped_; }
    const Image& getLinearImage() const { return linear_; }
    uint64_t getWidth() const { return width_; }
    uint64_t getHeight() const { return height_; }

private:
    uint64_t width_ = 0, height_ = 0;
    Image linear_, toneMapped_;
};

// TracerState (TracerState.h:13-26)
struct TracerState {
    TracerState() : pixelsProcessed(0) {}
    Scene* scene = nullptr;
    Film* film = nullptr;
    uint64_t filmWidth = 0;
    uint64_t filmHeight = 0;
    uint64_t pixelStartOffset = 0;
    uint64_t pixelCount = 0;
    std::atomic<uint64_t> pixelsProcessed;
};

// The GPU tracer with CLTracer's lifecycle (CLTracer.h:23-90): the tracer owns
// every device buffer; run() accumulates into the device film until clear().
class HipTracer {
public:
    explicit HipTracer(int device = 0) { check(rc_create(device, &ctx_)); }
    ~HipTracer() {
        releaseImageBuffers();
        releaseBuffers();
        rc_destroy(ctx_);
    }
    HipTracer(const HipTracer&) = delete;
    HipTracer& operator=(const HipTracer&) = delete;

    // Kernels are compiled ahead of time for gfx950; nothing to do.
    void initializeKernels() {}
    void releaseKernels() {}

    void initializeImageBuffers(uint64_t width, uint64_t height) {
        releaseImageBuffers();
        check(rc_film_create(ctx_, uint32_t(width), uint32_t(height), &film_));
        width_ = width;
        height_ = height;
        frame_ = 0;
    }
    void releaseImageBuffers() {
        if (film_) rc_film_free(film_);
        film_ = nullptr;
    }

    // CLTracer::initializeBuffers: reads the scene (built root BVH required).
    void initializeBuffers(const Scene& scene) {
        if (!scene.rootBVH.bvh.hasBeenBuilt) throw RaycerError(RC_ERR_INVALID, "scene is not initialized");
        releaseBuffers();
        std::vector<rc_material> mats(scene.materials.size());
        for (size_t i = 0; i < mats.size(); ++i) mats[i] = to_rc_material(scene.materials[i]);
        const auto& bvh = scene.rootBVH.bvh;
        check(rc_scene_upload(ctx_, scene.flatVertices().data(), scene.flatVertices().size() / 9,
                              scene.materialIds().data(), mats.data(), mats.size(), bvh.flatNodes.data(),
                              bvh.flatNodes.size(), bvh.orderedPrimitiveIds.data(), &scene_));
    }
    void releaseBuffers() {
        if (scene_) rc_scene_free(scene_);
        scene_ = nullptr;
    }

    // Tracer::run: renders one frame of the scene's samples into the device
    // film, polling `interrupted` between passes and publishing progress in
    // state.pixelsProcessed.
    void run(TracerState& state, std::atomic<bool>& interrupted) {
        if (!film_ || !scene_) throw RaycerError(RC_ERR_INVALID, "HipTracer buffers are not initialized");
        if (state.filmWidth != width_ || state.filmHeight != height_)
            throw RaycerError(RC_ERR_INVALID, "TracerState film size differs from the image buffers");
        const Scene& sc = *state.scene;
        rc_render_cfg c{};
        c.width = uint32_t(width_);
        c.height = uint32_t(height_);
        c.sampler = to_rc(sc.general.multiSamplerType);
        if (sc.general.multiSamplerType != SamplerType::RANDOM)
            c.spp = uint32_t(sc.general.multiSampleCountSqrt * sc.general.multiSampleCountSqrt);
        else
            c.spp = uint32_t(sc.general.randomSampleCount);
        c.max_path_length = uint32_t(sc.general.maxPathLength);
        c.rr_start = uint32_t(sc.general.russianRouletteStart);
        c.filter = to_rc(sc.general.multiSamplerFilterType);
        c.filter_radius = float(sc.general.filterRadius);
        if (sc.general.multiSamplerFilterType == FilterType::GAUSSIAN) {
            c.filter_param[0] = float(sc.general.gaussianStandardDeviation);
        } else if (sc.general.multiSamplerFilterType == FilterType::MITCHELL) {
            c.filter_param[0] = float(sc.general.mitchellB);
            c.filter_param[1] = float(sc.general.mitchellC);
        }
        c.estimator = sc.general.lambertEstimator ? RC_EST_LAMBERT : RC_EST_REFERENCE;
        c.ray_start_offset = float(sc.general.rayStartOffset);
        c.sky[0] = float(sc.general.skyRadiance.r);
        c.sky[1] = float(sc.general.skyRadiance.g);
        c.sky[2] = float(sc.general.skyRadiance.b);
        c.sun_dir[0] = float(sc.general.sunDirection.x);
        c.sun_dir[1] = float(sc.general.sunDirection.y);
        c.sun_dir[2] = float(sc.general.sunDirection.z);
        c.sun_irradiance[0] = float(sc.general.sunIrradiance.r);
        c.sun_irradiance[1] = float(sc.general.sunIrradiance.g);
        c.sun_irradiance[2] = float(sc.general.sunIrradiance.b);
        c.seed = sc.general.seed;
        c.frame = frame_++;
        rc_camera cam{{float(sc.camera.position.x), float(sc.camera.position.y), float(sc.camera.position.z)},
                      float(sc.camera.orientation.pitch), float(sc.camera.orientation.yaw),
                      float(sc.camera.orientation.roll), float(sc.camera.fov)};
        if (sc.general.tracerType == TracerType::RAY) {
            // Raytracer (Raytracer.cpp:21-597) over the same camera samples
            if (interrupted.load()) return;
            RayInputs in = rayInputs(sc);
            check(rc_raytrace(ctx_, scene_, film_, &cam, &c, &in.settings, in.materials.data(),
                              uint32_t(in.materials.size()), in.lights.data(), uint32_t(in.lights.size())));
            state.pixelsProcessed = state.pixelCount ? state.pixelCount : width_ * height_;
            return;
        }
        struct Poll {
            TracerState* state;
            std::atomic<bool>* interrupted;
        } poll{&state, &interrupted};
        auto cb = [](void* u, uint64_t done) -> int {
            auto* p = static_cast<Poll*>(u);
            p->state->pixelsProcessed = done;
            return p->interrupted->load() ? 1 : 0;
        };
        int rc = rc_render_progress(ctx_, scene_, film_, &cam, &c, cb, &poll);
        if (rc == RC_ERR_INTERRUPTED) return;  // the film keeps the passes done, as Tracer::run
        check(rc);
        state.pixelsProcessed = state.pixelCount ? state.pixelCount : width_ * height_;
    }

    // The ray tracer's inputs from the scene: Scene::General's ray-tracer
    // fields, the lights, SimpleFog and the materials' ray-tracing fields.
    struct RayInputs {
        rc_ray_settings settings{};
        std::vector<rc_ray_material> materials;
        std::vector<rc_light> lights;
    };
    static RayInputs rayInputs(const Scene& sc) {
        RayInputs in;
        rc_ray_settings& r = in.settings;
        auto put3 = [](float* d, const Color& c) { d[0] = float(c.r); d[1] = float(c.g); d[2] = float(c.b); };
        auto putv = [](float* d, const Vector3& v) { d[0] = float(v.x); d[1] = float(v.y); d[2] = float(v.z); };
        put3(r.background, sc.general.backgroundColor);
        r.max_ray_iterations = uint32_t(sc.general.maxRayIterations);
        const AmbientLight& a = sc.lights.ambientLight;
        put3(r.ambient_color, a.color);
        r.ambient_intensity = float(a.intensity);
        r.ao_enabled = a.enableAmbientOcclusion ? 1u : 0u;
        r.ao_sampler = to_rc(a.ambientOcclusionSamplerType);
        r.ao_n = uint32_t(a.ambientOcclusionSampleCountSqrt);
        r.ao_max_distance = float(a.ambientOcclusionMaxSampleDistance);
        r.ao_distribution = float(a.ambientOcclusionSampleDistribution);
        r.visualize_depth = sc.general.visualizeDepth ? 1u : 0u;
        r.visualize_depth_max = float(sc.general.visualizeDepthMaxDistance);
        const Scene::SimpleFog& f = sc.simpleFog;
        r.fog_enabled = f.enabled ? 1u : 0u;
        put3(r.fog_color, f.color);
        r.fog_distance = float(f.distance);
        r.fog_steepness = float(f.steepness);
        r.fog_height_dispersion = f.heightDispersion ? 1u : 0u;
        r.fog_height = float(f.height);
        r.fog_height_steepness = float(f.heightSteepness);
        for (const Material& m : sc.materials) in.materials.push_back(to_rc_ray_material(m));
        auto point = [&](const PointLight& p, uint32_t type) {
            rc_light l{};
            l.type = type;
            put3(l.color, p.color);
            l.intensity = float(p.intensity);
            putv(l.position, p.position);
            l.max_distance = float(p.maxDistance);
            l.attenuation = float(p.attenuation);
            l.area_light = p.enableAreaLight ? 1u : 0u;
            l.area_sampler = to_rc(p.areaLightSamplerType);
            l.area_n = uint32_t(p.areaLightSampleCountSqrt);
            l.area_radius = float(p.areaLightRadius);
            return l;
        };
        for (const DirectionalLight& d : sc.lights.directionalLights) {
            rc_light l{};
            l.type = RC_LIGHT_DIRECTIONAL;
            put3(l.color, d.color);
            l.intensity = float(d.intensity);
            putv(l.direction, d.direction);
            in.lights.push_back(l);
        }
        for (const PointLight& p : sc.lights.pointLights) in.lights.push_back(point(p, RC_LIGHT_POINT));
        for (const SpotLight& sp : sc.lights.spotLights) {
            rc_light l = point(sp, RC_LIGHT_SPOT);
            putv(l.direction, sp.direction);
            l.angle = float(sp.angle);
            l.side_attenuation = float(sp.sideAttenuation);
            in.lights.push_back(l);
        }
        return in;
    }

    // The one-light raycast of the 1080p Phong workload: primary ray through a
    // jittered point of each pixel, ambient + Phong + hard shadow for the
    // scene's first point light (Raytracer.cpp:312-407, 524-551).  Returns the
    // linear RGBA image; the film is not touched.
    Image raycastPhong(const Scene& sc, uint64_t seed = 1) {
        if (!scene_) throw RaycerError(RC_ERR_INVALID, "HipTracer buffers are not initialized");
        if (sc.lights.pointLights.empty()) throw RaycerError(RC_ERR_INVALID, "raycastPhong needs a point light");
        const PointLight& p = sc.lights.pointLights.front();
        const AmbientLight& a = sc.lights.ambientLight;
        rc_point_light pl{{float(p.position.x), float(p.position.y), float(p.position.z)},
                          {float(p.color.r), float(p.color.g), float(p.color.b)},
                          float(p.intensity),
                          float(p.maxDistance),
                          float(p.attenuation),
                          {float(a.color.r), float(a.color.g), float(a.color.b)},
                          float(a.intensity)};
        Image img = blank();
        rc_camera cam{{float(sc.camera.position.x), float(sc.camera.position.y), float(sc.camera.position.z)},
                      float(sc.camera.orientation.pitch), float(sc.camera.orientation.yaw),
                      float(sc.camera.orientation.roll), float(sc.camera.fov)};
        // pixel (x, y) at y * width + x with y = 0 at the bottom, as Image
        check(rc_raycast_phong(ctx_, scene_, &cam, uint32_t(width_), uint32_t(height_), seed,
                               float(sc.general.rayStartOffset), &pl, nullptr, nullptr, img.pixels.data()));
        return img;
    }

    // Film::generateToneMappedImage on the device with the scene's tone mapper,
    // then download (CLTracer::downloadImage).  Reinhard's moving average of the
    // maximum luminance lives with the film, as it lives with the reference's.
    Image downloadImage(const Scene::ToneMapper& tm = Scene::ToneMapper()) {
        Image img = blank();
        rc_tonemap t{};
        t.key = float(tm.key);
        t.gamma = float(tm.gamma);
        t.clamp = tm.shouldClamp ? 1u : 0u;
        t.apply_gamma = tm.applyGamma ? 1u : 0u;
        t.fast_pow = tm.fastPow ? 1u : 0u;
        t.type = to_rc(tm.type);
        t.exposure = float(tm.exposure);
        t.enable_averaging = tm.enableAveraging ? 1u : 0u;
        t.averaging_alpha = float(tm.averagingAlpha);
        check(rc_film_tonemap(ctx_, film_, &t, img.pixels.data(), nullptr, nullptr));
        return img;
    }
    Image downloadLinearImage() {
        Image img = blank();
        check(rc_film_download_linear(ctx_, film_, img.pixels.data()));
        return img;
    }
    // CLTracer::clear: restart progressive accumulation.
    void clear() {
        check(rc_film_clear(ctx_, film_));
        frame_ = 0;
    }

    rc_ctx* context() { return ctx_; }
    rc_stats stats() {
        rc_stats s;
        check(rc_get_stats(ctx_, &s));
        return s;
    }

private:
    Image blank() const {
        Image img;
        img.width = width_;
        img.height = height_;
        img.pixels.resize(width_ * height_ * 4);
        return img;
    }
    rc_ctx* ctx_ = nullptr;
    rc_scene* scene_ = nullptr;
    rc_film* film_ = nullptr;
    uint64_t width_ = 0, height_ = 0;
    uint32_t frame_ = 0;
};

inline void Scene::saveToFile(const std::string& path) const {
    rch_scene_file_data d{};
    d.tracer_type = general.tracerType == TracerType::RAY ? 0u : 1u;
    d.camera = rc_camera{{float(camera.position.x), float(camera.position.y), float(camera.position.z)},
                         float(camera.orientation.pitch), float(camera.orientation.yaw),
                         float(camera.orientation.roll), float(camera.fov)};
    rc_render_cfg& c = d.cfg;
    c.sampler = to_rc(general.multiSamplerType);
    c.spp = uint32_t(general.multiSampleCountSqrt * general.multiSampleCountSqrt);
    c.filter = to_rc(general.multiSamplerFilterType);
    c.filter_radius = float(general.filterRadius);
    c.max_path_length = uint32_t(general.maxPathLength);
    c.ray_start_offset = float(general.rayStartOffset);
    d.tonemap.type = to_rc(toneMapper.type);
    d.tonemap.apply_gamma = toneMapper.applyGamma ? 1u : 0u;
    d.tonemap.clamp = toneMapper.shouldClamp ? 1u : 0u;
    d.tonemap.gamma = float(toneMapper.gamma);
    d.tonemap.exposure = float(toneMapper.exposure);
    d.tonemap.key = float(toneMapper.key);
    d.tonemap.enable_averaging = toneMapper.enableAveraging ? 1u : 0u;
    d.tonemap.averaging_alpha = float(toneMapper.averagingAlpha);
    HipTracer::RayInputs in = HipTracer::rayInputs(*this);
    d.ray = in.settings;
    std::vector<rc_material> mats;
    std::vector<uint64_t> ids;
    for (size_t i = 0; i < materials.size(); ++i) {
        mats.push_back(to_rc_material(materials[i]));
        ids.push_back(i + 1);
    }
    std::vector<float> v(primitives.triangles.size() * 9);
    std::vector<uint32_t> tm(primitives.triangles.size());
    std::vector<uint64_t> tid(primitives.triangles.size());
    for (size_t i = 0; i < primitives.triangles.size(); ++i) {
        const Triangle& t = primitives.triangles[i];
        for (int k = 0; k < 3; ++k) {
            v[9 * i + 3 * k] = float(t.vertices[k].x);
            v[9 * i + 3 * k + 1] = float(t.vertices[k].y);
            v[9 * i + 3 * k + 2] = float(t.vertices[k].z);
        }
        tm[i] = uint32_t(t.materialId);
        tid[i] = i + 1;
    }
    std::vector<rch_scene_model> ms(models.size());
    for (size_t i = 0; i < models.size(); ++i) {
        const ModelLoaderInfo& mi = models[i];
        rch_scene_model& m = ms[i];
        m = rch_scene_model{};
        if (mi.modelFilePath.size() >= sizeof(m.path)) throw RaycerError(RC_ERR_INVALID, "model path too long");
        std::memcpy(m.path, mi.modelFilePath.c_str(), mi.modelFilePath.size() + 1);
        rch_obj_default_info(&m.info);
        m.info.scale[0] = mi.scale.x; m.info.scale[1] = mi.scale.y; m.info.scale[2] = mi.scale.z;
        m.info.rotate[0] = mi.rotate.pitch; m.info.rotate[1] = mi.rotate.yaw; m.info.rotate[2] = mi.rotate.roll;
        m.info.translate[0] = mi.translate.x; m.info.translate[1] = mi.translate.y; m.info.translate[2] = mi.translate.z;
        m.info.ignore_materials = mi.ignoreMaterials ? 1u : 0u;
    }
    d.nmat = mats.size();
    d.materials = mats.data();
    d.ray_materials = in.materials.data();
    d.material_ids = ids.data();
    d.ntri = tm.size();
    d.vertices = v.data();
    d.tri_material = tm.data();
    d.tri_ids = tid.data();
    d.nlights = in.lights.size();
    d.lights = in.lights.data();
    d.nmodels = ms.size();
    d.models = ms.data();
    // a built tree is written only when it covers exactly these triangles
    std::vector<rc_flat_node> nodes;
    std::vector<uint64_t> ordered;
    if (rootBVH.bvh.hasBeenBuilt && models.empty() && rootBVH.bvh.orderedPrimitiveIds.size() == tm.size()) {
        nodes = rootBVH.bvh.flatNodes;
        ordered = rootBVH.bvh.orderedPrimitiveIds;
        d.nnodes = nodes.size();
        d.nodes = nodes.data();
        d.ordered = ordered.data();
    }
    check(rch_scene_file_save(path.c_str(), &d));
}

}  // namespace raycer
