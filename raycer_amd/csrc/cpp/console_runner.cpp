// raycer_render: non-interactive renderer, the counterpart of raycer's
// ConsoleRunner (src/Runners/ConsoleRunner.cpp:22-175).  Builds a seeded test
// scene, initialises it (root BVH), renders on a render thread while the main
// thread polls TracerState::pixelsProcessed every 100 ms, tone maps and saves.
// SIGINT interrupts the render; the passes done are still saved, as the
// reference saves partial_result.png (ConsoleRunner.cpp:63-65).
//
//   raycer_render --scene box|sphere|terrain|soup|cornell [--seed N] [--level N] [--grid N]
//                 [--tris N] [--mesh cornellbox.obj] [--tracer ray|path|phong] [--width W] [--height H]
//                 [--spp N] [--frames F] [--linear out.pfm] [--tonemapped out.ppm]
//                 [--sampler center|random|regular|jittered|cmj|poisson_disc]
//                 [--filter box|tent|bell|gaussian|mitchell|lanczos_sinc]
//                 [--tonemapper passthrough|linear|simple|reinhard]
//                 [--scene-file scene.xml|scene.json] [--save-scene out.xml|out.json]
//
// --scene-file reads a cereal XML or JSON scene (Scene::loadFromFile) instead
// of a generated one; --save-scene writes the scene that was rendered.
// cornell is raycer's TestScene3 (the Cornell box OBJ, by default the copy in
// tests/golden/meshes next to this program's package); its default tracer is
// the RAY (Whitted) tracer with the scene's ambient and point light, --tracer
// path renders it with the path tracer, lit by the box's emissive ceiling
// quad, and --tracer phong runs the one-light raycast of the Phong workload.
#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <thread>

#include "raycer.hpp"

namespace {
std::atomic<bool> g_interrupted(false);
void on_sigint(int) { g_interrupted = true; }

uint32_t scene_kind(const std::string& s) {
    if (s == "box") return RCH_SCENE_BOX;
    if (s == "sphere") return RCH_SCENE_SPHERE_GROUND;
    if (s == "terrain") return RCH_SCENE_TERRAIN;
    if (s == "soup") return RCH_SCENE_SOUP;
    throw raycer::RaycerError(RC_ERR_INVALID, "unknown scene " + s);
}

// the reference's SamplerType / FilterType / ToneMapperType names, lower case
template <typename E, size_t N>
E parse_enum(const std::string& s, const char* const (&names)[N], const char* what) {
    for (size_t i = 0; i < N; ++i)
        if (s == names[i]) return E(i);
    throw raycer::RaycerError(RC_ERR_INVALID, std::string("unknown ") + what + " " + s);
}
const char* const kSamplers[] = {"center", "random", "regular", "jittered", "cmj", "poisson_disc"};
const char* const kFilters[] = {"box", "tent", "bell", "gaussian", "mitchell", "lanczos_sinc"};
const char* const kToneMappers[] = {"passthrough", "linear", "simple", "reinhard"};
}  // namespace

int main(int argc, char** argv) {
    std::string scene_name = "box", linear_out, tm_out, mesh, tracer_name, sampler_name, filter_name, tm_name;
    std::string scene_file, save_scene;
    uint32_t seed = 0, level = 0, grid = 0, width = 0, height = 0, spp = 0, frames = 1;
    uint64_t tris = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw raycer::RaycerError(RC_ERR_INVALID, "missing value for " + a);
            return argv[++i];
        };
        try {
            if (a == "--scene") scene_name = next();
            else if (a == "--seed") seed = std::stoul(next());
            else if (a == "--level") level = std::stoul(next());
            else if (a == "--grid") grid = std::stoul(next());
            else if (a == "--tris") tris = std::stoull(next());
            else if (a == "--width") width = std::stoul(next());
            else if (a == "--height") height = std::stoul(next());
            else if (a == "--spp") spp = std::stoul(next());
            else if (a == "--frames") frames = std::stoul(next());
            else if (a == "--linear") linear_out = next();
            else if (a == "--tonemapped") tm_out = next();
            else if (a == "--mesh") mesh = next();
            else if (a == "--tracer") tracer_name = next();
            else if (a == "--sampler") sampler_name = next();
            else if (a == "--filter") filter_name = next();
            else if (a == "--tonemapper") tm_name = next();
            else if (a == "--scene-file") scene_file = next();
            else if (a == "--save-scene") save_scene = next();
            else {
                std::fprintf(stderr, "unknown option %s\n", a.c_str());
                return 2;
            }
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 2;
        }
    }
    std::signal(SIGINT, on_sigint);
    try {
        auto t0 = std::chrono::steady_clock::now();
        raycer::Scene scene;
        if (!scene_file.empty()) {
            scene = raycer::Scene::loadFromFile(scene_file);
            scene_name = scene_file;
        } else if (scene_name == "cornell") {
            if (mesh.empty()) {
                std::string exe(argv[0]);
                size_t slash = exe.find_last_of('/');
                mesh = (slash == std::string::npos ? std::string(".") : exe.substr(0, slash)) +
                       "/../tests/golden/meshes/cornellbox.obj";
            }
            scene = raycer::Scene::createTestScene3(mesh);
            scene.general.maxPathLength = 8;
            scene.general.russianRouletteStart = 3;
        } else {
            scene = raycer::Scene::createTestScene(scene_kind(scene_name), seed, level, grid, tris);
        }
        if (!sampler_name.empty()) {
            scene.general.multiSamplerType = parse_enum<raycer::SamplerType>(sampler_name, kSamplers, "sampler");
            if (scene.general.multiSamplerType == raycer::SamplerType::RANDOM && scene.general.randomSampleCount == 0)
                scene.general.randomSampleCount = scene.general.multiSampleCountSqrt * scene.general.multiSampleCountSqrt;
        }
        if (!filter_name.empty()) {
            scene.general.multiSamplerFilterType = parse_enum<raycer::FilterType>(filter_name, kFilters, "filter");
            // the reference filters' default sizes (BoxFilter.h, TentFilter.h, LanczosSincFilter.h)
            if (scene.general.multiSamplerFilterType == raycer::FilterType::BOX) scene.general.filterRadius = 0.5;
            if (scene.general.multiSamplerFilterType == raycer::FilterType::TENT) scene.general.filterRadius = 1.0;
            if (scene.general.multiSamplerFilterType == raycer::FilterType::LANCZOS_SINC) scene.general.filterRadius = 2.0;
        }
        if (!tm_name.empty()) scene.toneMapper.type = parse_enum<raycer::ToneMapperType>(tm_name, kToneMappers, "tone mapper");
        const bool phong = tracer_name == "phong";
        if (tracer_name == "ray") scene.general.tracerType = raycer::TracerType::RAY;
        else if (tracer_name == "path") scene.general.tracerType = raycer::TracerType::PATH;
        else if (!tracer_name.empty() && !phong) throw raycer::RaycerError(RC_ERR_INVALID, "unknown tracer " + tracer_name);
        if (spp) {
            if (scene.general.multiSamplerType == raycer::SamplerType::RANDOM) {
                scene.general.randomSampleCount = spp;
            } else {
                uint64_t n = 1;
                while ((n + 1) * (n + 1) <= spp) ++n;
                scene.general.multiSampleCountSqrt = n;
            }
        }
        scene.initialize();
        if (!save_scene.empty()) scene.saveToFile(save_scene);
        // a scene file carries no film size: 1280 x 800 unless given
        uint32_t w = width ? width : (scene.defaultWidth ? scene.defaultWidth : 1280),
                 h = height ? height : (scene.defaultHeight ? scene.defaultHeight : 800);
        raycer::Film film;
        film.resize(w, h);
        raycer::HipTracer tracer(0);
        tracer.initializeKernels();
        tracer.initializeImageBuffers(w, h);
        tracer.initializeBuffers(scene);
        raycer::TracerState state;
        state.scene = &scene;
        state.film = &film;
        state.filmWidth = w;
        state.filmHeight = h;
        state.pixelCount = uint64_t(w) * h;
        double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

        if (phong) {
            // the one-light raycast of the 1080p Phong workload
            auto r0 = std::chrono::steady_clock::now();
            raycer::Image img = tracer.raycastPhong(scene);
            double render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
            bool ok = true;
            if (!linear_out.empty()) ok &= img.savePFM(linear_out);
            if (!tm_out.empty()) ok &= img.savePPM(tm_out);
            std::printf("{\"scene\": \"%s\", \"tracer\": \"phong\", \"triangles\": %zu, \"bvh_ms\": %.1f, "
                        "\"setup_ms\": %.1f, \"render_ms\": %.1f}\n",
                        scene_name.c_str(), scene.primitives.triangles.size(), scene.rootBVH.bvh.buildMilliseconds,
                        setup_ms, render_ms);
            return ok ? 0 : 1;
        }
        std::exception_ptr render_error = nullptr;
        auto r0 = std::chrono::steady_clock::now();
        std::thread render([&]() {
            try {
                for (uint32_t f = 0; f < frames && !g_interrupted; ++f) tracer.run(state, g_interrupted);
            } catch (...) {
                render_error = std::current_exception();
                g_interrupted = true;
            }
        });
        uint64_t last = ~0ull;
        while (true) {
            std::this_thread::sleep_for(std::chrono::milliseconds(100));
            uint64_t p = state.pixelsProcessed.load();
            if (p != last) {
                std::fprintf(stderr, "\rpixels %llu / %llu", (unsigned long long)p, (unsigned long long)state.pixelCount);
                last = p;
            }
            if (p >= state.pixelCount || g_interrupted) break;
        }
        render.join();
        std::fprintf(stderr, "\n");
        if (render_error) std::rethrow_exception(render_error);
        double render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
        film.setLinearImage(tracer.downloadLinearImage());
        film.setToneMappedImage(tracer.downloadImage(scene.toneMapper));
        bool ok = true;
        if (!linear_out.empty()) ok &= film.getLinearImage().savePFM(linear_out);
        if (!tm_out.empty()) ok &= film.getToneMappedImage().savePPM(tm_out);
        uint64_t samples_per_pixel = scene.general.multiSamplerType == raycer::SamplerType::RANDOM
                                         ? scene.general.randomSampleCount
                                         : scene.general.multiSampleCountSqrt * scene.general.multiSampleCountSqrt;
        double msamples = double(w) * h * samples_per_pixel * frames / (render_ms * 1e3);
        std::printf("{\"scene\": \"%s\", \"tracer\": \"%s\", \"triangles\": %zu, \"bvh_ms\": %.1f, \"setup_ms\": %.1f, "
                    "\"render_ms\": %.1f, \"frames\": %u, \"spp\": %llu, \"msamples_per_s\": %.2f, "
                    "\"interrupted\": %s}\n",
                    scene_name.c_str(), scene.general.tracerType == raycer::TracerType::RAY ? "ray" : "path",
                    scene.primitives.triangles.size(), scene.rootBVH.bvh.buildMilliseconds,
                    setup_ms, render_ms, frames, (unsigned long long)samples_per_pixel, msamples,
                    g_interrupted ? "true" : "false");
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
