/*
 * raycer_hip.h — C ABI of the MI355X (gfx950) triangle path tracer.
 *
 * This is the device boundary that replaces raycer's OpenCL GPU tracer
 * (src/OpenCL/CLTracer.h:23-90, CLScene.h:23-33).  Conventions:
 *   - every function returns int status, RC_OK == 0; rc_last_error() gives the
 *     message of the last failure on the calling thread (the counterpart of
 *     CLManager::checkError, src/OpenCL/CLManager.cpp:220-226, which throws);
 *   - handles are opaque; host pointers passed in are only read during the call;
 *   - one context per host thread, or externally synchronised; calls block until
 *     their GPU work has finished unless stated otherwise;
 *   - there is no CPU fallback: without a usable gfx950 device rc_create fails.
 */
#ifndef RAYCER_HIP_H
#define RAYCER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RC_OK = 0,
    RC_ERR_INVALID = 1,     /* bad argument or inconsistent input */
    RC_ERR_HIP = 2,         /* a HIP runtime call failed */
    RC_ERR_ALLOC = 3,       /* device allocation failed */
    RC_ERR_INTERRUPTED = 4, /* *interrupt became nonzero; the film holds the passes done */
    RC_ERR_NO_DEVICE = 5    /* no gfx950 device */
};

typedef struct rc_ctx rc_ctx;
typedef struct rc_scene rc_scene;
typedef struct rc_film rc_film;
typedef struct rc_cache rc_cache;

/* Ray (32 bytes): Ray {origin, direction, minDistance, maxDistance}
 * (src/Raytracing/Ray.h:12-29).  tmax may be +inf. */
typedef struct rc_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} rc_ray;

/* Closest hit (16 bytes): distance, barycentrics and the ORIGINAL triangle
 * index (-1 = miss).  Replaces Intersection {wasFound, distance, primitive}
 * (src/Raytracing/Intersection.h:17-27). */
typedef struct rc_hit {
    float t;
    float u;
    float v;
    int32_t tri;
} rc_hit;

/* FlatBVHNode exactly as the reference lays it out in host memory
 * (src/Raytracing/Primitives/FlatBVH.h:38-53): preorder, left child at i + 1,
 * right child at i + right_offset, right_offset == 0 marks a leaf whose
 * triangles are ordered_ids[start_offset .. start_offset + primitive_count). */
typedef struct rc_flat_node {
    double aabb_min[3];
    double aabb_max[3];
    int64_t right_offset;
    uint64_t start_offset;
    uint64_t primitive_count;
} rc_flat_node;

/* Material (Material.h:16-120 subset): Lambert reflectance, emittance,
 * Phong ambient/specular terms; flags bit0 isEmissive, bit1 nonShadowing. */
typedef struct rc_material {
    float diffuse[3];
    float emittance[3];
    float ambient[3];
    float specular[3];
    float shininess;
    uint32_t flags;
    uint32_t pad[2];
} rc_material;

enum { RC_MAT_EMISSIVE = 1u, RC_MAT_NON_SHADOWING = 2u };

/* Camera: position, EulerAngle (pitch, yaw, roll in degrees) and the
 * horizontal field of view in degrees (src/Raytracing/Camera.cpp:29-34,160,308-331). */
typedef struct rc_camera {
    float position[3];
    float pitch, yaw, roll;
    float fov;
} rc_camera;

/* Filter::getFilter's FilterType (src/Rendering/Filters/Filter.h:12) with the
 * reference's shapes: box and tent take filter_radius (BoxFilter.cpp:15-37,
 * TentFilter.cpp:12-44); bell has radius 1.5 (BellFilter.cpp:12-28); gaussian
 * takes the standard deviation filter_param[0] and has radius 7.43384 sigma / 2
 * (GaussianFilter.cpp:15-23); mitchell takes B = filter_param[0], C =
 * filter_param[1] and has radius 2 (MitchellFilter.cpp:12-37; the reference's
 * default is B = C = 1/3); lanczos_sinc takes the integer radius filter_radius
 * (LanczosSincFilter.cpp:12-40; the reference's default is 2). */
enum {
    RC_FILTER_BOX = 0, RC_FILTER_TENT = 1, RC_FILTER_BELL = 2, RC_FILTER_GAUSSIAN = 3,
    RC_FILTER_MITCHELL = 4, RC_FILTER_LANCZOS_SINC = 5
};
/* Sampler::getSampler's SamplerType (src/Rendering/Samplers/Sampler.h:23) as
 * Tracer::generateMultiSamples uses it (Tracer.cpp:115-144): every sampler but
 * random needs spp = n*n and draws sample s = y*n + x of the n x n grid; one
 * sample is the pixel centre with weight 1.  center: (0.5, 0.5)
 * (CenterSampler.cpp:19-33); regular: ((x + 0.5)/n, (y + 0.5)/n)
 * (RegularSampler.cpp:19-34); cmj: correlated multi-jittered with one 64-bit
 * permutation per pixel and frame (CMJSampler.cpp:12-77); poisson_disc: the
 * point (0, 0), which is what PoissonDiscSampler::getSample2D returns
 * (PoissonDiscSampler.cpp:22-32): every sample then sits at offset (-r, -r)
 * from the pixel, r the filter radius, as in the reference. */
enum {
    RC_SAMPLER_JITTERED = 0, RC_SAMPLER_RANDOM = 1, RC_SAMPLER_CENTER = 2, RC_SAMPLER_REGULAR = 3,
    RC_SAMPLER_CMJ = 4, RC_SAMPLER_POISSON_DISC = 5
};
enum { RC_EST_REFERENCE = 0, RC_EST_LAMBERT = 1 };

/* Render settings: Scene::General names where they exist (Scene.h:89-131). */
typedef struct rc_render_cfg {
    uint32_t width, height;
    uint32_t spp;               /* samples per pixel; all samplers but random need spp = n*n (multiSampleCountSqrt = n) */
    uint32_t sampler;           /* RC_SAMPLER_* (multiSamplerType) */
    uint32_t max_path_length;   /* maxPathLength: rays per path including the camera ray */
    uint32_t rr_start;          /* Russian roulette at path vertex k >= rr_start; 0 = off */
    uint32_t filter;            /* RC_FILTER_* (multiSamplerFilterType) */
    uint32_t estimator;         /* RC_EST_REFERENCE: throughput *= 2 rho |n.w|; RC_EST_LAMBERT: *= rho */
    uint32_t cache_vertex;      /* with a radiance cache: path vertex that ends in the cache (0 = off) */
    float filter_radius;        /* box, tent: radius; lanczos_sinc: integer radius */
    float filter_param[2];      /* gaussian: standard deviation; mitchell: B, C */
    float ray_start_offset;     /* rayStartOffset */
    float sky[3];               /* radiance on a miss (the reference returns black) */
    float sun_dir[3];           /* unit vector towards a directional sun */
    float sun_irradiance[3];    /* 0 = no sun */
    uint32_t frame;             /* RNG stream for progressive frames */
    uint64_t seed;
} rc_render_cfg;

/* Tone mappers (ToneMapper::getToneMapper, src/Rendering/ToneMappers/ToneMapper.cpp:13-23).
 * reinhard: ReinhardToneMapper.cpp:18-86; linear: c * 2^exposure
 * (LinearToneMapper.cpp:12-35); simple: c' / (1 + c') with c' = c * 2^exposure
 * (SimpleToneMapper.cpp:12-37); passthrough: the linear image as it is
 * (PassthroughToneMapper.cpp:12-22).  Reinhard is 0 so that a zeroed struct
 * selects it. */
enum { RC_TONEMAP_REINHARD = 0, RC_TONEMAP_LINEAR = 1, RC_TONEMAP_SIMPLE = 2, RC_TONEMAP_PASSTHROUGH = 3 };

/* Tone-mapper settings (Scene::ToneMapper, Scene.h:137-160). */
typedef struct rc_tonemap {
    float key;                  /* reinhard */
    float gamma;
    uint32_t clamp;             /* shouldClamp */
    uint32_t apply_gamma;       /* applyGamma */
    uint32_t fast_pow;          /* 1 = MathUtils::fastPow for gamma and 2^exposure (reference), 0 = pow */
    uint32_t type;              /* RC_TONEMAP_* */
    float exposure;             /* linear, simple */
    uint32_t enable_averaging;  /* reinhard: moving average of the maximum luminance, kept by the film */
    float averaging_alpha;      /* in (0, 1] */
    uint32_t pad;
} rc_tonemap;

/* Point light + ambient for the Phong raycast (Lights.h PointLight/AmbientLight). */
typedef struct rc_point_light {
    float position[3];
    float color[3];
    float intensity;
    float max_distance;
    float attenuation;
    float ambient_color[3];
    float ambient_intensity;
} rc_point_light;

/* Launch tuning: threads per block of the ray-query kernels, pass size in
 * samples per pixel (0, the default: each frame takes as many samples per pass
 * as keep a pass within 2^26 paths, 2^21 for rc_raytrace; else 1..64), grid
 * limit of the ray-query kernels in blocks per CU, the traversal stack
 * entries held on chip (registers and LDS) before a ray is handed to the
 * deep-stack kernel, which walks it again with a stack in global memory
 * (0, the default: all the build holds; 1..that depth: fewer, which sends
 * more rays down the deep-stack path and changes no result), and the
 * traversal steps a path tracer's extension ray takes in the first kernel
 * of a bounce before it is parked and finished by a second kernel whose
 * waves hold only such long traversals (a wave with at least 8 lanes still
 * traversing goes on, up to twice that; 0: never park; 1..1024; the results
 * are the same either way).  Defaults 256, 0, 64, 0, RC_PARK_STEPS_DEFAULT. */
#define RC_PARK_STEPS_DEFAULT 20
typedef struct rc_tuning {
    uint32_t block;
    uint32_t spp_per_pass;
    uint32_t blocks_per_cu;
    uint32_t stack_limit;
    uint32_t park_steps;
} rc_tuning;

typedef struct rc_stats {
    uint64_t extension_rays;     /* closest-hit rays traced by the last render */
    uint64_t shadow_rays;        /* any-hit rays traced by the last render */
    uint64_t samples;            /* camera samples of the last render */
    uint64_t node_records;       /* 32-byte node records fetched (counting runs only) */
    uint64_t tri_tests;          /* 48-byte triangle records tested (counting runs only) */
    uint64_t overflow_rays;      /* rays re-walked by the deep-stack kernel, summed over all launches of the call */
    uint64_t extend_launches;
    double extend_ms;            /* sum of closest-hit kernel durations (profiling on) */
    double shadow_ms;
    double shade_ms;
    double other_ms;
    double render_ms;            /* wall time of the last render call */
    uint64_t shadow_launches;
    double cache_ms;             /* radiance-cache inference kernels (profiling on) */
    double train_ms;             /* optimizer steps after the last cached frame (GPU time on the cache's stream) */
    uint64_t cache_queries;      /* paths that ended into the cache */
    uint64_t train_samples;      /* training samples of the last cached render */
    double graph_ms;             /* last batch call: capture and instantiation or update (0: graph relaunched) */
    uint64_t graph_relaunched;   /* last batch call: 1 when it relaunched the executable graph as it was */
    uint64_t parked_rays;        /* extension traversals parked and finished by the resume kernel (rc_tuning.park_steps) */
} rc_stats;

/* ------------------------------------------------------------ context */
int rc_last_error(char* buf, size_t len);
/* CLManager::initialize + CLTracer::initializeKernels (CLTracer.cpp:45-71). */
int rc_create(int device, rc_ctx** out);
int rc_destroy(rc_ctx* ctx);
int rc_device_name(rc_ctx* ctx, char* buf, size_t len, int* cu_count);
/* profiling: 1 = time every launch with HIP events on the context's stream */
int rc_set_profiling(rc_ctx* ctx, int enable);
/* counting: 1 = use the traversal variant that counts node/triangle fetches */
int rc_set_counting(rc_ctx* ctx, int enable);
/* How a ray-triangle test is decided (Triangle::intersect, Triangle.cpp:51-121,
 * and its FP32 OpenCL form, data/opencl/intersections.cl:27-68):
 * RC_HIT_FP32 (the default): Möller–Trumbore in FP32 with inclusive
 * barycentric bounds and det == 0 rejected; RC_HIT_FP64: an FP32 filter with
 * widened bounds, then the candidates decided in FP64 on the same FP32 inputs
 * (distances to FP64 accuracy at grazing angles).  Applies to every traversal
 * of the context. */
enum { RC_HIT_FP32 = 0, RC_HIT_FP64 = 1 };
int rc_set_hit_decision(rc_ctx* ctx, int mode);
int rc_set_tuning(rc_ctx* ctx, const rc_tuning* t);
int rc_get_stats(rc_ctx* ctx, rc_stats* out);
int rc_synchronize(rc_ctx* ctx);
/* Achievable-bandwidth probe (no reference counterpart): a float4 streaming
 * copy of `bytes` bytes, `iters` times, at four grid sizes; returns the best
 * read+write GB/s. */
int rc_copy_bandwidth(rc_ctx* ctx, size_t bytes, int iters, double* gbps);
/* Read-bandwidth probe of one cache level (no reference counterpart): a
 * working set of `bytes` read `iters` times per launch by threads that
 * re-read the same addresses, so a set that fits L2 (4 MiB per XCD) or the
 * Infinity Cache (256 MiB) is served from it; returns the best read GB/s over
 * three grid sizes. */
int rc_read_bandwidth(rc_ctx* ctx, size_t bytes, int iters, double* gbps);
/* Scattered-read probe (no reference counterpart): `records` threads each
 * read one 64-byte record (a wide BVH node's size) at distinct pseudo-random
 * records of a buffer of the largest power-of-two record count within
 * `bytes`, `iters` launches; returns the requested GB/s.  Its rocprofv3
 * FETCH_SIZE against the records' 64 bytes shows how the counter scales for
 * scattered 64-byte reads. */
int rc_gather_bandwidth(rc_ctx* ctx, size_t bytes, size_t records, int iters, double* gbps);

/* ------------------------------------------------------------ scene */
/* CLTracer::initializeBuffers / CLScene::readSceneFull (CLScene.cpp:72-263):
 * the host FlatBVH (nodes + orderedPrimitiveIds) is validated, collapsed into
 * 4-wide nodes of four 32-byte child records (leaves of 32 or more triangles
 * are split over extra wide nodes) and 48-byte triangle records, and uploaded.
 * vertices: ntri * 9 floats in original triangle order (ntri <= 2^26);
 * tri_material: ntri. */
int rc_scene_upload(rc_ctx* ctx, const float* vertices, size_t ntri, const uint32_t* tri_material,
                    const rc_material* materials, size_t nmat, const rc_flat_node* nodes, size_t nnodes,
                    const uint64_t* ordered_ids, rc_scene** out);
int rc_scene_free(rc_scene* scene);
int rc_scene_info(const rc_scene* scene, size_t* ntri, size_t* nrecords, size_t* device_bytes);

/* ------------------------------------------------------------ ray queries */
/* FlatBVH::intersect + Triangle::intersect for a ray batch (closest hit).
 * Host arrays in, host arrays out. */
int rc_trace_closest(rc_ctx* ctx, const rc_scene* scene, const rc_ray* rays, size_t n, rc_hit* hits);
/* Any hit (fastOcclusion, FlatBVH.cpp:49-50): occluded[i] = 1 if a triangle
 * whose material is not nonShadowing lies at t in [tmin, tmax]. */
int rc_trace_any(rc_ctx* ctx, const rc_scene* scene, const rc_ray* rays, size_t n, uint8_t* occluded);
/* Same queries on device-resident buffers (no copies, asynchronous on the
 * context's stream). */
int rc_trace_closest_device(rc_ctx* ctx, const rc_scene* scene, const rc_ray* d_rays, size_t n, rc_hit* d_hits);
int rc_trace_any_device(rc_ctx* ctx, const rc_scene* scene, const rc_ray* d_rays, size_t n, uint8_t* d_occluded);
/* rc_trace_closest_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory while
 * they run (code under HIP graph capture, tight serving loops).  The kernels'
 * scratch (overflow list, deep stack, counters) is a fixed 6.25 MiB array in
 * device memory that every GPU the library runs on holds until the process
 * ends; at most 2^20 rays per call (larger batches are rejected, split them).
 * Calls that overlap on one GPU share that array and race: serialise them. */
int rc_trace_closest_device_noalloc(rc_ctx* ctx, const rc_scene* scene, const rc_ray* d_rays, size_t n,
                                    rc_hit* d_hits);
/* Device buffers owned by the context's allocator, for the *_device calls. */
int rc_device_alloc(rc_ctx* ctx, size_t bytes, void** dptr);
int rc_device_free(rc_ctx* ctx, void* dptr);
/* Page-locked host memory (hipHostMalloc): the entry points that return
 * images or hit arrays to the host copy into such buffers at the host link's
 * full rate, into pageable memory through a staging copy. */
int rc_host_alloc(rc_ctx* ctx, size_t bytes, void** hptr);
int rc_host_free(rc_ctx* ctx, void* hptr);
int rc_memcpy_h2d(rc_ctx* ctx, void* dst, const void* src, size_t bytes);
int rc_memcpy_d2h(rc_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------ film */
/* Film (src/Rendering/Film.h:23-52): per pixel FP32 Kahan-compensated sums of
 * w*color and w.  Accumulates across rc_render calls until rc_film_clear. */
int rc_film_create(rc_ctx* ctx, uint32_t width, uint32_t height, rc_film** out);
int rc_film_free(rc_film* film);
int rc_film_clear(rc_ctx* ctx, rc_film* film);
/* linear image sum/weight as RGBA float (Film::generateToneMappedImage first half) */
int rc_film_download_linear(rc_ctx* ctx, const rc_film* film, float* rgba);
/* raw accumulators: per pixel sum.rgb, weight (4 floats) */
int rc_film_download_sums(rc_ctx* ctx, const rc_film* film, float* sums);
/* Film::generateToneMappedImage's tone map on the device (Film.cpp:61-69) with
 * tm->type's mapper.  For reinhard, log_avg / max_lum (either may be NULL) get
 * the log-average luminance and the maximum luminance used (after the moving
 * average when tm->enable_averaging; the film keeps the average, starting at
 * 1, as ReinhardToneMapper's maxLuminanceAverage does, and rc_film_clear does
 * not reset it); other mappers set them to 0. */
int rc_film_tonemap(rc_ctx* ctx, const rc_film* film, const rc_tonemap* tm, float* rgba, double* log_avg,
                    double* max_lum);

/* ------------------------------------------------------------ rendering */
/* Tracer::run + PathTracer (Tracer.cpp:58-113, PathTracer.cpp:13-75) as a
 * wavefront path tracer.  interrupt (may be NULL) is polled between passes;
 * pixels_processed (may be NULL) is advanced as passes finish. */
int rc_render(rc_ctx* ctx, const rc_scene* scene, rc_film* film, const rc_camera* cam, const rc_render_cfg* cfg,
              volatile const int* interrupt, uint64_t* pixels_processed);
/* Progress callback, called on the calling thread after every pass with the
 * pixels finished so far; a nonzero return interrupts the render
 * (TracerState::pixelsProcessed and the interrupted flag of Tracer::run,
 * Tracer.cpp:82-95). */
typedef int (*rc_progress_fn)(void* user, uint64_t pixels_processed);
int rc_render_progress(rc_ctx* ctx, const rc_scene* scene, rc_film* film, const rc_camera* cam,
                       const rc_render_cfg* cfg, rc_progress_fn progress, void* user);
/* Primary ray + Phong + one point-light hard shadow (Raytracer.cpp:312-407,
 * 524-551).  The camera ray of pixel (x, y) passes through (x + u0, y + u1),
 * (u0, u1) the jitter draw of that pixel.  hit_ids / t / rgba are host arrays
 * of width*height (rgba: 4 floats per pixel); any may be NULL. */
int rc_raycast_phong(rc_ctx* ctx, const rc_scene* scene, const rc_camera* cam, uint32_t width, uint32_t height,
                     uint64_t seed, float ray_start_offset, const rc_point_light* light, int32_t* hit_ids,
                     float* t, float* rgba);

/* ------------------------------------------------------------ Whitted ray tracer
 * Raytracer::traceRecursive (src/Raytracing/Tracers/Raytracer.cpp:27-310) with
 * calculateLightColor, Phong shading, ambient occlusion and hard or area-light
 * shadows (:312-597), run as Tracer::run runs it: the camera samples of
 * rc_render_cfg (sampler, filter, spp) accumulate into the film.  The settings
 * are Scene::General's ray-tracer fields, Scene::Lights, Scene::SimpleFog
 * (Scene.h:89-200, Lights.h) and Material's ray-tracing fields (Material.h:20-66);
 * the scene's rc_material gives the ambient, diffuse and specular reflectances
 * and the shininess.  Textures, normal maps and volumetric fog are not
 * supported. */
enum { RC_LIGHT_DIRECTIONAL = 0, RC_LIGHT_POINT = 1, RC_LIGHT_SPOT = 2 };

typedef struct rc_light {
    uint32_t type;              /* RC_LIGHT_* */
    float color[3];
    float intensity;
    float position[3];          /* point, spot */
    float direction[3];         /* directional: direction of travel; spot: axis */
    float max_distance;         /* point, spot: 1 - min(1, d / max_distance)^attenuation */
    float attenuation;
    uint32_t area_light;        /* point, spot: soft shadows from a disc (enableAreaLight) */
    uint32_t area_sampler;      /* RC_SAMPLER_* (areaLightSamplerType) */
    uint32_t area_n;            /* areaLightSampleCountSqrt, 1..16 */
    float area_radius;
    float angle;                /* spot: cone angle in degrees */
    float side_attenuation;     /* spot */
    uint32_t pad;
} rc_light;

/* Material's ray-tracing fields, one per scene material, in the scene's order. */
typedef struct rc_ray_material {
    float ray_reflectance;
    float ray_transmittance;
    float refractive_index;
    uint32_t fresnel_reflection;
    uint32_t skip_lighting;     /* the colour is the diffuse reflectance */
    uint32_t refl_sampler;      /* rayReflectanceGlossinessSamplerType */
    uint32_t refl_n;            /* rayReflectanceGlossinessSampleCountSqrt, 0 = mirror, at most 16 */
    float refl_glossiness;      /* rayReflectanceGlossiness (hemisphere exponent) */
    uint32_t trans_sampler;
    uint32_t trans_n;
    float trans_glossiness;
    uint32_t enable_attenuation;    /* enableRayTransmissionAttenuation */
    float attenuation_factor;
    float attenuation_color[3];
} rc_ray_material;

typedef struct rc_ray_settings {
    float background[3];        /* backgroundColor */
    uint32_t max_ray_iterations;    /* at most 16 */
    float ambient_color[3];
    float ambient_intensity;
    uint32_t ao_enabled;        /* enableAmbientOcclusion */
    uint32_t ao_sampler;
    uint32_t ao_n;              /* ambientOcclusionSampleCountSqrt, 1..16 when enabled */
    float ao_max_distance;
    float ao_distribution;
    uint32_t visualize_depth;
    float visualize_depth_max;
    uint32_t fog_enabled;       /* SimpleFog */
    float fog_color[3];
    float fog_distance;
    float fog_steepness;
    uint32_t fog_height_dispersion;
    float fog_height;
    float fog_height_steepness;
} rc_ray_settings;

/* Renders one frame with the Whitted tracer into the film (accumulating, as
 * rc_render does).  mats has one entry per scene material (nmat must equal the
 * scene's material count); at most 64 lights.  Each pixel's rays form a tree
 * (reflection and transmission children, glossy ones n*n per hit) traced level
 * by level; a level that would outgrow the buffers is processed in parts. */
int rc_raytrace(rc_ctx* ctx, const rc_scene* scene, rc_film* film, const rc_camera* cam, const rc_render_cfg* cfg,
                const rc_ray_settings* settings, const rc_ray_material* mats, uint32_t nmat, const rc_light* lights,
                uint32_t nlights);

/* ------------------------------------------------------------ radiance cache
 * No counterpart in the reference.  An MLP (81 -> 64 x 4 ReLU -> 3 exp,
 * 17,923 parameters) with BF16 weights on MFMA and FP32 accumulation, trained
 * online with Adam on radiance samples the path tracer produces; a render with
 * cfg.cache_vertex = k ends paths at vertex k into the cache.
 * Parameters are FP32 masters, packed per layer as W (out x in, row-major) then b.
 * A query is 12 floats: position (scene units), direction, normal, albedo;
 * the position is scaled to [0,1]^3 by the cache's bounds before the 12-frequency
 * sin/cos encoding. */
enum { RC_CACHE_PARAMS = 17923, RC_CACHE_QUERY_FLOATS = 12 };
typedef struct rc_cache_cfg {
    float lr;            /* Adam learning rate (1e-3) */
    float beta1, beta2;  /* 0.9, 0.99 */
    float adam_eps;      /* 1e-8 */
    float loss_eps;      /* relative L2 denominator epsilon (1e-2) */
    uint32_t seed;       /* weight initialisation */
    uint32_t max_batch;  /* largest training batch (65,536) */
    uint32_t train_stride;  /* renderer: one training path per train_stride slots; 0 = adaptive, steering the
                               training samples per frame towards 0.9 max_batch */
} rc_cache_cfg;

int rc_cache_create(rc_ctx* ctx, const rc_cache_cfg* cfg, const float bounds_min[3], const float bounds_max[3],
                    rc_cache** out);
int rc_cache_free(rc_cache* cache);
int rc_cache_set_params(rc_cache* cache, const float* params);
int rc_cache_get_params(rc_cache* cache, float* params);
/* queries: n * 12 floats (host); radiance: n * 3 floats (host) */
int rc_cache_infer(rc_ctx* ctx, rc_cache* cache, const float* queries, size_t n, float* radiance);
/* One Adam step on a batch (host arrays); loss (may be NULL) gets the batch loss. */
int rc_cache_train_step(rc_ctx* ctx, rc_cache* cache, const float* queries, const float* targets, size_t n,
                        float* loss);
/* Gradient of the batch loss without an update (for checks): grad has 17,923 floats. */
int rc_cache_gradient(rc_ctx* ctx, rc_cache* cache, const float* queries, const float* targets, size_t n,
                      float* grad, float* loss);
/* rc_render with paths ending into the cache at cfg->cache_vertex (1 <=
 * cache_vertex < max_path_length - 1): a path that reaches that vertex without
 * emission traces its sun next-event ray as usual, adds throughput x
 * cache(query) for the indirect light and stops.  Paths of one slot in
 * train_stride run to the end instead and give up to max_batch training
 * samples (target = radiance gathered after the vertex's next-event ray /
 * throughput at the vertex);
 * with train = k > 0, k (at most 16) Adam steps run after the frame, each on
 * an interleaved 1/k of those samples, and loss (may be NULL) gets their mean
 * loss. */
int rc_render_cached(rc_ctx* ctx, const rc_scene* scene, rc_film* film, const rc_camera* cam,
                     const rc_render_cfg* cfg, rc_cache* cache, int train, float* loss);

/* Many scenes in one captured HIP graph: scenes[i] is rendered into
 * films[i] with cams[i] and cfgs[i].  The executable graph is kept: a call
 * with the same batch (scenes, films, cameras, configurations, plans, context
 * settings) relaunches it without capturing; a changed batch is captured
 * again and the executable graph updated in place (instantiated anew when its
 * shape changed).  rc_stats.graph_ms / graph_relaunched tell which. */
int rc_render_batch(rc_ctx* ctx, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                    const rc_render_cfg* cfgs, size_t nscenes, const rc_tuning* tuning);
/* The same with samples per pass chosen per scene (spp_per_pass[i] in [1, 64]). */
int rc_render_batch_plan(rc_ctx* ctx, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                         const rc_render_cfg* cfgs, size_t nscenes, const uint32_t* spp_per_pass);
/* Launch plan of one scene of a batch: the wavefront queue length (samples
 * per pass: a pass's queue holds pixels x spp_per_pass paths) and the block
 * shapes of its traversal kernels. */
typedef struct rc_batch_plan {
    uint32_t spp_per_pass;   /* 1..64 */
    uint32_t ext_block;      /* threads per block of the extension kernel: 128 or 256 (0 = 256) */
    uint32_t shadow_block;   /* of the shadow kernel: 64, 128 or 256 (0 = 128) */
    uint32_t pad;
} rc_batch_plan;
/* rc_render_batch with one plan per scene and the frames spread over
 * `branches` (1..8) concurrent graph branches, each with its own wavefront
 * buffers: frames are dealt longest first (estimated cost pixels x spp x
 * log2 triangles) to the least loaded branch, and run in that order. */
int rc_render_batch_tuned(rc_ctx* ctx, const rc_scene* const* scenes, rc_film* const* films, const rc_camera* cams,
                          const rc_render_cfg* cfgs, size_t nscenes, const rc_batch_plan* plans, uint32_t branches);
/* Autotuner for rc_render_batch_plan: per size class (ceil(log2(pixels))) it
 * renders one scene of the class with 1, 2, 4, ... 64 samples per pass (at
 * most the frame's spp and 2^25 paths per pass) and keeps the fastest.  With a
 * cache_path, choices are read from and added to that JSON file, keyed
 * "<device>/<class>".  Writes one entry per scene into spp_per_pass. */
int rc_autotune_batch(rc_ctx* ctx, const rc_scene* const* scenes, const rc_camera* cams, const rc_render_cfg* cfgs,
                      size_t nscenes, const char* cache_path, uint32_t* spp_per_pass);
/* Autotuner for rc_render_batch_tuned: per size class, the samples per pass
 * as rc_autotune_batch chooses them, then the block shapes (extension 128 or
 * 256, shadow 64, 128 or 256 threads) timed at those samples per pass; then,
 * for the whole batch, the branch count (1, 2, 4, 8) timed as relaunches of
 * the captured graph.  A candidate replaces the default only when it is 2%
 * faster.  Choices are read from and added to the JSON cache (keys
 * "<device>/<class>/spp|ext|shadow" and "<device>/branches/<nscenes>").
 * Writes plans[nscenes] and *branches. */
int rc_autotune_batch_plan(rc_ctx* ctx, const rc_scene* const* scenes, const rc_camera* cams,
                           const rc_render_cfg* cfgs, size_t nscenes, const char* cache_path, rc_batch_plan* plans,
                           uint32_t* branches);

#ifdef __cplusplus
}
#endif
#endif
