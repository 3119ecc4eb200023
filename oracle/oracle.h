/*
 * FP64 CPU oracle for the raycer triangle path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are compared against this library: only the test suite,
 * the smoke check and the cpu_baseline leg of bench.py load it.  The product
 * library never links, calls or falls back to it.
 *
 * It restates, in double precision, the reference CPU path of mikoro/raycer:
 *   - FlatBVH::build / calculateSAHSplit / calculateSAHScore /
 *     calculateMedianPoint      (src/Raytracing/Primitives/FlatBVH.cpp:89-342)
 *   - FlatBVH::intersect        (FlatBVH.cpp:22-75, unordered, no t-culling)
 *   - AABB::intersects          (src/Raytracing/AABB.cpp:56-80)
 *   - Triangle::initialize / intersect (src/Raytracing/Primitives/Triangle.cpp:18-121)
 *   - Camera::getRay (perspective) + setImagePlaneSize + imagePlaneDistance
 *                               (src/Raytracing/Camera.cpp:29-34,160,279-331)
 *   - EulerAngle::getDirection  (src/Math/EulerAngle.cpp:77-87)
 *   - ONB::fromNormal           (src/Math/ONB.cpp:32-39)
 *   - Tracer::generateMultiSamples + JitteredSampler (Tracer.cpp:115-144,
 *                               JitteredSampler.cpp:21-34)
 *   - mapToHemisphere           (src/Rendering/Samplers/Sampler.cpp:50-64)
 *   - PathTracer::traceRecursive (src/Raytracing/Tracers/PathTracer.cpp:25-75)
 *   - the six filters           (src/Rendering/Filters/)
 *   - the six samplers as generateMultiSamples uses them, CMJ's permute
 *                               (src/Rendering/Samplers/, CMJSampler.cpp:12-77)
 *   - Film::addSample / generateToneMappedImage (src/Rendering/Film.cpp:54-69)
 *   - ReinhardToneMapper::apply (ReinhardToneMapper.cpp:18-86), Linear,
 *     Simple and Passthrough tone mappers (src/Rendering/ToneMappers/)
 *   - MathUtils::fastPow        (src/Math/MathUtils.cpp:56-68)
 *   - Raytracer Phong + point-light shadow (Raytracer.cpp:312-407, 524-551)
 *   - the Whitted ray tracer: Raytracer::traceRecursive with reflection,
 *     refraction, Fresnel, glossy sampling, attenuation, simple fog, ambient
 *     occlusion and directional / point / spot lights with area-light shadows
 *     (Raytracer.cpp:27-597)
 * plus the estimator terms the reference does not have (Russian roulette, sky
 * radiance on miss, directional-sun next-event estimation), the counter-based
 * RNG (Philox4x32-10, Salmon et al., SC'11) and an FP64 MLP (forward, backward,
 * Adam) for the radiance cache.  Every function is plain C ABI for ctypes.
 */
#ifndef RAYCER_ORACLE_H
#define RAYCER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- math pins */
double or_fastpow(double a, double b);
void or_normalize(const double in[3], double out[3]);
void or_cross(const double a[3], const double b[3], double out[3]);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* uniform in [0,1): (x >> 8) * 2^-24, the exact value the GPU computes in FP32 */
double or_uniform(uint32_t x);

/* ---------------------------------------------------------------- BVH */
typedef struct or_bvh or_bvh;
/* Flat node exactly as FlatBVHNode (FlatBVH.h:38-53): 72 bytes. */
typedef struct or_flat_node {
    double aabb_min[3];
    double aabb_max[3];
    int64_t right_offset;   /* 0 => leaf */
    uint64_t start_offset;
    uint64_t primitive_count;
} or_flat_node;

/* vertices: ntri * 9 floats (v0.xyz v1.xyz v2.xyz), original triangle order. */
int or_bvh_build(const float* vertices, size_t ntri, uint64_t max_leaf_size, or_bvh** out);
size_t or_bvh_node_count(const or_bvh* b);
void or_bvh_nodes(const or_bvh* b, or_flat_node* out);
void or_bvh_ordered_ids(const or_bvh* b, uint64_t* out);
void or_bvh_free(or_bvh* b);

/* Möller–Trumbore exactly as Triangle::intersect, without the
 * "t > intersection.distance" check.  Returns 1 on hit. */
int or_intersect_triangle(const double v[9], const double org[3], const double dir[3],
                          double min_distance, double max_distance,
                          double* t, double* u, double* v_out);

/* rays: n * 8 floats (org.xyz, tmin, dir.xyz, tmax) */
/* Reference traversal (unordered, no culling, last-tested wins on ties). */
void or_trace_closest_ref(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                          double* t_out, int64_t* id_out, int nthreads);
/* Ordered, t-culled FP64 traversal with near-tie diagnostics.
 * flags bit0: near tie (another candidate within rel_tie of the best t, or a
 * candidate inside the barycentric band closer than the best);
 * bit1: best hit inside the barycentric band (grazes an edge);
 * bit2: |det| of the best hit inside the FP32 rounding band.
 * cond_out (may be NULL): conditioning |det| / (|e1| |e2| |d|) of the best
 * hit (1 for a miss), which bounds the FP32 error of its decision. */
void or_trace_closest_diag(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                           double rel_tie, double bary_band,
                           double* t_out, int64_t* id_out, uint32_t* flags_out, double* cond_out, int nthreads);
/* Brute force over every triangle (no BVH): the BVH-free pin. */
void or_trace_closest_brute(const float* vertices, size_t ntri, const float* rays, size_t n,
                            double* t_out, int64_t* id_out, int nthreads);
/* any-hit: occluded[i] = 1 if some triangle is hit with t in [tmin, tmax]
 * (FlatBVH::intersect with fastOcclusion, FlatBVH.cpp:49-50). flags bit0 =
 * decision within the barycentric band. */
void or_trace_any(const or_bvh* b, const float* vertices, const float* rays, size_t n,
                  double bary_band, uint8_t* occluded_out, uint8_t* flags_out, int nthreads);

/* ---------------------------------------------------------------- camera */
/* cam: position.xyz, pitch, yaw, roll (degrees), fov (horizontal, degrees) */
void or_camera_ray(const double cam[7], uint32_t width, uint32_t height,
                   double px, double py, double org[3], double dir[3]);

/* ---------------------------------------------------------------- render */
typedef struct or_render_cfg {
    uint32_t width, height;
    uint32_t spp;               /* samples per pixel (jittered: n*n) */
    uint32_t sampler;           /* 0 jittered, 1 random, 2 center, 3 regular, 4 CMJ, 5 Poisson disc */
    uint32_t max_path_length;   /* maxPathLength */
    uint32_t rr_start;          /* RR at path vertex k >= rr_start (0 = off) */
    uint32_t filter;            /* 0 box, 1 tent, 2 bell, 3 gaussian, 4 mitchell, 5 lanczos sinc */
    uint32_t estimator;         /* 0 reference 2*rho*|cos|, 1 lambert rho */
    uint32_t traversal;         /* 0 reference unordered BVH, 1 ordered culled BVH */
    double filter_radius;
    double ray_start_offset;
    double sky[3];
    double sun_dir[3];          /* unit vector towards the sun */
    double sun_irradiance[3];   /* zero => no sun */
    uint64_t seed;
    uint32_t frame;
    uint32_t cache_vertex;      /* unused by the oracle renderer (no cache) */
    double filter_param[2];     /* gaussian: standard deviation; mitchell: B, C */
} or_render_cfg;

/* materials: nmat * 16 doubles:
 *   diffuse.rgb, emittance.rgb, ambient.rgb, specular.rgb, shininess, flags, pad, pad
 *   flags bit0 emissive, bit1 nonShadowing */
enum { OR_MAT_STRIDE = 16 };

/* Path-trace the pixels listed in pixel_ids (or all if pixel_ids == NULL).
 * stats_out: per listed pixel 14 doubles:
 *   sum_w, sum_w2, sum_wL[3], sum_w2L[3], sum_w2L2[3], extension_rays,
 *   node visits and triangle tests of those rays (the reference's unordered
 *   traversal, cfg->traversal 0; zero with the ordered one)
 * Returns 0, or nonzero when *interrupt becomes nonzero. */
int or_render(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
              const double* materials, size_t nmat, const double cam[7],
              const or_render_cfg* cfg, const uint32_t* pixel_ids, size_t npix,
              double* stats_out, int nthreads, volatile int* interrupt);

/* Split-tail estimate of the variance a radiance cache at cfg->cache_vertex
 * (> 0) could remove, over the listed pixels: per pixel 7 doubles
 *   sum_w, sum_w La[3], sum_w^2 (La - Lb)^2 [3]
 * where La and Lb are two paths of the same sample sharing everything up to
 * that vertex's next-event ray.  Returns nonzero for cache_vertex 0. */
int or_render_split_tail(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                         const double* materials, const double cam[7], const or_render_cfg* cfg,
                         const uint32_t* pixel_ids, size_t npix, double* out, int nthreads);

/* One path (pixel, sample) with a trace of every vertex, for replay checks.
 * vertices_out: max_path_length * 8 doubles (hit t, tri id, thr.rgb, L.rgb).
 * Returns number of traced rays. */
int or_trace_path(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                  const double* materials, const double cam[7], const or_render_cfg* cfg,
                  uint32_t pixel, uint32_t sample, double L_out[3], double* weight_out,
                  double* vertices_out);

/* ---------------------------------------------------------------- Whitted ray tracer */
/* Same layouts as the product's rc_light / rc_ray_material / rc_ray_settings. */
typedef struct or_light {
    uint32_t type;              /* 0 directional, 1 point, 2 spot */
    float color[3];
    float intensity;
    float position[3];
    float direction[3];
    float max_distance;
    float attenuation;
    uint32_t area_light;
    uint32_t area_sampler;
    uint32_t area_n;
    float area_radius;
    float angle;
    float side_attenuation;
    uint32_t pad;
} or_light;
typedef struct or_ray_material {
    float ray_reflectance, ray_transmittance, refractive_index;
    uint32_t fresnel_reflection, skip_lighting;
    uint32_t refl_sampler, refl_n;
    float refl_glossiness;
    uint32_t trans_sampler, trans_n;
    float trans_glossiness;
    uint32_t enable_attenuation;
    float attenuation_factor;
    float attenuation_color[3];
} or_ray_material;
typedef struct or_ray_settings {
    float background[3];
    uint32_t max_ray_iterations;
    float ambient_color[3];
    float ambient_intensity;
    uint32_t ao_enabled, ao_sampler, ao_n;
    float ao_max_distance, ao_distribution;
    uint32_t visualize_depth;
    float visualize_depth_max;
    uint32_t fog_enabled;
    float fog_color[3];
    float fog_distance, fog_steepness;
    uint32_t fog_height_dispersion;
    float fog_height, fog_height_steepness;
} or_ray_settings;

/* Raytracer::traceRecursive (Raytracer.cpp:27-597) under Tracer::run's camera
 * sampling, recursively, with the ray tree's keyed draws.  stats_out: per
 * listed pixel the 12 sums of or_render (the last is the number of rays). */
void or_raytrace(const or_bvh* b, const float* vertices, size_t ntri, const uint32_t* tri_mat,
                 const double* materials, const double cam[7], const or_render_cfg* cfg,
                 const or_ray_settings* settings, const or_ray_material* rmats, const or_light* lights,
                 uint32_t nlights, const uint32_t* pixel_ids, size_t npix, double* stats_out, int nthreads);

/* ---------------------------------------------------------------- raycast (Phong) */
/* light: position.xyz, color.rgb, intensity, maxDistance, attenuation
 * ambient: color.rgb, intensity */
void or_raycast_phong(const or_bvh* b, const float* vertices, const uint32_t* tri_mat,
                      const double* materials, const double cam[7], uint32_t width, uint32_t height,
                      uint64_t seed, double ray_start_offset, const double light[9],
                      const double ambient[4], const uint32_t* pixel_ids, size_t npix,
                      int64_t* hit_out, double* t_out, double* rgb_out, int nthreads);

/* ---------------------------------------------------------------- tone map */
/* linear: npix * 4 floats (RGBA).  out: npix * 4 floats.
 * params: key, gamma, clamp(0/1), apply_gamma(0/1), use_fastpow(0/1) */
void or_reinhard(const float* linear, size_t npix, const double params[5], float* out,
                 double* log_avg_out, double* max_lum_out);

/* general tone map: params = type (0 reinhard, 1 linear, 2 simple, 3 passthrough),
 * key, gamma, clamp, apply_gamma, use_fastpow, exposure, enable_averaging,
 * averaging_alpha.  *max_lum_average is the moving average's state (read and,
 * with averaging, updated; start it at 1). */
void or_tonemap(const float* linear, size_t npix, const double params[9], float* out, double* log_avg_out,
                double* max_lum_out, double* max_lum_average);

/* ---------------------------------------------------------------- filters */
/* p0, p1: gaussian standard deviation; mitchell B, C (ignored otherwise) */
double or_filter_weight(uint32_t filter, double radius, double p0, double p1, double x, double y);
double or_filter_radius(uint32_t filter, double radius, double p0, double p1);
uint64_t or_cmj_permute(uint64_t i, uint64_t l, uint64_t p);
void or_cmj_sample(uint64_t x, uint64_t y, uint64_t n, uint64_t p, double r0, double r1, double out[2]);
/* cosine hemisphere (Sampler.cpp:50-64) with ONB::fromNormal-style tangents of a
 * triangle without texcoords (Triangle.cpp:40-44). */
void or_hemisphere(const double n[3], double u1, double u2, double out[3]);

/* ---------------------------------------------------------------- MLP */
/* Layer sizes: in -> h x nh -> out. Parameters packed layer by layer:
 * W (rows = out features, cols = in features, row-major) then b. */
size_t or_mlp_param_count(uint32_t in_dim, uint32_t hidden, uint32_t n_hidden, uint32_t out_dim);
/* x: batch * in_dim. y: batch * out_dim (exp activation on output). */
void or_mlp_forward(const double* params, uint32_t in_dim, uint32_t hidden, uint32_t n_hidden,
                    uint32_t out_dim, const double* x, size_t batch, double* y);
/* Loss = mean over batch and channels of relative L2: (y - t)^2 / (sg(y)^2 + eps)
 * with stop-gradient denominator when rel != 0, else plain (y - t)^2.
 * grad_out: parameter gradient. Returns the loss. */
double or_mlp_loss_grad(const double* params, uint32_t in_dim, uint32_t hidden, uint32_t n_hidden,
                        uint32_t out_dim, const double* x, const double* target, size_t batch,
                        int rel, double eps, double* grad_out);
/* One Adam step (Kingma & Ba), t = step number starting at 1. */
void or_adam_step(double* params, double* m, double* v, const double* grad, size_t n,
                  double lr, double beta1, double beta2, double eps, uint32_t t);
/* Radiance-cache input encoding: pos (3, scaled to [0,1] by the scene box),
 * 12 frequencies sin/cos per axis (72), dir (3), normal (3), albedo (3) = 81. */
void or_cache_encode(const double pos01[3], const double dir[3], const double nrm[3],
                     const double albedo[3], double out81[81]);

#ifdef __cplusplus
}
#endif
#endif
