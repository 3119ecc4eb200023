/*
 * raycer_host.h — C ABI of the host half of the renderer: the FP64 SAH BVH
 * builder (Scene::initialize -> FlatBVH::build) and the seeded scene
 * generators used by the benchmarks and tests (the counterpart of raycer's
 * Scene::createTestScene, src/Raytracing/Scene.h:60-87).
 *
 * Results are written into caller-owned buffers after a size query.  Same
 * status and error conventions as raycer_hip.h.
 */
#ifndef RAYCER_HOST_H
#define RAYCER_HOST_H

#include "raycer_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BVHBuildInfo (src/Raytracing/Primitives/FlatBVH.h:19-36); of the
 * reference's options only the SAH path (useSAH = true, regularSAHSplits = 0)
 * is provided.  method selects the split rule:
 *   RCH_BVH_REFERENCE  the reference's 6-candidate SAH (node-for-node parity);
 *   RCH_BVH_BINNED     a 32-bin SAH with a leaf-cost test (no counterpart in
 *                      the reference), in the same flat format.
 * Closest-hit and any-hit results do not depend on the method. */
enum { RCH_BVH_REFERENCE = 0, RCH_BVH_BINNED = 1 };
typedef struct rch_bvh_build_info {
    uint64_t max_leaf_size;  /* maxLeafSize, default 5 */
    uint32_t threads;        /* 0 = all hardware threads */
    uint32_t method;         /* RCH_BVH_REFERENCE (0) or RCH_BVH_BINNED */
    float tri_cost;          /* RCH_BVH_BINNED: cost of a triangle test relative to a node visit (0 = 1) */
    uint32_t pad;
} rch_bvh_build_info;

typedef struct rch_bvh rch_bvh;

/* FlatBVH::build (src/Raytracing/Primitives/FlatBVH.cpp:89-204) over the
 * triangles' FP64 bounding boxes, with the reference's 6-candidate SAH split
 * (FlatBVH.cpp:241-342).  Independent subtrees are built on separate threads;
 * the node order is the reference's preorder regardless of thread count. */
int rch_bvh_build(const float* vertices, size_t ntri, const rch_bvh_build_info* info, rch_bvh** out);
size_t rch_bvh_node_count(const rch_bvh* bvh);
int rch_bvh_get(const rch_bvh* bvh, rc_flat_node* nodes, uint64_t* ordered_ids);
double rch_bvh_build_ms(const rch_bvh* bvh);
uint32_t rch_bvh_depth(const rch_bvh* bvh);
int rch_bvh_free(rch_bvh* bvh);

/* A built BVH on disk (the reference serializes FlatBVH's flatNodes and
 * orderedPrimitiveIds, FlatBVH.h:99-105, and restores them on load instead of
 * rebuilding, FlatBVH.cpp:206-214, Scene.cpp:487-488).  Loading fails unless
 * the triangles hash to what was saved.  The upload validates the tree again. */
int rch_bvh_save(const char* path, const float* vertices, size_t ntri, const rc_flat_node* nodes, size_t nnodes,
                 const uint64_t* ordered_ids);
int rch_bvh_file_counts(const char* path, size_t* nnodes, size_t* ntri);
int rch_bvh_load(const char* path, const float* vertices, size_t ntri, rc_flat_node* nodes, uint64_t* ordered_ids);

/* ------------------------------------------------------------ scenes */
enum {
    RCH_SCENE_BOX = 1,          /* closed box + two rotated inner boxes + ceiling light, 36 triangles */
    RCH_SCENE_SPHERE_GROUND = 2,/* displaced icosphere on a jittered ground grid (1,179,648 triangles at level 7 / grid 512) */
    RCH_SCENE_TERRAIN = 3,      /* ridged-fBm heightfield, 2 (n-1)^2 triangles (2,093,058 at n = 1024) */
    RCH_SCENE_SOUP = 4          /* seeded random triangle soup (batched multi-scene workload) */
};

typedef struct rch_scene_params {
    uint32_t kind;
    uint32_t seed;
    uint32_t level;      /* SPHERE_GROUND: icosphere subdivision level (7) */
    uint32_t grid;       /* SPHERE_GROUND: ground quads per side (512); TERRAIN: vertices per side (1024) */
    uint64_t tri_count;  /* SOUP: 0 = draw log-uniformly in [1e3, 5e5] from the seed */
} rch_scene_params;

typedef struct rch_scene_desc {
    size_t ntri;
    size_t nmat;
    rc_camera camera;
    rc_render_cfg cfg;        /* the workload's render settings */
    rc_point_light light;     /* SPHERE_GROUND only */
    float bounds_min[3];
    float bounds_max[3];
} rch_scene_desc;

/* Fills desc (sizes, camera, settings) for the given parameters. */
int rch_scene_describe(const rch_scene_params* p, rch_scene_desc* desc);
/* Writes ntri * 9 vertices, ntri material indices and nmat materials. */
int rch_scene_generate(const rch_scene_params* p, float* vertices, uint32_t* tri_material, rc_material* materials);

/* ------------------------------------------------------------ OBJ models
 * ModelLoader::readObjFile (src/Utils/ObjModelLoader.cpp:332-482) for
 * triangle scenes: v, f (i, i/t, i//n, i/t/n, negative indices, fan
 * triangulation), mtllib / usemtl, and Ka / Kd / Ks / Ke / Ns from the MTL
 * file; numbers read as StringUtils::parseDouble reads them.  Material 0 is
 * the base material (faces before any usemtl, or with an unknown name); the
 * MTL materials follow in file order.  Texture maps, groups and instances are
 * not supported. */
typedef struct rch_obj_info {  /* ModelLoaderInfo (src/Utils/ModelLoader.h:27-46) */
    double scale[3];           /* default 1, 1, 1 */
    double rotate[3];          /* EulerAngle pitch, yaw, roll (degrees), applied as rotateXYZ */
    double translate[3];
    rc_material base;          /* baseMaterial: every MTL material starts from it */
    uint32_t ignore_materials;
    uint32_t pad;
} rch_obj_info;
typedef struct rch_obj rch_obj;

void rch_obj_default_info(rch_obj_info* info);
int rch_obj_load(const char* path, const rch_obj_info* info /* NULL = defaults */, rch_obj** out);
/* ModelLoader::readPlyFile (src/Utils/PlyModelLoader.cpp): ascii or binary
 * PLY with a vertex element (x, y, z) and a face element (vertex_indices
 * lists, fanned); one material, the base material. */
int rch_ply_load(const char* path, const rch_obj_info* info /* NULL = defaults */, rch_obj** out);
size_t rch_obj_triangle_count(const rch_obj* model);
size_t rch_obj_material_count(const rch_obj* model);
/* vertices: 9 floats per triangle; tri_material: one index per triangle */
int rch_obj_get(const rch_obj* model, float* vertices, uint32_t* tri_material, rc_material* materials);
int rch_obj_free(rch_obj* model);

/* ------------------------------------------------------------ scene files
 * Scene::loadFromFile / saveToFile (src/Raytracing/Scene.cpp:67-183) for the
 * cereal XML (.xml), JSON (.json) and binary (.bin) forms of Scene::serialize (Scene.h:330-345):
 * the settings, camera, tone mapper, simple fog, lights, materials, triangle
 * primitives, the model list and a built root BVH.  Other primitive types
 * (planes, spheres, boxes, ...) are counted and skipped, as are textures and
 * the volumetric fog.  The binary form (.bin) holds triangle scenes only. */
typedef struct rch_scene_model {   /* ModelLoaderInfo (src/Utils/ModelLoader.h:27-85) */
    char path[1024];               /* modelFilePath */
    rch_obj_info info;             /* scale, rotate, translate, baseMaterial, ignoreMaterials */
    uint32_t invisible_triangles;
    uint32_t pad;
} rch_scene_model;

typedef struct rch_scene_file_data {
    uint32_t tracer_type;          /* 0 ray tracer, 1 path tracer (TracerType) */
    uint32_t pad;
    rc_camera camera;
    rc_render_cfg cfg;             /* sampler, filter, spp = multiSampleCountSqrt^2, max_path_length,
                                      ray_start_offset; width and height 0 (kept outside the scene) */
    rc_tonemap tonemap;
    rc_ray_settings ray;
    size_t nmat;
    rc_material* materials;
    rc_ray_material* ray_materials;
    uint64_t* material_ids;        /* Material::id of each entry */
    size_t ntri;                   /* visible triangles */
    float* vertices;               /* 9 floats per triangle */
    uint32_t* tri_material;        /* index into materials (resolved from Primitive::materialId) */
    uint64_t* tri_ids;             /* Primitive::id of each triangle */
    size_t nlights;
    rc_light* lights;              /* directional, then point, then spot lights */
    size_t nmodels;
    rch_scene_model* models;
    size_t nnodes;                 /* 0 unless the file holds a built BVH over exactly these triangles */
    rc_flat_node* nodes;
    uint64_t* ordered;             /* orderedPrimitiveIds as indices into the triangle arrays */
    size_t skipped_primitives;     /* non-triangle primitives and invisible triangles in the file */
} rch_scene_file_data;

/* Reads a scene file (.xml, .json or .bin); the result is freed with rch_scene_file_free. */
int rch_scene_file_load(const char* path, rch_scene_file_data** out);
/* Parses a scene from memory: format RCH_SCENE_XML (0), RCH_SCENE_JSON (1)
 * or RCH_SCENE_BINARY (2, cereal's BinaryOutputArchive: Scene.cpp:88-90,
 * 144-146; files holding textures or non-triangle primitives are rejected). */
enum { RCH_SCENE_XML = 0, RCH_SCENE_JSON = 1, RCH_SCENE_BINARY = 2 };
int rch_scene_file_parse(const char* text, size_t len, int format, rch_scene_file_data** out);
/* Writes a scene (.xml, .json or .bin by the path's extension) in cereal's layout. */
int rch_scene_file_save(const char* path, const rch_scene_file_data* scene);
void rch_scene_file_free(rch_scene_file_data* scene);

#ifdef __cplusplus
}
#endif
#endif
