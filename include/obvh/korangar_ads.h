/*
 * korangar_ads.h -- the intersector (ADS) plugin interface that the
 * reference's application loads with dlopen("./lib<name>.so") and
 * dlsym("korangar_ads_action") (src/common/ads.c:6-33), implemented here by
 * obvh_amd/liblbvh_gpu.so and obvh_amd/libatrbvh_gpu.so on top of
 * libobvh.so.
 *
 * The types below keep the reference's layouts byte for byte (plain float
 * vectors, 4-byte aligned; kr_size and pointers 8 bytes), so the reference's
 * host code passes its own structs in unchanged.  Sizes are checked by the
 * static asserts at the bottom.
 *
 * Reference interfaces (paths relative to the reference repository root):
 *   kr_descriptor / kr_descriptor_container  src/common/korangar.h:17-26
 *   kr_camera_type / kr_camera               src/common/korangar.h:55-91
 *   kr_object_type                           src/common/korangar.h:70-76
 *   kr_mesh / kr_sphere / kr_object          src/common/korangar.h:106-127
 *   kr_object_instance                       src/common/korangar.h:129-132
 *   kr_scene                                 src/common/korangar.h:171-198
 *   kr_transform                             src/common/vecmath.h:201-204
 *   kr_ads_action                            src/common/ads.h:6-26
 *   kr_query_type                            src/common/ads.h:34-40
 *   kr_ads, kr_ads_callback                  src/common/ads.h:42-62
 *   kr_intersection_query                    src/common/ads.h:98-114
 *   korangar_ads_action                      src/ads/lbvh_gpu/module.c:6-41,
 *                                            src/ads/atrbvh_gpu/module.c:6-41
 */
#ifndef OBVH_KORANGAR_ADS_H
#define OBVH_KORANGAR_ADS_H

#include "obvh/obvh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef const void* kr_handle;
typedef kr_size kr_index;
typedef struct { u32 x, y; } kr_uvec2;                     /* 8 B */

typedef struct {
    const char* key;
    const char* value;
} kr_descriptor;

typedef struct {
    kr_descriptor* descriptors;
    kr_size descriptor_count;
} kr_descriptor_container;

typedef struct {
    kr_mat4 to;
    kr_mat4 from;
} kr_transform;                                            /* 128 B */

typedef enum { KR_CAMERA_NONE, KR_CAMERA_PINHOLE, KR_CAMERA_ORTHO, KR_CAMERA_MAX } kr_camera_type;

typedef struct {
    kr_transform projection;
    kr_scalar fov;
} kr_camera_pinhole;

typedef struct {
    kr_camera_type type;
    kr_transform view;
    kr_uvec4 viewport;
    union {
        kr_camera_pinhole as_pinhole;
    };
} kr_camera;                                               /* 280 B */

typedef enum { KR_OBJECT_NONE, KR_OBJECT_MESH, KR_OBJECT_SPHERE, KR_OBJECT_AABB, KR_OBJECT_MAX } kr_object_type;

typedef struct {
    kr_vec3* vertices;
    kr_vec3* normals;
    kr_vec2* uvs;
    kr_uvec4* faces;
    kr_size face_count;
    kr_size attr_count;  /* number of vertices */
} kr_mesh;

typedef struct {
    kr_vec3 center;
    kr_scalar radius;
} kr_sphere;

typedef struct {
    kr_aabb3 aabb;
    kr_object_type type;
    union {
        kr_sphere as_sphere;
        kr_mesh as_mesh;
    };
} kr_object;                                               /* 80 B */

typedef struct {
    kr_transform model;
    kr_index object_id;
} kr_object_instance;                                      /* 136 B */

/* Only objects, instances and benchmark_file are read by the plugins; the
 * rest is carried for the layout. */
typedef struct {
    kr_camera camera;
    kr_aabb3 aabb;
    kr_object* objects;
    kr_size object_count;
    kr_object_instance* instances;
    kr_size instance_count;
    void* cameras;
    kr_size camera_count;
    void* materials;
    kr_size material_count;
    void* textures;
    kr_size texture_count;
    void* samplers;
    kr_size sampler_count;
    kr_uvec2 frame;
    kr_uvec2 tile;
    const char* benchmark_file;
} kr_scene;                                                /* 424 B */

typedef enum {
    KR_ADS_ACTION_NONE,
    KR_ADS_ACTION_INTERSECTION_QUERY,
    KR_ADS_ACTION_OCCLUSION_QUERY,
    KR_ADS_ACTION_PREFERRED_QUERY_TYPE,
    KR_ADS_ACTION_GET_BLAS,
    KR_ADS_ACTION_GET_TLAS,
    KR_ADS_ACTION_CREATE,
    KR_ADS_ACTION_COMMIT,
    KR_ADS_ACTION_INIT,
    KR_ADS_ACTION_UPDATE,
    KR_ADS_ACTION_SETTINGS_UPDATE,
    KR_ADS_ACTION_DESTROY,
    KR_ADS_ACTION_CSV_EXPORT,
    KR_ADS_ACTION_CSV_IMPORT,
    KR_ADS_ACTION_BVH_MODEL_EXPORT,
    KR_ADS_ACTION_BINARY_EXPORT,
    KR_ADS_ACTION_BINARY_IMPORT,
    KR_ADS_ACTION_NAME,
    KR_ADS_ACTION_MAX
} kr_ads_action;

typedef enum {
    KR_QUERY_TYPE_NONE,
    KR_QUERY_TYPE_CPU,
    KR_QUERY_TYPE_CUDA,
    KR_QUERY_TYPE_CUDA_MANAGED,
    KR_QUERY_TYPE_MAX
} kr_query_type;

typedef struct kr_ads kr_ads;
typedef kr_error (*kr_ads_callback)(kr_ads*, kr_handle, kr_ads_action);
struct kr_ads {
    kr_ads_callback callback;
    kr_handle library;
    kr_handle context;  /* what CREATE returned */
    kr_handle runtime;
};

/* Device rays in, device hit records out (the "cuda" member of the
 * reference's query union; the cpu member has the same layout). */
typedef struct {
    const kr_ray* rays;
    kr_intersection* isects;
    kr_size ray_count;
} kr_intersection_query_cuda;

typedef struct {
    kr_query_type type;
    union {
        kr_intersection_query_cuda cuda;
    };
} kr_intersection_query;                                   /* 32 B */

/* The plugin entry point.  Actions (as the reference's lbvh_gpu and
 * atrbvh_gpu modules, which the two libraries replace):
 *   CREATE                ads may be NULL; returns the new context.
 *   INIT                  descriptor: kr_descriptor_container* of settings
 *                         (collapse, use_obbs, use_persistent_kernel,
 *                         intersect_bounds, intersect_bounds_level, ct, ci,
 *                         save_construction_information; "y" = on).
 *   COMMIT                descriptor: kr_scene*.  Builds the tree of every
 *                         mesh object (through the instances when there are
 *                         any), keeping the last one as the reference does:
 *                         upload, LBVH or ATRBVH (treelet 9, 2 passes),
 *                         OBB transform (use_obbs), collapse (collapse),
 *                         AoS packing + SoA OBVH (use_persistent_kernel).
 *                         With scene->benchmark_file set, appends the SAH
 *                         figures and one line of build timings (ms).
 *   INTERSECTION_QUERY    descriptor: kr_intersection_query* with device
 *                         rays and hit records; closest hit through the SoA
 *                         OBVH (use_obbs + use_persistent_kernel).  Without
 *                         either it returns success and writes nothing, as
 *                         the reference's non-OBB path does.
 *   NAME                  "LBVH GPU" / "ATRBVH".
 *   PREFERRED_QUERY_TYPE  KR_QUERY_TYPE_CUDA.
 *   DESTROY               frees the context and its device arrays (the
 *                         reference leaves this action empty).
 *   anything else         success, nothing done.
 * Errors are static strings (NULL = success).  The reference's other query
 * modes (bounds visualisation, the AoS and AABB traversals) are not built:
 * INIT accepts their settings, COMMIT or the query then returns an error
 * naming the missing mode.                                                */
kr_error korangar_ads_action(kr_ads* ads, kr_handle descriptor, kr_ads_action action);

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus)
static_assert(sizeof(kr_transform) == 128, "kr_transform");
static_assert(sizeof(kr_camera) == 280, "kr_camera");
static_assert(sizeof(kr_object) == 80, "kr_object");
static_assert(sizeof(kr_object_instance) == 136, "kr_object_instance");
static_assert(sizeof(kr_scene) == 424, "kr_scene");
static_assert(sizeof(kr_ads) == 32, "kr_ads");
static_assert(sizeof(kr_intersection_query) == 32, "kr_intersection_query");
static_assert(sizeof(kr_descriptor_container) == 16, "kr_descriptor_container");
#else
_Static_assert(sizeof(kr_transform) == 128, "kr_transform");
_Static_assert(sizeof(kr_camera) == 280, "kr_camera");
_Static_assert(sizeof(kr_object) == 80, "kr_object");
_Static_assert(sizeof(kr_object_instance) == 136, "kr_object_instance");
_Static_assert(sizeof(kr_scene) == 424, "kr_scene");
_Static_assert(sizeof(kr_ads) == 32, "kr_ads");
_Static_assert(sizeof(kr_intersection_query) == 32, "kr_intersection_query");
_Static_assert(sizeof(kr_descriptor_container) == 16, "kr_descriptor_container");
#endif

#endif /* OBVH_KORANGAR_ADS_H */
