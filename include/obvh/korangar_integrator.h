/*
 * korangar_integrator.h -- the integrator plugin interface that the
 * reference's application loads with dlopen("./lib<name>.so") and
 * dlsym("korangar_action") (src/common/integrator.c:8-38), implemented here
 * by obvh_amd/libbench.so (the reference's `bench` integrator,
 * src/integrators/bench/module.c:6-36, bench.cu:20-135) on top of
 * libobvh.so.  Layouts are the reference's, byte for byte.
 *
 * Reference interfaces (paths relative to the reference repository root):
 *   kr_action, kr_integrator, kr_callback    src/common/integrator.h:6-32
 *   kr_target_type, kr_target                src/common/korangar.h:28-52
 *   kr_ads_action_payload                    src/common/ads.h:126-129
 *   korangar_action                          src/integrators/bench/module.c:6-36
 */
#ifndef OBVH_KORANGAR_INTEGRATOR_H
#define OBVH_KORANGAR_INTEGRATOR_H

#include "obvh/korangar_ads.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    KR_ACTION_NONE,
    KR_ACTION_RENDER,
    KR_ACTION_CREATE,
    KR_ACTION_INIT,
    KR_ACTION_COMMIT,
    KR_ACTION_CAMERA_UPDATE,
    KR_ACTION_UPDATE,
    KR_ACTION_SETTINGS_UPDATE,
    KR_ACTION_DESTROY,
    KR_ACTION_ADS_ACTION,
    KR_ACTION_MAX
} kr_action;

typedef struct kr_integrator kr_integrator;
typedef kr_error (*kr_callback)(kr_integrator*, kr_handle, kr_action);
struct kr_integrator {
    kr_callback callback;
    kr_handle library;
    kr_handle context;  /* what CREATE returned */
};                                                        /* 24 B */

typedef enum {
    KR_TARGET_NONE,
    KR_TARGET_CPU_BUFFER,
    KR_TARGET_CPU_BUFFER_F32,
    KR_TARGET_GL_TEXTURE_2D,
    KR_TARGET_MAX
} kr_target_type;

typedef struct {
    kr_target_type type;
    union {
        struct { u32 width, height; void* data; } as_cpu_buffer;
        struct { u32 width, height; void* data; } as_cpu_buffer_f32;
        struct { u32 width, height; u32 handle; } as_gl_texture_2d;
    };
} kr_target;                                              /* 24 B */

typedef struct {
    kr_ads_action action;
    kr_handle data;
} kr_ads_action_payload;                                  /* 16 B */

/* What RENDER returns when the benchmark has run: the loop stops on a
 * non-NULL result (src/integrators/bench/bench.cu:91). */
#define KR_RENDER_DONE ((kr_error)"Simulation is done")

/* The bench integrator.  Actions:
 *   CREATE      returns the new context (integrator may be NULL).
 *   INIT        descriptor: kr_descriptor_container* of settings (the
 *               reference's keys are accepted; none changes the DiTO run).
 *   COMMIT      descriptor: kr_scene*, kept for RENDER.
 *   RENDER      descriptor: kr_target* (unused).  The DiTO benchmark: the
 *               vertices of the first instance's mesh through kr_points_obb
 *               (host) and kr_cuda_points_obb (GPU), both boxes printed as
 *               the reference prints them; returns KR_RENDER_DONE, or an
 *               error string (no scene, no instance, not a mesh, GPU error).
 *   ADS_ACTION  descriptor: kr_ads_action_payload*, forwarded to the
 *               attached intersector; none is attached (the reference's
 *               bench never attaches one), so it returns an error string.
 *   DESTROY     frees the context.
 *   anything else: success.                                              */
kr_error korangar_action(kr_integrator* integrator, kr_handle descriptor, kr_action action);

#ifdef __cplusplus
}
#endif

#if defined(__cplusplus)
static_assert(sizeof(kr_integrator) == 24, "kr_integrator");
static_assert(sizeof(kr_target) == 24, "kr_target");
static_assert(sizeof(kr_ads_action_payload) == 16, "kr_ads_action_payload");
#else
_Static_assert(sizeof(kr_integrator) == 24, "kr_integrator");
_Static_assert(sizeof(kr_target) == 24, "kr_target");
_Static_assert(sizeof(kr_ads_action_payload) == 16, "kr_ads_action_payload");
#endif

#endif /* OBVH_KORANGAR_INTEGRATOR_H */
