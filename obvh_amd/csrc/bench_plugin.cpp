// bench_plugin.cpp -- libbench.so, the reference's `bench` integrator
// (src/integrators/bench/module.c:6-36, bench.cu:20-135) over libobvh.so:
// RENDER fits the first instance's mesh vertices with the host DiTO
// (kr_points_obb) and the GPU DiTO (kr_cuda_points_obb) and prints both
// boxes in the reference's format.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "obvh/korangar_integrator.h"

namespace {

struct Bench {
    kr_scene* scene = nullptr;
    kr_ads* ads = nullptr;  // the attached intersector (never set, as in the reference)
    u32 sample_count = 1000000, ray_sample_count = 1, repeat_count = 1, indirect_bounces = 0;
    bool dito = false;
    const char* benchmark_file = nullptr;
    const char* benchmark_type = nullptr;
    const char* intersector = "lbvh";
    float cpu_ms = 0.0f, gpu_ms = 0.0f;
};

bool key_is(const kr_descriptor& d, const char* k) { return d.key && strcmp(d.key, k) == 0; }
bool value_yes(const kr_descriptor& d) { return d.value && strcmp(d.value, "y") == 0; }
u32 value_u32(const kr_descriptor& d) { return d.value ? (u32)atoi(d.value) : 0u; }

// The settings the reference's bench reads (bench.cu:40-83), key spellings included.
kr_error init(Bench* b, const kr_descriptor_container* s) {
    if (!s) return "bench: INIT needs a settings container";
    for (kr_size i = 0; i < s->descriptor_count; ++i) {
        const kr_descriptor& d = s->descriptors[i];
        if (key_is(d, "benhcmark_sample_count")) b->sample_count = value_u32(d);
        else if (key_is(d, "benhcmark_ray_sample_count")) b->ray_sample_count = value_u32(d);
        else if (key_is(d, "benchmark_dito")) b->dito = value_yes(d);
        else if (key_is(d, "benchmark_file")) b->benchmark_file = d.value;
        else if (key_is(d, "benchmark_type")) b->benchmark_type = d.value;
        else if (key_is(d, "intersector")) b->intersector = d.value;
        else if (key_is(d, "benchmark_repeat_count")) b->repeat_count = value_u32(d);
        else if (key_is(d, "benchmark_indirect_bounces")) b->indirect_bounces = value_u32(d);
    }
    return nullptr;
}

void print_box(const char* title, const kr_obb3& o) {
    printf("Computed OBB %s:\n", title);
    printf("Midpoint: %f %f %f\n", o.mid.x, o.mid.y, o.mid.z);
    printf("v0: %f %f %f\n", o.v0.x, o.v0.y, o.v0.z);
    printf("v1: %f %f %f\n", o.v1.x, o.v1.y, o.v1.z);
    printf("v2: %f %f %f\n", o.v2.x, o.v2.y, o.v2.z);
    printf("ext: %f %f %f\n", o.ext.x, o.ext.y, o.ext.z);
    printf("Area: %f\n", 2 * (o.ext.x * o.ext.y + o.ext.x * o.ext.z + o.ext.y * o.ext.z));
}

// bench.cu:93-132: the vertices of instance 0's object, host fit then GPU fit.
kr_error render(Bench* b) {
    const kr_scene* sc = b->scene;
    if (!sc) return "bench: RENDER before COMMIT";
    if (sc->instance_count == 0 || !sc->instances) return "bench: the scene has no instances";
    const kr_index oid = sc->instances[0].object_id;
    if (oid >= sc->object_count || !sc->objects) return "bench: instance 0 names a missing object";
    const kr_object& obj = sc->objects[oid];
    if (obj.type != KR_OBJECT_MESH) return "bench: instance 0 is not a mesh";
    const kr_size n = obj.as_mesh.attr_count;
    if (n > 0 && !obj.as_mesh.vertices) return "bench: the mesh has no vertices";
    std::vector<kr_vec3> points(obj.as_mesh.vertices, obj.as_mesh.vertices + n);
    const kr_obb3 cpu = kr_points_obb(points.data(), n, &b->cpu_ms);
    print_box("CPU", cpu);
    const kr_obb3 gpu = kr_cuda_points_obb(points.data(), n, &b->gpu_ms);
    print_box("GPU", gpu);
    fflush(stdout);
    if (n > 0 && gpu.ext.x != gpu.ext.x) {  // all-NaN box: the GPU fit failed
        const char* e = obvh_last_error();
        return e ? e : "bench: kr_cuda_points_obb failed";
    }
    return KR_RENDER_DONE;
}

}  // namespace

extern "C" __attribute__((visibility("default"))) kr_error korangar_action(kr_integrator* integrator,
                                                                         kr_handle descriptor, kr_action action) {
    if (action == KR_ACTION_CREATE) return (kr_error) new Bench();
    Bench* b = integrator ? (Bench*)integrator->context : nullptr;
    switch (action) {
        case KR_ACTION_INIT:
            if (!b) return "bench: no context (CREATE first)";
            return init(b, (const kr_descriptor_container*)descriptor);
        case KR_ACTION_COMMIT:
            if (!b) return "bench: no context (CREATE first)";
            b->scene = (kr_scene*)descriptor;
            return nullptr;
        case KR_ACTION_RENDER:
            if (!b) return "bench: no context (CREATE first)";
            return render(b);
        case KR_ACTION_ADS_ACTION: {
            if (!b) return "bench: no context (CREATE first)";
            const kr_ads_action_payload* p = (const kr_ads_action_payload*)descriptor;
            if (!p) return "bench: ADS_ACTION needs a payload";
            if (!b->ads || !b->ads->callback) return "bench: no intersector attached";
            return b->ads->callback(b->ads, p->data, p->action);
        }
        case KR_ACTION_DESTROY:
            delete b;
            if (integrator) integrator->context = nullptr;
            return nullptr;
        default:
            return nullptr;
    }
}
