// ads_plugin.hip -- the lbvh_gpu and atrbvh_gpu intersector plugins
// (korangar_ads_action; src/ads/lbvh_gpu/module.c:6-41,
// src/ads/atrbvh_gpu/module.c:6-41) on top of libobvh.so.  Host code only.
// Built twice: OBVH_ADS_ATRBVH=0 gives liblbvh_gpu.so, =1 libatrbvh_gpu.so.
//
// COMMIT follows the plugins' commit (lbvh.cu:1547-1762 with 1764-1858,
// atrbvh.cu:250-481 with 483-550): upload the mesh, build the tree, then the
// OBB transform, the collapse and the OBVH packing the settings ask for.
// The mesh, the collapse's primitive slices and the SoA OBVH stay for the
// queries; every other array is freed before COMMIT returns, as there.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "obvh/korangar_ads.h"

#ifndef OBVH_ADS_ATRBVH
#error "build with -DOBVH_ADS_ATRBVH=0 (liblbvh_gpu) or -DOBVH_ADS_ATRBVH=1 (libatrbvh_gpu)"
#endif

namespace {

constexpr bool kAtr = OBVH_ADS_ATRBVH != 0;
// AgglomerativeTreeletOptimizer(9, 2) in the ATRBVH plugin (atrbvh.cu:330)
constexpr int kTreeletSize = 9;
constexpr int kIterations = 2;
constexpr const char* kName = kAtr ? "ATRBVH" : "LBVH GPU";

struct Blas {
    kr_vec3* V = nullptr;
    kr_uvec4* F = nullptr;
    u32* primitives = nullptr;
    kr_SoA_OBVH* soa = nullptr;
};

struct Context {
    Blas blas;
    kr_scene* scene = nullptr;
    kr_scalar ct = 0.0f, ci = 0.0f;
    int32_t intersect_bounds_level = 0;
    bool save_construction_information = false, collapse = false, use_obbs = false, intersect_bounds = false,
         use_persistent_kernel = false;
};

// Build timings (ms) of one COMMIT, the reference's benchmark line.
struct Metrics {
    double build = 0, obb = 0, mark = 0, collapse = 0;
};

// KR_EQUALS_LITERAL (src/common/util.h:6): strncmp over the literal and its
// terminating NUL, i.e. an exact match.
template <size_t n>
bool equals(const char* s, const char (&lit)[n]) {
    return s != nullptr && std::strncmp(s, lit, n) == 0;
}

// A device array freed on scope exit unless released.
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    kr_error alloc(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return "korangar_ads_action: out of device memory";
        }
        return nullptr;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    template <class T>
    T* release() {
        T* q = static_cast<T*>(p);
        p = nullptr;
        return q;
    }
};

void free_blas(Blas& b) {
    if (b.soa) (void)obvh_soa_obvh_free(b.soa);
    if (b.V) (void)hipFree(b.V);
    if (b.F) (void)hipFree(b.F);
    if (b.primitives) (void)hipFree(b.primitives);
    b = Blas{};
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

#define ADS_TRY(expr)                    \
    do {                                 \
        kr_error e_ = (expr);            \
        if (e_ != nullptr) return e_;    \
    } while (0)
#define ADS_HIP(expr)                                                     \
    do {                                                                  \
        if ((expr) != hipSuccess) {                                       \
            (void)hipGetLastError();                                      \
            return "korangar_ads_action: HIP call failed during COMMIT";  \
        }                                                                 \
    } while (0)

template <class T>
kr_error to_host(std::vector<T>& h, const void* d, size_t n) {
    h.resize(n);
    if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return "korangar_ads_action: HIP call failed during COMMIT";
    }
    return nullptr;
}

kr_error init(Context& c, const kr_descriptor_container* settings) {
    if (settings == nullptr) return "korangar_ads_action: INIT needs a kr_descriptor_container";
    for (kr_size i = 0; i < settings->descriptor_count; ++i) {
        const char* key = settings->descriptors[i].key;
        const char* value = settings->descriptors[i].value;
        if (equals(key, "save_construction_information")) c.save_construction_information = equals(value, "y");
        else if (equals(key, "collapse")) c.collapse = equals(value, "y");
        else if (equals(key, "ct")) c.ct = value ? (kr_scalar)std::atof(value) : 0.0f;
        else if (equals(key, "ci")) c.ci = value ? (kr_scalar)std::atof(value) : 0.0f;
        else if (equals(key, "use_obbs")) c.use_obbs = equals(value, "y");
        else if (equals(key, "intersect_bounds_level")) c.intersect_bounds_level = value ? std::atoi(value) : 0;
        else if (equals(key, "intersect_bounds")) c.intersect_bounds = equals(value, "y");
        else if (equals(key, "use_persistent_kernel")) c.use_persistent_kernel = equals(value, "y");
    }
    return nullptr;
}

// One object's tree (lbvh_cuda_blas_commit / atrbvh_cuda_blas_commit).
kr_error blas_commit(Context& c, const kr_object* obj, Metrics& m) {
    if (obj->type != KR_OBJECT_MESH) return "korangar_ads_action: only mesh objects are supported";
    const kr_mesh& mesh = obj->as_mesh;
    if (mesh.face_count == 0) return "korangar_ads_action: a mesh object has no faces";
    if (mesh.face_count > 0x7FFFFFFFu) return "korangar_ads_action: a mesh object has too many faces";
    if (mesh.vertices == nullptr || mesh.faces == nullptr || mesh.attr_count == 0)
        return "korangar_ads_action: a mesh object has no vertices or faces";
    if (c.use_persistent_kernel && !c.use_obbs)
        return "korangar_ads_action: use_persistent_kernel without use_obbs (the AABB traversal) is not built";
    const u32 N = (u32)mesh.face_count, I = N - 1, total = 2 * N - 1;
    const char* file = c.scene->benchmark_file;
    free_blas(c.blas);

    DevBuf V, F, prims, nodes, parents, counts, costs, ws;
    ADS_TRY(V.alloc(mesh.attr_count * sizeof(kr_vec3)));
    ADS_TRY(F.alloc((size_t)N * sizeof(kr_uvec4)));
    ADS_TRY(prims.alloc((size_t)N * sizeof(u32)));
    ADS_TRY(nodes.alloc((size_t)total * sizeof(kr_bvh_node)));
    ADS_TRY(parents.alloc((size_t)total * sizeof(u32)));
    ADS_TRY(counts.alloc((size_t)total * sizeof(u32)));
    ADS_TRY(costs.alloc((size_t)total * sizeof(kr_scalar)));
    ADS_HIP(hipMemcpy(V.p, mesh.vertices, mesh.attr_count * sizeof(kr_vec3), hipMemcpyHostToDevice));
    ADS_HIP(hipMemcpy(F.p, mesh.faces, (size_t)N * sizeof(kr_uvec4), hipMemcpyHostToDevice));
    ADS_HIP(hipMemset(prims.p, 0, (size_t)N * sizeof(u32)));

    auto t0 = std::chrono::steady_clock::now();
    if (kAtr) {
        const size_t wb = obvh_atrbvh_workspace_size(N);
        ADS_TRY(ws.alloc(wb));
        ADS_TRY(obvh_atrbvh_build(V.as<kr_vec3>(), F.as<kr_uvec4>(), N, kTreeletSize, kIterations, c.ci, c.ct,
                                  nodes.as<kr_bvh_node>(), parents.as<u32>(), counts.as<u32>(),
                                  costs.as<kr_scalar>(), ws.p, wb, nullptr));
    } else {
        const size_t wb = obvh_lbvh_workspace_size(N);
        ADS_TRY(ws.alloc(wb));
        ADS_TRY(obvh_lbvh_build(V.as<kr_vec3>(), F.as<kr_uvec4>(), N, nodes.as<kr_bvh_node>(), parents.as<u32>(),
                                counts.as<u32>(), costs.as<kr_scalar>(), nullptr, ws.p, wb, nullptr));
    }
    ADS_HIP(hipDeviceSynchronize());
    m.build += ms_since(t0);

    // the LBVH plugin appends the SAH of each stage to the benchmark file
    const bool sah = !kAtr && file != nullptr;
    std::vector<kr_bvh_node> h_tree;
    std::vector<kr_obb3> h_obbs;
    if (sah) {
        ADS_TRY(to_host(h_tree, nodes.p, total));
        kr_bvh_sah(h_tree.data(), 0, c.ci, c.ct, nullptr, file);
    }
    DevBuf obbs, mats;
    if (c.use_obbs) {
        ADS_TRY(obbs.alloc((size_t)total * sizeof(kr_obb3)));
        ADS_TRY(mats.alloc((size_t)total * sizeof(kr_mat4)));
        t0 = std::chrono::steady_clock::now();
        ADS_TRY(kr_cuda_bvh_obb_tree(nodes.as<kr_bvh_node>(), parents.as<u32>(), counts.as<u32>(),
                                     costs.as<kr_scalar>(), V.as<kr_vec3>(), F.as<kr_uvec4>(), prims.as<u32>(),
                                     obbs.as<kr_obb3>(), mats.as<kr_mat4>(), c.ci, c.ct, 1.0f, N, I, 0xFFFFFFFFu));
        m.obb += ms_since(t0);
        if (sah) {
            ADS_TRY(to_host(h_obbs, obbs.p, total));
            kr_bvh_sah(h_tree.data(), 0, c.ci, c.ct, h_obbs.data(), file);
        }
    }
    DevBuf collapsed;
    if (c.collapse) {
        ADS_TRY(collapsed.alloc((size_t)total * sizeof(kr_bvh_node)));
        kr_ads_collapse_metrics cm = {0.0f, 0.0f};
        ADS_TRY(kr_cuda_bvh_collapse(nodes.as<kr_bvh_node>(), parents.as<u32>(), collapsed.as<kr_bvh_node>(),
                                     counts.as<u32>(), prims.as<u32>(), costs.as<kr_scalar>(), c.ci, c.ct, N, I,
                                     &cm));
        m.mark += cm.collapse_mark_time;
        m.collapse += cm.collapse_time;
        if (sah) {
            ADS_TRY(to_host(h_tree, collapsed.p, total));
            kr_bvh_sah(h_tree.data(), 0, c.ci, c.ct, nullptr, file);
            if (c.use_obbs) kr_bvh_sah(h_tree.data(), 0, c.ci, c.ct, h_obbs.data(), file);
        }
    }
    if (c.use_persistent_kernel) {  // lbvh_cuda_blas_prepare_persistent, lbvh.cu:1516-1545
        const kr_bvh_node* tree = c.collapse ? collapsed.as<kr_bvh_node>() : nodes.as<kr_bvh_node>();
        DevBuf packed, pairs;
        ADS_TRY(packed.alloc((size_t)total * sizeof(kr_bvh_node_packed)));
        ADS_TRY(pairs.alloc((size_t)I * sizeof(kr_bvh_transformation_pair)));
        ADS_TRY(kr_cuda_obvh_persistent(tree, mats.as<kr_mat4>(), packed.as<kr_bvh_node_packed>(),
                                        pairs.as<kr_bvh_transformation_pair>(), total));
        kr_SoA_OBVH* soa = nullptr;
        ADS_TRY(kr_cuda_soa_obvh(packed.as<kr_bvh_node_packed>(), pairs.as<kr_bvh_transformation_pair>(), &soa, I,
                                 total));
        c.blas.soa = soa;
    }
    c.blas.V = V.release<kr_vec3>();
    c.blas.F = F.release<kr_uvec4>();
    c.blas.primitives = prims.release<u32>();
    return nullptr;
}

// lbvh_cuda_commit (lbvh.cu:1764-1858) / atrbvh_cuda_commit: every mesh in
// turn (through the instances when there are any); the last one stays.
kr_error commit(Context& c, kr_scene* scene) {
    if (scene == nullptr) return "korangar_ads_action: COMMIT needs a kr_scene";
    c.scene = scene;
    std::vector<const kr_object*> objs;
    kr_size primitive_count = 0;
    if (scene->instance_count > 0) {
        for (kr_size i = 0; i < scene->instance_count; ++i) {
            const kr_index id = scene->instances[i].object_id;
            if (id >= scene->object_count) return "korangar_ads_action: an instance names a missing object";
            const kr_object* o = &scene->objects[id];
            if (o->type == KR_OBJECT_AABB) primitive_count += 1;
            else if (o->type == KR_OBJECT_MESH) primitive_count += o->as_mesh.face_count;
            objs.push_back(o);
        }
    } else {
        for (kr_size i = 0; i < scene->object_count; ++i) {
            primitive_count += scene->objects[i].as_mesh.face_count;
            objs.push_back(&scene->objects[i]);
        }
    }
    if (primitive_count == 0) return "The scene has no intersectable primitives";
    Metrics m;
    for (const kr_object* o : objs) ADS_TRY(blas_commit(c, o, m));
    if (scene->benchmark_file) {
        // the reference's 14 columns: morton, sort, radix tree, aabbs,
        // collapse mark, collapse, obb projection, candidates, refit,
        // finalize, then the totals.  The builder and the transform run as
        // one call each here: their time is reported in the totals and the
        // per-stage columns are 0.
        FILE* out = std::fopen(scene->benchmark_file, "a");
        if (out == nullptr) return "korangar_ads_action: cannot open the benchmark file";
        const double coll = m.mark + m.collapse;
        std::fprintf(out, "%f, %f, %f, %f, %f, %f, %f, %f, %f, %f, %f, %f, %f, %f\n", 0.0, 0.0, 0.0, 0.0, m.mark,
                     m.collapse, 0.0, 0.0, 0.0, 0.0, m.build, m.obb, coll, m.build + m.obb + coll);
        std::fclose(out);
    }
    return nullptr;
}

// lbvh_cuda_query_intersection (lbvh.cu:46-131).
kr_error query(const Context& c, const kr_intersection_query* q) {
    if (q == nullptr) return "korangar_ads_action: INTERSECTION_QUERY needs a kr_intersection_query";
    if (c.intersect_bounds) return "korangar_ads_action: intersect_bounds (bounds visualisation) is not built";
    if (!c.use_obbs) return nullptr;  // the reference's non-persistent AABB path returns without work
    if (!c.use_persistent_kernel) return "korangar_ads_action: the AoS OBVH traversal (use_persistent_kernel=n) is not built";
    if (c.blas.soa == nullptr) return "korangar_ads_action: INTERSECTION_QUERY before a successful COMMIT";
    if (q->cuda.ray_count > 0xFFFFFFFFu) return "korangar_ads_action: too many rays in one query";
    return kr_cuda_soa_obvh_intersect(c.blas.soa, c.blas.V, c.blas.F, c.blas.primitives, q->cuda.rays,
                                      q->cuda.isects, (u32)q->cuda.ray_count, 0, 0, nullptr);
}

}  // namespace

extern "C" __attribute__((visibility("default"))) kr_error korangar_ads_action(kr_ads* ads, kr_handle descriptor,
                                                                              kr_ads_action action) {
    Context* c = ads != nullptr ? (Context*)ads->context : nullptr;
    try {
        switch (action) {
        case KR_ADS_ACTION_CREATE:
            return (kr_error) new (std::nothrow) Context();
        case KR_ADS_ACTION_NAME:
            return (kr_error)kName;
        case KR_ADS_ACTION_PREFERRED_QUERY_TYPE:
            return (kr_error)(uintptr_t)KR_QUERY_TYPE_CUDA;
        case KR_ADS_ACTION_INIT:
            if (c == nullptr) return "korangar_ads_action: no context (CREATE first)";
            return init(*c, (const kr_descriptor_container*)descriptor);
        case KR_ADS_ACTION_COMMIT:
            if (c == nullptr) return "korangar_ads_action: no context (CREATE first)";
            return commit(*c, (kr_scene*)descriptor);
        case KR_ADS_ACTION_INTERSECTION_QUERY:
            if (c == nullptr) return "korangar_ads_action: no context (CREATE first)";
            return query(*c, (const kr_intersection_query*)descriptor);
        case KR_ADS_ACTION_DESTROY:
            if (c != nullptr) {
                free_blas(c->blas);
                delete c;
                ads->context = nullptr;
            }
            return nullptr;
        default:
            return nullptr;
        }
    } catch (const std::bad_alloc&) {
        return "korangar_ads_action: out of host memory";
    }
}
