// abi.hip -- the extern "C" entry points declared in include/obvh/obvh.h.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "atrbvh.hpp"
#include "collapse.hpp"
#include "common.hpp"
#include "dito.hpp"
#include "lbvh.hpp"
#include "transform.hpp"
#include "traverse.hpp"
#include "upload.hpp"

namespace {

thread_local char g_err[512];
thread_local const char* g_last = nullptr;

struct DeviceCache {
    void* pts = nullptr;
    size_t pts_bytes = 0;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* dws = nullptr;   // workspace of kr_cuda_points_obb (its completion counters stay zero between calls)
    size_t dws_bytes = 0;
    bool dws_clear = true;  // counters not yet zeroed on this allocation
    hipEvent_t ev[2] = {nullptr, nullptr};  // timing of kr_cuda_points_obb
    void* out = nullptr;  // kr_obb3 + error word
    void* trav = nullptr;  // traversal workspace of kr_cuda_soa_obvh_intersect
    size_t trav_bytes = 0;
};
std::mutex g_mu;
std::vector<DeviceCache> g_cache;

// Device kr_SoA_OBVH structs made by kr_cuda_soa_obvh: a host copy of the
// struct and the array sizes, so that the traversal needs no read-back and
// can check indices.
struct SoaInfo {
    kr_SoA_OBVH host;
    uint32_t internal_count, node_count;
};
std::unordered_map<const kr_SoA_OBVH*, SoaInfo> g_soa;

DeviceCache* cache_for_current(int* dev) {
    if (hipGetDevice(dev) != hipSuccess) return nullptr;
    if ((int)g_cache.size() <= *dev) g_cache.resize(*dev + 1);
    return &g_cache[*dev];
}

hipError_t grow(void** p, size_t* cur, size_t need) {
    if (*cur >= need && *p) return hipSuccess;
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
        *p = nullptr;
        *cur = 0;
    }
    size_t n = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(p, n);
    if (e == hipSuccess) *cur = n;
    return e;
}

kr_error fail(const char* msg) {
    g_last = msg;
    return msg;
}

// Workspace of obvh_dito_soa_noalloc.  Each device holds its own copy at its
// own address until the process ends; in host code `scratch` names a host
// stand-in, so kernels get the device address from hipGetSymbolAddress.
// 1 MiB holds the single-set DiTO workspace of up to
// OBVH_DITO_NOALLOC_MAX_POINTS points (at most 4098 tiles of 140 B).
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // device address of `scratch`, per device
std::vector<char> g_scratch_used;  // the first call on a device zeroes the fit's completion counters

// Address of the current device's `scratch`, looked up once per device;
// *first is true for the first call on the device.
kr_error noalloc_scratch(void** out, bool* first) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return obvh_hip_error(e, "obvh_dito_soa_noalloc: hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) {
        g_scratch_ptr.resize(dev + 1, nullptr);
        g_scratch_used.resize(dev + 1, 0);
    }
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return obvh_hip_error(e, "obvh_dito_soa_noalloc: hipGetSymbolAddress");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    *first = !g_scratch_used[dev];
    return nullptr;
}

// The two-launch fit ran on the current device's `scratch`: its counters are zero from now on.
void noalloc_scratch_used() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_used.size() > dev) g_scratch_used[dev] = 1;
}

}  // namespace

void obvh_set_error(const char* msg) { g_last = msg; }

const char* obvh_hip_error(hipError_t e, const char* where) {
    snprintf(g_err, sizeof(g_err), "HIP error %d (%s) at %s", (int)e, hipGetErrorString(e), where);
    g_last = g_err;
    return g_err;
}

#define TRY(expr)                                                   \
    do {                                                            \
        kr_error r__ = (expr);                                      \
        if (r__) { g_last = (const char*)r__; return r__; }         \
    } while (0)

extern "C" {

const char* obvh_last_error(void) { return g_last; }
const char* obvh_version(void) { return "obvh-mi355x 0.1 (gfx950)"; }

// ------------------------------------------------------------ DiTO --

size_t obvh_dito_workspace_size(kr_size count) {
    obvh::DitoWs ws;
    uint32_t tile = obvh::dito_auto_tile(1, count);
    return obvh::dito_carve(&ws, nullptr, 0, 1, count, tile);
}

kr_error obvh_dito_soa(const float* x, const float* y, const float* z, kr_size count, kr_obb3* out_obb,
                       obvh_dito_trace* trace, void* workspace, size_t workspace_bytes, void* stream) {
    if (!out_obb) return fail("obvh_dito_soa: out_obb is NULL");
    if (count > 0 && (!x || !y || !z)) return fail("obvh_dito_soa: NULL point array");
    if (count > 0xFFFFFFF0ull) return fail("obvh_dito_soa: more than 2^32 - 16 points");
    if (count == 0) {  // the empty set's box, written by the set-up kernel
        TRY(obvh::dito_launch(x, y, z, nullptr, nullptr, count, 1, count, out_obb, trace, 0, 0, workspace,
                              workspace_bytes, (hipStream_t)stream));
        return nullptr;
    }
    TRY(obvh::dito_one_launch(x, y, z, nullptr, count, out_obb, trace, workspace, workspace_bytes, true,
                              (hipStream_t)stream));
    return nullptr;
}

// The allocation-free single-set fit: obvh_dito_soa's kernels with the
// workspace in the file-scope device array `scratch` (its address per device
// from hipGetSymbolAddress, taken once per device on the first call).
kr_error obvh_dito_soa_noalloc(const float* x, const float* y, const float* z, kr_size count, kr_obb3* out_obb,
                               obvh_dito_trace* trace, void* stream) {
    if (!out_obb) return fail("obvh_dito_soa_noalloc: out_obb is NULL");
    if (count > 0 && (!x || !y || !z)) return fail("obvh_dito_soa_noalloc: NULL point array");
    if (count > OBVH_DITO_NOALLOC_MAX_POINTS)
        return fail("obvh_dito_soa_noalloc: more than OBVH_DITO_NOALLOC_MAX_POINTS points");
    obvh::DitoWs wsd;
    const size_t need = obvh::dito_carve(&wsd, nullptr, 0, 1, count, obvh::dito_auto_tile(1, count));
    if (need > kNoAllocScratchBytes) return fail("obvh_dito_soa_noalloc: the static scratch array is too small");
    void* ws = nullptr;
    bool first = false;
    TRY(noalloc_scratch(&ws, &first));
    if (count == 0) {
        TRY(obvh::dito_launch(x, y, z, nullptr, nullptr, count, 1, count, out_obb, trace, 0, 0, ws,
                              kNoAllocScratchBytes, (hipStream_t)stream));
        return nullptr;
    }
    // the completion counters keep a fixed place in the array; the launches
    // leave them zero, so only the first call on a device clears them
    TRY(obvh::dito_one_launch(x, y, z, nullptr, count, out_obb, trace, ws, kNoAllocScratchBytes, first,
                              (hipStream_t)stream));
    if (first) noalloc_scratch_used();
    return nullptr;
}

size_t obvh_dito_batch_workspace_size(u32 nsets, kr_size total_points, u32 tile_points) {
    obvh::DitoWs ws;
    uint32_t tile = tile_points ? tile_points : obvh::dito_auto_tile(nsets, total_points);
    return obvh::dito_carve(&ws, nullptr, 0, nsets, total_points, tile);
}

kr_error obvh_dito_batch(const float* x, const float* y, const float* z, const kr_size* offsets, u32 nsets,
                         kr_size total_points, kr_obb3* out_obbs, obvh_dito_trace* traces, u32 tile_points,
                         u32 workers, void* workspace, size_t workspace_bytes, void* stream) {
    if (nsets == 0) return nullptr;
    if (!offsets || !out_obbs) return fail("obvh_dito_batch: NULL offsets or output");
    if (total_points > 0 && (!x || !y || !z)) return fail("obvh_dito_batch: NULL point array");
    TRY(obvh::dito_launch(x, y, z, nullptr, (const uint64_t*)offsets, 0, nsets, total_points, out_obbs, traces,
                          tile_points, workers, workspace, workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

// Upload of many host point sets into one SoA batch: one 2-D copy per set
// (its x, y and z rows into the three rows of the (3, row_stride) array).
kr_error obvh_upload_sets(const float* const* host_sets, const kr_size* offsets, u32 nsets, float* soa,
                          kr_size row_stride, void* stream) {
    if (nsets == 0) return nullptr;
    if (!host_sets || !offsets || !soa) return fail("obvh_upload_sets: NULL argument");
    if (offsets[nsets] > row_stride) return fail("obvh_upload_sets: offsets exceed the row stride");
    for (u32 s = 0; s < nsets; ++s) {
        if (offsets[s + 1] < offsets[s]) return fail("obvh_upload_sets: offsets must not decrease");
        const size_t n = (size_t)(offsets[s + 1] - offsets[s]);
        if (n == 0) continue;
        if (!host_sets[s]) return fail("obvh_upload_sets: NULL set");
        OBVH_HIP_TRY(hipMemcpy2DAsync(soa + offsets[s], row_stride * sizeof(float), host_sets[s], n * sizeof(float),
                                      n * sizeof(float), 3, hipMemcpyHostToDevice, (hipStream_t)stream));
    }
    return nullptr;
}

kr_error obvh_gather_sets(const float* const* host_sets, const float* const* set_table, const kr_size* offsets,
                          const kr_size* offsets_table, u32 nsets, float* soa, kr_size row_stride, void* stream) {
    if (nsets == 0) return nullptr;
    if (!host_sets || !set_table || !offsets || !offsets_table || !soa) return fail("obvh_gather_sets: NULL argument");
    if (offsets[0] != 0) return fail("obvh_gather_sets: offsets[0] must be 0");
    if (offsets[nsets] > row_stride) return fail("obvh_gather_sets: offsets exceed the row stride");
    for (u32 s = 0; s < nsets; ++s) {
        if (offsets[s + 1] < offsets[s]) return fail("obvh_gather_sets: offsets must not decrease");
        if (offsets[s + 1] == offsets[s]) continue;
        // the kernel reads the set over PCIe: it must be pinned (device-accessible) host memory
        hipPointerAttribute_t attr;
        if (!host_sets[s] || hipPointerGetAttributes(&attr, host_sets[s]) != hipSuccess ||
            (attr.type != hipMemoryTypeHost && attr.type != hipMemoryTypeDevice)) {
            (void)hipGetLastError();
            return fail("obvh_gather_sets: a set is not pinned host or device memory");
        }
    }
    TRY(obvh::gather_sets_launch(set_table, (const uint64_t*)offsets_table, nsets, soa, row_stride, offsets[nsets],
                                 (hipStream_t)stream));
    return nullptr;
}

// Legacy drop-in: host AoS points in, box by value (dito.cuh:12-13).
kr_obb3 kr_cuda_points_obb(const kr_vec3* points, kr_size count, kr_scalar* elapsed_time) {
    kr_obb3 nan_box;
    float* nb = reinterpret_cast<float*>(&nan_box);
    for (int i = 0; i < 15; ++i) nb[i] = NAN;
    if (elapsed_time) *elapsed_time = 0.0f;
    if (count == 0) {  // DiTO of no points + extent floor
        kr_obb3 o;
        o.mid = kr_vec3{0, 0, 0};
        o.v0 = kr_vec3{1, 0, 0};
        o.v1 = kr_vec3{0, 1, 0};
        o.v2 = kr_vec3{0, 0, 1};
        o.ext = kr_vec3{0.001f, 0.001f, 0.001f};
        return o;
    }
    if (!points || count > 0xFFFFFFF0ull) { fail("kr_cuda_points_obb: bad arguments"); return nan_box; }
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    DeviceCache* c = cache_for_current(&dev);
    if (!c) { fail("kr_cuda_points_obb: no HIP device"); return nan_box; }
    obvh::DitoWs wsd;
    const uint32_t tile = obvh::dito_auto_tile(1, count);
    const size_t wsb = obvh::dito_carve(&wsd, nullptr, 0, 1, count, tile);
    hipError_t e;
    const bool realloc_ws = c->dws_bytes < wsb || !c->dws;
    if ((e = grow(&c->pts, &c->pts_bytes, count * sizeof(kr_vec3))) != hipSuccess ||
        (e = grow(&c->dws, &c->dws_bytes, wsb)) != hipSuccess) {
        obvh_hip_error(e, "kr_cuda_points_obb: hipMalloc");
        return nan_box;
    }
    if (realloc_ws) c->dws_clear = true;
    if (!c->out && (e = hipMalloc(&c->out, 256)) != hipSuccess) {
        obvh_hip_error(e, "kr_cuda_points_obb: hipMalloc");
        return nan_box;
    }
    for (auto& ev : c->ev) {
        if (!ev && (e = hipEventCreate(&ev)) != hipSuccess) {
            obvh_hip_error(e, "kr_cuda_points_obb: hipEventCreate");
            return nan_box;
        }
    }
    if ((e = hipMemcpy(c->pts, points, count * sizeof(kr_vec3), hipMemcpyHostToDevice)) != hipSuccess) {
        obvh_hip_error(e, "kr_cuda_points_obb: hipMemcpy H2D");
        return nan_box;
    }
    hipEventRecord(c->ev[0], nullptr);
    // two launches; a fresh workspace first gets its completion counters zeroed
    kr_error r = obvh::dito_one_launch(nullptr, nullptr, nullptr, (const float*)c->pts, count, (kr_obb3*)c->out,
                                       nullptr, c->dws, c->dws_bytes, c->dws_clear, nullptr);
    hipEventRecord(c->ev[1], nullptr);
    kr_obb3 o = nan_box;
    if (!r) {
        c->dws_clear = false;
        if ((e = hipMemcpy(&o, c->out, sizeof(kr_obb3), hipMemcpyDeviceToHost)) != hipSuccess) {
            obvh_hip_error(e, "kr_cuda_points_obb: hipMemcpy D2H");
            o = nan_box;
        } else if (elapsed_time) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
            *elapsed_time = ms;
        }
    } else {
        g_last = (const char*)r;
    }
    return o;
}

// ------------------------------------------------------- transform --

size_t obvh_obb_tree_workspace_size(u32 primitive_count, u32 internal_count) {
    (void)internal_count;
    obvh::TreeWs ws;
    return obvh::tree_carve(&ws, nullptr, 0, primitive_count);
}

kr_error obvh_obb_tree_async(const kr_bvh_node* nodes, const u32* parents, const kr_vec3* vertices,
                             const kr_uvec4* faces, kr_obb3* obbs, kr_mat4* transforms, u32* kdop_ids,
                             u32 primitive_count, u32 internal_count, void* workspace, size_t workspace_bytes,
                             void* stream) {
    if (primitive_count == 0) return nullptr;
    if (internal_count != primitive_count - 1)
        return fail("obvh_obb_tree: internal_count must be primitive_count - 1");
    if (primitive_count > 0x7FFFFFFFu) return fail("obvh_obb_tree: too many primitives");
    if (!nodes || !parents || !vertices || !faces || !obbs || !transforms)
        return fail("obvh_obb_tree: NULL array");
    TRY(obvh::tree_launch(nodes, parents, vertices, faces, obbs, transforms, kdop_ids, primitive_count, workspace,
                          workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

// Legacy drop-in (transforms.cuh:52-65): synchronous, scratch owned here.
kr_error kr_cuda_bvh_obb_tree(kr_bvh_node* nodes, const u32* parents, const u32* primitive_counts,
                              const kr_scalar* costs, const kr_vec3* vertices, const kr_uvec4* faces,
                              const u32* primitives, kr_obb3* obbs, kr_mat4* transforms, kr_scalar ci,
                              kr_scalar ct, kr_scalar obb_cost, u32 primitive_count, u32 internal_count,
                              u32 primitive_threshold) {
    (void)primitive_counts; (void)costs; (void)primitives; (void)ci; (void)ct; (void)obb_cost;
    (void)primitive_threshold;
    if (primitive_count == 0) return nullptr;
    if (internal_count != primitive_count - 1)
        return fail("kr_cuda_bvh_obb_tree: internal_count must be primitive_count - 1");
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    DeviceCache* c = cache_for_current(&dev);
    if (!c) return fail("kr_cuda_bvh_obb_tree: no HIP device");
    obvh::TreeWs wsd;
    size_t need = obvh::tree_carve(&wsd, nullptr, 0, primitive_count);
    OBVH_HIP_TRY(grow(&c->ws, &c->ws_bytes, need));
    TRY(obvh::tree_launch(nodes, parents, vertices, faces, obbs, transforms, nullptr, primitive_count, c->ws,
                          c->ws_bytes, nullptr));
    obvh::TreeWs ws;
    obvh::tree_carve(&ws, c->ws, c->ws_bytes, primitive_count);
    uint32_t err = 0;
    OBVH_HIP_TRY(hipMemcpy(&err, ws.error, sizeof(err), hipMemcpyDeviceToHost));  // synchronises
    if (err & 2u) return fail("kr_cuda_bvh_obb_tree: parent/child links disagree");
    if (err & 1u) return fail("kr_cuda_bvh_obb_tree: parent index out of range or cycle");
    if (err & 4u) return fail("kr_cuda_bvh_obb_tree: leaf face id out of range");
    if (err & 8u) return fail("kr_cuda_bvh_obb_tree: internal node never completed (malformed tree)");
    return nullptr;
}

kr_error obvh_obb_tree_status_async(const void* workspace, size_t workspace_bytes, u32 primitive_count,
                                    u32* status, void* stream) {
    if (!status) return fail("obvh_obb_tree_status_async: status is NULL");
    if (primitive_count == 0) return fail("obvh_obb_tree_status_async: no tree has zero primitives");
    TRY(obvh::tree_status_launch(workspace, workspace_bytes, primitive_count, status, (hipStream_t)stream));
    return nullptr;
}

kr_error obvh_obb_tree_status_string(u32 status) { return obvh::tree_status_string(status); }

size_t obvh_obb_tree_batch_workspace_size(u32 ntrees, u32 total_primitives) {
    (void)ntrees;
    if (total_primitives == 0) return 256;
    return obvh::batch_workspace_size(total_primitives);
}

kr_error obvh_obb_tree_batch(const kr_bvh_node* nodes, const u32* parents, const kr_vec3* vertices,
                             const kr_uvec4* faces, const u32* primitive_offsets, const u32* vertex_offsets,
                             u32 ntrees, u32 total_primitives, kr_obb3* obbs, kr_mat4* transforms, void* workspace,
                             size_t workspace_bytes, void* stream) {
    if (ntrees == 0 || total_primitives == 0) return nullptr;
    if (total_primitives > 0x7FFFFFFFu) return fail("obvh_obb_tree_batch: too many primitives");
    if (!nodes || !parents || !vertices || !faces || !primitive_offsets || !vertex_offsets || !obbs || !transforms)
        return fail("obvh_obb_tree_batch: NULL array");
    TRY(obvh::tree_batch_launch(nodes, parents, vertices, faces, primitive_offsets, vertex_offsets, ntrees,
                                total_primitives, obbs, transforms, workspace, workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

kr_error obvh_obb_tree_batch_status_async(const void* workspace, size_t workspace_bytes, u32 total_primitives,
                                          u32* status, void* stream) {
    if (!status) return fail("obvh_obb_tree_batch_status_async: status is NULL");
    if (total_primitives == 0) return fail("obvh_obb_tree_batch_status_async: no batch has zero primitives");
    TRY(obvh::tree_batch_status_launch(workspace, workspace_bytes, total_primitives, status, (hipStream_t)stream));
    return nullptr;
}

kr_error obvh_validate_tree_host(const kr_bvh_node* nodes, const u32* parents, u32 primitive_count,
                                 u32 internal_count) {
    if (primitive_count == 0) return nullptr;
    if (internal_count != primitive_count - 1) return fail("tree: internal_count != primitive_count - 1");
    const uint32_t I = internal_count, total = I + primitive_count;
    if (parents[0] != 0xFFFFFFFFu) return fail("tree: root (node 0) has a parent");
    for (uint32_t i = 0; i < I; ++i) {
        if (nodes[i].nPrimitives != 0) return fail("tree: internal node with primitives");
        uint32_t l = nodes[i].u.children.left, r = nodes[i].u.children.right;
        if (l >= total || r >= total || l == 0 || r == 0 || l == r) return fail("tree: bad child index");
        if (parents[l] != i || parents[r] != i) return fail("tree: parent/child links disagree");
    }
    for (uint32_t i = I; i < total; ++i)
        if (nodes[i].nPrimitives == 0) return fail("tree: leaf without primitives");
    // every node reachable exactly once from the root
    std::vector<uint32_t> stack;
    std::vector<uint8_t> seen(total, 0);
    stack.push_back(0);
    uint32_t count = 0;
    while (!stack.empty()) {
        uint32_t v = stack.back();
        stack.pop_back();
        if (seen[v]) return fail("tree: node reached twice");
        seen[v] = 1;
        ++count;
        if (v < I) {
            stack.push_back(nodes[v].u.children.left);
            stack.push_back(nodes[v].u.children.right);
        }
    }
    if (count != total) return fail("tree: unreachable nodes");
    return nullptr;
}

// ------------------------------------------------------------ LBVH --

size_t obvh_lbvh_workspace_size(u32 face_count) {
    obvh::LbvhWs ws;
    return obvh::lbvh_carve(&ws, nullptr, 0, face_count);
}

kr_error obvh_lbvh_build(const kr_vec3* vertices, const kr_uvec4* faces, u32 face_count, kr_bvh_node* nodes,
                         u32* parents, u32* primitive_counts, kr_scalar* costs, u32* mortons_sorted,
                         void* workspace, size_t workspace_bytes, void* stream) {
    if (!vertices || !faces || !nodes || !parents || !primitive_counts || !costs)
        return fail("obvh_lbvh_build: NULL array");
    if (face_count > 0x7FFFFFFFu) return fail("obvh_lbvh_build: too many faces");
    TRY(obvh::lbvh_launch(vertices, faces, face_count, nodes, parents, primitive_counts, costs, mortons_sorted,
                          workspace, workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

// -------------------------------------------------------- collapse --

size_t obvh_collapse_workspace_size(u32 primitive_count) {
    obvh::CollapseWs ws;
    return obvh::collapse_carve(&ws, nullptr, 0, primitive_count);
}

kr_error obvh_collapse_async(const kr_bvh_node* nodes, const u32* parents, kr_bvh_node* collapsed_nodes,
                             u32* primitive_counts, u32* primitives, kr_scalar* costs, kr_scalar ci, kr_scalar ct,
                             u32 primitive_count, void* workspace, size_t workspace_bytes, void* stream) {
    if (primitive_count == 0) return nullptr;
    if (primitive_count > 0x7FFFFFFFu) return fail("obvh_collapse: too many primitives");
    if (!nodes || !parents || !collapsed_nodes || !primitive_counts || !primitives || !costs)
        return fail("obvh_collapse: NULL array");
    TRY(obvh::collapse_launch(nodes, parents, collapsed_nodes, primitive_counts, primitives, costs, ci, ct,
                              primitive_count, workspace, workspace_bytes, (hipStream_t)stream, nullptr));
    return nullptr;
}

// Legacy drop-in (transforms.cuh:40-50): synchronous, scratch owned here,
// metrics (nullable) receive the mark and collapse times in ms.
kr_error kr_cuda_bvh_collapse(const kr_bvh_node* nodes, const u32* parents, kr_bvh_node* collapsed_nodes,
                              u32* primitive_counts, u32* primitives, kr_scalar* costs, kr_scalar ci, kr_scalar ct,
                              u32 leaf_count, u32 internal_count, kr_ads_collapse_metrics* metrics) {
    if (leaf_count == 0) return nullptr;
    if (internal_count != leaf_count - 1)
        return fail("kr_cuda_bvh_collapse: internal_count must be leaf_count - 1");
    if (leaf_count > 0x7FFFFFFFu) return fail("kr_cuda_bvh_collapse: too many primitives");
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    DeviceCache* c = cache_for_current(&dev);
    if (!c) return fail("kr_cuda_bvh_collapse: no HIP device");
    obvh::CollapseWs wsd;
    const size_t need = obvh::collapse_carve(&wsd, nullptr, 0, leaf_count);
    OBVH_HIP_TRY(grow(&c->ws, &c->ws_bytes, need));
    hipEvent_t ev[3];
    for (auto& e : ev) OBVH_HIP_TRY(hipEventCreate(&e));
    OBVH_HIP_TRY(hipEventRecord(ev[0], nullptr));
    kr_error r = obvh::collapse_launch(nodes, parents, collapsed_nodes, primitive_counts, primitives, costs, ci, ct,
                                       leaf_count, c->ws, c->ws_bytes, nullptr, ev[1]);
    hipEventRecord(ev[2], nullptr);
    uint32_t err = 0;
    hipError_t he = hipSuccess;
    if (!r) {
        obvh::CollapseWs ws;
        obvh::collapse_carve(&ws, c->ws, c->ws_bytes, leaf_count);
        he = hipMemcpy(&err, ws.error, sizeof(err), hipMemcpyDeviceToHost);  // synchronises
        if (he == hipSuccess && metrics) {
            float a = 0.0f, b = 0.0f;
            hipEventElapsedTime(&a, ev[0], ev[1]);
            hipEventElapsedTime(&b, ev[1], ev[2]);
            metrics->collapse_mark_time += a;
            metrics->collapse_time += b;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    TRY(r);
    if (he != hipSuccess) return obvh_hip_error(he, "kr_cuda_bvh_collapse: hipMemcpy D2H");
    if (err) return fail("kr_cuda_bvh_collapse: malformed subtree below a collapsed node");
    return nullptr;
}

// kr_bvh_sah (src/ads/common/transforms.cpp:250-303), host arrays: a LIFO
// walk from the root summing ci*area over internal nodes and ct*n*area over
// leaves (the AABB area for one-primitive leaves even when OBBs are given),
// skipping nodes marked 0xFFFF, divided by the root area; the quality is
// appended as "%f," to outputFile when it is not NULL.
kr_scalar kr_bvh_sah(const kr_bvh_node* nodes, kr_i32 root_idx, kr_scalar ci, kr_scalar ct, const kr_obb3* obbs,
                     const char* outputFile) {
    auto aabb_area = [](const kr_bvh_node& n) {
        const float dx = n.bounds.max.x - n.bounds.min.x, dy = n.bounds.max.y - n.bounds.min.y,
                    dz = n.bounds.max.z - n.bounds.min.z;
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    };
    auto obb_area = [](const kr_obb3& o) {
        const float dx = 2.0f * o.ext.x, dy = 2.0f * o.ext.y, dz = 2.0f * o.ext.z;
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    };
    std::vector<int32_t> stack;
    const kr_bvh_node* root = &nodes[root_idx];
    float sah = 0.0f;
    const float root_area = obbs ? obb_area(obbs[root_idx]) : aabb_area(*root);
    if (root->nPrimitives > 0) sah += ct * root_area * root->nPrimitives;
    else stack.push_back(root_idx);
    while (!stack.empty()) {
        const int32_t idx = stack.back();
        stack.pop_back();
        const kr_bvh_node& node = nodes[idx];
        float area = obbs ? obb_area(obbs[idx]) : aabb_area(node);
        if (node.nPrimitives > 0) {
            if (node.nPrimitives == 0xFFFF) continue;
            if (node.nPrimitives == 1) area = aabb_area(node);
            sah += ct * node.nPrimitives * area;
        } else {
            sah += ci * area;
            stack.push_back((int32_t)node.u.children.left);
            stack.push_back((int32_t)node.u.children.right);
        }
    }
    const float quality = sah / root_area;
    if (outputFile) {
        FILE* f = fopen(outputFile, "a");
        if (f) {
            fprintf(f, "%f,", quality);
            fclose(f);
        }
    }
    return quality;
}

// ---------------------------------------------------------- ATRBVH --

size_t obvh_atrbvh_workspace_size(u32 face_count) {
    obvh::AtrWs ws;
    return obvh::atr_carve(&ws, nullptr, 0, face_count);
}

kr_error obvh_atrbvh_build(const kr_vec3* vertices, const kr_uvec4* faces, u32 face_count, int treelet_size,
                           int iterations, kr_scalar ci, kr_scalar ct, kr_bvh_node* nodes, u32* parents,
                           u32* primitive_counts, kr_scalar* costs, void* workspace, size_t workspace_bytes,
                           void* stream) {
    if (!vertices || !faces || !nodes || !parents || !primitive_counts || !costs)
        return fail("obvh_atrbvh_build: NULL array");
    if (face_count > 0x7FFFFFFFu) return fail("obvh_atrbvh_build: too many faces");
    TRY(obvh::atr_launch(vertices, faces, face_count, treelet_size, iterations, ci, ct, nodes, parents,
                         primitive_counts, costs, workspace, workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

// ------------------------------------------------------- OBVH --

// Legacy drop-in (transforms.cuh:32-38): AoS packing, synchronous.
kr_error kr_cuda_obvh_persistent(const kr_bvh_node* nodes, const kr_mat4* transforms, kr_bvh_node_packed* packed_nodes,
                                 kr_bvh_transformation_pair* packed_transforms, u32 node_count) {
    if (node_count == 0) return nullptr;
    if (!nodes || !transforms || !packed_nodes || !packed_transforms)
        return fail("kr_cuda_obvh_persistent: NULL array");
    if (node_count > 0x7FFFFFFEu) return fail("kr_cuda_obvh_persistent: too many nodes");
    TRY(obvh::aos_pack_launch(nodes, transforms, packed_nodes, packed_transforms, node_count, nullptr));
    OBVH_HIP_TRY(hipStreamSynchronize(nullptr));
    return nullptr;
}

// Legacy drop-in (transforms.cuh:72-77): allocates the SoA arrays and a
// device copy of the struct, converts on the device, synchronous.
kr_error kr_cuda_soa_obvh(const kr_bvh_node_packed* packed_bvh, const kr_bvh_transformation_pair* packed_transforms,
                          kr_SoA_OBVH** soa_bvh, u32 internal_count, u32 node_count) {
    if (!soa_bvh) return fail("kr_cuda_soa_obvh: soa_bvh is NULL");
    if (!packed_bvh || !packed_transforms) return fail("kr_cuda_soa_obvh: NULL array");
    if (node_count > 0x7FFFFFFEu || internal_count > node_count)
        return fail("kr_cuda_soa_obvh: bad sizes");
    kr_SoA_OBVH h{};
    kr_SoA_OBVH* d = nullptr;
    auto cleanup = [&]() {
        hipFree(h.left_T);
        hipFree(h.right_T);
        hipFree(h.metadata);
        hipFree(d);
    };
    hipError_t e = hipMalloc((void**)&h.left_T, sizeof(kr_mat43) * (internal_count ? internal_count : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&h.right_T, sizeof(kr_mat43) * (internal_count ? internal_count : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&h.metadata, sizeof(kr_ivec4) * (node_count ? node_count : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&d, sizeof(kr_SoA_OBVH));
    if (e != hipSuccess) {
        cleanup();
        return obvh_hip_error(e, "kr_cuda_soa_obvh: hipMalloc");
    }
    kr_error r = obvh::aos_to_soa_launch(packed_bvh, packed_transforms, internal_count, node_count, h.left_T,
                                         h.right_T, h.metadata, nullptr);
    if (!r) {
        e = hipMemcpy(d, &h, sizeof(h), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) r = obvh_hip_error(e, "kr_cuda_soa_obvh");
    }
    if (r) {
        cleanup();
        TRY(r);
    }
    {
        std::lock_guard<std::mutex> lock(g_mu);
        g_soa[d] = SoaInfo{h, internal_count, node_count};
    }
    *soa_bvh = d;
    return nullptr;
}

kr_error obvh_soa_obvh_free(kr_SoA_OBVH* soa_bvh) {
    if (!soa_bvh) return nullptr;
    kr_SoA_OBVH h{};
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_soa.find(soa_bvh);
        if (it == g_soa.end()) return fail("obvh_soa_obvh_free: not made by kr_cuda_soa_obvh");
        h = it->second.host;
        g_soa.erase(it);
    }
    hipError_t e = hipFree(h.left_T);
    if (e == hipSuccess) e = hipFree(h.right_T);
    if (e == hipSuccess) e = hipFree(h.metadata);
    if (e == hipSuccess) e = hipFree(soa_bvh);
    if (e != hipSuccess) return obvh_hip_error(e, "obvh_soa_obvh_free");
    return nullptr;
}

// Legacy drop-in (intersectors.cuh:81-91).  For a struct made by
// kr_cuda_soa_obvh the array sizes are known and every index is checked;
// for any other device struct it is read back and only the face/vertex
// checks are off (sizes unknown), with a visit cap against cycles.
kr_error kr_cuda_soa_obvh_intersect(const kr_SoA_OBVH* bvh, const kr_vec3* vertices, const kr_uvec4* faces,
                                    const u32* primitives, const kr_ray* rays, kr_intersection* isects, u32 ray_count,
                                    u32 block_size, u32 grid_size, u32* warp_counter) {
    (void)block_size;
    (void)grid_size;
    (void)warp_counter;
    if (ray_count == 0) return nullptr;
    if (!bvh || !vertices || !faces || !rays || !isects) return fail("kr_cuda_soa_obvh_intersect: NULL argument");
    std::lock_guard<std::mutex> lock(g_mu);
    obvh::TravCounts counts{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 1u << 24};
    kr_SoA_OBVH h{};
    auto it = g_soa.find(bvh);
    if (it != g_soa.end()) {
        h = it->second.host;
        counts.internal_count = it->second.internal_count;
        counts.node_count = it->second.node_count;
        counts.max_steps = it->second.node_count;
    } else {
        OBVH_HIP_TRY(hipMemcpy(&h, bvh, sizeof(h), hipMemcpyDeviceToHost));
    }
    int dev = 0;
    DeviceCache* c = cache_for_current(&dev);
    if (!c) return fail("kr_cuda_soa_obvh_intersect: no HIP device");
    obvh::TravWs ws;
    OBVH_HIP_TRY(grow(&c->trav, &c->trav_bytes, obvh::trav_carve(&ws, nullptr, 0, ray_count)));
    TRY(obvh::intersect_launch(h.left_T, h.right_T, h.metadata, vertices, faces, primitives, rays, isects, ray_count,
                               counts, c->trav, c->trav_bytes, nullptr));
    return nullptr;
}

kr_error obvh_soa_pack_async(const kr_bvh_node* nodes, const kr_mat4* transforms, u32 node_count, u32 internal_count,
                             kr_mat43* left_T, kr_mat43* right_T, kr_ivec4* metadata, void* stream) {
    if (node_count == 0) return nullptr;
    if (!nodes || !transforms || !metadata || (internal_count && (!left_T || !right_T)))
        return fail("obvh_soa_pack: NULL array");
    if (node_count > 0x7FFFFFFEu || internal_count > node_count) return fail("obvh_soa_pack: bad sizes");
    TRY(obvh::soa_pack_launch(nodes, transforms, node_count, internal_count, left_T, right_T, metadata,
                              (hipStream_t)stream));
    return nullptr;
}

size_t obvh_soa_intersect_workspace_size(u32 ray_count) {
    obvh::TravWs ws;
    return obvh::trav_carve(&ws, nullptr, 0, ray_count);
}

kr_error obvh_soa_intersect_async(const kr_mat43* left_T, const kr_mat43* right_T, const kr_ivec4* metadata,
                                  u32 internal_count, u32 node_count, const kr_vec3* vertices, u32 vertex_count,
                                  const kr_uvec4* faces, const u32* primitives, u32 face_count, const kr_ray* rays,
                                  kr_intersection* isects, u32 ray_count, void* workspace, size_t workspace_bytes,
                                  void* stream) {
    if (ray_count == 0) return nullptr;
    if (!left_T || !right_T || !metadata || !vertices || !faces || !rays || !isects)
        return fail("obvh_soa_intersect: NULL argument");
    if (internal_count == 0 || node_count == 0) return fail("obvh_soa_intersect: the tree needs an internal root");
    obvh::TravCounts counts{internal_count, node_count, face_count, vertex_count, node_count};
    TRY(obvh::intersect_launch(left_T, right_T, metadata, vertices, faces, primitives, rays, isects, ray_count, counts,
                               workspace, workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

kr_error obvh_records_pack_async(const kr_bvh_node* nodes, const kr_mat4* transforms, const u32* primitives,
                                 const kr_vec3* vertices, u32 vertex_count, const kr_uvec4* faces, u32 face_count,
                                 obvh_node_record* node_records, obvh_tri_record* tri_records, void* stream) {
    if (!nodes || !transforms || !vertices || !faces || !node_records || !tri_records)
        return fail("obvh_records_pack: NULL array");
    TRY(obvh::rec_pack_launch(nodes, transforms, primitives, vertices, faces, face_count, vertex_count, node_records,
                              tri_records, (hipStream_t)stream));
    return nullptr;
}

kr_error obvh_records_intersect_async(const obvh_node_record* node_records, const obvh_tri_record* tri_records,
                                      u32 face_count, const kr_ray* rays, kr_intersection* isects, u32 ray_count,
                                      void* workspace, size_t workspace_bytes, void* stream) {
    if (ray_count == 0) return nullptr;
    if (!node_records || !tri_records || !rays || !isects) return fail("obvh_records_intersect: NULL argument");
    TRY(obvh::rec_intersect_launch(node_records, tri_records, face_count, rays, isects, ray_count, workspace,
                                   workspace_bytes, (hipStream_t)stream));
    return nullptr;
}

}  // extern "C"
