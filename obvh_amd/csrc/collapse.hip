// collapse.hip -- BVH collapse on gfx950 (kr_cuda_bvh_collapse,
// src/ads/common/cuda/transforms.cu:18-211,1061-1125).
//
//  collapse_mark    bottom-up (phased arrival protocol, phased.hpp): leaf
//                   cost ct * area; an internal node with count = counts of
//                   its children, node_cost = ci*area + cost(left) +
//                   cost(right), leaf_cost = ct*area*count is collapsed when
//                   leaf_cost < node_cost and count <= 8 (cost = leaf_cost,
//                   the node and both children flagged, both children get
//                   nPrimitives = 0xFFFF in the collapsed copy).
//  collapse_roots   one lane per node: a flagged node with more than one
//                   primitive whose parent is not flagged becomes a leaf of
//                   the collapsed tree; its primitive count goes to a scan.
//  (scan)           exclusive scan of those counts: each new leaf's slice of
//                   `primitives`, in node-index order.  The reference takes
//                   slices from an atomic counter in arrival order; the ids
//                   inside a slice are the same.
//  collapse_gather  one lane per new leaf: the reference's right-first
//                   stack walk over the subtree collecting distinct
//                   primitive ids (bvh_node_collapse, :18-61).
#include <float.h>

#include <rocprim/device/device_scan.hpp>

#include "collapse.hpp"
#include "common.hpp"

namespace obvh {

constexpr uint16_t kCollapsed = 0xFFFF;
constexpr uint32_t kMaxLeafSize = 8;

// kr_aabb_surface_area3, vecmath.h:1583-1596.
__device__ __forceinline__ float aabb_area(const kr_aabb3& b) {
    const float dx = b.max.x - b.min.x, dy = b.max.y - b.min.y, dz = b.max.z - b.min.z;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

__global__ void collapse_clear(CollapseWs ws, uint32_t nf) {
    phase_clear(ws.ph, nf);
    if (blockIdx.x == 0 && threadIdx.x == 0) *ws.error = 0;
    const uint32_t total = 2 * nf - 1;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
        ws.flag[i] = 0;
}

struct CollapseUp {
    const kr_bvh_node* nodes;
    kr_bvh_node* out;
    uint32_t* counts;
    float* costs;
    uint8_t* flag;
    float ci, ct;
    __device__ void leaf(uint32_t cur) { costs[cur] = ct * aabb_area(nodes[cur].bounds); }
    __device__ void merge(uint32_t p, uint32_t l, uint32_t r) {
        const uint32_t count = counts[l] + counts[r];
        counts[p] = count;
        const float area = aabb_area(nodes[p].bounds);
        const float node_cost = ci * area + costs[l] + costs[r];
        const float leaf_cost = ct * area * (float)count;
        if (leaf_cost < node_cost && count <= kMaxLeafSize) {
            costs[p] = leaf_cost;
            flag[p] = 1;
            flag[l] = 1;
            flag[r] = 1;
            out[l].nPrimitives = kCollapsed;
            out[r].nPrimitives = kCollapsed;
        } else {
            costs[p] = node_cost;
        }
    }
};

__global__ __launch_bounds__(1024) void collapse_mark(const kr_bvh_node* nodes, const uint32_t* parents,
                                                      kr_bvh_node* out, uint32_t* counts, float* costs, float ci,
                                                      float ct, uint32_t nf, CollapseWs ws, uint32_t phase,
                                                      uint32_t chunk) {
    CollapseUp op{nodes, out, counts, costs, ws.flag, ci, ct};
    phased_up(nodes, parents, nf, ws.ph, phase, chunk, op);
}

__global__ void collapse_roots(const uint32_t* parents, const uint32_t* counts, CollapseWs ws, uint32_t total) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const uint32_t p = parents[c];
    const bool root = p < total && !ws.flag[p] && ws.flag[c] && counts[c] > 1;
    ws.rootcnt[c] = root ? counts[c] : 0u;
}

__global__ void collapse_gather(const kr_bvh_node* nodes, kr_bvh_node* out, uint32_t* counts, uint32_t* primitives,
                                CollapseWs ws, uint32_t total) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total || ws.rootcnt[c] == 0) return;
    const uint32_t offset = ws.offset[c];
    const uint32_t cap = ws.rootcnt[c];
    uint32_t* buf = primitives + offset;
    uint32_t stack[2 * kMaxLeafSize];
    int sp = 0;
    stack[sp++] = nodes[c].u.children.left;
    stack[sp++] = nodes[c].u.children.right;
    uint32_t position = 0;
    bool bad = false;
    while (sp > 0) {
        const uint32_t u = stack[--sp];
        if (u >= total) { bad = true; break; }
        const kr_bvh_node n = nodes[u];
        if (n.nPrimitives > 0) {
            const uint32_t data = n.u.primitivesOffset;
            bool unique = true;
            for (uint32_t i = 0; i < position; ++i)
                if (buf[i] == data) { unique = false; break; }
            if (unique) {
                if (position >= cap) { bad = true; break; }
                buf[position++] = data;
            }
        } else {
            if (sp + 2 > (int)(2 * kMaxLeafSize)) { bad = true; break; }
            stack[sp++] = n.u.children.left;
            stack[sp++] = n.u.children.right;
        }
    }
    if (bad) atomicOr(ws.error, 1u);
    counts[c] = position;
    out[c].nPrimitives = (uint16_t)position;
    out[c].u.primitivesOffset = offset;
}

// ------------------------------------------------------------- host --

static size_t collapse_scan_bytes(uint32_t total) {
    size_t bytes = 0;
    rocprim::exclusive_scan((void*)nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, 0u, total,
                            rocprim::plus<uint32_t>());
    return bytes;
}

size_t collapse_carve(CollapseWs* ws, void* base, size_t cap, uint32_t nf) {
    Carver c{(char*)base, 0, cap};
    const size_t total = nf ? 2 * (size_t)nf - 1 : 0;
    phase_carve(&ws->ph, c, nf);
    ws->flag = c.take<uint8_t>(total);
    ws->rootcnt = c.take<uint32_t>(total);
    ws->offset = c.take<uint32_t>(total);
    ws->error = c.take<uint32_t>(4);
    ws->scan_bytes = collapse_scan_bytes((uint32_t)total);
    ws->scan_tmp = c.take<char>(ws->scan_bytes);
    return c.used;
}

kr_error collapse_launch(const kr_bvh_node* nodes, const uint32_t* parents, kr_bvh_node* out, uint32_t* counts,
                         uint32_t* primitives, float* costs, float ci, float ct, uint32_t nf, void* workspace,
                         size_t ws_bytes, hipStream_t stream, hipEvent_t mark_done) {
    if (nf == 0) return nullptr;
    CollapseWs ws;
    const size_t need = collapse_carve(&ws, workspace, ws_bytes, nf);
    if (!workspace || ws_bytes < need) return "obvh_collapse: workspace too small";
    const uint32_t total = 2 * nf - 1;
    auto blocks = [](uint32_t n, uint32_t b) { return dim3((n + b - 1) / b); };
    OBVH_HIP_TRY(hipMemcpyAsync(out, nodes, sizeof(kr_bvh_node) * total, hipMemcpyDeviceToDevice, stream));
    uint32_t gb = (total + 255) / 256;
    if (gb > 2048) gb = 2048;
    hipLaunchKernelGGL(collapse_clear, dim3(gb), dim3(256), 0, stream, ws, nf);
    phased_launch(collapse_mark, nf, stream, nodes, parents, out, counts, costs, ci, ct, nf, ws);
    if (mark_done) OBVH_HIP_TRY(hipEventRecord(mark_done, stream));
    hipLaunchKernelGGL(collapse_roots, blocks(total, 256), dim3(256), 0, stream, parents, (const uint32_t*)counts, ws,
                       total);
    size_t sb = ws.scan_bytes;
    OBVH_HIP_TRY(rocprim::exclusive_scan(ws.scan_tmp, sb, ws.rootcnt, ws.offset, 0u, total,
                                         rocprim::plus<uint32_t>(), stream));
    hipLaunchKernelGGL(collapse_gather, blocks(total, 256), dim3(256), 0, stream, nodes, out, counts, primitives, ws,
                       total);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
