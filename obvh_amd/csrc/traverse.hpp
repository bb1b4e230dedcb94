// traverse.hpp -- internal interface of the OBVH packing and closest-hit
// traversal (traverse.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"

namespace obvh {

constexpr int32_t kTravSentinel = 0x76543210;  // the reference's empty-stack marker
constexpr int kTravThreads = 256;              // 4 waves per workgroup
constexpr int kTravLdsSlots = 24;              // stack slots 1..24 per lane in LDS
constexpr int kTravSlots = 63;                 // slots 1..63 (slot 0 is the sentinel)
constexpr int kTravDeepSlots = kTravSlots - kTravLdsSlots;  // 25..63 in global memory
constexpr uint32_t kTravMaxBlocks = 2048;      // persistent grid cap (256 CUs x 8)
constexpr uint32_t kTravAborted = 0xFFFFFFFFu; // kr_intersection.instance of an aborted ray
constexpr int kTravRefill = 16;                // idle lanes that trigger a refill of the wave
// Every lane asks the counter for at most one ray past the end, so it never
// exceeds ray_count + 64 x (waves of the grid); this bound keeps it in range.
constexpr uint32_t kTravMaxRays = 0xFFFFFFFFu - (1u << 20);

struct TravWs {
    uint32_t* counter;  // rays handed out so far
    int32_t* deep;      // overflow stack slots, slot-major: [kTravDeepSlots][blocks * 256]
    uint32_t blocks;    // grid size the deep slots were carved for
};

size_t trav_carve(TravWs* ws, void* base, size_t cap, uint32_t ray_count);

kr_error soa_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, uint32_t node_count,
                         uint32_t internal_count, kr_mat43* left_T, kr_mat43* right_T, kr_ivec4* meta,
                         hipStream_t stream);
kr_error aos_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, kr_bvh_node_packed* packed,
                         kr_bvh_transformation_pair* pairs, uint32_t node_count, hipStream_t stream);
kr_error aos_to_soa_launch(const kr_bvh_node_packed* packed, const kr_bvh_transformation_pair* pairs,
                           uint32_t internal_count, uint32_t node_count, kr_mat43* left_T, kr_mat43* right_T,
                           kr_ivec4* meta, hipStream_t stream);
// Array sizes the traversal checks every index against (0xFFFFFFFF: not
// known, unchecked) and the node + leaf visits after which a ray is
// abandoned (a valid tree never needs more than its node count).
struct TravCounts {
    uint32_t internal_count, node_count, face_count, vertex_count, max_steps;
};

kr_error intersect_launch(const kr_mat43* left_T, const kr_mat43* right_T, const kr_ivec4* meta,
                          const kr_vec3* V, const kr_uvec4* F, const uint32_t* prims, const kr_ray* rays,
                          kr_intersection* out, uint32_t ray_count, const TravCounts& counts, void* workspace,
                          size_t ws_bytes, hipStream_t stream);
kr_error rec_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, const uint32_t* prims, const kr_vec3* V,
                         const kr_uvec4* F, uint32_t face_count, uint32_t vertex_count, obvh_node_record* rec,
                         obvh_tri_record* tri, hipStream_t stream);
kr_error rec_intersect_launch(const obvh_node_record* rec, const obvh_tri_record* tri, uint32_t face_count,
                              const kr_ray* rays, kr_intersection* out, uint32_t ray_count, void* workspace,
                              size_t ws_bytes, hipStream_t stream);

}  // namespace obvh
