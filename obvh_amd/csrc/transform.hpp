// transform.hpp -- internal interface of the transform kernels (transform.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"

namespace obvh {

// Candidate axes of an internal node (48 B).
struct Cand {
    float b[9];
    int32_t degen;  // 1: the node's box is its AABB
    float aq;       // half area of the node's AABB
    int32_t pad;
};

constexpr int kMaxPhases = 4;

struct TreeWs {
    uint64_t* visit;   // per internal node: arrivals (low word) + sum of arrival tags (high word)
    uint32_t* nlist;   // per node list: number of entries
    uint32_t* phase_done;  // workgroups of climb phase 1 that finished
    uint32_t* list;    // per climb phase: nodes deferred to the next phase
    uint2* span;       // per node: first and last leaf position below it (bit 31 of .y: the
                       // leaves below are not every position from first to last)
    uint32_t* count;   // per internal node with bit 31 of its span set: leaves below it
    uint32_t* error;   // 1: parent index out of range / too deep, 2: parent/child mismatch,
                       // 4: face id out of range, 8: a node did not complete
    float4* kdop;      // 7 per node: {min, min vertex id, max, max vertex id}
    Cand* cand;        // per internal node
    int32_t* range;    // 6 per internal node: order-preserving keys, min b0..b2, max b0..b2
    // Per refit tile (kRTile leaf positions): the lowest node holding all of
    // the tile's leaves (span covers it, no gap) and its ancestors, at most
    // kTChain of them (tchain), and {count, next ancestor beyond them or
    // kInvalid} (tmeta).  Written by the leading workgroups of the candidate
    // kernel, read by the refit.
    uint32_t* tchain;
    uint2* tmeta;
};
constexpr uint32_t kRTile = 512;  // leaves per refit workgroup (tiles aligned to multiples of it)
constexpr uint32_t kTChain = 64;
static_assert(2 + kMaxPhases <= 8, "error words 2 .. 2 + kMaxPhases hold the phase list counts");

size_t tree_carve(TreeWs* ws, void* base, size_t cap, uint32_t N);
struct BatchMap;
kr_error tree_launch(const kr_bvh_node* nodes, const uint32_t* parents, const kr_vec3* V,
                     const kr_uvec4* F, kr_obb3* obbs, kr_mat4* mats, uint32_t* ids, uint32_t N,
                     void* workspace, size_t ws_bytes, hipStream_t stream, const BatchMap* batch = nullptr);

// Copy the call's device error word (0 = success) into `status` (device or
// pinned host memory) on `stream`, and its meaning.
kr_error tree_status_launch(const void* workspace, size_t ws_bytes, uint32_t N, uint32_t* status,
                            hipStream_t stream);
const char* tree_status_string(uint32_t err);

// Many trees in one transform (transform.hip, "batch").
struct BatchWs;
size_t batch_carve(BatchWs* w, void* base, size_t cap, uint32_t Nsum);
size_t batch_workspace_size(uint32_t Nsum);
kr_error tree_batch_launch(const kr_bvh_node* nodes, const uint32_t* parents, const kr_vec3* V, const kr_uvec4* F,
                           const uint32_t* prim_off, const uint32_t* vert_off, uint32_t K, uint32_t Nsum,
                           kr_obb3* obbs, kr_mat4* mats, void* workspace, size_t ws_bytes, hipStream_t stream);
kr_error tree_batch_status_launch(const void* workspace, size_t ws_bytes, uint32_t Nsum, uint32_t* status,
                                  hipStream_t stream);

}  // namespace obvh
