// lbvh.hpp -- internal interface of the LBVH builder (lbvh.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"
#include "phased.hpp"

namespace obvh {

struct LbvhWs {
    int32_t* cbox;       // centroid box as order-preserving keys (min xyz, max xyz)
    PhaseWs ph;          // bottom-up pass state
    uint32_t* keys_in;   // Morton codes by face
    uint32_t* keys_out;  // sorted codes
    uint32_t* vals_in;   // face ids
    uint32_t* vals_out;  // face ids in sorted order
    char* sort_tmp;      // radix-sort scratch
    size_t sort_bytes;
};

size_t lbvh_carve(LbvhWs* ws, void* base, size_t cap, uint32_t nf);
kr_error lbvh_launch(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, kr_bvh_node* nodes, uint32_t* parents,
                     uint32_t* counts, float* costs, uint32_t* mortons, void* workspace, size_t ws_bytes,
                     hipStream_t stream);

}  // namespace obvh
