// atrbvh.hpp -- internal interface of the ATRBVH builder (atrbvh.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"
#include "phased.hpp"

namespace obvh {

constexpr int kMaxTreelet = 20;  // treelet sizes 3..20 (distance matrix in LDS)

struct AtrWs {
    PhaseWs ph;          // bottom-up pass state
    int32_t* sbox;       // scene box, order-preserving keys (min xyz, max xyz)
    uint64_t* keys_in;   // 64-bit Morton codes by face
    uint64_t* keys_out;  // sorted codes
    uint32_t* vals_in;   // face ids
    uint32_t* vals_out;  // face ids in code order
    float4* bmin;        // per node: working box (fminf/fmaxf semantics, -0 below +0)
    float4* bmax;
    float* area;         // per node: surface area of the working box
    int32_t* tri;        // per node: triangles below (optimiser pass)
    float* sah;          // per node: SAH cost of the subtree (optimiser pass)
    char* sort_tmp;
    size_t sort_bytes;
};

size_t atr_carve(AtrWs* ws, void* base, size_t cap, uint32_t nf);
kr_error atr_launch(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, int treelet_size, int iterations, float ci,
                    float ct, kr_bvh_node* nodes, uint32_t* parents, uint32_t* counts, float* costs, void* workspace,
                    size_t ws_bytes, hipStream_t stream);

}  // namespace obvh
