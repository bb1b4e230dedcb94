// collapse.hpp -- internal interface of the BVH collapse (collapse.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"
#include "phased.hpp"

namespace obvh {

struct CollapseWs {
    PhaseWs ph;           // bottom-up pass state
    uint8_t* flag;        // per node: collapsed (the node or an ancestor's leaf holds it)
    uint32_t* rootcnt;    // per node: primitives if it becomes a leaf of the collapsed tree, else 0
    uint32_t* offset;     // exclusive scan of rootcnt
    uint32_t* error;      // 1: a collapsed subtree was malformed
    char* scan_tmp;
    size_t scan_bytes;
};

size_t collapse_carve(CollapseWs* ws, void* base, size_t cap, uint32_t nf);
kr_error collapse_launch(const kr_bvh_node* nodes, const uint32_t* parents, kr_bvh_node* out, uint32_t* counts,
                         uint32_t* primitives, float* costs, float ci, float ct, uint32_t nf, void* workspace,
                         size_t ws_bytes, hipStream_t stream, hipEvent_t mark_done);

}  // namespace obvh
