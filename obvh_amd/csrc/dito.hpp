// dito.hpp -- internal interface of the DiTO kernels (dito.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"

namespace obvh {

// Per-set state handed from pass 1 (candidate search) to pass 2.
struct SetState {
    float ax[9];      // b0, b1, b2
    float almid[3];   // axis-aligned mid/len from slabs 0-2
    float allen[3];
    float alval;      // quality of the axis-aligned box
    int32_t mode;     // 0 compare with AABB, 1 AABB (no pass 2), 2 line-aligned
    int32_t pad;
};

// Workspace pieces (carved from one caller buffer, see dito_carve).
struct DitoWs {
    uint64_t* offsets_single;  // [0, count] for the single-set entry
    uint32_t* done;            // 2 completion counters of the two-launch single-set fit
    uint32_t* first_tile;      // nsets + 1 (exclusive scan of tiles per set)
    uint32_t* tile_set;        // tile -> set
    float* part1_v;            // 14 per tile
    uint32_t* part1_i;         // 14 per tile (index within the set)
    float* part2;              // 6 per tile
    SetState* state;           // per set
    obvh_dito_trace* trace_tmp;  // per set, used when the caller passes none
};

size_t dito_max_tiles(uint32_t nsets, uint64_t total, uint32_t tile);
size_t dito_carve(DitoWs* ws, void* base, size_t cap, uint32_t nsets, uint64_t total, uint32_t tile);
uint32_t dito_auto_tile(uint32_t nsets, uint64_t total);

// Enqueue a DiTO of `nsets` sets.  Exactly one of (x,y,z) / xyz is used.
// offsets == nullptr means one set of `single_count` points.
kr_error dito_launch(const float* x, const float* y, const float* z, const float* xyz,
                     const uint64_t* offsets, uint64_t single_count, uint32_t nsets, uint64_t total,
                     kr_obb3* out, obvh_dito_trace* trace, uint32_t tile, uint32_t workers,
                     void* workspace, size_t ws_bytes, hipStream_t stream);

// One set of n > 0 points in two launches (slab pass + candidate search,
// projection pass + box), each combined by the last workgroup to finish.
// `clear` first zeroes the workspace's completion counters (needed before
// the first call on a workspace; the launches leave them zero).
kr_error dito_one_launch(const float* x, const float* y, const float* z, const float* xyz, uint64_t n,
                         kr_obb3* out, obvh_dito_trace* trace, void* workspace, size_t ws_bytes, bool clear,
                         hipStream_t stream);

}  // namespace obvh
