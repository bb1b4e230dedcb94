/*
 * oracle.h -- CPU restatement of the reference's OBB algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; it is the checker,
 * never the product path.
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, x86-64 SSE math,
 * correctly rounded / and sqrt), see oracle/Makefile.
 *
 * Contents and the reference code each part restates (paths relative to
 * the reference repository root):
 *   dito.c           DiTO_14<float> + kr_points_obb
 *                    src/common/algorithm/dito/dito.cpp:56-538,
 *                    src/common/geometry.cpp:232-252
 *   transform.c      the GPU transform method K1-K4
 *                    src/ads/common/cuda/transforms.cu:213-964
 *   lbvh.c           LBVH build (centroid box, 30-bit Morton, stable sort,
 *                    Karras split, bottom-up AABBs/counts/costs)
 *                    src/ads/lbvh_gpu/lbvh.cu:252-457,1547-1692,
 *                    src/ads/common/cuda/transforms.cu:1127-1250
 *   cpu_transform.c  the reference CPU transform kr_bvh_obb_tree (full
 *                    DiTO per node), src/ads/common/transforms.cpp:349-481
 *   atrbvh.c         ATRBVH builder (64-bit Morton LBVH + agglomerative
 *                    treelet optimiser), src/ads/atrbvh_gpu/
 *   collapse.c       BVH collapse and SAH, src/ads/common/cuda/
 *                    transforms.cu:18-211,1061-1125,
 *                    src/ads/common/transforms.cpp:250-303
 *   traverse.c       OBVH packing and closest-hit traversal,
 *                    src/ads/common/cuda/transforms.cu:1298-1449,
 *                    src/ads/common/cuda/intersectors.cu:801-972
 */
#ifndef OBVH_ORACLE_H
#define OBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } ov3;
typedef struct { ov3 mid, v0, v1, v2, ext; } oobb;
typedef struct { float v[16]; } omat4;
typedef struct {
    ov3 bmin, bmax;
    uint32_t a, b;          /* leaf: a = face id; internal: a = left, b = right */
    uint16_t nprim;
    uint8_t axis, pad;
} onode;                     /* 36 B, same layout as kr_bvh_node */

typedef struct {
    uint32_t slab_min_idx[7];
    uint32_t slab_max_idx[7];
    int32_t pair_k, p2_sel, q0_sel, q1_sel, q0_valid, q1_valid;
    int32_t degenerate, best, result_aabb, pad_;
} odito_trace;               /* 96 B, same layout as obvh_dito_trace */

/* DiTO_14<float> on AoS xyz (n points).  trace may be NULL. */
void oracle_dito14(const float* xyz, int n, oobb* out, odito_trace* trace);
/* kr_points_obb: DiTO_14 followed by ext = max(ext, 0.001). */
void oracle_points_obb(const float* xyz, int n, oobb* out, odito_trace* trace);
/* Many sets back to back: set s = points [off[s], off[s+1]). */
void oracle_points_obb_batch(const float* xyz, const uint64_t* off, int nsets,
                             oobb* out, odito_trace* traces);
/* The same on `threads` host threads (sets in parallel, same results). */
void oracle_points_obb_batch_mt(const float* xyz, const uint64_t* off, int nsets,
                                oobb* out, odito_trace* traces, int threads);

/* Parallel for (par.c): fn(ctx, i, thread id) for i in [0, n), chunks of
 * `chunk` indices handed out dynamically to `threads` POSIX threads
 * (thread ids 0..threads-1; 1 or fewer runs inline).  Returns the number
 * of threads that ran. */
#define OPAR_MAX_THREADS 256
typedef void (*opar_fn)(void* ctx, uint32_t i, int tid);
int oracle_parallel_for(uint32_t n, uint32_t chunk, int threads, opar_fn fn, void* ctx);

/* The GPU transform method.  nodes/parents: 2N-1 entries (N = nprim);
 * outputs obbs/mats for every node.  kdop_ids (nullable): 14 ids per node
 * [min0..min6, max0..max6].  refit_ranges (nullable): 6 floats per node
 * (min b0..b2, max b0..b2 after refit; internal nodes only).  cand_axes
 * (nullable): 11 floats per internal node, the candidate axes b0, b1, b2,
 * 1.0 if the node is degenerate (its box is its AABB) and the winning
 * candidate id (0..2 base triangle, 3..11 upper, 12..20 lower tetrahedron,
 * -1 the AABB's axes).
 * Returns 0 on success, -1 on a malformed tree. */
int oracle_obb_tree(const onode* nodes, const uint32_t* parents,
                    const float* vertices, const uint32_t* faces /* uvec4 */,
                    uint32_t nprim, oobb* obbs, omat4* mats,
                    uint32_t* kdop_ids, float* refit_ranges, float* cand_axes);
/* The same on `threads` host threads: candidates, refit + finalize and leaf
 * boxes run node-parallel; outputs are identical for every thread count. */
int oracle_obb_tree_mt(const onode* nodes, const uint32_t* parents,
                       const float* vertices, const uint32_t* faces,
                       uint32_t nprim, oobb* obbs, omat4* mats,
                       uint32_t* kdop_ids, float* refit_ranges, float* cand_axes, int threads);

/* LBVH build.  vertices AoS xyz, faces uvec4.  Outputs 2N-1 nodes, parents,
 * primitive_counts, costs; mortons_sorted (N entries, nullable). */
int oracle_lbvh_build(const float* vertices, const uint32_t* faces, uint32_t nfaces,
                      onode* nodes, uint32_t* parents, uint32_t* counts,
                      float* costs, uint32_t* mortons_sorted);

/* Reference CPU transform kr_bvh_obb_tree (DiTO of every subtree).
 * Processes leaves [leaf_begin, leaf_end) only (the reference climbs from
 * every leaf; a sub-range bounds the CPU time of a timed sample).
 * Returns the number of internal nodes it fitted. */
int64_t oracle_cpu_obb_tree(onode* nodes, const uint32_t* parents,
                            const float* vertices, const uint32_t* faces,
                            uint32_t nprim, oobb* obbs, omat4* mats,
                            uint32_t leaf_begin, uint32_t leaf_end);

/* ATRBVH build (treelet_size 3..20, `iterations` optimiser passes, gamma =
 * treelet_size doubled per pass).  Outputs as oracle_lbvh_build. */
int oracle_atrbvh_build(const float* vertices, const uint32_t* faces, uint32_t nfaces, int treelet_size,
                        int iterations, float ci, float ct, onode* nodes, uint32_t* parents,
                        uint32_t* counts, float* costs);

/* kr_cuda_bvh_collapse.  nodes/parents: 2N-1 entries.  counts: leaf entries
 * read (primitives per leaf), internal entries and collapsed leaves written.
 * costs (2N-1), collapsed (2N-1, a copy of nodes with the collapse edits),
 * primitives (N), table (2N-1 flags, nullable).  Returns the number of
 * primitive slots used, or -1 on a malformed tree. */
int64_t oracle_collapse(const onode* nodes, const uint32_t* parents, uint32_t* counts, float* costs,
                        onode* collapsed, uint32_t* primitives, uint8_t* table, float ci, float ct,
                        uint32_t nprim);

/* kr_bvh_sah over `total` nodes; obbs nullable (AABB areas). */
float oracle_bvh_sah(const onode* nodes, int32_t root_idx, float ci, float ct, const oobb* obbs,
                     uint32_t total);

typedef struct { ov3 origin; float tmin; ov3 direction; float tmax; } oray;       /* 32 B, kr_ray */
typedef struct { ov3 normal; float bx, by; uint32_t primitive, instance; } oisect; /* 28 B, kr_intersection */

/* obvh_persistent_prepare + kr_cuda_soa_obvh: metadata (4 int32 per node,
 * node_count), left_T / right_T (12 floats per internal slot).  Returns -1 on
 * an internal node outside [0, internal_count) or a child out of range. */
int oracle_obvh_pack(const onode* nodes, const omat4* mats, uint32_t node_count, uint32_t internal_count,
                     float* left_T, float* right_T, int32_t* meta);

/* soa_obvh_persistent_intersect, one ray after another.  stats (nullable):
 * per ray {internal nodes visited, triangles tested, deepest stack slot used
 * (0xFFFFFFFF on overflow)}.  Returns -1 if a ray overflowed the stack. */
int oracle_obvh_intersect(const float* left_T, const float* right_T, const int32_t* meta, const float* V,
                          const uint32_t* F, const uint32_t* prims, const oray* rays, uint32_t n, oisect* out,
                          uint32_t* stats);
/* The same on `threads` host threads (rays in parallel, same results). */
int oracle_obvh_intersect_mt(const float* left_T, const float* right_T, const int32_t* meta, const float* V,
                             const uint32_t* F, const uint32_t* prims, const oray* rays, uint32_t n, oisect* out,
                             uint32_t* stats, int threads);

/* kr_ray_unit_aabb_intersect -> (t0, t1); kr_ray_triangle_intersect -> (b1, b2, t). */
void oracle_ray_unit_box(const float* rd, const float* ro, float* out2);
void oracle_ray_triangle(const float* dir, const float* org, const float* v1, const float* v2, const float* v3,
                         float tmax, float* out3);

/* 4x4 helpers exposed for tests (vecmath.h:946-1075,1214-1234). */
void oracle_minverse4(const omat4* m, omat4* out);
void oracle_mmul4(const omat4* l, const omat4* r, omat4* out);

#ifdef __cplusplus
}
#endif
#endif
