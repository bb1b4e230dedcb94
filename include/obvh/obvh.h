/*
 * obvh.h -- C ABI of the MI355X (gfx950) OBB engine.
 *
 * One shared library, obvh_amd/libobvh.so, exports every symbol below.  The
 * first group keeps the exact signatures and synchronous semantics of the
 * reference's GPU entry points, so a plugin that used to compile the CUDA
 * sources in links this library instead.  The second group is new: device-
 * resident, stream-ordered entry points for SoA point sets, batches of sets,
 * the LBVH builder and the transform with a caller-owned workspace.
 *
 * All types are plain C with fixed layouts; sizes are checked by
 * static asserts at the bottom of this file.
 *
 * Reference interfaces (paths relative to the reference repository root):
 *   kr_vec3 / kr_uvec4 / kr_obb3 / kr_mat4   src/common/vecmath.h:136,172,156,158
 *   kr_bvh_node                              src/common/ads.h:131-142
 *   kr_error (NULL = success)                src/common/vecmath.h:73,21-24
 *   kr_ads_collapse_metrics                  src/common/ads.h:210-214
 *   kr_vec2 / kr_mat43 / kr_ivec4            src/common/vecmath.h:135,159,188
 *   kr_ray / kr_intersection                 src/common/ads.h:64-76
 *   kr_bvh_node_packed                       src/common/ads.h:144-164
 *   kr_bvh_transformation_pair / kr_SoA_OBVH src/common/ads.h:166-190
 */
#ifndef OBVH_OBVH_H
#define OBVH_OBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Types (layouts identical to the reference)                              */
/* ------------------------------------------------------------------------ */

typedef uint32_t u32;
typedef int32_t kr_i32;
typedef uint64_t kr_size;
typedef float kr_scalar;
typedef const void* kr_error; /* NULL on success, else a static message */

typedef struct { float x, y, z; } kr_vec3;                 /* 12 B */
typedef struct { u32 x, y, z, w; } kr_uvec4;               /* 16 B, face = x,y,z; w unused */
typedef struct { kr_vec3 mid, v0, v1, v2, ext; } kr_obb3;  /* 60 B */
typedef struct { float v[16]; } kr_mat4;                   /* 64 B, column major */
typedef struct { kr_vec3 min, max; } kr_aabb3;             /* 24 B */

/* 36 B BVH node.  Node array layout: internal nodes [0, N-1), leaves
 * [N-1, 2N-1).  Leaf: nPrimitives > 0, primitivesOffset = face id.
 * Internal: nPrimitives == 0, left/right = child node indices.       */
typedef struct {
    kr_aabb3 bounds;
    union {
        u32 primitivesOffset;
        struct { u32 left, right; } children;
    } u;
    uint16_t nPrimitives;
    uint8_t axis;
    uint8_t pad_;
} kr_bvh_node;

typedef struct { float x, y; } kr_vec2;                    /* 8 B */
typedef struct { float v[12]; } kr_mat43;                  /* 48 B, 4 columns of 3 */
typedef struct { int32_t x, y, z, w; } kr_ivec4;           /* 16 B */

/* Packed OBVH node (64 B): child boxes, then metadata.  Internal node:
 * metadata.x / .y = left / right child, each as its bitwise NOT when the
 * child is a leaf; leaf: x = primitivesOffset, y = x + nPrimitives;
 * z = axis; w unused.                                                   */
typedef struct {
    kr_aabb3 lbounds, rbounds;
    kr_ivec4 metadata;
} kr_bvh_node_packed;
typedef struct { kr_mat4 l, r; } kr_bvh_transformation_pair;  /* 128 B */
/* Structure-of-arrays OBVH: per internal node the left / right child's
 * world -> unit-cube matrix (3x4), per node the metadata above.          */
typedef struct {
    kr_mat43* left_T;
    kr_mat43* right_T;
    kr_ivec4* metadata;
} kr_SoA_OBVH;

typedef struct {
    kr_vec3 origin;
    kr_scalar tmin;
    kr_vec3 direction;
    kr_scalar tmax;
} kr_ray;                                                  /* 32 B */
typedef struct {
    kr_vec3 geom_normal;
    kr_vec2 barys;
    u32 primitive;   /* 0xFFFFFFFF: no hit */
    u32 instance;    /* 0; 0xFFFFFFFF if the ray's traversal stack overflowed */
} kr_intersection;                                         /* 28 B */

/* Native OBVH layout (new): one 128-byte record (one cache line) per
 * internal node holding both children's world -> unit-cube matrices and
 * their codes -- an internal child's index, or for a leaf child
 * ~(first_slot << 4 | count - 1) -- and one 48-byte record per triangle
 * slot, a leaf's triangles in consecutive slots.                         */
typedef struct {
    kr_mat43 left_T, right_T;
    int32_t left, right;
    int32_t pad_[6];
} obvh_node_record;                                        /* 128 B */
typedef struct {
    kr_vec3 v0, v1, v2;
    u32 primitive;
    u32 pad_[2];
} obvh_tri_record;                                         /* 48 B */

/* Collapse timings (ms), accumulated by kr_cuda_bvh_collapse. */
typedef struct {
    kr_scalar collapse_mark_time;
    kr_scalar collapse_time;
} kr_ads_collapse_metrics;

/* ------------------------------------------------------------------------ */
/* Legacy entry points (drop-in)                                            */
/* ------------------------------------------------------------------------ */

/* DiTO-14 OBB of a host AoS point set.  Replaces kr_cuda_points_obb
 * (src/common/algorithm/dito/dito.cuh:12-13, body dito.cu:614-744).
 * `points` is host memory, borrowed; returns by value; `elapsed_time`
 * (nullable) receives the device time of the fit in ms.  The result equals
 * the reference CPU path kr_points_obb (src/common/geometry.cpp:232-252),
 * i.e. DiTO_14<float> followed by the 0.001 extent floor.  On an internal
 * error the returned box is all NaN.                                       */
kr_obb3 kr_cuda_points_obb(const kr_vec3* points, kr_size count, kr_scalar* elapsed_time);

/* The same fit on the host: replaces the reference's CPU kr_points_obb
 * (src/common/geometry.h:13-14, body geometry.cpp:232-252): DiTO_14<float>
 * (src/common/algorithm/dito/dito.cpp:478-538) plus the 0.001 extent floor,
 * single thread; `elapsed_time` (nullable) receives its wall time in ms.
 * Equal to kr_cuda_points_obb's box for finite input.                      */
kr_obb3 kr_points_obb(const kr_vec3* points, kr_size count, kr_scalar* elapsed_time);

/* AABB-BVH -> OBB-tree transform.  Replaces kr_cuda_bvh_obb_tree
 * (src/ads/common/cuda/transforms.cuh:52-65, body transforms.cu:966-1059).
 * All pointers are device pointers owned by the caller.  Writes obbs[] and
 * transforms[] for all primitive_count + internal_count nodes; never writes
 * nodes[].  primitive_counts, costs, primitives, ci, ct, obb_cost and
 * primitive_threshold are accepted and ignored, as in the reference.
 * Synchronous: returns after the device work has finished.  Returns NULL
 * on success, else an error string (bad sizes, malformed tree, HIP error). */
kr_error kr_cuda_bvh_obb_tree(
    kr_bvh_node* nodes, const u32* parents, const u32* primitive_counts,
    const kr_scalar* costs, const kr_vec3* vertices, const kr_uvec4* faces,
    const u32* primitives, kr_obb3* obbs, kr_mat4* transforms,
    kr_scalar ci, kr_scalar ct, kr_scalar obb_cost,
    u32 primitive_count, u32 internal_count, u32 primitive_threshold);

/* ------------------------------------------------------------------------ */
/* Device-resident, stream-ordered entry points (new; no reference          */
/* counterpart).  `stream` is a hipStream_t passed as void* (NULL = the     */
/* default stream).  None of these synchronise the host.                    */
/* ------------------------------------------------------------------------ */

/* Per-fit trace: which points DiTO selected.  Indices into the input set
 * (slab extremes) or into the candidate point list `sel` (see below).
 * sel = [min0..min6, max0..max6] for count > 14, else the input order.     */
typedef struct {
    u32 slab_min_idx[7];
    u32 slab_max_idx[7];
    int32_t pair_k;        /* furthest slab pair, -1 if not reached        */
    int32_t p2_sel;        /* base-triangle apex, index into sel, or -1    */
    int32_t q0_sel;        /* max point along the base normal, or -1       */
    int32_t q1_sel;        /* min point along the base normal, or -1       */
    int32_t q0_valid, q1_valid;
    int32_t degenerate;    /* 0 ok, 1 AABB (pair too close), 2 line-aligned, 3 empty,
                              4 invalid offsets (decreasing or past total_points: NaN box) */
    int32_t best;          /* winning candidate 0..20, -1 = initial AABB axes */
    int32_t result_aabb;   /* 1 if the returned box is the axis-aligned one */
    int32_t pad_;
} obvh_dito_trace;       /* 96 B */

/* Workspace bytes needed by obvh_dito_soa for `count` points. */
size_t obvh_dito_workspace_size(kr_size count);

/* DiTO-14 (+ 0.001 extent floor) of one SoA point set in device memory.
 * out_obb: device kr_obb3; trace: device obvh_dito_trace or NULL.          */
kr_error obvh_dito_soa(const float* x, const float* y, const float* z, kr_size count,
                       kr_obb3* out_obb, obvh_dito_trace* trace,
                       void* workspace, size_t workspace_bytes, void* stream);

/* obvh_dito_soa without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  Its
 * workspace is a static 1 MiB device array inside libobvh.so, one per GPU,
 * held until the process ends; count must not exceed
 * OBVH_DITO_NOALLOC_MAX_POINTS (larger counts return an error).  Calls
 * that overlap on one GPU share that array and race: serialise them (one
 * stream, or events).  Results equal obvh_dito_soa's.                    */
#define OBVH_DITO_NOALLOC_MAX_POINTS ((kr_size)1 << 28)
kr_error obvh_dito_soa_noalloc(const float* x, const float* y, const float* z, kr_size count,
                               kr_obb3* out_obb, obvh_dito_trace* trace, void* stream);

/* Batched DiTO over `nsets` sets stored back to back in SoA arrays;
 * set s is points [offsets[s], offsets[s+1]).  offsets: device, nsets+1
 * entries.  tile_points: points per work tile (multiple of 512; 0 = auto).
 * workers: workgroups per pass, each looping over tiles (0 = one per tile,
 * at most 65536).  One launch per pass; the whole call is graph-capturable. */
size_t obvh_dito_batch_workspace_size(u32 nsets, kr_size total_points, u32 tile_points);
kr_error obvh_dito_batch(const float* x, const float* y, const float* z,
                         const kr_size* offsets, u32 nsets, kr_size total_points,
                         kr_obb3* out_obbs, obvh_dito_trace* traces,
                         u32 tile_points, u32 workers,
                         void* workspace, size_t workspace_bytes, void* stream);

/* Upload of nsets host point sets for obvh_dito_batch: set s is three rows
 * (x, y, z) of offsets[s+1] - offsets[s] floats at host_sets[s] (pinned
 * memory for an asynchronous copy); it lands at column offsets[s] of the
 * device array soa with rows of row_stride floats (x = soa, y = soa +
 * row_stride, z = soa + 2 row_stride).  offsets: host, nsets + 1 entries.
 * One 2-D copy per set on `stream`.                                      */
kr_error obvh_upload_sets(const float* const* host_sets, const kr_size* offsets, u32 nsets,
                          float* soa, kr_size row_stride, void* stream);

/* The same upload as one kernel that reads the sets over PCIe: host_sets
 * and offsets are the host copies (checked: every non-empty set must be
 * pinned host or device memory of 3 x its count floats), set_table and
 * offsets_table the same arrays in device-accessible memory (device or
 * pinned), which the kernel reads.                                       */
kr_error obvh_gather_sets(const float* const* host_sets, const float* const* set_table,
                          const kr_size* offsets, const kr_size* offsets_table, u32 nsets,
                          float* soa, kr_size row_stride, void* stream);

/* Transform with caller workspace, stream ordered.  The kernels compute
 * what kr_cuda_bvh_obb_tree computes; the returned value reports argument
 * and launch errors only.  A malformed tree (the errors
 * kr_cuda_bvh_obb_tree returns) is found on the device: it leaves NaN
 * boxes and a nonzero status word in the workspace, which
 * obvh_obb_tree_status_async copies out.  `kdop_ids` (device, nullable)
 * receives for every node the 14 selected extremal vertex ids
 * [min0..min6, max0..max6].                                              */
size_t obvh_obb_tree_workspace_size(u32 primitive_count, u32 internal_count);
kr_error obvh_obb_tree_async(
    const kr_bvh_node* nodes, const u32* parents,
    const kr_vec3* vertices, const kr_uvec4* faces,
    kr_obb3* obbs, kr_mat4* transforms, u32* kdop_ids,
    u32 primitive_count, u32 internal_count,
    void* workspace, size_t workspace_bytes, void* stream);

/* Status of the last obvh_obb_tree_async call on `workspace` (same
 * primitive_count): a kernel on `stream` copies its device error word into
 * *status (device or pinned host memory; 0 = success).
 * obvh_obb_tree_status_string maps a nonzero word to the error message
 * kr_cuda_bvh_obb_tree would return (NULL for 0).                        */
kr_error obvh_obb_tree_status_async(const void* workspace, size_t workspace_bytes,
                                    u32 primitive_count, u32* status, void* stream);
kr_error obvh_obb_tree_status_string(u32 status);

/* Many trees in one transform (no reference counterpart; the batched form
 * of kr_cuda_bvh_obb_tree, transforms.cu:966-1059, for serving loops).
 * Tree t has N_t = primitive_offsets[t+1] - primitive_offsets[t] >= 1
 * primitives; its 2 N_t - 1 nodes and parents start at index
 * 2 primitive_offsets[t] - t of `nodes` / `parents` (the single-tree layout,
 * indices local to the tree, root at its local node 0), its N_t faces at
 * primitive_offsets[t] of `faces` with vertex ids local to its vertices,
 * which start at vertex_offsets[t] of `vertices`.  Both offset arrays are
 * device arrays of ntrees + 1 entries; total_primitives is the sum of N_t
 * (host).  obbs / transforms receive every tree's boxes and matrices in
 * the layout of `nodes`, each equal to what kr_cuda_bvh_obb_tree computes
 * for that tree alone.  Stream ordered, no allocation, graph capturable.
 * Errors in the trees (as kr_cuda_bvh_obb_tree reports them, plus an empty
 * tree, indices outside their tree) reach the caller through
 * obvh_obb_tree_batch_status_async; obvh_obb_tree_status_string names them. */
size_t obvh_obb_tree_batch_workspace_size(u32 ntrees, u32 total_primitives);
kr_error obvh_obb_tree_batch(
    const kr_bvh_node* nodes, const u32* parents,
    const kr_vec3* vertices, const kr_uvec4* faces,
    const u32* primitive_offsets, const u32* vertex_offsets,
    u32 ntrees, u32 total_primitives,
    kr_obb3* obbs, kr_mat4* transforms,
    void* workspace, size_t workspace_bytes, void* stream);
kr_error obvh_obb_tree_batch_status_async(const void* workspace, size_t workspace_bytes,
                                          u32 total_primitives, u32* status, void* stream);

/* Host-side tree validation (parents/children consistent, single root at
 * node 0, every node reachable once).  Inputs are HOST arrays.  Returns
 * NULL if the tree is well formed.  Used before launching the bottom-up
 * kernels so that a malformed tree is an error instead of a hang.        */
kr_error obvh_validate_tree_host(const kr_bvh_node* nodes, const u32* parents,
                                 u32 primitive_count, u32 internal_count);

/* LBVH builder (30-bit Morton codes, stable sort, Karras radix tree,
 * bottom-up AABBs/counts/costs), the input side of the transform.
 * Replaces the build in src/ads/lbvh_gpu/lbvh.cu:1547-1692.
 * All pointers are device pointers; nodes/parents/primitive_counts/costs
 * have 2*face_count-1 entries.                                            */
size_t obvh_lbvh_workspace_size(u32 face_count);
kr_error obvh_lbvh_build(const kr_vec3* vertices, const kr_uvec4* faces, u32 face_count,
                         kr_bvh_node* nodes, u32* parents, u32* primitive_counts,
                         kr_scalar* costs, u32* mortons_sorted,
                         void* workspace, size_t workspace_bytes, void* stream);

/* ATRBVH builder (the atrbvh_gpu plugin's build, src/ads/atrbvh_gpu/
 * atrbvh.cu:300-430): 63-bit Morton LBVH over the triangle-box centres
 * normalised by the scene box, then `iterations` passes of the agglomerative
 * treelet optimiser (treelet_size 3..20 leaves, gamma = treelet_size doubled
 * per pass, costs ci/ct), then bounds / counts / costs as the plugin's
 * calculate_aabbs_kernel.  Device pointers; nodes/parents/primitive_counts/
 * costs have 2*face_count-1 entries.  Stream ordered, caller workspace.   */
size_t obvh_atrbvh_workspace_size(u32 face_count);
kr_error obvh_atrbvh_build(const kr_vec3* vertices, const kr_uvec4* faces, u32 face_count,
                           int treelet_size, int iterations, kr_scalar ci, kr_scalar ct,
                           kr_bvh_node* nodes, u32* parents, u32* primitive_counts, kr_scalar* costs,
                           void* workspace, size_t workspace_bytes, void* stream);

/* BVH collapse, replaces kr_cuda_bvh_collapse
 * (src/ads/common/cuda/transforms.cuh:40-50, body transforms.cu:18-211,
 * 1061-1125).  Device pointers, 2N-1 entries each except primitives (N):
 * bottom-up, an internal node with count = counts[l] + counts[r] <= 8 and
 * ct*area*count < ci*area + costs[l] + costs[r] is collapsed; every
 * collapsed node whose parent is not collapsed becomes a leaf of
 * collapsed_nodes (a copy of nodes) holding its subtree's distinct primitive
 * ids in primitives[primitivesOffset ...]; children of collapsed nodes get
 * nPrimitives = 0xFFFF.  primitive_counts: leaf entries are read, internal
 * entries and new leaves written; costs: written for every node.  The
 * leaves' slices of primitives follow increasing node index (the reference
 * assigns them in arrival order).  Synchronous; metrics may be NULL.        */
kr_error kr_cuda_bvh_collapse(const kr_bvh_node* nodes, const u32* parents, kr_bvh_node* collapsed_nodes,
                              u32* primitive_counts, u32* primitives, kr_scalar* costs,
                              kr_scalar ci, kr_scalar ct, u32 leaf_count, u32 internal_count,
                              kr_ads_collapse_metrics* metrics);

/* The same collapse, stream ordered, with a caller-owned workspace.     */
size_t obvh_collapse_workspace_size(u32 primitive_count);
kr_error obvh_collapse_async(const kr_bvh_node* nodes, const u32* parents, kr_bvh_node* collapsed_nodes,
                             u32* primitive_counts, u32* primitives, kr_scalar* costs,
                             kr_scalar ci, kr_scalar ct, u32 primitive_count,
                             void* workspace, size_t workspace_bytes, void* stream);

/* Surface-area heuristic of a host tree, replaces kr_bvh_sah
 * (src/ads/common/transforms.h:38-42, body transforms.cpp:250-303): with
 * obbs != NULL the OBB surface areas are used (AABB for one-primitive
 * leaves); nodes marked 0xFFFF are skipped; the quality is appended as
 * "%f," to outputFile when it is not NULL.  Host code, no GPU.            */
#ifdef __cplusplus
kr_scalar kr_bvh_sah(const kr_bvh_node* nodes, kr_i32 root_idx, kr_scalar ci, kr_scalar ct,
                     const kr_obb3* obbs = nullptr, const char* outputFile = nullptr);
#else
kr_scalar kr_bvh_sah(const kr_bvh_node* nodes, kr_i32 root_idx, kr_scalar ci, kr_scalar ct,
                     const kr_obb3* obbs, const char* outputFile);
#endif

/* ------------------------------------------------------------------------ */
/* OBVH packing and closest-hit traversal                                   */
/* ------------------------------------------------------------------------ */

/* Packs a (plain or collapsed) tree and its transforms into AoS form.
 * Replaces kr_cuda_obvh_persistent (src/ads/common/cuda/transforms.cuh:
 * 32-38, body transforms.cu:1298-1364).  Device pointers; packed_bvh has
 * node_count entries, packed_transforms one per internal node (indexed by
 * node).  Nodes marked 0xFFFF and the slots they would fill are left
 * untouched, as in the reference.  Synchronous.                          */
kr_error kr_cuda_obvh_persistent(const kr_bvh_node* nodes, const kr_mat4* transforms,
                                 kr_bvh_node_packed* packed_nodes,
                                 kr_bvh_transformation_pair* packed_transforms, u32 node_count);

/* Builds the structure-of-arrays OBVH from the packed form.  Replaces
 * kr_cuda_soa_obvh (src/ads/common/cuda/transforms.cuh:72-77, body
 * transforms.cu:1366-1449), whose third parameter is a kr_SoA_OBVH*&: the
 * same argument passing as the pointer-to-pointer here.  Allocates the
 * arrays and a device copy of the kr_SoA_OBVH struct and stores that
 * device pointer in *soa_bvh.  The conversion runs on the device (the
 * reference copies to the host and back).  Synchronous.                  */
kr_error kr_cuda_soa_obvh(const kr_bvh_node_packed* packed_bvh,
                          const kr_bvh_transformation_pair* packed_transforms,
                          kr_SoA_OBVH** soa_bvh, u32 internal_count, u32 node_count);

/* Closest-hit traversal of ray_count rays.  Replaces
 * kr_cuda_soa_obvh_intersect (src/ads/common/cuda/intersectors.cuh:81-91,
 * body intersectors.cu:801-972,1452-1472).  bvh is the device struct made
 * by kr_cuda_soa_obvh; primitives is read for leaves of more than one
 * primitive (the collapsed tree's slices).  block_size, grid_size and
 * warp_counter are accepted and ignored (the reference ignores the first
 * two; rays are handed out from the library's own counter).  Stream
 * ordered on the default stream; returns after the launch.  A ray whose
 * traversal stack would need a 64th entry stops with its hit so far and
 * instance = 0xFFFFFFFF (the reference writes past its stack there).     */
kr_error kr_cuda_soa_obvh_intersect(const kr_SoA_OBVH* bvh, const kr_vec3* vertices,
                                    const kr_uvec4* faces, const u32* primitives, const kr_ray* rays,
                                    kr_intersection* isects, u32 ray_count, u32 block_size,
                                    u32 grid_size, u32* warp_counter);

/* New: frees what kr_cuda_soa_obvh allocated (the reference never does). */
kr_error obvh_soa_obvh_free(kr_SoA_OBVH* soa_bvh);

/* New: one-pass device packing from the tree and its transforms straight
 * into caller-owned SoA arrays (left_T / right_T: internal_count entries,
 * 16-byte aligned; metadata: node_count).  Equals kr_cuda_obvh_persistent
 * followed by kr_cuda_soa_obvh, except that slots the reference leaves
 * unwritten (0xFFFF nodes, metadata.w, matrices of non-internal slots)
 * are zero.  Stream ordered.                                             */
kr_error obvh_soa_pack_async(const kr_bvh_node* nodes, const kr_mat4* transforms, u32 node_count,
                             u32 internal_count, kr_mat43* left_T, kr_mat43* right_T,
                             kr_ivec4* metadata, void* stream);

/* New: the traversal with the SoA arrays passed directly, their sizes
 * (left_T / right_T: internal_count, metadata: node_count, faces and
 * primitives: face_count, vertices: vertex_count) and a caller workspace
 * (ray counter and overflow stack slots); stream ordered.  Every index the
 * walk meets is checked against those sizes: a ray that meets one outside
 * them (a malformed tree) stops with instance = 0xFFFFFFFF, as does a ray
 * that overflows the 63-entry stack.                                     */
size_t obvh_soa_intersect_workspace_size(u32 ray_count);
kr_error obvh_soa_intersect_async(const kr_mat43* left_T, const kr_mat43* right_T,
                                  const kr_ivec4* metadata, u32 internal_count, u32 node_count,
                                  const kr_vec3* vertices, u32 vertex_count,
                                  const kr_uvec4* faces, const u32* primitives, u32 face_count,
                                  const kr_ray* rays, kr_intersection* isects, u32 ray_count,
                                  void* workspace, size_t workspace_bytes, void* stream);

/* New: packs a (plain or collapsed) tree into the native record layout:
 * node_records (face_count - 1 entries, one per internal node slot),
 * tri_records (2 * face_count slots: a leaf node v >= face_count - 1 owns
 * slot face_count + (v - face_count + 1), a collapsed leaf owns its slice
 * of primitives).  A leaf may hold at most 16 primitives.  All pointers
 * device, records 16-byte aligned; stream ordered.                       */
kr_error obvh_records_pack_async(const kr_bvh_node* nodes, const kr_mat4* transforms, const u32* primitives,
                                 const kr_vec3* vertices, u32 vertex_count, const kr_uvec4* faces,
                                 u32 face_count, obvh_node_record* node_records,
                                 obvh_tri_record* tri_records, void* stream);

/* New: the closest-hit traversal over the record layout (same hits as the
 * SoA traversal, bit for bit); workspace as obvh_soa_intersect_async.    */
kr_error obvh_records_intersect_async(const obvh_node_record* node_records,
                                      const obvh_tri_record* tri_records, u32 face_count,
                                      const kr_ray* rays, kr_intersection* isects, u32 ray_count,
                                      void* workspace, size_t workspace_bytes, void* stream);

/* Last error string of the calling host thread (NULL if none). */
const char* obvh_last_error(void);

/* Library version string. */
const char* obvh_version(void);

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(kr_vec3) == 12, "kr_vec3 must be 12 bytes");
static_assert(sizeof(kr_uvec4) == 16, "kr_uvec4 must be 16 bytes");
static_assert(sizeof(kr_obb3) == 60, "kr_obb3 must be 60 bytes");
static_assert(sizeof(kr_mat4) == 64, "kr_mat4 must be 64 bytes");
static_assert(sizeof(kr_bvh_node) == 36, "kr_bvh_node must be 36 bytes");
static_assert(sizeof(obvh_dito_trace) == 96, "obvh_dito_trace must be 96 bytes");
static_assert(sizeof(kr_mat43) == 48, "kr_mat43 must be 48 bytes");
static_assert(sizeof(kr_bvh_node_packed) == 64, "kr_bvh_node_packed must be 64 bytes");
static_assert(sizeof(kr_bvh_transformation_pair) == 128, "kr_bvh_transformation_pair must be 128 bytes");
static_assert(sizeof(kr_ray) == 32, "kr_ray must be 32 bytes");
static_assert(sizeof(kr_intersection) == 28, "kr_intersection must be 28 bytes");
static_assert(sizeof(obvh_node_record) == 128, "obvh_node_record must be 128 bytes");
static_assert(sizeof(obvh_tri_record) == 48, "obvh_tri_record must be 48 bytes");
#endif

#endif /* OBVH_OBVH_H */
