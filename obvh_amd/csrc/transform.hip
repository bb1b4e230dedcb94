// transform.hip -- AABB-BVH -> OBB-tree transform on gfx950.
//
// Method of the reference GPU transform (src/ads/common/cuda/transforms.cu
// :213-964), restructured for CDNA4:
//
//  climb_leaves  one lane per leaf, 512-leaf workgroups: the leaf's 7-slab
//                k-DOP (3 vertices, first vertex wins ties), then a bottom-up climb
//                in which the second child to arrive at a parent merges the
//                two k-DOPs.  Arrivals at parents inside the workgroup's range
//                are counted in LDS; the rest use a 64-bit global arrival
//                word tagged with the workgroup, and parents whose children
//                finished in different workgroups are deferred.  Then the
//                leaf OBBs and matrices (transforms.cu:803-834).
//  climb_phase   the deferred parents, in 32768-leaf workgroups whose last
//                one to finish then climbs the rest of the tree; arrivals
//                within a workgroup are counted, and k-DOPs handed over, in
//                LDS.  The merge takes the right child's extreme only if
//                strictly better, so ties go to the left child and the
//                selected vertex ids do not depend on scheduling.  k-DOPs
//                carry vertex ids, not positions (7 x {min, id, max, id} =
//                112 B per node); spans carry an exact gap flag from leaf
//                counts.
//  candidates    one lane per internal node: gather the 14 extremal vertices
//                and run the base-triangle + ditetrahedron search in
//                registers (transforms.cu:413-756); the nodes the tile climb
//                completed on a second stream beside climb_phase, the rest
//                after it, together with each refit tile's ancestor chain.
//  cand_redo     the search in reference mode for nodes with non-finite or
//                huge extremal points.
//  refit         exact projection range of every subtree vertex on each
//                node's candidate axes (transforms.cu:836-900): per-lane walks
//                below each wave's common ancestor, DPP wave reductions above
//                it (both reading the tile's own nodes from an LDS copy), a
//                tile-wide pass over the tile's distinct vertices for the
//                tile's ancestor chain, LDS keys per workgroup, global
//                atomics.
//  finalize      one lane per internal node: OBB vs AABB by half area, extent
//                floor, inverse(T * R * S) (transforms.cu:905-964).
//
// Tie rule, degenerate nodes and floating point follow the reference with
// the decisions documented in the oracle (oracle/transform.c).
#include <float.h>

#include <mutex>

#include "common.hpp"
#include "transform.hpp"

namespace obvh {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kChunk0 = 512;     // leaves per workgroup, first climb phase
constexpr uint32_t kChunk1 = 32768;   // second phase

struct Mat4 {
    float v[16];
};

__device__ __forceinline__ f3 ldv(const kr_vec3* V, uint32_t i) {
    kr_vec3 a = V[i];
    return f3{a.x, a.y, a.z};
}

// kr_minverse4, vecmath.h:946-1075.
__device__ __forceinline__ void minverse4(const float* m, float* out) {
    float r[16];
    r[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
           m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    r[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
           m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    r[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
           m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    r[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
            m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    r[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
           m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    r[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
           m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    r[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
           m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    r[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
            m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    r[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
           m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    r[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
           m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    r[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
            m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    r[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
            m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    r[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
           m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    r[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
           m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    r[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
            m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    r[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
            m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * r[0] + m[1] * r[4] + m[2] * r[8] + m[3] * r[12];
    if (det == 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        return;
    }
    det = 1.0f / det;
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = r[i] * det;
}

// kr_mmul4, vecmath.h:1214-1234.
__device__ __forceinline__ void mmul4(const float* l, const float* r, float* o) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[4 * c + i] = l[i] * r[4 * c] + l[4 + i] * r[4 * c + 1] + l[8 + i] * r[4 * c + 2] +
                           l[12 + i] * r[4 * c + 3];
}

// inverse(T(mid) * (R(v0,v1,v2) * S(2 ext))), transforms.cu:832-834,939-942.
__device__ __forceinline__ void obb_matrix(f3 v0, f3 v1, f3 v2, f3 mid, f3 ext, float* out) {
    float R[16] = {v0.x, v0.y, v0.z, 0, v1.x, v1.y, v1.z, 0, v2.x, v2.y, v2.z, 0, 0, 0, 0, 1};
    float S[16] = {ext.x * 2.0f, 0, 0, 0, 0, ext.y * 2.0f, 0, 0, 0, 0, ext.z * 2.0f, 0, 0, 0, 0, 1};
    float T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, mid.x, mid.y, mid.z, 1};
    float RS[16], TRS[16];
    mmul4(R, S, RS);
    mmul4(T, RS, TRS);
    minverse4(TRS, out);
}

// One box (15 floats, kr_obb3 order) and one matrix (16 floats) into
// staging rows.
__device__ __forceinline__ void obb_row(float* d, f3 mid, f3 v0, f3 v1, f3 v2, f3 ext) {
    d[0] = mid.x; d[1] = mid.y; d[2] = mid.z;
    d[3] = v0.x; d[4] = v0.y; d[5] = v0.z;
    d[6] = v1.x; d[7] = v1.y; d[8] = v1.z;
    d[9] = v2.x; d[10] = v2.y; d[11] = v2.z;
    d[12] = ext.x; d[13] = ext.y; d[14] = ext.z;
}

// ------------------------------------------------------ batch map --
//
// Numbering of a batch of trees joined into one combined tree (see the batch
// section below).

struct BatchMap {
    const uint32_t* prim_off;  // K + 1 leaf offsets L_t
    uint32_t K;
    uint32_t Nsum;             // sum of N_t
};

// Tree of input node g: the last t with 2 L_t - t <= g.
__device__ __forceinline__ uint32_t batch_tree_of_node(const BatchMap& m, uint32_t g) {
    uint32_t lo = 0, hi = m.K;  // nb(lo) <= g < nb(hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (2u * m.prim_off[mid] - mid <= g) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t batch_tree_of_leaf(const BatchMap& m, uint32_t f) {
    uint32_t lo = 0, hi = m.K;  // L(lo) <= f < L(hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (m.prim_off[mid] <= f) lo = mid; else hi = mid;
    }
    return lo;
}
// Combined index of local node j of tree t (N leaves, first leaf Lt); j must be < 2N - 1.
__device__ __forceinline__ uint32_t batch_map(const BatchMap& m, uint32_t t, uint32_t Lt, uint32_t N, uint32_t j) {
    if (j + 1 < N) return Lt + j + (t + 1 < m.K ? 1u : 0u);
    return (m.Nsum - 1) + Lt + (j - (N - 1));
}
// Output index of combined node c (a box the caller receives), or kInvalid
// for an unused slot.  K == 0: no batch, the identity.
__device__ __forceinline__ uint32_t batch_out_index(const BatchMap& m, uint32_t c) {
    if (m.K == 0) return c;
    const uint32_t Ic = m.Nsum - 1;
    if (c >= Ic) {  // leaf position c - Ic
        const uint32_t pos = c - Ic;
        const uint32_t t = batch_tree_of_leaf(m, pos);
        const uint32_t Lt = m.prim_off[t], N = m.prim_off[t + 1] - Lt;
        return 2 * Lt - t + (N - 1) + (pos - Lt);
    }
    const uint32_t t = batch_tree_of_leaf(m, c);
    const uint32_t Lt = m.prim_off[t];
    if (t + 1 < m.K) {
        if (c == Lt) return kInvalid;  // unused slot
        return 2 * Lt - t + (c - Lt - 1);
    }
    return 2 * Lt - t + (c - Lt);
}

// Output staging for R consecutive nodes (R = the workgroup size): boxes at
// row stride 15 floats, matrices at row stride 17 (odd strides: no LDS bank
// conflicts), written back to HBM as contiguous dwords.
constexpr int kStage = 256;
template <int R>
struct StageT {
    float ob[R * 15];
    float mt[R * 17];
    uint32_t orow[R];  // output node of each row (batches)
};
using Stage = StageT<kStage>;

template <int R>
__device__ __forceinline__ void stage_nan(StageT<R>& st, uint32_t r) {
    const float nan = __builtin_nanf("");
#pragma unroll
    for (int q = 0; q < 15; ++q) st.ob[r * 15 + q] = nan;
#pragma unroll
    for (int q = 0; q < 16; ++q) st.mt[r * 17 + q] = nan;
}

// Dwords g[0 .. n) = at(j), written by a workgroup of R threads as 16-byte
// stores from the first 16-byte boundary of g on, the dwords before it and
// after the last full vector one by one.
template <int R, class At>
__device__ __forceinline__ void flush_dwords(float* g, uint32_t n, At at) {
    const uint32_t h = min((uint32_t)(((16u - ((uint32_t)(uintptr_t)g & 15u)) & 15u) >> 2), n);
    const uint32_t body = (n - h) >> 2;
    for (uint32_t j = threadIdx.x; j < h; j += R) g[j] = at(j);
    float4* gv = reinterpret_cast<float4*>(g + h);
    for (uint32_t v = threadIdx.x; v < body; v += R) {
        const uint32_t j = h + 4 * v;
        gv[v] = make_float4(at(j), at(j + 1), at(j + 2), at(j + 3));
    }
    for (uint32_t j = h + 4 * body + threadIdx.x; j < n; j += R) g[j] = at(j);
}

// Rows first .. first + cnt - 1 (combined indices) to their output nodes: the
// same node indices, or for a batch each tree's own layout (runs of
// consecutive rows stay consecutive; top nodes are dropped).
template <int R>
__device__ __forceinline__ void stage_flush(StageT<R>& st, kr_obb3* obbs, kr_mat4* mats, size_t first, uint32_t cnt,
                                            const BatchMap& m) {
    if (m.K == 0) {
        flush_dwords<R>(reinterpret_cast<float*>(obbs) + first * 15, cnt * 15, [&](uint32_t j) { return st.ob[j]; });
        flush_dwords<R>(reinterpret_cast<float*>(mats) + first * 16, cnt * 16,
                        [&](uint32_t j) { return st.mt[(j >> 4) * 17 + (j & 15)]; });
        return;
    }
    if (threadIdx.x < cnt) st.orow[threadIdx.x] = batch_out_index(m, (uint32_t)first + threadIdx.x);
    __syncthreads();
    float* go = reinterpret_cast<float*>(obbs);
    for (uint32_t j = threadIdx.x; j < cnt * 15; j += R) {
        const uint32_t r = j / 15, o = st.orow[r];
        if (o != kInvalid) go[(size_t)o * 15 + (j - r * 15)] = st.ob[j];
    }
    float* gm = reinterpret_cast<float*>(mats);
    for (uint32_t j = threadIdx.x; j < cnt * 16; j += R) {
        const uint32_t o = st.orow[j >> 4];
        if (o != kInvalid) gm[(size_t)o * 16 + (j & 15)] = st.mt[(j >> 4) * 17 + (j & 15)];
    }
}

// ------------------------------------------------------------ k-DOP --

struct Kdop {
    float mn[7], mx[7];
    uint32_t imn[7], imx[7];
};

__device__ __forceinline__ void kdop_store(float4* kd, uint32_t node, const Kdop& k) {
#pragma unroll
    for (int s = 0; s < 7; ++s)
        kd[(size_t)node * 7 + s] = make_float4(k.mn[s], __uint_as_float(k.imn[s]), k.mx[s], __uint_as_float(k.imx[s]));
}
__device__ __forceinline__ void kdop_load(const float4* kd, uint32_t node, Kdop& k) {
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        float4 a = kd[(size_t)node * 7 + s];
        k.mn[s] = a.x; k.imn[s] = __float_as_uint(a.y);
        k.mx[s] = a.z; k.imx[s] = __float_as_uint(a.w);
    }
}

// Leaf OBB from the triangle frame, transforms.cu:803-834.
__device__ __forceinline__ void leaf_obb(f3 v0, f3 v1, f3 v2, float* ob, float* mt) {
    f3 e0 = sub3(v1, v0);
    f3 e1 = sub3(v2, v0);
    f3 a2 = vm::cross(e0, e1);
    f3 a0 = vm::normalize(e0);
    f3 a1 = vm::normalize(vm::cross(a0, a2));
    float R[16] = {a0.x, a0.y, a0.z, 0, a1.x, a1.y, a1.z, 0, a2.x, a2.y, a2.z, 0, 0, 0, 0, 1};
    float Ri[16];
    minverse4(R, Ri);
    auto xf = [&](f3 r) {  // kr_vtransform3
        return f3{Ri[0] * r.x + Ri[4] * r.y + Ri[8] * r.z + Ri[12],
                  Ri[1] * r.x + Ri[5] * r.y + Ri[9] * r.z + Ri[13],
                  Ri[2] * r.x + Ri[6] * r.y + Ri[10] * r.z + Ri[14]};
    };
    f3 l0 = xf(v0), l1 = xf(v1), l2 = xf(v2);
    f3 mx = vm::vmax3(l0, vm::vmax3(l1, l2));
    f3 mn = vm::vmin3(l0, vm::vmin3(l1, l2));
    f3 mid = vm::mul1(add3(mx, mn), 0.5f);
    f3 ext = vm::vmax3(vm::mul1(sub3(mx, mn), 0.5f), f3{0.001f, 0.001f, 0.001f});
    f3 m = vm::mul1(a0, mid.x);
    m = add3(m, vm::mul1(a1, mid.y));
    m = add3(m, vm::mul1(a2, mid.z));
    obb_row(ob, m, a0, a1, a2, ext);
    float M[16];
    obb_matrix(a0, a1, a2, m, ext, M);
#pragma unroll
    for (int q = 0; q < 16; ++q) mt[q] = M[q];
}

// Arrival word of an unused internal slot of a batch: complete, tagged so
// that the per-node kernels skip it (arrival tags never sum to this).
constexpr uint32_t kHoleTag = 0xFFFFFFFFu;
__device__ __forceinline__ bool hole(uint64_t v) { return (uint32_t)(v >> 32) == kHoleTag; }

__global__ void batch_holes(TreeWs ws, BatchMap m) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k + 1 < m.K; k += gridDim.x * blockDim.x) {
        const uint32_t idx = m.prim_off[k];
        ws.visit[idx] = ((uint64_t)kHoleTag << 32) | 2ull;
        ws.span[idx] = make_uint2(1u, 0u);  // empty: holds no refit tile
    }
}

__global__ void tree_clear(uint64_t* visit, uint32_t n, TreeWs ws) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *ws.error = 0;
        for (int k = 0; k < kMaxPhases; ++k) ws.nlist[k] = 0;
        *ws.phase_done = 0;
    }
    for (uint32_t i = t; i < n; i += gridDim.x * blockDim.x) visit[i] = 0;
}

// Arrival word of an internal node: low 32 bits count the children that
// finished, high 32 bits sum their tags.  tag = (phase << 24) | (owner + 1)
// identifies the kernel phase and the workgroup that completed the child.
__device__ __forceinline__ bool complete(uint64_t v) { return (uint32_t)v == 2u; }
__device__ __forceinline__ uint32_t make_tag(uint32_t phase, uint32_t owner) { return (phase << 24) | (owner + 1u); }

__device__ __forceinline__ void merge_left_first(const Kdop& L, const Kdop& R, Kdop& m) {
    m = L;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        if (R.mn[s] < m.mn[s]) { m.mn[s] = R.mn[s]; m.imn[s] = R.imn[s]; }
        if (R.mx[s] > m.mx[s]) { m.mx[s] = R.mx[s]; m.imx[s] = R.imx[s]; }
    }
}

// Leaf k-DOP of one triangle: 7 slabs, first vertex wins ties.
__device__ __forceinline__ void leaf_kdop(f3 v0, f3 v1, f3 v2, const uint32_t vid[3], Kdop& k) {
    const f3 vs[3] = {v0, v1, v2};
    float pr[7];
    slab7(vs[0], pr);
#pragma unroll
    for (int s = 0; s < 7; ++s) { k.mn[s] = k.mx[s] = pr[s]; k.imn[s] = k.imx[s] = vid[0]; }
#pragma unroll
    for (int j = 1; j < 3; ++j) {
        slab7(vs[j], pr);
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            if (pr[s] < k.mn[s]) { k.mn[s] = pr[s]; k.imn[s] = vid[j]; }
            if (pr[s] > k.mx[s]) { k.mx[s] = pr[s]; k.imx[s] = vid[j]; }
        }
    }
}

// Leaf span of a subtree: first and last leaf position below it.  Bit 31 of
// .y is set when the subtree does not hold every position from first to
// last (its leaf count is below the width; the count of such a node is kept
// beside the span); leaf positions are < 2^31.
constexpr uint32_t kGap = 0x80000000u;
__device__ __forceinline__ uint32_t span_hi(uint2 s) { return s.y & ~kGap; }
__device__ __forceinline__ bool span_gap(uint2 s) { return (s.y & kGap) != 0; }
__device__ __forceinline__ uint2 leaf_span(uint32_t pos) { return make_uint2(pos, pos); }
// Leaves below a node: the span's width, or the kept count if it has a gap.
__device__ __forceinline__ uint32_t span_count(uint2 s, uint32_t kept) {
    return span_gap(s) ? kept : span_hi(s) - s.x + 1u;
}
__device__ __forceinline__ uint2 span_join(uint2 a, uint32_t ca, uint2 b, uint32_t cb, uint32_t& c) {
    const uint32_t lo = min(a.x, b.x), hi = max(span_hi(a), span_hi(b));
    c = ca + cb;
    return make_uint2(lo, hi | (hi - lo + 1u != c ? kGap : 0u));
}
// Span and leaf count of node c from global memory (the count word is read
// unconditionally, beside the span, and used only for a gapped span).
__device__ __forceinline__ uint2 span_load(const TreeWs& ws, uint32_t c, uint32_t I, uint32_t& cnt) {
    if (c >= I) {
        cnt = 1;
        return leaf_span(c - I);
    }
    const uint2 s = ws.span[c];
    const uint32_t kept = ws.count[c];
    cnt = span_count(s, kept);
    return s;
}
__device__ __forceinline__ void span_store(TreeWs& ws, uint32_t c, uint2 s, uint32_t cnt) {
    ws.span[c] = s;
    if (span_gap(s)) ws.count[c] = cnt;
}

// Node lists (ws.list, counts in ws.nlist): 0 and 1 hold the nodes a climb
// phase defers to the next, kUpperList the nodes completed after the tile
// climb, kRedoList the nodes the candidate search hands to reference mode.
constexpr int kUpperList = 2;
constexpr int kRedoList = 3;
static_assert(kRedoList < kMaxPhases, "one list per slot");

// A deferred node is marked in bit 63 of its arrival word: its count already
// reads 2 (both children arrived), but its k-DOP is made by the next phase,
// so the candidate launch that runs beside the later phases must skip it.
constexpr uint64_t kDeferred = 1ull << 63;
__device__ __forceinline__ void defer(TreeWs& ws, uint32_t phase, uint32_t node, uint32_t I) {
    atomicOr(reinterpret_cast<unsigned long long*>(&ws.visit[node]), (unsigned long long)kDeferred);
    const uint32_t slot = atomicAdd(&ws.nlist[phase], 1u);
    if (slot < I) ws.list[(size_t)phase * (I + 1) + slot] = node;
}

// k-DOP of leaf node `leaf` from the mesh.
__device__ __forceinline__ bool leaf_kdop_global(const kr_bvh_node* nodes, const kr_vec3* V, const kr_uvec4* F,
                                                 uint32_t leaf, uint32_t N, Kdop& k) {
    const uint32_t f = nodes[leaf].u.primitivesOffset;
    if (f >= N) return false;
    const kr_uvec4 face = F[f];
    const uint32_t vid[3] = {face.x, face.y, face.z};
    leaf_kdop(ldv(V, face.x), ldv(V, face.y), ldv(V, face.z), vid, k);
    return true;
}

// Phase 0 of the bottom-up k-DOP pass.  Workgroup w owns leaf positions
// [w*kTile, (w+1)*kTile) and the internal nodes with the same indices (for a
// Karras-ordered tree these are the internal nodes whose leaves lie in the
// range).  The workgroup first loads the tile's topology into LDS (children
// and parents of its internal nodes, parents and vertex ids of its leaves)
// and checks that every child of its internal nodes points back to it.
// Then one lane per leaf climbs: at a parent in the tile the arrival is
// counted in LDS, the first child stops and the second merges, reading the
// sibling's k-DOP from LDS (internal node of the tile, or a leaf's k-DOP
// left in the parent's slot), recomputing it (a leaf that left none) or
// loading it (node completed outside the tile's index range).  At the
// first level two leaf siblings in one wave merge without an arrival count:
// the left one takes the right one's k-DOP by lane shuffles.  Parents outside the tile use the global arrival word with
// the workgroup's tag; such a parent is deferred to the next phase when its
// children finished in different workgroups.  Finally tile nodes reached by
// one child only are folded into the global arrival words, and the k-DOPs,
// spans and arrival words of tile nodes completed here are written back
// contiguously.  The merge keeps the left child's extreme on ties.  Last,
// the tile's leaf boxes (transforms.cu:803-834) from the triangles each lane
// loaded, staged in the same LDS and written back contiguously.
constexpr uint32_t kTile = kChunk0;

struct ClimbTile {
    float4 kd[kTile][7];        // k-DOP of internal node lo + s, once completed here
    uint2 span[kTile];
    uint32_t cnt[kTile];        // leaves below internal node lo + s
    uint32_t lid[kTile][3];     // vertex ids of leaf position lo + s
    uint32_t arrive[kTile];     // arrivals at internal node lo + s from this workgroup
    uint2 kids[kTile];          // children of internal node lo + s
    uint32_t par_int[kTile];    // parent of internal node lo + s
    uint32_t par_leaf[kTile];   // parent of the leaf at position lo + s
};
union ClimbLds {
    ClimbTile L;
    StageT<kTile> st;           // after the climb: the leaf boxes
};

__device__ __forceinline__ void tile_leaf_kdop(const ClimbTile& L, const kr_vec3* V, uint32_t s, Kdop& k) {
    const uint32_t vid[3] = {L.lid[s][0], L.lid[s][1], L.lid[s][2]};
    leaf_kdop(ldv(V, vid[0]), ldv(V, vid[1]), ldv(V, vid[2]), vid, k);
}

__global__ __launch_bounds__(kTile) void climb_leaves(const kr_bvh_node* __restrict__ nodes,
                                                      const uint32_t* __restrict__ parents,
                                                      const kr_vec3* __restrict__ V,
                                                      const kr_uvec4* __restrict__ F, TreeWs ws, uint32_t N,
                                                      kr_obb3* obbs, kr_mat4* mats, BatchMap bm) {
    __shared__ ClimbLds U;
    ClimbTile& L = U.L;
    const uint32_t I = N - 1, total = 2 * N - 1;
    const uint32_t t = threadIdx.x;
    const uint32_t lo = blockIdx.x * kTile;
    const uint32_t tag = make_tag(0, blockIdx.x);
    const uint32_t pos = lo + t;

    // tile topology and the lane's leaf
    Kdop k = {};
    bool climbing = false;
    f3 v0 = f3{0, 0, 0}, v1 = v0, v2 = v0;  // the lane's triangle, for its leaf box
    if (pos < N) {
        const uint32_t leaf = I + pos;
        L.par_leaf[t] = parents[leaf];
        const uint32_t f = nodes[leaf].u.primitivesOffset;
        if (f < N) {
            const kr_uvec4 face = F[f];
            const uint32_t vid[3] = {face.x, face.y, face.z};
            L.lid[t][0] = vid[0];
            L.lid[t][1] = vid[1];
            L.lid[t][2] = vid[2];
            v0 = ldv(V, vid[0]);
            v1 = ldv(V, vid[1]);
            v2 = ldv(V, vid[2]);
            leaf_kdop(v0, v1, v2, vid, k);
            climbing = true;
        } else {
            L.lid[t][0] = kInvalid;  // never climbs, never read as a sibling
            atomicOr(ws.error, 4u);  // faces has one entry per leaf
        }
    }
    if (pos < I) {
        const uint32_t l = nodes[pos].u.children.left, r = nodes[pos].u.children.right;
        L.kids[t] = make_uint2(l, r);
        L.par_int[t] = parents[pos];
        // (an unused slot of a batch has no children: no check)
        if ((l >= total || r >= total || parents[l] != pos || parents[r] != pos) && !hole(ws.visit[pos]))
            atomicOr(ws.error, 2u);
    }
    L.arrive[t] = 0;
    __syncthreads();

    // First level, all lanes together: two leaf siblings in one wave are
    // merged by the left one (the right one's k-DOP comes by a lane shuffle)
    // without an arrival count.  Any other leaf with a parent in the tile
    // leaves its k-DOP in the parent's LDS slot before it arrives (when its
    // sibling is internal, or it is the left one of two leaves), so that a
    // sibling arriving second reads it from LDS instead of recomputing it.
    uint32_t cur = I + pos;
    uint2 span = leaf_span(pos);
    uint32_t count = 1;
    uint32_t p = L.par_leaf[t];
    bool climb = climbing;  // the lane climbs on past the first level
    {
        const uint64_t climbers = __ballot(climbing);
        uint32_t src = (uint32_t)lane_id(), ps = p - lo, sib = kInvalid;
        bool pair_left = false, pair_right = false, stash = false;
        if (climbing && p < I && ps < kTile) {
            const uint2 kc = L.kids[ps];
            const bool am_left = kc.x == cur;
            if (am_left || kc.y == cur) {  // otherwise the climb below reports it
                sib = am_left ? kc.y : kc.x;
                const uint32_t sl = sib - I - (pos - (uint32_t)lane_id());  // the sibling leaf's lane
                if (sib >= I && sib < total && sl < (uint32_t)kWave && ((climbers >> sl) & 1ull)) {
                    pair_left = am_left;
                    pair_right = !am_left;
                    src = sl;
                } else {
                    stash = sib < I || am_left;
                }
            }
        }
        Kdop o;
#pragma unroll
        for (int q = 0; q < 7; ++q) {  // every lane of the wave takes part
            o.mn[q] = __shfl(k.mn[q], (int)src, kWave);
            o.mx[q] = __shfl(k.mx[q], (int)src, kWave);
            o.imn[q] = (uint32_t)__shfl((int)k.imn[q], (int)src, kWave);
            o.imx[q] = (uint32_t)__shfl((int)k.imx[q], (int)src, kWave);
        }
        if (pair_right) climb = false;  // the left sibling merges
        if (pair_left) {
            Kdop m;
            merge_left_first(k, o, m);
            k = m;
            span = span_join(span, count, leaf_span(sib - I), 1u, count);
#pragma unroll
            for (int q = 0; q < 7; ++q)
                L.kd[ps][q] = make_float4(k.mn[q], __uint_as_float(k.imn[q]), k.mx[q], __uint_as_float(k.imx[q]));
            L.span[ps] = span;
            L.cnt[ps] = count;
            L.arrive[ps] = 2u;  // both children: no other lane arrives here
            cur = p;
            p = L.par_int[ps];
        }
        if (stash) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                L.kd[ps][q] = make_float4(k.mn[q], __uint_as_float(k.imn[q]), k.mx[q], __uint_as_float(k.imx[q]));
        }
    }

    if (climb) {
        for (uint32_t step = 0; p != kInvalid; ++step) {
            if (p >= I || step >= N) { atomicOr(ws.error, 1u); break; }
            const uint32_t s = p - lo;
            uint32_t sib, pp;
            bool am_left;
            if (s < kTile) {  // parent in the tile: LDS arrival
                const uint2 kc = L.kids[s];  // topology: read before the arrival
                pp = L.par_int[s];
                const uint32_t cnt =
                    __hip_atomic_fetch_add(&L.arrive[s], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cnt == 0) break;  // first arrival: the sibling continues
                if (cnt != 1) { atomicOr(ws.error, 2u); break; }
                am_left = kc.x == cur;
                if (!am_left && kc.y != cur) { atomicOr(ws.error, 2u); break; }
                sib = am_left ? kc.y : kc.x;
            } else {  // parent outside the tile: publish, then global arrival
                kdop_store(ws.kdop, cur, k);
                span_store(ws, cur, span, count);
                const uint32_t l = nodes[p].u.children.left, rr = nodes[p].u.children.right;
                pp = parents[p];
                const uint64_t old = __hip_atomic_fetch_add(&ws.visit[p], ((uint64_t)tag << 32) | 1ull,
                                                            __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t cnt = (uint32_t)old;
                if (cnt == 0) break;
                if (cnt != 1) { atomicOr(ws.error, 2u); break; }
                if ((uint32_t)(old >> 32) != tag) { defer(ws, 0, p, I); break; }  // other workgroup
                am_left = l == cur;
                if (!am_left && rr != cur) { atomicOr(ws.error, 2u); break; }
                sib = am_left ? rr : l;
            }
            // the sibling finished in this workgroup
            Kdop o;
            uint2 so;
            uint32_t co;
            if (sib >= I) {
                const uint32_t ss = sib - I - lo;
                if (sib >= total || ss >= kTile) { atomicOr(ws.error, 2u); break; }
                // a leaf left its k-DOP in the parent's slot unless this lane
                // is its left leaf sibling (from another wave)
                const uint32_t from = s < kTile && (cur < I || !am_left) ? s : kInvalid;
                if (from != kInvalid) {
#pragma unroll
                    for (int q = 0; q < 7; ++q) {
                        const float4 a = L.kd[from][q];
                        o.mn[q] = a.x; o.imn[q] = __float_as_uint(a.y);
                        o.mx[q] = a.z; o.imx[q] = __float_as_uint(a.w);
                    }
                } else {
                    tile_leaf_kdop(L, V, ss, o);
                }
                so = leaf_span(sib - I);
                co = 1;
            } else if (sib - lo < kTile) {  // completed in this tile: k-DOP in LDS
                const uint32_t ss = sib - lo;
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    const float4 a = L.kd[ss][q];
                    o.mn[q] = a.x; o.imn[q] = __float_as_uint(a.y);
                    o.mx[q] = a.z; o.imx[q] = __float_as_uint(a.w);
                }
                so = L.span[ss];
                co = L.cnt[ss];
            } else {
                kdop_load(ws.kdop, sib, o);
                so = span_load(ws, sib, I, co);
            }
            Kdop m;
            if (am_left) merge_left_first(k, o, m); else merge_left_first(o, k, m);
            k = m;
            span = span_join(span, count, so, co, count);
            if (s < kTile) {
#pragma unroll
                for (int q = 0; q < 7; ++q)
                    L.kd[s][q] = make_float4(k.mn[q], __uint_as_float(k.imn[q]), k.mx[q], __uint_as_float(k.imx[q]));
                L.span[s] = span;
                L.cnt[s] = count;
            } else {
                kdop_store(ws.kdop, p, k);
                span_store(ws, p, span, count);
            }
            cur = p;
            p = pp;
        }
    }
    __syncthreads();

    // tile nodes reached by one child: publish a tile-leaf child, then the
    // global arrival (the other child finishes in another workgroup)
    if (pos < I && L.arrive[t] == 1u) {
        const uint2 kc = L.kids[t];
        const uint32_t c2[2] = {kc.x, kc.y};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t ss = c2[j] - I - lo;
            if (c2[j] >= I && c2[j] < total && ss < kTile && L.lid[ss][0] != kInvalid) {
                Kdop o;
                tile_leaf_kdop(L, V, ss, o);
                kdop_store(ws.kdop, c2[j], o);
                ws.span[c2[j]] = leaf_span(c2[j] - I);
            }
        }
        const uint64_t old = __hip_atomic_fetch_add(&ws.visit[pos], ((uint64_t)tag << 32) | 1ull,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((uint32_t)old == 1u) defer(ws, 0, pos, I);
        else if ((uint32_t)old != 0u) atomicOr(ws.error, 2u);
    }
    // write back the tile nodes completed here
    const uint32_t nin = lo < I ? min(kTile, I - lo) : 0u;
    float4* kd = ws.kdop + (size_t)lo * 7;
    for (uint32_t j = t; j < nin * 7; j += kTile) {
        const uint32_t ss = j / 7;
        if (L.arrive[ss] == 2u) kd[j] = L.kd[ss][j - ss * 7];
    }
    if (t < nin && L.arrive[t] == 2u) {
        span_store(ws, pos, L.span[t], L.cnt[t]);
        ws.visit[pos] = ((uint64_t)tag << 33) | 2ull;
    }
    // the leaf boxes (a bad face id gives a NaN row; the climb reported it)
    __syncthreads();
    if (pos < N) {
        if (climbing) leaf_obb(v0, v1, v2, &U.st.ob[t * 15], &U.st.mt[t * 17]);
        else stage_nan(U.st, t);
    }
    __syncthreads();
    stage_flush(U.st, obbs, mats, (size_t)I + lo, min(kTile, N - lo), bm);
}

// Phase >= 1 of the bottom-up k-DOP pass.  Workgroup w takes the nodes the
// previous phase deferred whose index lies in [w*chunk, (w+1)*chunk) (both
// children complete, k-DOPs in global memory), merges them and climbs.  At
// a parent p:
//   - one child of p completed in an earlier kernel (its arrival is in p's
//     arrival word, read a level ahead): the sibling's k-DOP is loaded and
//     the climb goes on;
//   - otherwise the arrival is counted in an LDS hash keyed by p.  Each
//     child first writes its k-DOP and span to its side of p's slot; the
//     first arrival stops, the second reads the other side and goes on.
//     Slots still holding one arrival at the end are pairs split between
//     workgroups: they are folded into the global arrival words, and the
//     second fold defers p to the next phase;
//   - when p finds no slot (LDS full), the global arrival word with the
//     workgroup's tag is used, as at the tile boundaries of phase 0.
// Every completed node's k-DOP and span go to global memory as they are
// made (the candidate search and the next phase read them).  The last phase
// is one workgroup and defers nothing.  Arrivals within the workgroup need
// workgroup-scope ordering only, on LDS.
constexpr int kPSlots = 600;   // LDS hash slots (both children's k-DOPs each)
constexpr int kPProbe = 64;
constexpr int kPQueue = 1024;  // entries taken from the previous phase's list
constexpr int kPhaseThreads = 512;  // workgroup size of climb_phase (the pipeline needs ~150 VGPRs)

struct PhaseTile {
    float4 kd[kPSlots][2][7];  // side 0: left child, side 1: right child
    uint2 span[kPSlots][2];
    uint32_t cnt[kPSlots][2];
    uint32_t key[kPSlots];
    uint32_t arrive[kPSlots];
    uint32_t queue[kPQueue];
    uint32_t done[kPQueue];    // nodes completed here, for the second candidate launch
    uint32_t nq, ndone, dbase;
};

// Slot of node p, inserted if absent; kInvalid when the probe window is full
// (keys are never removed, so both children of p get the same answer).
__device__ __forceinline__ uint32_t phase_slot(PhaseTile& P, uint32_t p) {
    uint32_t h = __umulhi(p * 2654435761u, (uint32_t)kPSlots);
    for (int probe = 0; probe < kPProbe; ++probe) {
        uint32_t key = __hip_atomic_load(&P.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (key == kInvalid) key = atomicCAS(&P.key[h], kInvalid, p);
        if (key == kInvalid || key == p) return h;
        h = (h + 1 == (uint32_t)kPSlots) ? 0u : h + 1;
    }
    return kInvalid;
}

// Children, parent and arrival word of node p (zero / kInvalid past the
// internal nodes).
struct Topo {
    uint32_t l, r, parent;
    uint64_t v;
};
__device__ __forceinline__ Topo topo_load(const kr_bvh_node* __restrict__ nodes, const uint32_t* __restrict__ parents,
                                          const TreeWs& ws, uint32_t p, uint32_t I) {
    Topo T = {0u, 0u, kInvalid, 0ull};
    if (p < I) {
        T.l = nodes[p].u.children.left;
        T.r = nodes[p].u.children.right;
        T.parent = parents[p];
        T.v = ws.visit[p];
    }
    return T;
}
// The k-DOP, span and leaf count of the other child of the node whose
// topology is T, read only when T's arrival word says that child finished
// in an earlier kernel (visible since the launch).
struct Sib {
    Kdop k;
    uint2 s;
    uint32_t c;
};
__device__ __forceinline__ void sib_load(const TreeWs& ws, const Topo& T, uint32_t child, uint32_t phase,
                                         uint32_t I, uint32_t total, Sib& S) {
    const uint32_t c0 = (uint32_t)T.v, t0 = (uint32_t)(T.v >> 32);
    const uint32_t sib = T.l == child ? T.r : T.l;
    if (c0 == 1u && (t0 >> 24) < phase && sib < total) {
        kdop_load(ws.kdop, sib, S.k);
        S.s = span_load(ws, sib, I, S.c);
    }
}

// One phase as workgroup `wg` of `nwg`.
__device__ __forceinline__ void phase_pass(PhaseTile& P, const kr_bvh_node* __restrict__ nodes,
                                           const uint32_t* __restrict__ parents, TreeWs& ws, uint32_t N,
                                           uint32_t phase, uint32_t chunk, uint32_t wg, uint32_t nwg) {
    const uint32_t I = N - 1, total = 2 * N - 1, t = threadIdx.x;
    const uint32_t tag = make_tag(phase, wg);
    const uint64_t lo = (uint64_t)wg * chunk, hi = lo + chunk;  // owned node indices
    const uint32_t nin = min(ws.nlist[phase - 1], I);
    const uint32_t* in = ws.list + (size_t)(phase - 1) * (I + 1);
    for (uint32_t h = t; h < (uint32_t)kPSlots; h += blockDim.x) {
        P.key[h] = kInvalid;
        P.arrive[h] = 0;
    }
    if (t == 0) {
        P.nq = 0;
        P.ndone = 0;
    }
    const bool last = nwg == 1;
    auto completed = [&](uint32_t node) {
        const uint32_t d = atomicAdd(&P.ndone, 1u);
        if (d < (uint32_t)kPQueue) {
            P.done[d] = node;
        } else {
            const uint32_t slot = atomicAdd(&ws.nlist[kUpperList], 1u);
            if (slot < I) ws.list[(size_t)kUpperList * (I + 1) + slot] = node;
        }
    };
    __syncthreads();
    // the owned entries, four independent reads per lane and round
    for (uint32_t r0 = t; r0 < nin; r0 += 4 * blockDim.x) {
        uint32_t e[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t r = r0 + j * blockDim.x;
            e[j] = r < nin ? in[r] : kInvalid;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (e[j] == kInvalid || e[j] < lo || e[j] >= hi) continue;
            const uint32_t q = atomicAdd(&P.nq, 1u);
            if (q < (uint32_t)kPQueue) P.queue[q] = e[j];
        }
    }
    __syncthreads();
    const uint32_t nq = P.nq;
    const bool queued = nq <= (uint32_t)kPQueue;
    // more entries than the queue holds: every lane scans the list again
    for (uint32_t r = t; r < (queued ? nq : nin); r += blockDim.x) {
        uint32_t cur = queued ? P.queue[r] : in[r];
        if (!queued && (cur < lo || cur >= hi)) continue;
        if (cur >= I) { atomicOr(ws.error, 1u); continue; }
        const uint32_t l0 = nodes[cur].u.children.left, r0 = nodes[cur].u.children.right;
        if (l0 >= total || r0 >= total) { atomicOr(ws.error, 2u); continue; }
        Kdop k;
        uint32_t count, c0l, c0r;
        const uint2 s0l = span_load(ws, l0, I, c0l), s0r = span_load(ws, r0, I, c0r);
        uint2 span = span_join(s0l, c0l, s0r, c0r, count);
        {
            Kdop a, b;
            kdop_load(ws.kdop, l0, a);
            kdop_load(ws.kdop, r0, b);
            merge_left_first(a, b, k);
        }
        kdop_store(ws.kdop, cur, k);
        span_store(ws, cur, span, count);
        completed(cur);
        // Pipeline: the topology and arrival word of a node are read two
        // levels ahead (A: the parent p, B: its parent), and where the
        // sibling at the next level finished in an earlier kernel its k-DOP,
        // span and count are read as soon as this level's merge is done
        // (SA), so that a level costs about one memory round trip.
        uint32_t p = parents[cur];
        Topo A = topo_load(nodes, parents, ws, p, I);
        Sib SA;
        sib_load(ws, A, cur, phase, I, total, SA);
        Topo B = topo_load(nodes, parents, ws, A.parent, I);
        for (uint32_t step = 0; p != kInvalid; ++step) {
            if (p >= I || step >= N) { atomicOr(ws.error, 1u); break; }
            const bool am_left = (A.l == cur);
            const uint32_t sib = am_left ? A.r : A.l;
            if (sib >= total || (!am_left && A.r != cur)) { atomicOr(ws.error, 2u); break; }
            const Topo C = topo_load(nodes, parents, ws, B.parent, I);
            const uint32_t c0 = (uint32_t)A.v, t0 = (uint32_t)(A.v >> 32);
            Kdop o;
            uint2 so;
            uint32_t co;
            if (c0 == 1u && (t0 >> 24) < phase) {  // the sibling finished in an earlier kernel
                o = SA.k;
                so = SA.s;
                co = SA.c;
                ws.visit[p] = A.v + (((uint64_t)tag << 32) | 1ull);
            } else if (c0 > 1u) {
                atomicOr(ws.error, 2u);
                break;
            } else if (const uint32_t h = phase_slot(P, p); h != kInvalid) {
                const int side = am_left ? 0 : 1;
#pragma unroll
                for (int q = 0; q < 7; ++q)
                    P.kd[h][side][q] = make_float4(k.mn[q], __uint_as_float(k.imn[q]), k.mx[q], __uint_as_float(k.imx[q]));
                P.span[h][side] = span;
                P.cnt[h][side] = count;
                const uint32_t a = __hip_atomic_fetch_add(&P.arrive[h], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (a == 0u) break;  // first arrival: the sibling continues (or the end fold defers p)
                if (a != 1u) { atomicOr(ws.error, 2u); break; }
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    const float4 x = P.kd[h][side ^ 1][q];
                    o.mn[q] = x.x; o.imn[q] = __float_as_uint(x.y);
                    o.mx[q] = x.z; o.imx[q] = __float_as_uint(x.w);
                }
                so = P.span[h][side ^ 1];
                co = P.cnt[h][side ^ 1];
                ws.visit[p] = ((uint64_t)tag << 33) | 2ull;
            } else {  // no LDS slot: the global arrival word (cur's k-DOP is already stored)
                const uint64_t old = __hip_atomic_fetch_add(&ws.visit[p], ((uint64_t)tag << 32) | 1ull,
                                                            __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t cnt = (uint32_t)old, t1 = (uint32_t)(old >> 32);
                if (cnt == 0) break;
                if (cnt != 1) { atomicOr(ws.error, 2u); break; }
                if (t1 != tag && (t1 >> 24) >= phase) {  // the other child finished in another workgroup
                    if (last) atomicOr(ws.error, 2u);
                    else defer(ws, phase, p, I);
                    break;
                }
                kdop_load(ws.kdop, sib, o);
                so = span_load(ws, sib, I, co);
            }
            Kdop m;
            if (am_left) merge_left_first(k, o, m); else merge_left_first(o, k, m);
            k = m;
            span = span_join(span, count, so, co, count);
            kdop_store(ws.kdop, p, k);
            span_store(ws, p, span, count);
            completed(p);
            sib_load(ws, B, p, phase, I, total, SA);  // the next level's sibling
            cur = p;
            p = A.parent;
            A = B;
            B = C;
        }
    }
    __syncthreads();
    // pairs split between workgroups: fold the single arrival into the
    // global word; the second fold defers the node
    for (uint32_t h = t; h < (uint32_t)kPSlots; h += blockDim.x) {
        const uint32_t p = P.key[h];
        if (p == kInvalid || P.arrive[h] != 1u) continue;
        const uint64_t old = __hip_atomic_fetch_add(&ws.visit[p], ((uint64_t)tag << 32) | 1ull, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)old == 1u && !last) defer(ws, phase, p, I);
        else if ((uint32_t)old != 0u) atomicOr(ws.error, 2u);
    }
    // the completed nodes go to the upper list in one block
    if (t == 0) P.dbase = atomicAdd(&ws.nlist[kUpperList], min(P.ndone, (uint32_t)kPQueue));
    __syncthreads();
    const uint32_t nd = min(P.ndone, (uint32_t)kPQueue);
    for (uint32_t i = t; i < nd; i += blockDim.x)
        if (P.dbase + i < I) ws.list[(size_t)kUpperList * (I + 1) + P.dbase + i] = P.done[i];
}

// Phase 1 on `gridDim.x` workgroups; the last workgroup to finish (one
// agent-scope release per workgroup, an acquire in the last) then runs the
// last phase over the whole tree on the CU it holds, so that the last phase
// does not wait for a CU to drain while the side stream fills the chip.  A
// launch with one workgroup runs its phase as the last.
__global__ __launch_bounds__(kPhaseThreads) void climb_phase(const kr_bvh_node* __restrict__ nodes,
                                                    const uint32_t* __restrict__ parents, TreeWs ws, uint32_t N,
                                                    uint32_t phase, uint32_t chunk) {
    __shared__ PhaseTile P;
    __shared__ uint32_t finished_last;
    uint32_t wg = blockIdx.x, nwg = gridDim.x;
    for (;;) {
        phase_pass(P, nodes, parents, ws, N, phase, chunk, wg, nwg);
        if (nwg == 1) return;
        __syncthreads();
        if (threadIdx.x == 0)
            finished_last =
                __hip_atomic_fetch_add(ws.phase_done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == nwg - 1;
        __syncthreads();
        if (!finished_last) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ++phase;
        chunk = N;
        wg = 0;
        nwg = 1;
    }
}

// ------------------------------------------------------- candidates --

struct Best {
    float q;
    f3 b0, b1, b2;
    f3 dmin, dmax;
};

// (min, max) of 14 projections (transforms.cu:418-434).  v_min / v_max give the same value as the
// reference's compare-and-select rule (kmin / kmax) except for the sign of a
// zero result and for NaN inputs; a zero result, or a node whose points are
// not all finite and bounded (`sane`), takes the compare-and-select chain.
__device__ __forceinline__ void minmax14(const float (&p)[14], bool sane, float& mn, float& mx) {
    float a = fminf(p[0], p[1]), b = fmaxf(p[0], p[1]);
#pragma unroll
    for (int i = 2; i < 14; ++i) {
        a = fminf(a, p[i]);
        b = fmaxf(b, p[i]);
    }
    if (!sane || a == 0.0f || b == 0.0f) {
        a = p[0];
        b = p[0];
#pragma unroll
        for (int i = 1; i < 14; ++i) {
            a = kmin(a, p[i]);
            b = kmax(b, p[i]);
        }
    }
    mn = a;
    mx = b;
}

__device__ __forceinline__ void proj14(const f3* P, f3 d, bool sane, float& mn, float& mx) {
    float a = dot3(P[0], d), b = a;
#pragma unroll
    for (int i = 1; i < 14; ++i) {
        const float p = dot3(P[i], d);
        a = fminf(a, p);
        b = fmaxf(b, p);
    }
    if (!sane || a == 0.0f || b == 0.0f) {  // same projections, reference rule
        a = dot3(P[0], d);
        b = a;
#pragma unroll
        for (int i = 1; i < 14; ++i) {
            const float p = dot3(P[i], d);
            a = kmin(a, p);
            b = kmax(b, p);
        }
    }
    mn = a;
    mx = b;
}

// The candidate search over a point list P of NP points (NP a compile-time
// count).  Reference mode (kRef, NP = 14): the node's 14 extremal points in
// the reference's interleaved order, each projection range with the
// compare-and-select fallback above.  Distinct mode: the node's distinct
// extremal points, first occurrences in the interleaved order, padded to NP
// with copies of the first point.  Every search over the list is a
// first-index strict compare (a later copy of a point never wins) and every
// range a min / max (the same value for a multiset and its set), so both
// modes choose the same candidate with the same axes and lengths.  They can
// differ only in the sign of a zero range end, which no decision uses
// (lengths, qualities and their compares treat -0 as +0) and which the refit
// replaces: the refit's range is the key minimum / maximum over every
// subtree vertex, the extremal points included.  Distinct mode needs finite,
// bounded points (v_min / v_max); other nodes take reference mode.
template <int NP, bool kRef>
struct Pts {
    const f3* P;
    bool sane;  // reference mode: the points are finite and bounded

    __device__ __forceinline__ void proj(f3 d, float& mn, float& mx) const {
        if (kRef) {
            proj14(P, d, sane, mn, mx);
            return;
        }
        float a = dot3(P[0], d), b = a;
#pragma unroll
        for (int i = 1; i < NP; ++i) {
            const float p = dot3(P[i], d);
            a = fminf(a, p);
            b = fmaxf(b, p);
        }
        mn = a;
        mx = b;
    }
    __device__ __forceinline__ float2 projv(f3 d) const {
        float mn, mx;
        proj(d, mn, mx);
        return make_float2(mn, mx);
    }
};

__device__ __forceinline__ f3 pick(bool c, f3 a, f3 b) { return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ float2 pick2(bool c, float2 a, float2 b) {
    return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}

// One face of the ditetrahedron, transforms.cu:484-544: the normal's
// projection range, then for each edge e[k] (its range pe[k] computed by
// the caller: an edge is shared by up to three faces, its range is computed
// once with identical arithmetic) the third axis m = e x n; a candidate
// replaces the best one when its quality is strictly lower.
template <int NP, bool kRef>
__device__ __forceinline__ void face(const Pts<NP, kRef>& P, f3 n, float nmn, float nmx, f3 e0, f3 e1, f3 e2,
                                     float2 pe0, float2 pe1, float2 pe2, Best& b) {
    const float leny = nmx - nmn;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 e = k == 0 ? e0 : (k == 1 ? e1 : e2);
        const float2 pe = k == 0 ? pe0 : (k == 1 ? pe1 : pe2);
        const f3 m = vm::cross(e, n);
        float mmn, mmx;
        P.proj(m, mmn, mmx);
        const float lenx = pe.y - pe.x;
        const float lenz = mmx - mmn;
        const float q = quality(f3{lenx, leny, lenz});
        if (q < b.q) {
            b.q = q;
            b.dmax = f3{pe.y, nmx, mmx};
            b.dmin = f3{pe.x, nmn, mmn};
            b.b0 = e;
            b.b1 = n;
            b.b2 = m;
        }
    }
}

// Farthest slab pair (transforms.cu:437-456) over the interleaved list,
// first pair wins ties: its index s.
__device__ __forceinline__ int farthest_pair(const f3 (&P)[14]) {
    int at = 0;
    float d2b = vm::dist2(P[0], P[1]);
#pragma unroll
    for (int s = 1; s < 7; ++s) {
        const float d2 = vm::dist2(P[2 * s], P[2 * s + 1]);
        const bool c = d2 > d2b;
        d2b = c ? d2 : d2b;
        at = c ? s : at;
    }
    return at;
}

// Base triangle on the farthest pair (p0, p1), then the ditetrahedron
// (transforms.cu:546-674), over the list P.  The apex faces run as loops
// (two apexes, three faces each, in the reference's candidate order), so
// that the code per point count stays small.
template <int NP, bool kRef>
__device__ __forceinline__ void cand_search(const Pts<NP, kRef>& L, f3 p0, f3 p1, Best& b, int& degen) {
    const float eps = 0.000001f;
    const f3* P = L.P;
    degen = 0;
    if (vm::dist2(p0, p1) < eps) {
        degen = 1;
        return;
    }
    const f3 e0 = vm::normalize(sub3(p0, p1));
    float dm = vm::dist_inf_edge(P[0], p0, e0);
    f3 p2 = P[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) {
        const float d = vm::dist_inf_edge(P[i], p0, e0);
        const bool c = d > dm;
        dm = c ? d : dm;
        p2 = pick(c, P[i], p2);
    }
    if (dm < eps) {
        degen = 1;
        return;
    }
    const f3 e1 = vm::normalize(sub3(p1, p2));
    const f3 e2 = vm::normalize(sub3(p2, p0));
    const f3 nn = vm::normalize(vm::cross(e1, e0));
    const float2 pe0 = L.projv(e0), pe1 = L.projv(e1), pe2 = L.projv(e2);
    // the normal's projections: its range, then the apexes (extreme points
    // along n, first index wins, transforms.cu:609-628)
    float pn[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) pn[i] = dot3(P[i], nn);
    float nmn, nmx;
    if (kRef) {
        minmax14(*reinterpret_cast<const float(*)[14]>(pn), L.sane, nmn, nmx);
    } else {
        nmn = pn[0];
        nmx = pn[0];
#pragma unroll
        for (int i = 1; i < NP; ++i) {
            nmn = fminf(nmn, pn[i]);
            nmx = fmaxf(nmx, pn[i]);
        }
    }
    // the apexes' validity first: a node without one is degenerate and
    // keeps b at the AABB (its range then holds the AABB for finalize)
    float lo = pn[0], hi = pn[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) {
        lo = pn[i] < lo ? pn[i] : lo;
        hi = pn[i] > hi ? pn[i] : hi;
    }
    const float tri = dot3(p0, nn);
    const bool q0v = hi - eps > tri;
    const bool q1v = lo + eps < tri;
    if (!q0v && !q1v) {
        degen = 1;
        return;
    }
    face(L, nn, nmn, nmx, e0, e1, e2, pe0, pe1, pe2, b);
    lo = pn[0];
    hi = pn[0];
    f3 q1 = P[0], q0 = P[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) {
        const bool cl = pn[i] < lo, ch = pn[i] > hi;
        lo = cl ? pn[i] : lo;
        q1 = pick(cl, P[i], q1);
        hi = ch ? pn[i] : hi;
        q0 = pick(ch, P[i], q0);
    }
    // apex a = 0: q0 (candidates 3..11), a = 1: q1 (12..20); on apex q the
    // faces j = 0, 1, 2 have edges (e_j, f_{j+1}, f_j) with f_i = q - p_i
    // normalised and normal cross(f_{j+1}, e_j) normalised (transforms.cu:641-700)
#pragma unroll 1
    for (int a = 0; a < 2; ++a) {
        if (!(a == 0 ? q0v : q1v)) continue;
        const f3 q = pick(a == 0, q0, q1);
        const f3 f0 = vm::normalize(sub3(q, p0));
        const f3 f1 = vm::normalize(sub3(q, p1));
        const f3 f2 = vm::normalize(sub3(q, p2));
        const float2 pf0 = L.projv(f0), pf1 = L.projv(f1), pf2 = L.projv(f2);
#pragma unroll 1
        for (int j = 0; j < 3; ++j) {
            const f3 ej = pick(j == 0, e0, pick(j == 1, e1, e2));
            const float2 pej = pick2(j == 0, pe0, pick2(j == 1, pe1, pe2));
            const f3 fa = pick(j == 0, f1, pick(j == 1, f2, f0));
            const float2 pfa = pick2(j == 0, pf1, pick2(j == 1, pf2, pf0));
            const f3 fb = pick(j == 0, f0, pick(j == 1, f1, f2));
            const float2 pfb = pick2(j == 0, pf0, pick2(j == 1, pf1, pf2));
            const f3 nf = vm::normalize(vm::cross(fa, ej));
            float fmn, fmx;
            L.proj(nf, fmn, fmx);
            face(L, nf, fmn, fmx, ej, fa, fb, pej, pfa, pfb, b);
        }
    }
}

// Per internal node the candidate search (transforms.cu:677-756).  The
// search's loops run over the node's distinct extremal points: on average
// 5.4 of the 14 on a mesh LBVH (two triangles sharing an edge have four).  A
// wave runs the loops for the largest distinct count among its lanes, so the
// workgroup first sorts its 512 nodes by that count (LDS counting sort), lane
// t takes the t-th node in that order, and the wave runs the search compiled
// for 4, 6, 8 or 14 points.
constexpr int kCandBlock = 256;

struct CandTile {
    uint32_t cid[kCandBlock][15];   // distinct extremal vertex ids in list order (odd stride)
    uint32_t pair[kCandBlock][2];   // vertex ids of the farthest slab pair
    uint8_t cnt[kCandBlock];        // number of distinct ids; 0 = nothing to search; 15 = reference mode
    uint16_t order[kCandBlock];
    uint32_t node[kCandBlock];
    uint32_t hist[16];
};

__device__ __forceinline__ void kdop_ids(const Kdop& k, uint32_t (&id)[14]) {
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        id[2 * s] = k.imn[s];
        id[2 * s + 1] = k.imx[s];
    }
}

// Finite and below 2^61 in magnitude: no projection can be NaN or overflow.
__device__ __forceinline__ bool sane_points(const f3* P, int n) {
    uint32_t ex = 0;
    for (int i = 0; i < n; ++i) {
        ex = max(ex, __float_as_uint(P[i].x) & 0x7F800000u);
        ex = max(ex, __float_as_uint(P[i].y) & 0x7F800000u);
        ex = max(ex, __float_as_uint(P[i].z) & 0x7F800000u);
    }
    return ex <= ((127u + 60u) << 23);
}

// aq: half area of the node's AABB (the k-DOP's first three slabs), kept
// so that finalize reads the k-DOP only where the AABB wins.  The range
// starts as the result's projection range: for a degenerate node (which
// the refit skips and finalize boxes by its AABB) the search leaves b at
// the AABB, so finalize reads that from the range, not from the k-DOP.
__device__ __forceinline__ void store_cand(TreeWs& ws, uint32_t node, const Best& b, int degen, float aq) {
    Cand c;
    c.b[0] = b.b0.x; c.b[1] = b.b0.y; c.b[2] = b.b0.z;
    c.b[3] = b.b1.x; c.b[4] = b.b1.y; c.b[5] = b.b1.z;
    c.b[6] = b.b2.x; c.b[7] = b.b2.y; c.b[8] = b.b2.z;
    c.degen = degen;
    c.aq = aq;
    c.pad = 0;
    ws.cand[node] = c;
    int32_t* rk = &ws.range[(size_t)node * 6];
    rk[0] = f2ord(b.dmin.x); rk[1] = f2ord(b.dmin.y); rk[2] = f2ord(b.dmin.z);
    rk[3] = f2ord(b.dmax.x); rk[4] = f2ord(b.dmax.y); rk[5] = f2ord(b.dmax.z);
}

__device__ __forceinline__ Best aabb_best(const Kdop& k) {
    const f3 alen = f3{k.mx[0] - k.mn[0], k.mx[1] - k.mn[1], k.mx[2] - k.mn[2]};
    return Best{quality(alen), f3{1, 0, 0}, f3{0, 1, 0}, f3{0, 0, 1}, f3{k.mn[0], k.mn[1], k.mn[2]},
                f3{k.mx[0], k.mx[1], k.mx[2]}};
}

// Reference mode over the node's 14 points in list order, for the rare
// nodes with non-finite or huge extremal points (list kRedoList), run by
// cand_redo.

__device__ __forceinline__ void cand_ref(const kr_vec3* __restrict__ V, TreeWs& ws, uint32_t node) {
    Kdop k;
    kdop_load(ws.kdop, node, k);
    uint32_t id[14];
    kdop_ids(k, id);
    f3 Q[14];
#pragma unroll
    for (int i = 0; i < 14; ++i) Q[i] = ldv(V, id[i]);
    const int at = farthest_pair(Q);
    f3 p0 = Q[0], p1 = Q[1];
#pragma unroll
    for (int s = 1; s < 7; ++s)
        if (at == s) { p0 = Q[2 * s]; p1 = Q[2 * s + 1]; }
    Best b = aabb_best(k);
    const float aq = b.q;
    int degen = 0;
    const Pts<14, true> L{Q, sane_points(Q, 14)};
    cand_search(L, p0, p1, b, degen);
    store_cand(ws, node, b, degen, aq);
}

// Distinct-mode search of one node with the list padded to NP points.
template <int NP>
__device__ __forceinline__ void cand_distinct(const kr_vec3* __restrict__ V, TreeWs& ws, const CandTile& T,
                                              uint32_t s, uint32_t node, int cnt) {
    f3 P[NP];
    P[0] = ldv(V, T.cid[s][0]);
#pragma unroll
    for (int i = 1; i < NP; ++i) P[i] = i < cnt ? ldv(V, T.cid[s][i]) : P[0];  // copies of the first point pad the list
    const float4* kd = &ws.kdop[(size_t)node * 7];
    const float4 s0 = kd[0], s1 = kd[1], s2 = kd[2];
    const f3 alen = f3{s0.z - s0.x, s1.z - s1.x, s2.z - s2.x};
    const float aq = quality(alen);
    Best b = {aq, f3{1, 0, 0}, f3{0, 1, 0}, f3{0, 0, 1}, f3{s0.x, s1.x, s2.x}, f3{s0.z, s1.z, s2.z}};
    const f3 p0 = ldv(V, T.pair[s][0]), p1 = ldv(V, T.pair[s][1]);
    int degen = 0;
    const Pts<NP, false> L{P, true};
    cand_search(L, p0, p1, b, degen);
    store_cand(ws, node, b, degen, aq);
}

__device__ __forceinline__ bool holds(uint2 sp, uint32_t first, uint32_t last) {
    return sp.x <= first && sp.y >= last && !(sp.y & kGap);
}

// The ancestors of refit tile `tile` that hold all of its leaves: from the
// tile's first leaf up to the lowest node whose span covers the tile without
// a gap, then that node and its ancestors (at most kTChain; the next one is
// kept for a walk).
__device__ void tile_chain(const uint32_t* __restrict__ parents, TreeWs& ws, uint32_t N, uint32_t tile) {
    const uint32_t I = N - 1, lo = tile * kRTile, thi = min(lo + kRTile, N) - 1;
    uint32_t p = parents[I + lo];
    for (uint32_t step = 0; p < I && step < N && !holds(ws.span[p], lo, thi); ++step) p = parents[p];
    uint32_t n = 0;
    while (p < I && n < kTChain) {
        ws.tchain[(size_t)tile * kTChain + n] = p;
        ++n;
        p = parents[p];
    }
    ws.tmeta[tile] = make_uint2(n, p < I ? p : kInvalid);
}

// Arrival word of a node completed by the tile climb: both tags are phase 0
// (tag = owner + 1 < 2^23), so their sum stays below 2^24; later phases add
// a tag of at least 2^24, a deferred node carries bit 63 (kDeferred) and a
// hole's word is all ones.
__device__ __forceinline__ bool tile_complete(uint64_t v) { return complete(v) && (uint32_t)(v >> 32) < (1u << 24); }

// One workgroup's nodes (node[t] for lane t, kInvalid for none).
__device__ __forceinline__ void cand_block(const kr_vec3* __restrict__ V, TreeWs& ws, uint32_t I, CandTile& T,
                                           uint32_t node, bool bad) {
    const uint32_t t = threadIdx.x;
    if (t < 16) T.hist[t] = 0;
    T.node[t] = node;
    __syncthreads();
    // 1. the lane's own node: distinct ids and the farthest slab pair
    {
        uint32_t cnt = 0;
        if (node < I) {
            if (hole(ws.visit[node]) || bad) {
                Cand c = {};
                c.degen = 1;
                ws.cand[node] = c;
            } else {
                Kdop k;
                kdop_load(ws.kdop, node, k);
                uint32_t id[14];
                kdop_ids(k, id);
                f3 P[14];
#pragma unroll
                for (int i = 0; i < 14; ++i) P[i] = ldv(V, id[i]);
                if (!sane_points(P, 14)) {  // reference mode, after this kernel
                    const uint32_t slot = atomicAdd(&ws.nlist[kRedoList], 1u);
                    if (slot < I) ws.list[(size_t)kRedoList * (I + 1) + slot] = node;
                } else {
                    const int at = farthest_pair(P);
                    uint32_t a = id[0], b = id[1];
#pragma unroll
                    for (int s = 1; s < 7; ++s)
                        if (at == s) { a = id[2 * s]; b = id[2 * s + 1]; }
                    T.pair[t][0] = a;
                    T.pair[t][1] = b;
#pragma unroll
                    for (int i = 0; i < 14; ++i) {
                        bool dup = false;
#pragma unroll
                        for (int j = 0; j < i; ++j) dup |= id[j] == id[i];
                        if (!dup) T.cid[t][cnt++] = id[i];
                    }
                }
            }
        }
        T.cnt[t] = (uint8_t)cnt;
        // counting sort by distinct count (rank within a count from an LDS
        // atomic: the order among equal counts does not matter)
        const uint32_t r = atomicAdd(&T.hist[cnt], 1u);
        __syncthreads();
        if (t == 0) {
            uint32_t acc = 0;
            for (int c = 0; c < 16; ++c) {
                const uint32_t h = T.hist[c];
                T.hist[c] = acc;
                acc += h;
            }
        }
        __syncthreads();
        T.order[T.hist[cnt] + r] = (uint16_t)t;
        __syncthreads();
    }
    // 2. the t-th node in count order, searched over the wave's point count
    const uint32_t s = T.order[t];
    const int cnt = T.cnt[s];
    const int n = wave_max_i32(cnt);  // every lane is active here
    if (cnt != 0) {
        const uint32_t nd = T.node[s];
        if (n <= 4) cand_distinct<4>(V, ws, T, s, nd, cnt);
        else if (n <= 6) cand_distinct<6>(V, ws, T, s, nd, cnt);
        else if (n <= 8) cand_distinct<8>(V, ws, T, s, nd, cnt);
        else cand_distinct<14>(V, ws, T, s, nd, cnt);
    }
    __syncthreads();  // T is reused by the workgroup's next round
}

// The candidate search, in two launches.  kUpper = false: the nodes the
// tile climb completed (index order, one round per workgroup), launched
// on the side stream beside the top-of-tree climb phases, which leave most
// of the chip idle; holes get an AABB record.  kUpper = true: the nodes the
// later phases completed (their list), and workgroups below `cblocks` find
// the refit tiles' ancestor chains.
template <bool kUpper>
__global__ __launch_bounds__(kCandBlock, 4) void candidates(const kr_vec3* __restrict__ V,
                                                            const uint32_t* __restrict__ parents, TreeWs ws,
                                                            uint32_t N, uint32_t cblocks) {
    __shared__ CandTile T;
    const uint32_t t = threadIdx.x, I = N - 1;
    if (kUpper && blockIdx.x < cblocks) {
        const uint32_t tile = blockIdx.x * kCandBlock + t;
        if (tile < (N + kRTile - 1) / kRTile) {
            if (*ws.error != 0) ws.tmeta[tile] = make_uint2(0u, kInvalid);
            else tile_chain(parents, ws, N, tile);
        }
        return;
    }
    const bool bad = *ws.error != 0;
    if (!kUpper) {
        const uint32_t node = blockIdx.x * kCandBlock + t;
        bool mine = false;
        if (node < I) {
            const uint64_t vw = ws.visit[node];
            mine = hole(vw) || tile_complete(vw);
        }
        cand_block(V, ws, I, T, mine ? node : kInvalid, bad);
        return;
    }
    const uint32_t nb = gridDim.x - cblocks, nup = min(ws.nlist[kUpperList], I);
    const uint32_t* up = ws.list + (size_t)kUpperList * (I + 1);
    for (uint32_t base = (blockIdx.x - cblocks) * kCandBlock; base < nup; base += nb * kCandBlock)  // uniform
        cand_block(V, ws, I, T, base + t < nup ? up[base + t] : kInvalid, bad);
}

// The candidate search in reference mode on the nodes the distinct-mode
// search handed back (a separate launch after candidates: its registers do
// not add to the distinct-mode search's).
constexpr uint32_t kRedoBlocks = 8;
__global__ __launch_bounds__(1024) void cand_redo(const kr_vec3* __restrict__ V, TreeWs ws, uint32_t I) {
    const uint32_t nredo = min(ws.nlist[kRedoList], I);
    const uint32_t* redo = ws.list + (size_t)kRedoList * (I + 1);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nredo; r += gridDim.x * blockDim.x)
        cand_ref(V, ws, redo[r]);
}

// ------------------------------------------------------------ refit --
//
// Exact projection range of every subtree vertex on each internal node's
// candidate axes (transforms.cu:836-900), as order-preserving int32 keys
// (-0 orders below +0), min/max over the subtree's vertices.  A workgroup
// takes a tile of 512 consecutive leaves and splits each leaf's ancestors
// in three tiers:
//   A  below the lowest ancestor holding all 64 leaves of the lane's wave:
//      each lane folds its triangle's keys into LDS (a slot per internal
//      node of the tile's index range, a small hash for other nodes);
//   B  from that node up to the lowest ancestor holding the whole tile: the
//      wave walks together, reduces the 64 lanes' keys with DPP and one lane
//      folds them into LDS;
//   C  the tile's common ancestor and everything above it: the tile's
//      distinct vertices are collected once (LDS hash set on vertex ids) and
//      each wave takes some of those ancestors, looping over the distinct
//      vertices instead of 3 per triangle.
// "Holds all leaves of a range" is read from the span: its first and last
// leaf cover the range and the leaves below are contiguous.  LDS results are
// folded into global memory at the end.  Without the
// tile-wide pass (too many distinct vertices, or no tile-common node) tier B
// continues to the root.
// The walks of tiers A and B read the parent, span and axes of nodes in the
// tile's own index range [lo, lo + 512) from an LDS copy made at the start
// (a Karras tree numbers every node below the tile-common node inside that
// range; other nodes are read from global memory).  The ancestors for tier C
// (the tile-common node and up to kTChain - 1 of its ancestors) were listed
// per tile by the candidate kernel's leading workgroups, so no workgroup
// walks the ~10-30 dependent parent loads itself; ancestors beyond the list
// (very deep trees) are walked by every wave with its lanes' triangles.

constexpr int kRHash = 256;      // LDS hash for ancestors outside the tile's index range
constexpr int kVSet = 2048;      // distinct vertex ids per tile (hash set)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int32_t kKeyHi = 0x7FFFFFFF, kKeyLo = (int32_t)0x80000000;

struct RefitNodes {              // internal nodes lo .. lo + 511 (tiers A and B)
    float b[9][kRTile];          // candidate axes
    uint2 span[kRTile];
    uint32_t parent[kRTile];
    uint8_t degen[kRTile];
};

struct RefitTile {
    int32_t tk[kRTile][6];        // keys of internal node lo + s: min b0..b2, max b0..b2
    uint32_t hkey[kRHash];
    int32_t hval[kRHash][6];
    uint32_t vkey[kVSet];        // distinct vertex ids of the tile's triangles
    union {
        RefitNodes nd;           // until tiers A and B end
        float upos[kVSet][3];    // then: the distinct vertices' positions, compacted
    };
    uint32_t chain[kTChain];     // tile-common node and its ancestors
    uint32_t nchain, nuniq, overflow, rest;
};

// Parent, span and candidate axes of internal node p for tiers A and B.
__device__ __forceinline__ void refit_node(const RefitTile& R, const TreeWs& ws, const uint32_t* parents,
                                           uint32_t lo, uint32_t p, uint32_t& parent, uint2& span, Cand& c) {
    const uint32_t s = p - lo;
    if (s < kRTile) {
        parent = R.nd.parent[s];
        span = R.nd.span[s];
#pragma unroll
        for (int j = 0; j < 9; ++j) c.b[j] = R.nd.b[j][s];
        c.degen = R.nd.degen[s];
    } else {
        parent = parents[p];
        span = ws.span[p];
        c = ws.cand[p];
    }
}

// Order-preserving key of a float (same map as f2ord, without a branch).
__device__ __forceinline__ int32_t okey(float f) {
    const int32_t i = __float_as_int(f);
    return i ^ (int32_t)((uint32_t)(i >> 31) >> 1);
}

// Unconditional: reading the current keys first to skip atomics made every
// fold wait for a load of a line other workgroups update (about a fifth of
// the refit's time on a 2M-leaf tree).
__device__ __forceinline__ void fold_global(int32_t* rk, const int32_t v[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        atomicMin(&rk[a], v[a]);
        atomicMax(&rk[3 + a], v[3 + a]);
    }
}

__device__ __forceinline__ void fold_lds(int32_t* slot, const int32_t v[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (v[a] < slot[a]) atomicMin(&slot[a], v[a]);
        if (v[3 + a] > slot[3 + a]) atomicMax(&slot[3 + a], v[3 + a]);
    }
}

__device__ __forceinline__ void fold_tile(RefitTile& R, int32_t* range, uint32_t lo, uint32_t p, const int32_t k[6]) {
    if (p - lo < kRTile) {
        fold_lds(R.tk[p - lo], k);
        return;
    }
    uint32_t h = (p * 2654435761u) >> 24;
    for (int probe = 0; probe < 16; ++probe) {  // linear probing
        uint32_t key = R.hkey[h];
        if (key == kEmpty) key = atomicCAS(&R.hkey[h], kEmpty, p);
        if (key == kEmpty || key == p) {
            fold_lds(R.hval[h], k);
            return;
        }
        h = (h + 1) & (kRHash - 1);
    }
    fold_global(&range[(size_t)p * 6], k);
}

// Keys of one triangle's projection range on the node's axes.  The
// reference reduces the three projections with kr_min / kr_max from vertex 0
// and lets the result enter the node's range only when strictly beyond it
// (transforms.cu:855-897).  For finite vertices (fin) the dot products are
// never NaN and that is the integer min / max of the keys (up to the sign of
// a zero).  Otherwise the compare-and-select chain is evaluated as written,
// and a NaN result gives the neutral key, i.e. never enters the range.
__device__ __forceinline__ void tri_keys(const Cand& c, f3 v0, f3 v1, f3 v2, bool fin, int32_t k[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const f3 b = f3{c.b[3 * a], c.b[3 * a + 1], c.b[3 * a + 2]};
        const float d0 = dot3(v0, b), d1 = dot3(v1, b), d2 = dot3(v2, b);
        if (fin) {
            const int32_t k0 = okey(d0), k1 = okey(d1), k2 = okey(d2);
            k[a] = min(min(k0, k1), k2);
            k[3 + a] = max(max(k0, k1), k2);
        } else {
            const float lo = kmin(kmin(d0, d1), d2), hi = kmax(kmax(d0, d1), d2);
            k[a] = lo == lo ? okey(lo) : kKeyHi;
            k[3 + a] = hi == hi ? okey(hi) : kKeyLo;
        }
    }
}

// All coordinates finite (exponent field not all ones).
__device__ __forceinline__ bool finite3(f3 v) {
    const uint32_t m = 0x7F800000u;
    return (__float_as_uint(v.x) & m) != m && (__float_as_uint(v.y) & m) != m && (__float_as_uint(v.z) & m) != m;
}

__device__ __forceinline__ void wave_keys(int32_t k[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        k[a] = wave_min_i32(k[a]);
        k[3 + a] = wave_max_i32(k[3 + a]);
    }
}

__global__ __launch_bounds__(kRTile) void refit(const kr_bvh_node* __restrict__ nodes,
                                               const uint32_t* __restrict__ parents,
                                               const kr_vec3* __restrict__ V,
                                               const kr_uvec4* __restrict__ F, TreeWs ws, uint32_t N) {
    __shared__ RefitTile R;
    __shared__ uint32_t bad;
    const uint32_t I = N - 1, t = threadIdx.x;
    const uint32_t lo = blockIdx.x * kRTile;
    const uint32_t thi = min(lo + kRTile, N) - 1;  // last leaf position of the tile
#pragma unroll
    for (int a = 0; a < 3; ++a) { R.tk[t][a] = kKeyHi; R.tk[t][3 + a] = kKeyLo; }
    if (t < kRHash) {
        R.hkey[t] = kEmpty;
#pragma unroll
        for (int a = 0; a < 3; ++a) { R.hval[t][a] = kKeyHi; R.hval[t][3 + a] = kKeyLo; }
    }
    for (uint32_t h = t; h < kVSet; h += kRTile) R.vkey[h] = kEmpty;
    if (lo + t < I) {  // the tile's own index range, for tiers A and B
        const uint32_t node = lo + t;
        const Cand c = ws.cand[node];
#pragma unroll
        for (int j = 0; j < 9; ++j) R.nd.b[j][t] = c.b[j];
        R.nd.degen[t] = (uint8_t)(c.degen != 0);
        R.nd.span[t] = ws.span[node];
        R.nd.parent[t] = parents[node];
    }
    if (t == 0) {
        bad = *ws.error;  // one read per workgroup: the exit is uniform
        const uint2 m = ws.tmeta[blockIdx.x];
        R.nchain = min(m.x, (uint32_t)kTChain);
        R.rest = m.y;
        R.nuniq = 0;
        R.overflow = 0;
    }
    if (t < kTChain) R.chain[t] = ws.tchain[(size_t)blockIdx.x * kTChain + t];  // entries past the count unused
    __syncthreads();
    if (bad != 0) return;

    const uint32_t pos = lo + t;
    const uint32_t wlo = pos - (uint32_t)lane_id();
    const uint32_t whi = min(wlo + kWave - 1, N - 1);
    const bool valid = pos < N;
    f3 v0 = f3{0, 0, 0}, v1 = v0, v2 = v0;
    bool fin = true;
    if (valid) {
        const kr_uvec4 face = F[nodes[I + pos].u.primitivesOffset];
        v0 = ldv(V, face.x);
        v1 = ldv(V, face.y);
        v2 = ldv(V, face.z);
        fin = finite3(v0) && finite3(v1) && finite3(v2);
        // the tile-wide pass projects vertices, not triangles: a tile with a
        // non-finite vertex keeps the per-triangle rule up to the root
        if (!fin) R.overflow = 1;
        const uint32_t vid[3] = {face.x, face.y, face.z};
#pragma unroll
        for (int j = 0; j < 3; ++j) {  // distinct vertex ids of the tile
            uint32_t h = (vid[j] * 2654435761u) >> 21;
            bool placed = false;
            for (int probe = 0; probe < 32 && !placed; ++probe) {
                uint32_t key = R.vkey[h];
                if (key == kEmpty) key = atomicCAS(&R.vkey[h], kEmpty, vid[j]);
                placed = key == kEmpty || key == vid[j];
                h = (h + 1) & (kVSet - 1);
            }
            if (!placed) R.overflow = 1;
        }
    }
    __syncthreads();
    // tier C needs the tile-common node (chain[0]) and a vertex set that fit
    const bool tile_pass = R.overflow == 0 && R.nchain != 0;
    const uint32_t q0 = tile_pass ? R.chain[0] : kInvalid;

    if (wlo < N) {  // wave-uniform
        // tier A
        uint32_t common = kInvalid;
        if (valid) {
            uint32_t p = parents[I + pos];
            for (uint32_t step = 0; p != kInvalid && p < I && step < N; ++step) {
                uint32_t pp;
                uint2 sp;
                Cand c;
                refit_node(R, ws, parents, lo, p, pp, sp, c);
                if (holds(sp, wlo, whi)) { common = p; break; }
                if (!c.degen) {
                    int32_t k[6];
                    tri_keys(c, v0, v1, v2, fin, k);
                    fold_tile(R, ws.range, lo, p, k);
                }
                p = pp;
            }
        }
        // tier B; lane 0 is valid (wlo < N) and a well-formed tree gives
        // every valid lane the same node
        const uint32_t c0 = __builtin_amdgcn_readfirstlane(common);
        if (__ballot(valid && common != c0) != 0) {
            if (lane_id() == 0) atomicOr(ws.error, 2u);
        } else {
            uint32_t q = c0;
            for (uint32_t step = 0; q != kInvalid && q < I && step < N; ++step) {  // uniform
                uint32_t qp;
                uint2 sp;
                Cand c;
                if (q == q0) break;  // tier C takes q and above
                refit_node(R, ws, parents, lo, q, qp, sp, c);
                if (!c.degen) {
                    int32_t k[6];
                    if (valid) {
                        tri_keys(c, v0, v1, v2, fin, k);
                    } else {
#pragma unroll
                        for (int a = 0; a < 3; ++a) { k[a] = kKeyHi; k[3 + a] = kKeyLo; }
                    }
                    wave_keys(k);
                    if (lane_id() == 0) fold_tile(R, ws.range, lo, q, k);
                }
                q = __builtin_amdgcn_readfirstlane(qp);
            }
        }
    }
    __syncthreads();

    if (tile_pass) {
        for (uint32_t h = t; h < kVSet; h += kRTile) {
            const uint32_t id = R.vkey[h];
            if (id == kEmpty) continue;
            const uint32_t u = atomicAdd(&R.nuniq, 1u);
            const f3 p = ldv(V, id);
            R.upos[u][0] = p.x;
            R.upos[u][1] = p.y;
            R.upos[u][2] = p.z;
        }
        __syncthreads();
        // tier C over the distinct vertices; each wave reads the axes and
        // current keys of its next ancestor while it works on this one (the
        // keys only filter the atomics: they only move outwards, so an early
        // read never skips a needed update)
        const uint32_t nu = R.nuniq, nc = R.nchain;
        const uint32_t step_c = kRTile / kWave;
        uint32_t j = (uint32_t)wave_id();
        uint32_t qc = 0;
        Cand c = {};
        int32_t cur[6] = {0, 0, 0, 0, 0, 0};
        if (j < nc) {
            qc = __builtin_amdgcn_readfirstlane(R.chain[j]);
            c = ws.cand[qc];
#pragma unroll
            for (int a = 0; a < 6; ++a) cur[a] = ws.range[(size_t)qc * 6 + a];
        }
        for (; j < nc; j += step_c) {
            uint32_t qn = 0;
            Cand cn = {};
            int32_t curn[6] = {0, 0, 0, 0, 0, 0};
            if (j + step_c < nc) {
                qn = __builtin_amdgcn_readfirstlane(R.chain[j + step_c]);
                cn = ws.cand[qn];
#pragma unroll
                for (int a = 0; a < 6; ++a) curn[a] = ws.range[(size_t)qn * 6 + a];
            }
            if (!c.degen) {
                int32_t k[6];
#pragma unroll
                for (int a = 0; a < 3; ++a) { k[a] = kKeyHi; k[3 + a] = kKeyLo; }
                for (uint32_t u = (uint32_t)lane_id(); u < nu; u += kWave) {
                    const f3 p = f3{R.upos[u][0], R.upos[u][1], R.upos[u][2]};
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const int32_t kk = okey(dot3(p, f3{c.b[3 * a], c.b[3 * a + 1], c.b[3 * a + 2]}));
                        k[a] = min(k[a], kk);
                        k[3 + a] = max(k[3 + a], kk);
                    }
                }
                wave_keys(k);
                if (lane_id() == 0) {
                    int32_t* rk = &ws.range[(size_t)qc * 6];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        if (k[a] < cur[a]) atomicMin(&rk[a], k[a]);
                        if (k[3 + a] > cur[3 + a]) atomicMax(&rk[3 + a], k[3 + a]);
                    }
                }
            }
            qc = qn;
            c = cn;
#pragma unroll
            for (int a = 0; a < 6; ++a) cur[a] = curn[a];
        }
        // ancestors beyond the recorded batch (very deep trees): every wave
        // walks them with its lanes' triangles
        uint32_t q = R.rest;
        if (wlo < N) {
            for (uint32_t step = 0; q != kInvalid && q < I && step < N; ++step) {  // uniform
                const Cand c = ws.cand[q];
                if (!c.degen) {
                    int32_t k[6];
                    if (valid) {
                        tri_keys(c, v0, v1, v2, fin, k);
                    } else {
#pragma unroll
                        for (int a = 0; a < 3; ++a) { k[a] = kKeyHi; k[3 + a] = kKeyLo; }
                    }
                    wave_keys(k);
                    if (lane_id() == 0) fold_global(&ws.range[(size_t)q * 6], k);
                }
                q = __builtin_amdgcn_readfirstlane(parents[q]);
            }
        }
    }
    __syncthreads();
    // fold the tile's LDS keys into global memory.  A node whose leaves all
    // lie in this tile has no contributor outside this workgroup, and within
    // it only the LDS slot (tier C, which folds with atomics, never takes
    // such a node except the tile-common node, whose span is the tile): its
    // keys are merged with a plain read-modify-write.  The others are merged
    // with atomics.
    if (lo + t < I && (R.tk[t][0] != kKeyHi || R.tk[t][3] != kKeyLo)) {
        const uint32_t node = lo + t;
        const uint2 sp = ws.span[node];
        int32_t* rk = &ws.range[(size_t)node * 6];
        const bool chain_node = node == q0;
        if (sp.x >= lo && span_hi(sp) <= thi && !chain_node) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                rk[a] = min(rk[a], R.tk[t][a]);
                rk[3 + a] = max(rk[3 + a], R.tk[t][3 + a]);
            }
        } else {
            fold_global(rk, R.tk[t]);
        }
    }
    if (t < kRHash && R.hkey[t] != kEmpty) fold_global(&ws.range[(size_t)R.hkey[t] * 6], R.hval[t]);
}

// --------------------------------------------------------- finalize --
//
// One lane per internal node (transforms.cu:905-964): OBB vs AABB by half
// area, extent floor, inverse(T * R * S).  Boxes and matrices are staged in
// LDS (odd row strides, no bank conflicts) and written back as contiguous
// dwords.

__global__ __launch_bounds__(kStage) void finalize(TreeWs ws, kr_obb3* obbs, kr_mat4* mats, uint32_t I,
                                                   BatchMap bm) {
    __shared__ Stage st;
    const uint32_t base = blockIdx.x * kStage;
    const uint32_t node = base + threadIdx.x;
    float* ob = &st.ob[threadIdx.x * 15];
    float* mt = &st.mt[threadIdx.x * 17];
    if (node < I) {
        const uint64_t vw = ws.visit[node];
        if (!complete(vw)) atomicOr(ws.error, 8u);  // a malformed tree: the node never completed
        if (*ws.error != 0 || !complete(vw) || hole(vw)) {
            stage_nan(st, threadIdx.x);
        } else {
            const Cand c = ws.cand[node];
            const int32_t* rk = &ws.range[(size_t)node * 6];
            const f3 rmn = f3{ord2f(rk[0]), ord2f(rk[1]), ord2f(rk[2])};
            const f3 rmx = f3{ord2f(rk[3]), ord2f(rk[4]), ord2f(rk[5])};
            f3 v0, v1, v2, mn, mx;
            if (!c.degen && quality(sub3(rmx, rmn)) < c.aq) {
                v0 = f3{c.b[0], c.b[1], c.b[2]};
                v1 = f3{c.b[3], c.b[4], c.b[5]};
                v2 = f3{c.b[6], c.b[7], c.b[8]};
                mn = rmn;
                mx = rmx;
            } else {  // the AABB: the k-DOP's first three slabs
                v0 = f3{1, 0, 0};
                v1 = f3{0, 1, 0};
                v2 = f3{0, 0, 1};
                if (c.degen) {  // the candidate step kept them in the range, which the refit skips
                    mn = rmn;
                    mx = rmx;
                } else {
                    const float4* kd = &ws.kdop[(size_t)node * 7];
                    const float4 s0 = kd[0], s1 = kd[1], s2 = kd[2];
                    mn = f3{s0.x, s1.x, s2.x};
                    mx = f3{s0.z, s1.z, s2.z};
                }
            }
            const f3 mid = vm::mul1(add3(mn, mx), 0.5f);
            const f3 ext = vm::vmax3(vm::mul1(sub3(mx, mn), 0.5f), f3{0.001f, 0.001f, 0.001f});
            f3 m = vm::mul1(v0, mid.x);
            m = add3(m, vm::mul1(v1, mid.y));
            m = add3(m, vm::mul1(v2, mid.z));
            obb_row(ob, m, v0, v1, v2, ext);
            float M[16];
            obb_matrix(v0, v1, v2, m, ext, M);
#pragma unroll
            for (int q = 0; q < 16; ++q) mt[q] = M[q];
        }
    }
    __syncthreads();
    stage_flush(st, obbs, mats, base, min((uint32_t)kStage, I - min(base, I)), bm);
}

__global__ void copy_ids(const kr_bvh_node* __restrict__ nodes, const kr_vec3* __restrict__ V,
                         const kr_uvec4* __restrict__ F, TreeWs ws, uint32_t* ids, uint32_t total) {
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= total) return;
    const uint32_t I = (total - 1) / 2, N = I + 1;
    Kdop k;
    bool ok = *ws.error == 0;
    if (ok) {
        if (node < I) {
            ok = complete(ws.visit[node]);
            if (ok) kdop_load(ws.kdop, node, k);
        } else {
            ok = leaf_kdop_global(nodes, V, F, node, N, k);
        }
    }
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        ids[(size_t)node * 14 + s] = ok ? k.imn[s] : 0xFFFFFFFFu;
        ids[(size_t)node * 14 + 7 + s] = ok ? k.imx[s] : 0xFFFFFFFFu;
    }
}

__global__ void copy_status(const uint32_t* error, uint32_t* status) { *status = *error; }

// ------------------------------------------------------------- host --

size_t tree_carve(TreeWs* ws, void* base, size_t cap, uint32_t N) {
    Carver c{(char*)base, 0, cap};
    const size_t I = N ? N - 1 : 0, total = N ? 2 * (size_t)N - 1 : 0;
    ws->visit = c.take<uint64_t>(I + 1);
    ws->error = c.take<uint32_t>(8);
    ws->nlist = ws->error ? ws->error + 2 : nullptr;  // error[2 .. 2 + kMaxPhases)
    ws->phase_done = ws->error ? ws->error + 6 : nullptr;
    ws->list = c.take<uint32_t>((size_t)kMaxPhases * (I + 1));
    ws->span = c.take<uint2>(total);
    ws->count = c.take<uint32_t>(I + 1);
    ws->kdop = c.take<float4>(total * 7);
    ws->cand = c.take<Cand>(I + 1);
    ws->range = c.take<int32_t>((I + 1) * 6);
    const size_t tiles = ((size_t)N + kRTile - 1) / kRTile;
    ws->tchain = c.take<uint32_t>(tiles * kTChain);
    ws->tmeta = c.take<uint2>(tiles);
    return c.used;
}

// Per device: a second stream (default priority) and the events that fork
// the first candidate launch onto it and join it back.  One lock covers a whole call's
// enqueue sequence, so calls from several host threads cannot interleave
// their record / wait pairs.
namespace {
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kMaxDevices = 64;
std::mutex g_side_mu;
SideStream g_side[kMaxDevices];

kr_error side_stream(SideStream** out) {
    int dev = 0;
    OBVH_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return "obvh_obb_tree: device ordinal out of range";
    SideStream& ss = g_side[dev];
    if (ss.s == nullptr) {
        OBVH_HIP_TRY(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
        OBVH_HIP_TRY(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
        OBVH_HIP_TRY(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
    }
    *out = &ss;
    return nullptr;
}
}  // namespace

kr_error tree_launch(const kr_bvh_node* nodes, const uint32_t* parents, const kr_vec3* V,
                     const kr_uvec4* F, kr_obb3* obbs, kr_mat4* mats, uint32_t* ids, uint32_t N,
                     void* workspace, size_t ws_bytes, hipStream_t stream, const BatchMap* batch) {
    const BatchMap bm = batch ? *batch : BatchMap{nullptr, 0, 0};
    if (N == 0) return nullptr;
    TreeWs ws;
    size_t need = tree_carve(&ws, workspace, ws_bytes, N);
    if (workspace == nullptr || ws_bytes < need) return "obvh_obb_tree: workspace too small";
    const uint32_t I = N - 1;
    auto blocks = [](uint32_t n, uint32_t b) { return dim3((n + b - 1) / b); };
    std::lock_guard<std::mutex> lock(g_side_mu);
    SideStream* side = nullptr;
    if (kr_error e = side_stream(&side)) return e;
    hipLaunchKernelGGL(tree_clear, dim3(1024), dim3(256), 0, stream, ws.visit, I + 1, ws);
    if (bm.K > 1)
        hipLaunchKernelGGL(batch_holes, dim3(min((bm.K + 255) / 256, 1024u)), dim3(256), 0, stream, ws, bm);
    // climb: 512-leaf tiles, then 32768-leaf workgroups, then one workgroup
    // for the rest
    hipLaunchKernelGGL(climb_leaves, blocks(N, kTile), dim3(kTile), 0, stream, nodes, parents, V, F, ws, N, obbs, mats,
                       bm);
    // The side stream takes the candidate search of the nodes the tile climb
    // completed (nearly all of them; VALU-bound) beside the top-of-tree climb
    // phases, which are latency-bound on a few dozen workgroups.  The launch
    // stream joins it before the refit.
    if (I > 0) {
        OBVH_HIP_TRY(hipEventRecord(side->fork, stream));
        OBVH_HIP_TRY(hipStreamWaitEvent(side->s, side->fork, 0));
        hipLaunchKernelGGL(candidates<false>, blocks(I, kCandBlock), dim3(kCandBlock), 0, side->s, V, parents, ws, N,
                           0u);
        OBVH_HIP_TRY(hipEventRecord(side->join, side->s));
    }
    if (N > kTile) {  // phase 1 (and its last workgroup phase 2), or one phase for a small tree
        const uint32_t c = N > kChunk1 ? kChunk1 : N;
        hipLaunchKernelGGL(climb_phase, blocks(N, c), dim3(kPhaseThreads), 0, stream, nodes, parents, ws, N, 1u, c);
    }
    if (I > 0) {
        // the nodes the phases completed, and the refit tiles' chains
        const uint32_t cb = (((N + kRTile - 1) / kRTile) + kCandBlock - 1) / kCandBlock;
        const uint32_t ub = min((I + kCandBlock - 1) / kCandBlock, 1024u);
        hipLaunchKernelGGL(candidates<true>, dim3(cb + ub), dim3(kCandBlock), 0, stream, V, parents, ws, N, cb);
        OBVH_HIP_TRY(hipStreamWaitEvent(stream, side->join, 0));
        hipLaunchKernelGGL(cand_redo, dim3(kRedoBlocks), dim3(1024), 0, stream, V, ws, I);
        hipLaunchKernelGGL(refit, blocks(N, kRTile), dim3(kRTile), 0, stream, nodes, parents, V, F, ws, N);
        hipLaunchKernelGGL(finalize, blocks(I, kStage), dim3(kStage), 0, stream, ws, obbs, mats, I, bm);
    }
    if (ids)
        hipLaunchKernelGGL(copy_ids, blocks(2 * N - 1, 256), dim3(256), 0, stream, nodes, V, F, ws, ids, 2 * N - 1);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

// ------------------------------------------------------------ batch --
//
// Many trees in one transform.  The batch's trees (tree t: N_t >= 1 leaves,
// nodes, parents, faces and vertices back to back, indices local to the
// tree) become one forest in the single-tree numbering, so that every
// kernel runs once for the whole batch.  Each node's k-DOP, candidate axes,
// refit range and box depend only on its own subtree, so the forest gives
// every tree's nodes exactly the per-tree results.  The numbering keeps each
// tree Karras-aligned: with L_t the first leaf position of tree t (prefix
// sum of N_s), its internal node j goes to L_t + j + 1 (t < K - 1) or
// L_t + j (the last tree) and its leaf i to (sum N - 1) + L_t + i; the
// internal slots L_k (k < K - 1) stay unused (holes: marked complete with
// tag kHoleTag, skipped by the per-node kernels).  Every tree root has no
// parent.  Tree t's nodes start at 2 L_t - t in the input layout.

// One lane per input node: write the combined node and parent.  Bad local
// indices (children, parents, face ids outside the tree, an empty tree) set
// the batch error word and are mapped out of range, where the transform's
// own checks report them.
__global__ void batch_gather_nodes(BatchMap m, const kr_bvh_node* __restrict__ nodes,
                                   const uint32_t* __restrict__ parents, kr_bvh_node* cn, uint32_t* cp,
                                   uint32_t* berr, uint32_t total_in) {
    const uint32_t I = m.Nsum - 1, total = 2 * m.Nsum - 1;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total_in; g += gridDim.x * blockDim.x) {
        const uint32_t t = batch_tree_of_node(m, g);
        const uint32_t Lt = m.prim_off[t], N = m.prim_off[t + 1] - Lt;
        if (N == 0) { atomicOr(berr, 1u); continue; }
        const uint32_t nb = 2 * Lt - t, j = g - nb, nn = 2 * N - 1;
        if (j >= nn) { atomicOr(berr, 1u); continue; }
        const uint32_t c = batch_map(m, t, Lt, N, j);
        kr_bvh_node o = nodes[g];
        if (j + 1 < N) {
            const uint32_t l = o.u.children.left, r = o.u.children.right;
            if (l >= nn || r >= nn) atomicOr(berr, 2u);
            o.u.children.left = l < nn ? batch_map(m, t, Lt, N, l) : total;
            o.u.children.right = r < nn ? batch_map(m, t, Lt, N, r) : total;
        } else {
            const uint32_t f = o.u.primitivesOffset;
            if (f >= N) atomicOr(berr, 4u);
            o.u.primitivesOffset = f < N ? Lt + f : 0xFFFFFFFFu;
        }
        cn[c] = o;
        const uint32_t p = parents[g];
        uint32_t cpar;
        if (p == kInvalid) cpar = j == 0 ? kInvalid : total;  // only the root has none
        else cpar = p < nn ? batch_map(m, t, Lt, N, p) : total;
        if ((p == kInvalid) != (j == 0) || (p != kInvalid && p >= nn)) atomicOr(berr, 2u);
        cp[c] = cpar;
    }
    // the unused internal slots L_k (k < K - 1) between the trees
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k + 1 < m.K; k += gridDim.x * blockDim.x) {
        const uint32_t idx = m.prim_off[k];
        cn[idx] = kr_bvh_node{};
        cp[idx] = kInvalid;
    }
    (void)I;
}

// One lane per face: vertex ids rebased to the batch's vertex array.
__global__ void batch_gather_faces(BatchMap m, const uint32_t* __restrict__ vert_off,
                                   const kr_uvec4* __restrict__ faces, kr_uvec4* cf, uint32_t* berr) {
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < m.Nsum; f += gridDim.x * blockDim.x) {
        const uint32_t t = batch_tree_of_leaf(m, f);
        const uint32_t vb = vert_off[t], nv = vert_off[t + 1] - vb;
        kr_uvec4 a = faces[f];
        if (a.x >= nv || a.y >= nv || a.z >= nv) {
            atomicOr(berr, 8u);  // rebased to vertex 0 of the batch: never read outside the array
            a.x = a.x < nv ? vb + a.x : 0u;
            a.y = a.y < nv ? vb + a.y : 0u;
            a.z = a.z < nv ? vb + a.z : 0u;
        } else {
            a.x += vb;
            a.y += vb;
            a.z += vb;
        }
        cf[f] = a;
    }
}

struct BatchWs {
    kr_bvh_node* nodes;
    uint32_t* parents;
    kr_uvec4* faces;
    uint32_t* berr;
    void* tree_ws;
    size_t tree_ws_bytes;
};

size_t batch_carve(BatchWs* w, void* base, size_t cap, uint32_t Nsum) {
    Carver c{(char*)base, 0, cap};
    const size_t total = 2 * (size_t)Nsum - 1;
    w->berr = c.take<uint32_t>(4);
    w->nodes = c.take<kr_bvh_node>(total);
    w->parents = c.take<uint32_t>(total);
    w->faces = c.take<kr_uvec4>(Nsum);
    TreeWs tw;
    w->tree_ws_bytes = tree_carve(&tw, nullptr, 0, Nsum);
    w->tree_ws = c.take<char>(w->tree_ws_bytes);
    return c.used;
}

size_t batch_workspace_size(uint32_t Nsum) {
    BatchWs w;
    return batch_carve(&w, nullptr, 0, Nsum);
}

__global__ void batch_clear(uint32_t* berr) {
    if (threadIdx.x == 0) *berr = 0;
}

kr_error tree_batch_launch(const kr_bvh_node* nodes, const uint32_t* parents, const kr_vec3* V, const kr_uvec4* F,
                           const uint32_t* prim_off, const uint32_t* vert_off, uint32_t K, uint32_t Nsum,
                           kr_obb3* obbs, kr_mat4* mats, void* workspace, size_t ws_bytes, hipStream_t stream) {
    if (K == 0 || Nsum == 0) return nullptr;
    if (Nsum < K) return "obvh_obb_tree_batch: every tree needs at least one primitive";
    BatchWs w;
    const size_t need = batch_carve(&w, workspace, ws_bytes, Nsum);
    if (workspace == nullptr || ws_bytes < need) return "obvh_obb_tree_batch: workspace too small";
    const BatchMap m{prim_off, K, Nsum};
    const uint32_t total_in = 2 * Nsum - K;
    auto grid = [](uint32_t n) { return dim3(min((n + 255) / 256, 65536u)); };
    hipLaunchKernelGGL(batch_clear, dim3(1), dim3(64), 0, stream, w.berr);
    hipLaunchKernelGGL(batch_gather_nodes, grid(max(total_in, K)), dim3(256), 0, stream, m, nodes, parents, w.nodes,
                       w.parents, w.berr, total_in);
    hipLaunchKernelGGL(batch_gather_faces, grid(Nsum), dim3(256), 0, stream, m, vert_off, F, w.faces, w.berr);
    OBVH_HIP_TRY(hipGetLastError());
    // climb_leaves and finalize write every node's box and matrix straight
    // to its place in the caller's per-tree layout
    return tree_launch(w.nodes, w.parents, V, w.faces, obbs, mats, nullptr, Nsum, w.tree_ws, w.tree_ws_bytes, stream,
                       &m);
}

__global__ void copy_batch_status(const uint32_t* berr, const uint32_t* error, uint32_t* status) {
    const uint32_t b = *berr;
    uint32_t s = *error;
    if (b & 1u) s |= 16u;  // an empty tree or a node outside its tree
    if (b & 2u) s |= 2u;
    if (b & 4u) s |= 4u;
    if (b & 8u) s |= 32u;
    *status = s;
}

kr_error tree_batch_status_launch(const void* workspace, size_t ws_bytes, uint32_t Nsum, uint32_t* status,
                                  hipStream_t stream) {
    BatchWs w;
    const size_t need = batch_carve(&w, const_cast<void*>(workspace), ws_bytes, Nsum);
    if (workspace == nullptr || ws_bytes < need) return "obvh_obb_tree_batch_status: workspace too small";
    TreeWs tw;
    tree_carve(&tw, w.tree_ws, w.tree_ws_bytes, Nsum);
    hipLaunchKernelGGL(copy_batch_status, dim3(1), dim3(1), 0, stream, w.berr, tw.error, status);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error tree_status_launch(const void* workspace, size_t ws_bytes, uint32_t N, uint32_t* status,
                            hipStream_t stream) {
    TreeWs ws;
    const size_t need = tree_carve(&ws, const_cast<void*>(workspace), ws_bytes, N);
    if (workspace == nullptr || ws_bytes < need) return "obvh_obb_tree_status: workspace too small";
    hipLaunchKernelGGL(copy_status, dim3(1), dim3(1), 0, stream, ws.error, status);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

const char* tree_status_string(uint32_t err) {
    if (err & 2u) return "obvh_obb_tree: parent/child links disagree";
    if (err & 1u) return "obvh_obb_tree: parent index out of range or cycle";
    if (err & 4u) return "obvh_obb_tree: leaf face id out of range";
    if (err & 8u) return "obvh_obb_tree: internal node never completed (malformed tree)";
    if (err & 16u) return "obvh_obb_tree_batch: empty tree in the batch or node outside its tree";
    if (err & 32u) return "obvh_obb_tree_batch: face vertex id outside its tree's vertices";
    if (err != 0u) return "obvh_obb_tree: unknown device error";
    return nullptr;
}

}  // namespace obvh
