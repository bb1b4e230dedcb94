// transform.hip -- AABB-BVH -> OBB-tree transform on gfx950.
//
// Method of the reference GPU transform (src/ads/common/cuda/transforms.cu
// :213-964), restructured for CDNA4:
//
//  K1 leaf_climb   one lane per leaf: the leaf's 7-slab k-DOP (3 vertices,
//                  first vertex wins ties), the leaf OBB and its matrix
//                  (transforms.cu:803-834), then a bottom-up climb with one
//                  agent-scope acq_rel counter per internal node: the first
//                  child to arrive stops, the second merges.  The merge takes
//                  the right child's extreme only if strictly better, so ties
//                  go to the left child and the selected vertex ids do not
//                  depend on scheduling.  k-DOPs carry vertex ids, not
//                  positions (7 x {min, id, max, id} = 112 B per node).
//  K2 candidates   one lane per internal node: gather the 14 extremal
//                  vertices and run the base-triangle + ditetrahedron search
//                  entirely in registers (transforms.cu:413-756).
//  K3 refit        one lane per leaf: for every ancestor, project the 3
//                  vertices on its axes and lower/raise the ancestor's range
//                  with integer atomics on order-preserving float keys,
//                  after a plain-load filter (transforms.cu:836-900).  The
//                  result is the exact range of every subtree vertex.
//  K4 finalize     one lane per internal node: OBB vs AABB by half area,
//                  extent floor, inverse(T * R * S) (transforms.cu:905-964).
//
// Tie rule, degenerate nodes and floating point follow the reference with
// the decisions documented in the oracle (oracle/transform.c).
#include <float.h>

#include "common.hpp"
#include "transform.hpp"

namespace obvh {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kChunk0 = 1024;    // leaves per workgroup, first climb phase
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

__device__ __forceinline__ void store_obb(kr_obb3* o, f3 mid, f3 v0, f3 v1, f3 v2, f3 ext) {
    float* d = reinterpret_cast<float*>(o);
    d[0] = mid.x; d[1] = mid.y; d[2] = mid.z;
    d[3] = v0.x; d[4] = v0.y; d[5] = v0.z;
    d[6] = v1.x; d[7] = v1.y; d[8] = v1.z;
    d[9] = v2.x; d[10] = v2.y; d[11] = v2.z;
    d[12] = ext.x; d[13] = ext.y; d[14] = ext.z;
}
__device__ __forceinline__ void store_mat(kr_mat4* m, const float* v) {
    float4* d = reinterpret_cast<float4*>(m);
#pragma unroll
    for (int i = 0; i < 4; ++i) d[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
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
__device__ __forceinline__ void leaf_obb(f3 v0, f3 v1, f3 v2, kr_obb3* obb, kr_mat4* mat) {
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
    store_obb(obb, m, a0, a1, a2, ext);
    float M[16];
    obb_matrix(a0, a1, a2, m, ext, M);
    store_mat(mat, M);
}

__global__ void tree_clear(uint64_t* visit, uint32_t n, TreeWs ws) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *ws.error = 0;
        *ws.vmax = 0;
        for (int k = 0; k < kMaxPhases; ++k) ws.nlist[k] = 0;
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

// Leaf boxes (transforms.cu:803-834) and the largest referenced vertex id.
__global__ __launch_bounds__(256) void leaf_boxes(const kr_bvh_node* __restrict__ nodes,
                                                  const kr_vec3* __restrict__ V,
                                                  const kr_uvec4* __restrict__ F, kr_obb3* obbs,
                                                  kr_mat4* mats, TreeWs ws, uint32_t N) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t vmax = 0;
    if (i < N) {
        const uint32_t cur = N - 1 + i;
        const uint32_t f = nodes[cur].u.primitivesOffset;
        if (f < N) {
            const kr_uvec4 face = F[f];
            vmax = max(face.x, max(face.y, face.z));
            leaf_obb(ldv(V, face.x), ldv(V, face.y), ldv(V, face.z), &obbs[cur], &mats[cur]);
        }
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) vmax = max(vmax, (uint32_t)__shfl_xor((int)vmax, m, kWave));
    if (lane_id() == 0) atomicMax(ws.vmax, vmax);
}

// One phase of the bottom-up k-DOP pass.  Workgroup w of phase `phase` owns
// the leaf positions [w*chunk, (w+1)*chunk).  Phase 0 starts from the leaves;
// later phases start from the nodes the previous phase deferred (those whose
// leaf span lies inside the workgroup's range).  At a parent, the first
// arriving child stops; the second continues if the first was completed in
// an earlier kernel (visible after the kernel boundary) or by its own
// workgroup (workgroup-scope acq_rel on the arrival word orders the data);
// otherwise the parent is deferred to the next phase.  The last phase owns
// every leaf and never defers.  No device-scope fence is issued.
__global__ __launch_bounds__(1024) void climb_phase(const kr_bvh_node* __restrict__ nodes,
                                                    const uint32_t* __restrict__ parents,
                                                    const kr_vec3* __restrict__ V,
                                                    const kr_uvec4* __restrict__ F, TreeWs ws, uint32_t N,
                                                    uint32_t phase, uint32_t chunk) {
    const uint32_t I = N - 1, total = 2 * N - 1;
    const uint32_t tag = make_tag(phase, blockIdx.x);
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk;  // owned leaf positions
    const uint32_t nin = phase == 0 ? 0 : min(ws.nlist[phase - 1], I);
    const uint32_t* in = phase == 0 ? nullptr : ws.list + (size_t)(phase - 1) * (I + 1);
    uint32_t* out = ws.list + (size_t)phase * (I + 1);
    const uint32_t nstart = phase == 0 ? chunk : nin;
    for (uint32_t r = threadIdx.x; r < nstart; r += blockDim.x) {
        uint32_t cur;
        Kdop k;
        uint2 span;
        if (phase == 0) {
            const uint64_t pos = lo + r;
            if (pos >= N) break;
            cur = I + (uint32_t)pos;
            const uint32_t f = nodes[cur].u.primitivesOffset;
            if (f >= N) { atomicOr(ws.error, 4u); continue; }  // faces has one entry per leaf
            const kr_uvec4 face = F[f];
            const uint32_t vid[3] = {face.x, face.y, face.z};
            float pr[7];
            slab7(ldv(V, vid[0]), pr);
#pragma unroll
            for (int s = 0; s < 7; ++s) { k.mn[s] = k.mx[s] = pr[s]; k.imn[s] = k.imx[s] = vid[0]; }
#pragma unroll
            for (int j = 1; j < 3; ++j) {
                slab7(ldv(V, vid[j]), pr);
#pragma unroll
                for (int s = 0; s < 7; ++s) {
                    if (pr[s] < k.mn[s]) { k.mn[s] = pr[s]; k.imn[s] = vid[j]; }
                    if (pr[s] > k.mx[s]) { k.mx[s] = pr[s]; k.imx[s] = vid[j]; }
                }
            }
            span = make_uint2((uint32_t)pos, (uint32_t)pos);
        } else {
            cur = in[r];
            if (cur >= I) { atomicOr(ws.error, 1u); continue; }
            const uint32_t l0 = nodes[cur].u.children.left, r0 = nodes[cur].u.children.right;
            if (l0 >= total || r0 >= total) { atomicOr(ws.error, 2u); continue; }
            const uint2 sa = ws.span[l0], sb = ws.span[r0];
            span = make_uint2(min(sa.x, sb.x), max(sa.y, sb.y));
            if (span.x < lo || span.x >= hi) continue;  // another workgroup owns its first leaf
            if (span.y >= hi) {                         // wider than this workgroup: next phase
                const uint32_t slot = atomicAdd(&ws.nlist[phase], 1u);
                if (slot < I) out[slot] = cur;
                continue;
            }
            Kdop a, b;
            kdop_load(ws.kdop, l0, a);
            kdop_load(ws.kdop, r0, b);
            merge_left_first(a, b, k);
        }
        kdop_store(ws.kdop, cur, k);
        ws.span[cur] = span;
        uint32_t p = parents[cur];
        for (uint32_t step = 0; p != kInvalid; ++step) {
            if (p >= I || step >= N) { atomicOr(ws.error, 1u); break; }
            const uint64_t old = __hip_atomic_fetch_add(&ws.visit[p], ((uint64_t)tag << 32) | 1ull,
                                                        __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t cnt = (uint32_t)old, t1 = (uint32_t)(old >> 32);
            if (cnt == 0) break;  // first arrival: the sibling continues
            if (cnt != 1) { atomicOr(ws.error, 2u); break; }
            if (t1 != tag && (t1 >> 24) >= phase) {  // sibling done by another workgroup of this phase
                const uint32_t slot = atomicAdd(&ws.nlist[phase], 1u);
                if (slot < I) out[slot] = p;
                break;
            }
            const uint32_t l = nodes[p].u.children.left, rr = nodes[p].u.children.right;
            const bool am_left = (l == cur);
            const uint32_t sib = am_left ? rr : l;
            if (sib >= total || (!am_left && rr != cur)) { atomicOr(ws.error, 2u); break; }
            Kdop o, m;
            kdop_load(ws.kdop, sib, o);
            if (am_left) merge_left_first(k, o, m); else merge_left_first(o, k, m);
            k = m;
            const uint2 so = ws.span[sib];
            span = make_uint2(min(span.x, so.x), max(span.y, so.y));
            kdop_store(ws.kdop, p, k);
            ws.span[p] = span;
            cur = p;
            p = parents[cur];
        }
    }
}

// ------------------------------------------------------- candidates --

struct Best {
    float q;
    f3 b0, b1, b2;
    f3 dmin, dmax;
};

__device__ __forceinline__ void proj14(const f3* P, f3 d, float& mn, float& mx) {
    float p = dot3(P[0], d);
    mn = p;
    mx = p;
#pragma unroll
    for (int i = 1; i < 14; ++i) {
        p = dot3(P[i], d);
        mn = kmin(mn, p);
        mx = kmax(mx, p);
    }
}

// transforms.cu:484-544
__device__ __forceinline__ void from_normal_edges(const f3* P, f3 n, f3 e0, f3 e1, f3 e2, Best& b) {
    float nmn, nmx;
    proj14(P, n, nmn, nmx);
    const float leny = nmx - nmn;
    const f3 e[3] = {e0, e1, e2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 m = vm::cross(e[k], n);
        float emn, emx, mmn, mmx;
        proj14(P, e[k], emn, emx);
        const float lenx = emx - emn;
        proj14(P, m, mmn, mmx);
        const float lenz = mmx - mmn;
        const float q = quality(f3{lenx, leny, lenz});
        if (q < b.q) {
            b.q = q;
            b.dmax = f3{emx, nmx, mmx};
            b.dmin = f3{emn, nmn, mmn};
            b.b0 = e[k];
            b.b1 = n;
            b.b2 = m;
        }
    }
}

__global__ __launch_bounds__(256) void candidates(const kr_vec3* __restrict__ V, TreeWs ws, uint32_t I) {
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= I) return;
    if (*ws.error != 0 || !complete(ws.visit[node])) {
        // malformed tree: never read the ids of a node that did not complete
        if (!complete(ws.visit[node])) atomicOr(ws.error, 8u);
        Cand c = {};
        c.degen = 1;
        ws.cand[node] = c;
        return;
    }
    const float eps = 0.000001f;
    Kdop k;
    kdop_load(ws.kdop, node, k);
    f3 P[14];
    const uint32_t vmax = *ws.vmax;  // ids never exceed the largest face vertex id
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        P[2 * s] = ldv(V, min(k.imn[s], vmax));
        P[2 * s + 1] = ldv(V, min(k.imx[s], vmax));
    }
    const f3 alen = f3{k.mx[0] - k.mn[0], k.mx[1] - k.mn[1], k.mx[2] - k.mn[2]};
    Best b;
    b.q = quality(alen);
    b.b0 = f3{1, 0, 0};
    b.b1 = f3{0, 1, 0};
    b.b2 = f3{0, 0, 1};
    b.dmin = f3{k.mn[0], k.mn[1], k.mn[2]};
    b.dmax = f3{k.mx[0], k.mx[1], k.mx[2]};
    int degen = 0;

    int at = 0;
    float d2b = vm::dist2(P[0], P[1]);
#pragma unroll
    for (int s = 1; s < 7; ++s) {
        float d2 = vm::dist2(P[2 * s], P[2 * s + 1]);
        if (d2 > d2b) { d2b = d2; at = s; }
    }
    f3 p0 = P[0], p1 = P[1];
#pragma unroll
    for (int s = 1; s < 7; ++s)
        if (at == s) { p0 = P[2 * s]; p1 = P[2 * s + 1]; }
    if (vm::dist2(p0, p1) < eps) {
        degen = 1;
    } else {
        const f3 e0 = vm::normalize(sub3(p0, p1));
        float dm = vm::dist_inf_edge(P[0], p0, e0);
        f3 p2 = P[0];
#pragma unroll
        for (int i = 1; i < 14; ++i) {
            float d = vm::dist_inf_edge(P[i], p0, e0);
            if (d > dm) { dm = d; p2 = P[i]; }
        }
        if (dm < eps) {
            degen = 1;
        } else {
            const f3 e1 = vm::normalize(sub3(p1, p2));
            const f3 e2 = vm::normalize(sub3(p2, p0));
            const f3 n = vm::normalize(vm::cross(e1, e0));
            from_normal_edges(P, n, e0, e1, e2, b);
            float pr = dot3(P[0], n);
            float lo = pr, hi = pr;
            f3 q1 = P[0], q0 = P[0];
#pragma unroll
            for (int i = 1; i < 14; ++i) {
                pr = dot3(P[i], n);
                if (pr < lo) { lo = pr; q1 = P[i]; }
                if (pr > hi) { hi = pr; q0 = P[i]; }
            }
            const float tri = dot3(p0, n);
            const bool q0v = hi - eps > tri;
            const bool q1v = lo + eps < tri;
            if (!q0v && !q1v) {
                degen = 1;
            } else {
                if (q0v) {
                    const f3 f0 = vm::normalize(sub3(q0, p0));
                    const f3 f1 = vm::normalize(sub3(q0, p1));
                    const f3 f2 = vm::normalize(sub3(q0, p2));
                    from_normal_edges(P, vm::normalize(vm::cross(f1, e0)), e0, f1, f0, b);
                    from_normal_edges(P, vm::normalize(vm::cross(f2, e1)), e1, f2, f1, b);
                    from_normal_edges(P, vm::normalize(vm::cross(f0, e2)), e2, f0, f2, b);
                }
                if (q1v) {
                    const f3 g0 = vm::normalize(sub3(q1, p0));
                    const f3 g1 = vm::normalize(sub3(q1, p1));
                    const f3 g2 = vm::normalize(sub3(q1, p2));
                    from_normal_edges(P, vm::normalize(vm::cross(g1, e0)), e0, g1, g0, b);
                    from_normal_edges(P, vm::normalize(vm::cross(g2, e1)), e1, g2, g1, b);
                    from_normal_edges(P, vm::normalize(vm::cross(g0, e2)), e2, g0, g2, b);
                }
            }
        }
    }
    Cand c;
    c.b[0] = b.b0.x; c.b[1] = b.b0.y; c.b[2] = b.b0.z;
    c.b[3] = b.b1.x; c.b[4] = b.b1.y; c.b[5] = b.b1.z;
    c.b[6] = b.b2.x; c.b[7] = b.b2.y; c.b[8] = b.b2.z;
    c.degen = degen;
    c.pad[0] = c.pad[1] = 0;
    ws.cand[node] = c;
    int32_t* rk = &ws.range[(size_t)node * 6];
    rk[0] = f2ord(b.dmin.x); rk[1] = f2ord(b.dmin.y); rk[2] = f2ord(b.dmin.z);
    rk[3] = f2ord(b.dmax.x); rk[4] = f2ord(b.dmax.y); rk[5] = f2ord(b.dmax.z);
}

// ------------------------------------------------------------ refit --
//
// Exact projection ranges of every subtree vertex on each internal node's
// candidate axes (transforms.cu:836-900).  A workgroup takes 1024
// consecutive leaves; each lane walks from its leaf to the root and folds its
// triangle's min/max on every ancestor's axes into an LDS hash table keyed by
// the ancestor (LDS integer min/max on order-preserving keys).  Consecutive
// leaves share most ancestors, so each workgroup then issues one filtered
// global atomic per distinct ancestor instead of one per (leaf, ancestor).
// If the table fills, the lane falls back to global atomics for that pair.

constexpr int kRefitLeaves = 1024;
constexpr int kHashSlots = 2048;  // power of two
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ void fold_global(int32_t* rk, const int32_t v[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (v[a] < rk[a]) atomicMin(&rk[a], v[a]);
        if (v[3 + a] > rk[3 + a]) atomicMax(&rk[3 + a], v[3 + a]);
    }
}

__global__ __launch_bounds__(1024) void refit(const kr_bvh_node* __restrict__ nodes,
                                              const uint32_t* __restrict__ parents,
                                              const kr_vec3* __restrict__ V,
                                              const kr_uvec4* __restrict__ F, TreeWs ws, uint32_t N) {
    __shared__ uint32_t hkey[kHashSlots];
    __shared__ int32_t hval[kHashSlots][6];
    if (*ws.error != 0) return;
    for (int h = threadIdx.x; h < kHashSlots; h += blockDim.x) {
        hkey[h] = kEmpty;
#pragma unroll
        for (int a = 0; a < 3; ++a) { hval[h][a] = 0x7FFFFFFF; hval[h][3 + a] = (int32_t)0x80000000; }
    }
    __syncthreads();
    const uint32_t I = N - 1;
    const uint64_t pos = (uint64_t)blockIdx.x * kRefitLeaves + threadIdx.x;
    if (pos < N) {
        const uint32_t cur = I + (uint32_t)pos;
        const kr_uvec4 face = F[nodes[cur].u.primitivesOffset];
        const f3 v0 = ldv(V, face.x), v1 = ldv(V, face.y), v2 = ldv(V, face.z);
        uint32_t p = parents[cur];
        for (uint32_t step = 0; p != kInvalid && p < I && step < N; ++step) {
            const Cand c = ws.cand[p];
            if (!c.degen) {
                int32_t k[6];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const f3 b = f3{c.b[3 * a], c.b[3 * a + 1], c.b[3 * a + 2]};
                    const float d0 = dot3(v0, b), d1 = dot3(v1, b), d2 = dot3(v2, b);
                    k[a] = f2ord(kmin(kmin(d0, d1), d2));
                    k[3 + a] = f2ord(kmax(kmax(d0, d1), d2));
                }
                // find / claim the slot of p (linear probing)
                uint32_t h = (p * 2654435761u) >> (32 - 11);
                int slot = -1;
                for (int probe = 0; probe < 64; ++probe) {
                    uint32_t key = hkey[h];
                    if (key == kEmpty) key = atomicCAS(&hkey[h], kEmpty, p);
                    if (key == kEmpty || key == p) { slot = (int)h; break; }
                    h = (h + 1) & (kHashSlots - 1);
                }
                if (slot >= 0) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        atomicMin(&hval[slot][a], k[a]);
                        atomicMax(&hval[slot][3 + a], k[3 + a]);
                    }
                } else {
                    fold_global(&ws.range[(size_t)p * 6], k);
                }
            }
            p = parents[p];
        }
    }
    __syncthreads();
    for (int h = threadIdx.x; h < kHashSlots; h += blockDim.x) {
        const uint32_t node = hkey[h];
        if (node == kEmpty) continue;
        int32_t k[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) k[a] = hval[h][a];
        fold_global(&ws.range[(size_t)node * 6], k);
    }
}

// --------------------------------------------------------- finalize --

__global__ __launch_bounds__(256) void finalize(TreeWs ws, kr_obb3* obbs, kr_mat4* mats, uint32_t I) {
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= I) return;
    if (*ws.error != 0 || !complete(ws.visit[node])) {
        const float nan = __builtin_nanf("");
        store_obb(&obbs[node], f3{nan, nan, nan}, f3{nan, nan, nan}, f3{nan, nan, nan}, f3{nan, nan, nan},
                  f3{nan, nan, nan});
        float M[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) M[q] = nan;
        store_mat(&mats[node], M);
        return;
    }
    const float4* kd = &ws.kdop[(size_t)node * 7];
    const float4 s0 = kd[0], s1 = kd[1], s2 = kd[2];
    const f3 amn = f3{s0.x, s1.x, s2.x}, amx = f3{s0.z, s1.z, s2.z};
    const f3 alen = sub3(amx, amn);
    const Cand c = ws.cand[node];
    const int32_t* rk = &ws.range[(size_t)node * 6];
    const f3 rmn = f3{ord2f(rk[0]), ord2f(rk[1]), ord2f(rk[2])};
    const f3 rmx = f3{ord2f(rk[3]), ord2f(rk[4]), ord2f(rk[5])};
    f3 v0, v1, v2, mn, mx;
    if (!c.degen && quality(sub3(rmx, rmn)) < quality(alen)) {
        v0 = f3{c.b[0], c.b[1], c.b[2]};
        v1 = f3{c.b[3], c.b[4], c.b[5]};
        v2 = f3{c.b[6], c.b[7], c.b[8]};
        mn = rmn;
        mx = rmx;
    } else {
        v0 = f3{1, 0, 0};
        v1 = f3{0, 1, 0};
        v2 = f3{0, 0, 1};
        mn = amn;
        mx = amx;
    }
    const f3 mid = vm::mul1(add3(mn, mx), 0.5f);
    const f3 ext = vm::vmax3(vm::mul1(sub3(mx, mn), 0.5f), f3{0.001f, 0.001f, 0.001f});
    f3 m = vm::mul1(v0, mid.x);
    m = add3(m, vm::mul1(v1, mid.y));
    m = add3(m, vm::mul1(v2, mid.z));
    store_obb(&obbs[node], m, v0, v1, v2, ext);
    float M[16];
    obb_matrix(v0, v1, v2, m, ext, M);
    store_mat(&mats[node], M);
}

__global__ void copy_ids(TreeWs ws, uint32_t* ids, uint32_t total) {
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= total) return;
    const uint32_t I = (total - 1) / 2;
    if (*ws.error != 0 || (node < I && !complete(ws.visit[node]))) {
#pragma unroll
        for (int s = 0; s < 14; ++s) ids[(size_t)node * 14 + s] = 0xFFFFFFFFu;
        return;
    }
    const float4* kd = &ws.kdop[(size_t)node * 7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        float4 a = kd[s];
        ids[(size_t)node * 14 + s] = __float_as_uint(a.y);
        ids[(size_t)node * 14 + 7 + s] = __float_as_uint(a.w);
    }
}

// ------------------------------------------------------------- host --

size_t tree_carve(TreeWs* ws, void* base, size_t cap, uint32_t N) {
    Carver c{(char*)base, 0, cap};
    const size_t I = N ? N - 1 : 0, total = N ? 2 * (size_t)N - 1 : 0;
    ws->visit = c.take<uint64_t>(I + 1);
    ws->error = c.take<uint32_t>(8);
    ws->vmax = ws->error ? ws->error + 1 : nullptr;
    ws->nlist = ws->error ? ws->error + 2 : nullptr;
    ws->list = c.take<uint32_t>((size_t)kMaxPhases * (I + 1));
    ws->span = c.take<uint2>(total);
    ws->kdop = c.take<float4>(total * 7);
    ws->cand = c.take<Cand>(I + 1);
    ws->range = c.take<int32_t>((I + 1) * 6);
    return c.used;
}

kr_error tree_launch(const kr_bvh_node* nodes, const uint32_t* parents, const kr_vec3* V,
                     const kr_uvec4* F, kr_obb3* obbs, kr_mat4* mats, uint32_t* ids, uint32_t N,
                     void* workspace, size_t ws_bytes, hipStream_t stream) {
    if (N == 0) return nullptr;
    TreeWs ws;
    size_t need = tree_carve(&ws, workspace, ws_bytes, N);
    if (workspace == nullptr || ws_bytes < need) return "obvh_obb_tree: workspace too small";
    const uint32_t I = N - 1;
    auto blocks = [](uint32_t n, uint32_t b) { return dim3((n + b - 1) / b); };
    hipLaunchKernelGGL(tree_clear, dim3(1024), dim3(256), 0, stream, ws.visit, I + 1, ws);
    hipLaunchKernelGGL(leaf_boxes, blocks(N, 256), dim3(256), 0, stream, nodes, V, F, obbs, mats, ws, N);
    // climb phases: 1024-leaf workgroups, then 32768-leaf workgroups, then one
    uint32_t phase = 0;
    for (uint32_t chunk : {kChunk0, kChunk1, 0u}) {
        const bool last = (chunk == 0) || chunk >= N;
        const uint32_t c = last ? N : chunk;
        hipLaunchKernelGGL(climb_phase, blocks(N, c), dim3(1024), 0, stream, nodes, parents, V, F, ws, N, phase, c);
        ++phase;
        if (last) break;
    }
    if (I > 0) {
        hipLaunchKernelGGL(candidates, blocks(I, 256), dim3(256), 0, stream, V, ws, I);
        hipLaunchKernelGGL(refit, blocks(N, kRefitLeaves), dim3(kRefitLeaves), 0, stream, nodes, parents, V, F, ws, N);
        hipLaunchKernelGGL(finalize, blocks(I, 256), dim3(256), 0, stream, ws, obbs, mats, I);
    }
    if (ids) hipLaunchKernelGGL(copy_ids, blocks(2 * N - 1, 256), dim3(256), 0, stream, ws, ids, 2 * N - 1);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
