// atrbvh.hip -- ATRBVH builder on gfx950: 64-bit Morton LBVH followed by the
// agglomerative treelet optimiser (src/ads/atrbvh_gpu; the flow of
// atrbvh.cu:258-470).
//
//  atr_scene      scene box over the triangle boxes (atrbvh.cu:262-290)
//  atr_codes      63-bit Morton code of each triangle-box centre normalised
//                 by the scene box (LBVH.cu:9-93, Commons.cuh:422-430,602-645)
//  (sort)         stable radix sort of (code, face) over 63 bits (rocPRIM),
//                 the equivalent of cub::DeviceRadixSort (CubWrapper.cu:8-46)
//  atr_tree       Karras radix tree, duplicates split by 64 + clz(i ^ j)
//                 (LBVH.cu:156-302), written in kr_bvh_node layout
//  atr_boxes      bottom-up working boxes and areas (LBVH.cu:414-446,551-665)
//  atr_treelets   one optimiser pass: bottom-up over the current tree; a node
//                 holding at least gamma triangles is optimised as a treelet
//                 root right after its subtree (AgglomerativeTreelet.cu:
//                 405-561, Treelet.cuh:55-127).  A half-wave (32 lanes) of
//                 the wave whose lane completes the node forms the treelet
//                 (largest-area expansion), builds the distance matrix
//                 (union box areas), merges the closest pair ts-1 times and
//                 writes the new topology when its SAH is lower; the two
//                 halves take two treelets at once.  Each half runs the
//                 reference's 32-lane __shfl_down tournaments exactly, so
//                 ties resolve as in the warp code.
//  (final)        leaf bounds from the working boxes, then bounds / counts /
//                 costs bottom-up with kr_aabb_expand (atrbvh.cu:95-179).
// Every bottom-up pass uses the phased arrival protocol of phased.hpp.
// Working-box min/max order -0 below +0 (as fminf/fmaxf on the reference
// GPU); the final bounds use kr_min/kr_max like kr_aabb_expand.
#include <float.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "atrbvh.hpp"
#include "common.hpp"

namespace obvh {

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

__device__ __forceinline__ float zmin(float a, float b) {
    return a < b ? a : (b < a ? b : (__builtin_signbit(a) ? a : b));
}
__device__ __forceinline__ float zmax(float a, float b) {
    return a > b ? a : (b > a ? b : (__builtin_signbit(a) ? b : a));
}
__device__ __forceinline__ f3 ldv3a(const kr_vec3* V, uint32_t i) {
    const kr_vec3 a = V[i];
    return f3{a.x, a.y, a.z};
}

// calculateBoundingBoxSurfaceArea (Commons.cuh:96-103)
__device__ __forceinline__ float box_area3(float x, float y, float z) { return 2 * (x * y + x * z + y * z); }
__device__ __forceinline__ float box_area(float4 mn, float4 mx) {
    return box_area3(mx.x - mn.x, mx.y - mn.y, mx.z - mn.z);
}
// calculateBoundingBoxAndSurfaceArea (Commons.cuh:173-181)
__device__ __forceinline__ float union_area(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return box_area3(zmax(amx.x, bmx.x) - zmin(amn.x, bmn.x), zmax(amx.y, bmx.y) - zmin(amn.y, bmn.y),
                     zmax(amx.z, bmx.z) - zmin(amn.z, bmn.z));
}

// calculateLeafBoundingBox (LBVH.cu:414-430)
__device__ __forceinline__ void tri_box(const kr_vec3* V, const kr_uvec4 f, float4& mn, float4& mx) {
    const f3 a = ldv3a(V, f.x), b = ldv3a(V, f.y), c = ldv3a(V, f.z);
    mn = make_float4(zmin(zmin(a.x, b.x), c.x), zmin(zmin(a.y, b.y), c.y), zmin(zmin(a.z, b.z), c.z), 0.0f);
    mx = make_float4(zmax(zmax(a.x, b.x), c.x), zmax(zmax(a.y, b.y), c.y), zmax(zmax(a.z, b.z), c.z), 0.0f);
}

// expandBits64 (Commons.cuh:602-613)
__device__ __forceinline__ uint64_t expand_bits64(uint64_t v) {
    v &= 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
__device__ __forceinline__ uint64_t discretize21(float x) {
    return (uint64_t)fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f);  // NaN -> 0
}

// ----------------------------------------------------------- build --

__global__ void atr_init(AtrWs ws, uint32_t nf) {
    phase_clear(ws.ph, nf);
    if (blockIdx.x == 0 && threadIdx.x < 6) ws.sbox[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : (int32_t)0x80000000;
}

__global__ __launch_bounds__(256) void atr_scene(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, AtrWs ws) {
    int32_t mn[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, mx[3] = {(int32_t)0x80000000, (int32_t)0x80000000,
                                                                   (int32_t)0x80000000};
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const kr_uvec4 face = F[f];
        const uint32_t vid[3] = {face.x, face.y, face.z};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const f3 p = ldv3a(V, vid[j]);
            const int32_t k[3] = {f2ord(p.x), f2ord(p.y), f2ord(p.z)};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                mn[a] = min(mn[a], k[a]);
                mx[a] = max(mx[a], k[a]);
            }
        }
    }
    // wave, then workgroup reduction; one filtered atomic per key and
    // workgroup (a few thousand same-address atomics would serialise)
    __shared__ int32_t part[256 / kWave][6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mn[a] = wave_min_i32(mn[a]);
        mx[a] = wave_max_i32(mx[a]);
    }
    if (lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            part[wave_id()][a] = mn[a];
            part[wave_id()][3 + a] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int32_t v = part[0][a];
        for (int w = 1; w < 256 / kWave; ++w) v = a < 3 ? min(v, part[w][a]) : max(v, part[w][a]);
        const int32_t cur = __hip_atomic_load(&ws.sbox[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a < 3 ? v < cur : v > cur) {
            if (a < 3) atomicMin(&ws.sbox[a], v);
            else atomicMax(&ws.sbox[a], v);
        }
    }
}

__global__ void atr_codes(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, AtrWs ws) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    float4 mn, mx;
    tri_box(V, F[f], mn, mx);
    const float smn[3] = {ord2f(ws.sbox[0]), ord2f(ws.sbox[1]), ord2f(ws.sbox[2])};
    const float smx[3] = {ord2f(ws.sbox[3]), ord2f(ws.sbox[4]), ord2f(ws.sbox[5])};
    const float cmn[3] = {mn.x, mn.y, mn.z}, cmx[3] = {mx.x, mx.y, mx.z};
    uint64_t d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = (cmx[a] + cmn[a]) * 0.5f;
        d[a] = expand_bits64(discretize21((c - smn[a]) / (smx[a] - smn[a])));
    }
    ws.keys_in[f] = (d[0] << 2) | (d[1] << 1) | d[2];
    ws.vals_in[f] = f;
}

// longestCommonPrefix (LBVH.cu:173-187)
__device__ __forceinline__ int lcp64(const uint64_t* k, int n, int i, int j, uint64_t ki) {
    if (j < 0 || j >= n) return -1;
    const uint64_t kj = k[j];
    if (ki == kj) return 64 + __clz(i ^ j);
    return __clzll((long long)(ki ^ kj));
}

__global__ void atr_tree(uint32_t nf, AtrWs ws, kr_bvh_node* nodes, uint32_t* parents) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = (int)nf;
    const uint32_t I = nf - 1;
    if (nf == 1) {
        if (i == 0) {
            kr_bvh_node L = {};
            L.u.primitivesOffset = ws.vals_out[0];
            L.nPrimitives = 1;
            nodes[0] = L;
            parents[0] = kNoNode;
        }
        return;
    }
    if (i >= n - 1) return;
    const uint64_t* k = ws.keys_out;
    const uint64_t ki = k[i];
    const int l1 = lcp64(k, n, i, i + 1, ki), l2 = lcp64(k, n, i, i - 1, ki);
    const int dir = (0 < l1 - l2) - (l1 - l2 < 0);
    const int min_lcp = lcp64(k, n, i, i - dir, ki);
    int lmax = 2;
    while (lcp64(k, n, i, i + lmax * dir, ki) > min_lcp) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lcp64(k, n, i, i + (l + t) * dir, ki) > min_lcp) l += t;
    const int j = i + l * dir;
    const int node_lcp = lcp64(k, n, i, j, ki);
    int s = 0, divisor = 2;
    const int max_div = 1 << (32 - __clz(l));
    while (divisor <= max_div) {
        const int t = (l + divisor - 1) / divisor;
        if (lcp64(k, n, i, i + (s + t) * dir, ki) > node_lcp) s += t;
        divisor <<= 1;
    }
    const int split = i + s * dir + min(dir, 0);
    uint32_t li = (uint32_t)split, ri = (uint32_t)split + 1;
    if (min(i, j) == split) {
        li = (uint32_t)split + I;
        kr_bvh_node L = {};
        L.u.primitivesOffset = ws.vals_out[split];
        L.nPrimitives = 1;
        nodes[li] = L;
    }
    if (max(i, j) == split + 1) {
        ri = (uint32_t)split + 1 + I;
        kr_bvh_node R = {};
        R.u.primitivesOffset = ws.vals_out[split + 1];
        R.nPrimitives = 1;
        nodes[ri] = R;
    }
    kr_bvh_node N = {};
    N.u.children.left = li;
    N.u.children.right = ri;
    nodes[i] = N;
    parents[li] = (uint32_t)i;
    parents[ri] = (uint32_t)i;
    if (i == 0) parents[0] = kNoNode;
}

struct AtrBoxUp {
    const kr_vec3* V;
    const kr_uvec4* F;
    const kr_bvh_node* nodes;
    AtrWs ws;
    __device__ void leaf(uint32_t cur) {
        float4 mn, mx;
        tri_box(V, F[nodes[cur].u.primitivesOffset], mn, mx);
        ws.bmin[cur] = mn;
        ws.bmax[cur] = mx;
        ws.area[cur] = box_area(mn, mx);
    }
    __device__ void merge(uint32_t p, uint32_t l, uint32_t r) {
        const float4 amn = ws.bmin[l], amx = ws.bmax[l], bmn = ws.bmin[r], bmx = ws.bmax[r];
        const float4 mn = make_float4(zmin(amn.x, bmn.x), zmin(amn.y, bmn.y), zmin(amn.z, bmn.z), 0.0f);
        const float4 mx = make_float4(zmax(amx.x, bmx.x), zmax(amx.y, bmx.y), zmax(amx.z, bmx.z), 0.0f);
        ws.bmin[p] = mn;
        ws.bmax[p] = mx;
        ws.area[p] = box_area(mn, mx);
    }
};

__global__ __launch_bounds__(1024) void atr_boxes(const kr_vec3* V, const kr_uvec4* F, const kr_bvh_node* nodes,
                                                  const uint32_t* parents, uint32_t nf, AtrWs ws, uint32_t phase,
                                                  uint32_t chunk) {
    AtrBoxUp op{V, F, nodes, ws};
    phased_up(nodes, parents, nf, ws.ph, phase, chunk, op);
}

__global__ void atr_clear(AtrWs ws, uint32_t nf) { phase_clear(ws.ph, nf); }

// ------------------------------------------------------- optimiser --

struct TreeletLds {
    uint32_t tin[kMaxTreelet];
    uint32_t tl[kMaxTreelet];
    float ta[kMaxTreelet];  // leaf areas, then the SAH of the merged nodes
    float dist[kMaxTreelet * (kMaxTreelet - 1) / 2];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// LOWER_TRM_ROW / LOWER_TRM_COL (TriangularMatrix.cuh:37-52)
__device__ __forceinline__ int trm_row(int index) { return (int)((-1 + sqrtf(8.0f * index + 1)) / 2) + 1; }
__device__ __forceinline__ int trm_index(int row, int col) { return ((row - 1) * row) / 2 + col; }

__device__ __forceinline__ float4 shfl4(float4 v, int src) {
    return make_float4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), 0.0f);
}

// Optimise the treelet rooted at `root` with one half of the wave: lanes
// 0-31 and 32-63 each take a treelet (`valid` false: that half idles), as
// the reference's 32-lane warps do; every tournament is a width-32
// shuffle-down within the half, so ties break as in the warp code.  All 64
// lanes execute (the loops depend only on ts).  Two treelets found in the
// same round of the climb have disjoint subtrees, so running them together
// equals running them one after the other.
__device__ void treelet_half(uint32_t root, bool valid, kr_bvh_node* nodes, uint32_t* parents, const AtrWs& ws,
                             uint32_t N, int ts, float ci, TreeletLds& S) {
    const int hl = lane_id() & 31;  // lane within the half
    const int hb = lane_id() & 32;  // first lane of the half
    const uint32_t I = N - 1;
    // formTreelet: expand the treelet leaf with the largest area ts-2 times
    if (valid && hl == 0) {
        const uint32_t L = nodes[root].u.children.left, R = nodes[root].u.children.right;
        S.tin[0] = root;
        S.tl[0] = L;
        S.tl[1] = R;
        S.ta[0] = ws.area[L];
        S.ta[1] = ws.area[R];
    }
    wave_sync();
    for (int it = 0; it < ts - 2; ++it) {
        int idx = -1;
        float a = -FLT_MAX;
        if (valid && hl < 2 + it) {
            idx = hl;
            a = S.tl[hl] >= I ? -FLT_MAX : S.ta[hl];
        }
        for (int n = 2 + it, amount = n / 2; n > 1; n = (n + 1) / 2, amount = n / 2) {
            const int oi = __shfl_down(idx, amount, 32);
            const float oa = __shfl_down(a, amount, 32);
            if (oa > a) { a = oa; idx = oi; }
        }
        if (valid && hl == 0) {
            const uint32_t rep = S.tl[idx];
            const uint32_t l2 = nodes[rep].u.children.left, r2 = nodes[rep].u.children.right;
            S.tin[it + 1] = rep;
            S.tl[idx] = l2;
            S.tl[it + 2] = r2;
            S.ta[idx] = ws.area[l2];
            S.ta[it + 2] = ws.area[r2];
        }
        wave_sync();
    }
    // lane state: node, subtree SAH, box of treelet leaf `hl`
    uint32_t tnode = 0;
    float tsah = 0.0f;
    float4 bmn = make_float4(0, 0, 0, 0), bmx = bmn;
    if (valid && hl < ts) {
        tnode = S.tl[hl];
        tsah = ws.sah[tnode];
        bmn = ws.bmin[tnode];
        bmx = ws.bmax[tnode];
    }
    // distance matrix: union areas of every pair (row > col)
    const int npairs = ts * (ts - 1) / 2;
    for (int base = 0; base < npairs; base += 32) {
        const int e = base + hl;
        int r = 0, c = 0;
        if (e < npairs) {
            r = trm_row(e);
            c = e - trm_index(r, 0);
        }
        const float4 amn = shfl4(bmn, hb + r), amx = shfl4(bmx, hb + r), cmn = shfl4(bmn, hb + c),
                     cmx = shfl4(bmx, hb + c);
        if (e < npairs) S.dist[e] = union_area(amn, amx, cmn, cmx);
    }
    wave_sync();
    for (int last = ts - 1; last > 0; --last) {
        // findMinimumDistance: 32 lanes, strided, then the shfl_down tournament
        const int size = last * (1 + last) / 2;
        float md = FLT_MAX;
        int mi = 0;
        for (int j = hl; j < size; j += 32) {
            const float d = S.dist[j];
            if (d < md) { md = d; mi = j; }
        }
        for (int n = 32, amount = 16; n > 1; n = (n + 1) / 2, amount = n / 2) {
            const int oi = __shfl_down(mi, amount, 32);
            const float od = __shfl_down(md, amount, 32);
            if (od < md) { md = od; mi = oi; }
        }
        const int min_index = __shfl(mi, hb);
        const int row = trm_row(min_index), col = min_index - trm_index(row, 0);
        // updateState
        const float s = __shfl(tsah, hb + row) + __shfl(tsah, hb + col);
        const uint32_t li = (uint32_t)__shfl((int)tnode, hb + row), ri = (uint32_t)__shfl((int)tnode, hb + col);
        float4 amn = shfl4(bmn, hb + row), amx = shfl4(bmx, hb + row);
        const float4 cmn = shfl4(bmn, hb + col), cmx = shfl4(bmx, hb + col);
        amn = make_float4(zmin(cmn.x, amn.x), zmin(cmn.y, amn.y), zmin(cmn.z, amn.z), 0.0f);
        amx = make_float4(zmax(cmx.x, amx.x), zmax(cmx.y, amx.y), zmax(cmx.z, amx.z), 0.0f);
        if (hl == col) {
            tnode = S.tin[last - 1];
            bmn = amn;
            bmx = amx;
            S.ta[last] = s + ci * box_area(amn, amx);
            tsah = S.ta[last];
        }
        const uint32_t last_node = (uint32_t)__shfl((int)tnode, hb + last);
        const float last_sah = __shfl(tsah, hb + last);
        const float4 lmn = shfl4(bmn, hb + last), lmx = shfl4(bmx, hb + last);
        if (hl == row) {
            tnode = last_node;
            tsah = last_sah;
            bmn = lmn;
            bmx = lmx;
        }
        if (hl == last) {
            tnode = li;
            S.tl[last] = ri;
            bmn = amn;
            bmx = amx;
        }
        wave_sync();
        // updateDistancesMatrix
        if (last > 1) {
            if (hl < last && hl != row) {
                const float v = S.dist[trm_index(last, hl)];
                S.dist[trm_index(max(row, hl), min(row, hl))] = v;
            }
            wave_sync();
            int dr = hl, dc = hl;
            if (hl < last && hl != col) {
                dr = max(col, hl);
                dc = min(col, hl);
            }
            const float4 xmn = shfl4(bmn, hb + (dr & 31)), xmx = shfl4(bmx, hb + (dr & 31));
            const float4 ymn = shfl4(bmn, hb + (dc & 31)), ymx = shfl4(bmx, hb + (dc & 31));
            if (hl < last && hl != col) S.dist[trm_index(dr, dc)] = union_area(xmn, xmx, ymn, ymx);
        }
        wave_sync();
    }
    // updateTreelet: keep the new topology when its SAH is lower
    if (valid && S.ta[1] < ws.sah[S.tin[0]]) {
        if (hl >= 1 && hl < ts) {
            const uint32_t v = S.tin[hl - 1];
            const uint32_t right = S.tl[hl];
            nodes[v].u.children.left = tnode;
            nodes[v].u.children.right = right;
            parents[tnode] = v;
            parents[right] = v;
            ws.sah[v] = S.ta[hl];
            ws.bmin[v] = bmn;
            ws.bmax[v] = bmx;
            ws.area[v] = box_area(bmn, bmx);
        }
    }
    wave_sync();
}

// One optimiser pass, one phase of the bottom-up protocol (phased.hpp), with
// a wave-uniform climb so that whole waves can stop to optimise treelets.
__global__ __launch_bounds__(1024) void atr_treelets(kr_bvh_node* nodes, uint32_t* parents, AtrWs ws, uint32_t nf,
                                                     int ts, int gamma, float ci, float ct, uint32_t phase,
                                                     uint32_t chunk) {
    __shared__ TreeletLds lds[2 * 1024 / kWave];
    TreeletLds& S = lds[2 * wave_id() + (lane_id() >> 5)];
    const uint32_t I = nf - 1;
    const uint32_t tag = (phase << 24) | (blockIdx.x + 1u);
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk;
    const uint32_t nin = phase == 0 ? 0 : min(ws.ph.nlist[phase - 1], I);
    const uint2* in = phase == 0 ? nullptr : ws.ph.list + (size_t)(phase - 1) * (I + 1);
    uint2* out = ws.ph.list + (size_t)phase * (I + 1);
    const uint32_t nstart = phase == 0 ? chunk : nin;
    for (uint32_t base = 0; base < nstart; base += blockDim.x) {  // uniform over the workgroup
        const uint32_t r0 = base + threadIdx.x;
        bool active = false, at_node = false;
        uint32_t cur = 0, p = kNoNode, steps = 0;
        uint2 span = make_uint2(0, 0);
        if (r0 < nstart) {
            if (phase == 0) {
                const uint64_t pos = lo + r0;
                if (pos < nf) {
                    cur = I + (uint32_t)pos;
                    ws.tri[cur] = 1;
                    ws.sah[cur] = ct * ws.area[cur];
                    span = make_uint2((uint32_t)pos, (uint32_t)pos);
                    ws.ph.span[cur] = span;
                    p = parents[cur];
                    active = true;
                }
            } else {
                // Ownership comes from the listed key only: once its owner has
                // completed the node and an ancestor, a treelet there may
                // rewire the node's children while other workgroups still
                // scan this list.
                const uint2 e = in[r0];
                cur = e.x;
                if (cur < I && e.y >= lo && e.y < hi) {
                    const uint32_t l = nodes[cur].u.children.left, r = nodes[cur].u.children.right;
                    const uint2 sa = ws.ph.span[l], sb = ws.ph.span[r];
                    span = make_uint2(min(sa.x, sb.x), max(sa.y, sb.y));
                    if (span.x < lo || span.y >= hi) {
                        const uint32_t slot = atomicAdd(&ws.ph.nlist[phase], 1u);
                        if (slot < I) out[slot] = make_uint2(cur, span.x);
                    } else {
                        active = true;
                        at_node = true;  // both children are complete
                    }
                }
            }
        }
        while (__ballot(active) != 0) {  // wave-uniform climb
            bool treelet = false;
            uint32_t node = cur;
            if (active && !at_node) {
                if (p == kNoNode || p >= I || ++steps > nf) {
                    active = false;
                } else {
                    const uint64_t old = __hip_atomic_fetch_add(&ws.ph.visit[p], ((uint64_t)tag << 32) | 1ull,
                                                                __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const uint32_t t1 = (uint32_t)(old >> 32);
                    if ((uint32_t)old == 0) {
                        active = false;  // first arrival: the sibling continues
                    } else if (t1 != tag && (t1 >> 24) >= phase) {
                        const uint32_t slot = atomicAdd(&ws.ph.nlist[phase], 1u);
                        if (slot < I) out[slot] = make_uint2(p, span.x);
                        active = false;
                    } else {
                        const uint32_t l = nodes[p].u.children.left, r = nodes[p].u.children.right;
                        span = make_uint2(min(span.x, ws.ph.span[l == cur ? r : l].x),
                                          max(span.y, ws.ph.span[l == cur ? r : l].y));
                        node = p;
                        at_node = true;
                    }
                }
            }
            if (active && at_node) {
                const uint32_t l = nodes[node].u.children.left, r = nodes[node].u.children.right;
                const int32_t tri = ws.tri[l] + ws.tri[r];
                ws.tri[node] = tri;
                ws.sah[node] = ci * ws.area[node] + ws.sah[l] + ws.sah[r];
                ws.ph.span[node] = span;
                treelet = tri >= gamma;
            }
            uint64_t vote = __ballot(treelet);
            while (vote) {  // two treelets at a time, one per half-wave
                const int k1 = __builtin_ctzll(vote);
                vote &= vote - 1;
                const int k2 = vote ? __builtin_ctzll(vote) : k1;
                if (vote) vote &= vote - 1;
                const uint32_t n1 = (uint32_t)__shfl((int)node, k1), n2 = (uint32_t)__shfl((int)node, k2);
                const bool upper = lane_id() >= 32;
                treelet_half(upper ? n2 : n1, !upper || k2 != k1, nodes, parents, ws, nf, ts, ci, S);
            }
            if (active && at_node) {
                cur = node;
                p = parents[cur];
                at_node = false;
            }
        }
    }
}

__global__ void atr_leaf_bounds(kr_bvh_node* nodes, AtrWs ws, uint32_t nf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const uint32_t v = nf - 1 + i;
    const float4 mn = ws.bmin[v], mx = ws.bmax[v];
    nodes[v].bounds.min = kr_vec3{mn.x, mn.y, mn.z};
    nodes[v].bounds.max = kr_vec3{mx.x, mx.y, mx.z};
}

// Bounds / counts / costs of the final tree (atrbvh.cu:124-179: ct = 1.0,
// ci = 1.2, kr_aabb_expand), the same pass as the LBVH builder's.
__device__ __forceinline__ f3 kmin3(f3 a, f3 b) { return f3{kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)}; }
__device__ __forceinline__ f3 kmax3(f3 a, f3 b) { return f3{kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)}; }
__device__ __forceinline__ float kr_area(f3 mn, f3 mx) {
    const f3 d = sub3(mx, mn);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
struct AtrFinalUp {
    kr_bvh_node* nodes;
    uint32_t* counts;
    float* costs;
    __device__ void leaf(uint32_t cur) {
        const kr_bvh_node L = nodes[cur];
        costs[cur] = 1.0f * kr_area(f3{L.bounds.min.x, L.bounds.min.y, L.bounds.min.z},
                                    f3{L.bounds.max.x, L.bounds.max.y, L.bounds.max.z});
        counts[cur] = 1;
    }
    __device__ void merge(uint32_t p, uint32_t l, uint32_t r) {
        const kr_bvh_node A = nodes[l], B = nodes[r];
        f3 mn{A.bounds.min.x, A.bounds.min.y, A.bounds.min.z}, mx{A.bounds.max.x, A.bounds.max.y, A.bounds.max.z};
        const f3 bmn{B.bounds.min.x, B.bounds.min.y, B.bounds.min.z}, bmx{B.bounds.max.x, B.bounds.max.y,
                                                                          B.bounds.max.z};
        mn = kmin3(mn, bmn);
        mx = kmax3(mx, bmn);
        mn = kmin3(mn, bmx);
        mx = kmax3(mx, bmx);
        counts[p] = counts[l] + counts[r];
        costs[p] = (float)1.2 * kr_area(mn, mx) + costs[r] + costs[l];
        nodes[p].bounds.min = kr_vec3{mn.x, mn.y, mn.z};
        nodes[p].bounds.max = kr_vec3{mx.x, mx.y, mx.z};
    }
};

__global__ __launch_bounds__(1024) void atr_final(kr_bvh_node* nodes, const uint32_t* parents, uint32_t* counts,
                                                  float* costs, uint32_t nf, AtrWs ws, uint32_t phase,
                                                  uint32_t chunk) {
    AtrFinalUp op{nodes, counts, costs};
    phased_up(nodes, parents, nf, ws.ph, phase, chunk, op);
}

// ------------------------------------------------------------- host --

static size_t atr_sort_bytes(uint32_t nf) {
    size_t bytes = 0;
    rocprim::radix_sort_pairs((void*)nullptr, bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                              (uint32_t*)nullptr, nf, 0, 63);
    return bytes;
}

size_t atr_carve(AtrWs* ws, void* base, size_t cap, uint32_t nf) {
    Carver c{(char*)base, 0, cap};
    const size_t total = nf ? 2 * (size_t)nf - 1 : 0;
    phase_carve(&ws->ph, c, nf);
    ws->sbox = c.take<int32_t>(8);
    ws->keys_in = c.take<uint64_t>(nf);
    ws->keys_out = c.take<uint64_t>(nf);
    ws->vals_in = c.take<uint32_t>(nf);
    ws->vals_out = c.take<uint32_t>(nf);
    ws->bmin = c.take<float4>(total);
    ws->bmax = c.take<float4>(total);
    ws->area = c.take<float>(total);
    ws->tri = c.take<int32_t>(total);
    ws->sah = c.take<float>(total);
    ws->sort_bytes = atr_sort_bytes(nf);
    ws->sort_tmp = c.take<char>(ws->sort_bytes);
    return c.used;
}

kr_error atr_launch(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, int treelet_size, int iterations, float ci,
                    float ct, kr_bvh_node* nodes, uint32_t* parents, uint32_t* counts, float* costs, void* workspace,
                    size_t ws_bytes, hipStream_t stream) {
    if (nf == 0) return "obvh_atrbvh_build: no faces";
    if (treelet_size < 3 || treelet_size > kMaxTreelet) return "obvh_atrbvh_build: treelet size must be 3..20";
    if (iterations < 0) return "obvh_atrbvh_build: negative iteration count";
    AtrWs ws;
    const size_t need = atr_carve(&ws, workspace, ws_bytes, nf);
    if (!workspace || ws_bytes < need) return "obvh_atrbvh_build: workspace too small";
    auto blocks = [](uint32_t n, uint32_t b) { return dim3((n + b - 1) / b); };
    uint32_t gb = (nf + 255) / 256;
    if (gb > 2048) gb = 2048;
    hipLaunchKernelGGL(atr_init, dim3(gb), dim3(256), 0, stream, ws, nf);
    hipLaunchKernelGGL(atr_scene, dim3(gb), dim3(256), 0, stream, V, F, nf, ws);
    hipLaunchKernelGGL(atr_codes, blocks(nf, 256), dim3(256), 0, stream, V, F, nf, ws);
    size_t sb = ws.sort_bytes;
    OBVH_HIP_TRY(rocprim::radix_sort_pairs(ws.sort_tmp, sb, ws.keys_in, ws.keys_out, ws.vals_in, ws.vals_out, nf, 0,
                                           63, stream));
    hipLaunchKernelGGL(atr_tree, blocks(nf, 256), dim3(256), 0, stream, nf, ws, nodes, parents);
    phased_launch(atr_boxes, nf, stream, (const kr_vec3*)V, (const kr_uvec4*)F, (const kr_bvh_node*)nodes,
                  (const uint32_t*)parents, nf, ws);
    if (nf > 1) {
        int gamma = treelet_size;
        for (int it = 0; it < iterations; ++it) {
            hipLaunchKernelGGL(atr_clear, dim3(gb), dim3(256), 0, stream, ws, nf);
            phased_launch(atr_treelets, nf, stream, nodes, parents, ws, nf, treelet_size, gamma, ci, ct);
            gamma *= 2;
        }
    }
    hipLaunchKernelGGL(atr_leaf_bounds, blocks(nf, 256), dim3(256), 0, stream, nodes, ws, nf);
    hipLaunchKernelGGL(atr_clear, dim3(gb), dim3(256), 0, stream, ws, nf);
    phased_launch(atr_final, nf, stream, nodes, (const uint32_t*)parents, counts, costs, nf, ws);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
