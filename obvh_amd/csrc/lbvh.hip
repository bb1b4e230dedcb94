// lbvh.hip -- LBVH builder on gfx950 (the input side of the transform).
//
// Follows the reference build (src/ads/lbvh_gpu/lbvh.cu:252-457 and the
// sequence at :1640-1692):
//  1. centroid box of the triangle-box centres.  The reference reduction
//     (src/ads/common/cuda/transforms.cu:1127-1190) seeds every LDS slot of a
//     256-thread block with vertices[0] and lets out-of-range threads leave
//     before reducing, so for face counts that are not a multiple of 256 the
//     box also covers vertices[0]; that is reproduced because the Morton codes
//     (and so the tree) depend on it.  Exact min/max via integer atomics on
//     order-preserving float keys.
//  2. 30-bit Morton code of the offset triangle-box centre, safe divide by
//     the centroid extent (lbvh.cu:252-282, vecmath.h:698-716,760).
//  3. stable LSD radix sort of (code, face) pairs over 30 bits (rocPRIM),
//     the equivalent of thrust::stable_sort_by_key (lbvh.cu:1670).
//  4. Karras radix tree, duplicate codes split by 32 + clz(i ^ j)
//     (lbvh.cu:284-396); leaves written in sorted order.
//  5. bottom-up bounds, primitive counts, SAH costs (ct = 1.0, ci = 1.2),
//     lbvh.cu:399-457, with the phased workgroup-local arrival protocol
//     described at lbvh_aabbs.
#include <float.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "lbvh.hpp"

namespace obvh {

constexpr uint32_t kInv = 0xFFFFFFFFu;

__device__ __forceinline__ f3 ldv3(const kr_vec3* V, uint32_t i) {
    kr_vec3 a = V[i];
    return f3{a.x, a.y, a.z};
}

struct Box {
    f3 mn, mx;
};
__device__ __forceinline__ Box expand3(Box b, f3 a) { return Box{vm::vmin3(b.mn, a), vm::vmax3(b.mx, a)}; }
__device__ __forceinline__ Box tri_box(const kr_vec3* V, kr_uvec4 f) {
    f3 a = ldv3(V, f.x);
    Box b{a, a};
    b = expand3(b, ldv3(V, f.y));
    b = expand3(b, ldv3(V, f.z));
    return b;
}
__device__ __forceinline__ f3 box_center(Box b) {
    return f3{(b.mn.x + b.mx.x) * 0.5f, (b.mn.y + b.mx.y) * 0.5f, (b.mn.z + b.mx.z) * 0.5f};
}
__device__ __forceinline__ float box_area(f3 mn, f3 mx) {
    f3 d = sub3(mx, mn);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

__global__ void lbvh_init(const kr_vec3* V, uint32_t nf, LbvhWs ws) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        if (nf % 256u != 0) {  // vertices[0] joins the centroid box (see header)
            f3 v = ldv3(V, 0);
            ws.cbox[0] = f2ord(v.x); ws.cbox[1] = f2ord(v.y); ws.cbox[2] = f2ord(v.z);
            ws.cbox[3] = f2ord(v.x); ws.cbox[4] = f2ord(v.y); ws.cbox[5] = f2ord(v.z);
        } else {
            for (int k = 0; k < 3; ++k) { ws.cbox[k] = 0x7FFFFFFF; ws.cbox[3 + k] = (int32_t)0x80000000; }
        }
    }
    phase_clear(ws.ph, nf);
}

__global__ __launch_bounds__(256) void lbvh_centroid_box(const kr_vec3* V, const kr_uvec4* F, uint32_t nf,
                                                         LbvhWs ws) {
    __shared__ float sh[6][4];
    const float INF = FLT_MAX * 2.0f;
    float mn[3] = {INF, INF, INF}, mx[3] = {-INF, -INF, -INF};
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        f3 c = box_center(tri_box(V, F[f]));
        mn[0] = fminf(mn[0], c.x); mn[1] = fminf(mn[1], c.y); mn[2] = fminf(mn[2], c.z);
        mx[0] = fmaxf(mx[0], c.x); mx[1] = fmaxf(mx[1], c.y); mx[2] = fmaxf(mx[2], c.z);
    }
    for (int k = 0; k < 3; ++k) { mn[k] = wave_min(mn[k]); mx[k] = wave_max(mx[k]); }
    if (lane_id() == 0)
        for (int k = 0; k < 3; ++k) { sh[k][wave_id()] = mn[k]; sh[3 + k][wave_id()] = mx[k]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float a = sh[k][0];
        for (int w = 1; w < 4; ++w) a = k < 3 ? fminf(a, sh[k][w]) : fmaxf(a, sh[k][w]);
        if (a == a && a != INF && a != -INF) {
            // filtered: only a workgroup that improves the current bound
            // issues the atomic (thousands of same-address atomics serialise)
            const int32_t key = f2ord(a);
            const int32_t cur = __hip_atomic_load(&ws.cbox[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (k < 3 ? key < cur : key > cur) {
                if (k < 3) atomicMin(&ws.cbox[k], key);
                else atomicMax(&ws.cbox[k], key);
            }
        }
    }
}

__device__ __forceinline__ uint32_t stretch3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
}
__device__ __forceinline__ uint32_t quant10(float v) { return (uint32_t)kmin(kmax(v * 1024.0f, 0.0f), 1023.0f); }
__device__ __forceinline__ float safe_div(float a, float b) { return (b != 0.0f) ? a / b : 0.0f; }

__global__ __launch_bounds__(256) void lbvh_codes(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, LbvhWs ws) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const f3 cmn = f3{ord2f(ws.cbox[0]), ord2f(ws.cbox[1]), ord2f(ws.cbox[2])};
    const f3 cmx = f3{ord2f(ws.cbox[3]), ord2f(ws.cbox[4]), ord2f(ws.cbox[5])};
    Box b = tri_box(V, F[f]);
    const f3 neg = f3{-cmn.x, -cmn.y, -cmn.z};
    Box o{add3(b.mn, neg), add3(b.mx, neg)};  // kr_aabb_offset3(bbox, -cell.min)
    f3 c = box_center(o);
    const f3 ext = sub3(cmx, cmn);
    c = f3{safe_div(c.x, ext.x), safe_div(c.y, ext.y), safe_div(c.z, ext.z)};
    ws.keys_in[f] = stretch3(quant10(c.x)) << 2 | stretch3(quant10(c.y)) << 1 | stretch3(quant10(c.z));
    ws.vals_in[f] = f;
}

__device__ __forceinline__ int delta(const uint32_t* c, int i, int j, int n) {
    if (j > n - 1 || j < 0) return -1;
    uint32_t a = c[i], b = c[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}
__device__ __forceinline__ int lid(int num) {
    if (num == 1) return 0;
    if (num % 2 == 1) return num / 2 + 1;
    return num / 2;
}

__global__ __launch_bounds__(256) void lbvh_tree(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, LbvhWs ws,
                                                 kr_bvh_node* nodes, uint32_t* parents, uint32_t* mortons) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf) return;
    const uint32_t I = nf - 1;
    // leaf t (sorted order)
    {
        const uint32_t prim = ws.vals_out[t];
        Box b = tri_box(V, F[prim]);
        kr_bvh_node L;
        L.bounds.min = kr_vec3{b.mn.x, b.mn.y, b.mn.z};
        L.bounds.max = kr_vec3{b.mx.x, b.mx.y, b.mx.z};
        L.u.children.left = prim;
        L.u.children.right = 0;
        L.nPrimitives = 1;
        L.axis = 0;
        L.pad_ = 0;
        nodes[I + t] = L;
        if (mortons) mortons[t] = ws.keys_out[t];
    }
    if (nf == 1) { parents[0] = kInv; return; }
    if (t >= I) return;
    const uint32_t* c = ws.keys_out;
    const int n = (int)nf, i = (int)t;
    const int da = delta(c, i, i + 1, n), db = delta(c, i, i - 1, n);
    const int d = (da - db > 0) - (da - db < 0);
    const int dmin = delta(c, i, i - d, n);
    int lmax = 2;
    while (delta(c, i, i + lmax * d, n) > dmin) lmax *= 2;
    int l = 0;
    for (int s = lmax / 2; s >= 1; s /= 2)
        if (delta(c, i, i + (l + s) * d, n) > dmin) l += s;
    const int j = i + l * d;
    const int dnode = delta(c, i, j, n);
    int s = 0;
    for (int q = lid(l); q >= 1; q = lid(q))
        if (delta(c, i, i + (s + q) * d, n) > dnode) s += q;
    const int split = i + s * d + (d < 0 ? d : 0);
    const bool lleaf = ((i < j ? i : j) == split);
    const bool rleaf = ((i > j ? i : j) == split + 1);
    const uint32_t li = (uint32_t)split + (lleaf ? I : 0);
    const uint32_t ri = (uint32_t)split + 1 + (rleaf ? I : 0);
    parents[li] = t;
    parents[ri] = t;
    kr_bvh_node N;
    N.bounds.min = kr_vec3{FLT_MAX, FLT_MAX, FLT_MAX};
    N.bounds.max = kr_vec3{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    N.u.children.left = li;
    N.u.children.right = ri;
    N.nPrimitives = 0;
    N.axis = 0;
    N.pad_ = 0;
    nodes[t] = N;
    if (t == 0) parents[0] = kInv;
}

// Bottom-up bounds / counts / costs (lbvh.cu:399-457) with the phased
// arrival protocol of phased.hpp.
struct LbvhUp {
    kr_bvh_node* nodes;
    uint32_t* counts;
    float* costs;
    __device__ void leaf(uint32_t cur) {
        const kr_bvh_node L = nodes[cur];
        costs[cur] = 1.0f * box_area(f3{L.bounds.min.x, L.bounds.min.y, L.bounds.min.z},
                                     f3{L.bounds.max.x, L.bounds.max.y, L.bounds.max.z});
        counts[cur] = 1;
    }
    __device__ void merge(uint32_t p, uint32_t l, uint32_t r) {
        const float ci = (float)1.2;
        const kr_bvh_node A = nodes[l], B = nodes[r];
        Box bn{f3{A.bounds.min.x, A.bounds.min.y, A.bounds.min.z},
               f3{A.bounds.max.x, A.bounds.max.y, A.bounds.max.z}};
        bn = expand3(bn, f3{B.bounds.min.x, B.bounds.min.y, B.bounds.min.z});  // kr_aabb_expand
        bn = expand3(bn, f3{B.bounds.max.x, B.bounds.max.y, B.bounds.max.z});
        counts[p] = counts[l] + counts[r];
        costs[p] = ci * box_area(bn.mn, bn.mx) + costs[r] + costs[l];
        nodes[p].bounds.min = kr_vec3{bn.mn.x, bn.mn.y, bn.mn.z};
        nodes[p].bounds.max = kr_vec3{bn.mx.x, bn.mx.y, bn.mx.z};
    }
};

__global__ __launch_bounds__(1024) void lbvh_aabbs(kr_bvh_node* nodes, const uint32_t* parents, uint32_t* counts,
                                                   float* costs, uint32_t nf, LbvhWs ws, uint32_t phase,
                                                   uint32_t chunk) {
    LbvhUp op{nodes, counts, costs};
    phased_up(nodes, parents, nf, ws.ph, phase, chunk, op);
}

size_t lbvh_sort_bytes(uint32_t nf) {
    size_t bytes = 0;
    rocprim::radix_sort_pairs((void*)nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                              (uint32_t*)nullptr, nf, 0, 30);
    return bytes;
}

size_t lbvh_carve(LbvhWs* ws, void* base, size_t cap, uint32_t nf) {
    Carver c{(char*)base, 0, cap};
    ws->cbox = c.take<int32_t>(8);
    phase_carve(&ws->ph, c, nf);
    ws->keys_in = c.take<uint32_t>(nf);
    ws->keys_out = c.take<uint32_t>(nf);
    ws->vals_in = c.take<uint32_t>(nf);
    ws->vals_out = c.take<uint32_t>(nf);
    ws->sort_bytes = lbvh_sort_bytes(nf);
    ws->sort_tmp = c.take<char>(ws->sort_bytes);
    return c.used;
}

kr_error lbvh_launch(const kr_vec3* V, const kr_uvec4* F, uint32_t nf, kr_bvh_node* nodes, uint32_t* parents,
                     uint32_t* counts, float* costs, uint32_t* mortons, void* workspace, size_t ws_bytes,
                     hipStream_t stream) {
    if (nf == 0) return "obvh_lbvh_build: no faces";
    LbvhWs ws;
    size_t need = lbvh_carve(&ws, workspace, ws_bytes, nf);
    if (!workspace || ws_bytes < need) return "obvh_lbvh_build: workspace too small";
    auto blocks = [](uint32_t n, uint32_t b) { return dim3((n + b - 1) / b); };
    uint32_t gb = (nf + 255) / 256;
    if (gb > 2048) gb = 2048;
    hipLaunchKernelGGL(lbvh_init, dim3(gb), dim3(256), 0, stream, V, nf, ws);
    hipLaunchKernelGGL(lbvh_centroid_box, dim3(gb), dim3(256), 0, stream, V, F, nf, ws);
    hipLaunchKernelGGL(lbvh_codes, blocks(nf, 256), dim3(256), 0, stream, V, F, nf, ws);
    size_t sb = ws.sort_bytes;
    OBVH_HIP_TRY(rocprim::radix_sort_pairs(ws.sort_tmp, sb, ws.keys_in, ws.keys_out, ws.vals_in, ws.vals_out, nf,
                                           0, 30, stream));
    hipLaunchKernelGGL(lbvh_tree, blocks(nf, 256), dim3(256), 0, stream, V, F, nf, ws, nodes, parents, mortons);
    phased_launch(lbvh_aabbs, nf, stream, nodes, (const uint32_t*)parents, counts, costs, nf, ws);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
