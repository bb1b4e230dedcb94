// traverse.hip -- OBVH packing and closest-hit ray traversal on gfx950.
//
// Packing (transforms.cu:1298-1449).  soa_pack goes from the (plain or
// collapsed) tree and the transform's world -> unit-cube matrices straight
// to the structure-of-arrays OBVH in one pass, one lane per node: metadata
// (children, bitwise NOT for leaf children, or a leaf's primitive slice;
// axis) and, per internal node, the two children's 3x4 matrices.  The two
// legacy steps (AoS packed nodes + mat4 pairs, then AoS -> SoA) are kept as
// aos_pack / aos_to_soa for the drop-in entry points; the reference's
// second step runs on the host, here it is a kernel.
//
// Traversal (soa_obvh_persistent_intersect, intersectors.cu:801-972).  One
// ray per lane, a persistent grid whose waves take 64 consecutive rays at a
// time from a counter.  Per ray: the reference's depth-first walk with one
// postponed leaf -- descend through internal nodes (both children tested
// against the unit cube in their own frames; one entered: go, none: pop,
// both: enter the nearer by t0, left on ties, push the other) until a leaf
// is found, then intersect that leaf's triangles and any leaves popped
// right after it.  Per lane the node/leaf sequence, every floating-point
// operation and the strict "t < hitT" update equal the reference's, so the
// hit (primitive, barycentrics, normal) is bit-identical.
//
// Stack: slot 0 is the implicit sentinel, slots 1..24 live in LDS
// (slot-major, [slot][lane]: every access is bank-conflict free), slots
// 25..63 in a global overflow area ([slot][thread], coalesced), 63 slots as
// the reference's 64-entry array.  A ray that would need a 64th slot, or
// that exceeds `max_steps` node + leaf visits (only a malformed tree can),
// stops with its hit so far and instance = 0xFFFFFFFF; so does a ray that
// meets a node, primitive or vertex index outside the arrays' counts.
//
// Data per internal node visit: metadata 8 B + two 48 B matrices; per
// triangle: face 16 B + three 12 B vertices.
#include "common.hpp"
#include "traverse.hpp"

namespace obvh {

constexpr uint16_t kCollapsedNode = 0xFFFF;

// ------------------------------------------------------------- packing --

// kr_m43from4 (vecmath.h:895-902) as three float4: v[0..2], v[4..6],
// v[8..10], v[12..14] of the column-major 4x4.
__device__ __forceinline__ void m43_from_mat4(const kr_mat4* m, float4 out[3]) {
    const float4* c = reinterpret_cast<const float4*>(m->v);
    const float4 c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    out[0] = make_float4(c0.x, c0.y, c0.z, c1.x);
    out[1] = make_float4(c1.y, c1.z, c2.x, c2.y);
    out[2] = make_float4(c2.z, c3.x, c3.y, c3.z);
}

__device__ __forceinline__ int32_t child_code(const kr_bvh_node* nodes, uint32_t c) {
    return nodes[c].nPrimitives > 0 ? ~(int32_t)c : (int32_t)c;
}

__global__ __launch_bounds__(256) void soa_pack(const kr_bvh_node* __restrict__ nodes, const kr_mat4* __restrict__ mats,
                                                uint32_t node_count, uint32_t internal_count, float4* __restrict__ LT,
                                                float4* __restrict__ RT, int4* __restrict__ meta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= node_count) return;
    const kr_bvh_node n = nodes[i];
    int4 m = make_int4(0, 0, 0, 0);
    float4 l[3] = {}, r[3] = {};
    if (n.nPrimitives != kCollapsedNode) {
        m.z = n.axis;
        if (n.nPrimitives == 0) {
            const uint32_t a = n.u.children.left, b = n.u.children.right;
            if (a < node_count && b < node_count && i < internal_count) {
                m.x = child_code(nodes, a);
                m.y = child_code(nodes, b);
                m43_from_mat4(&mats[a], l);
                m43_from_mat4(&mats[b], r);
            }
        } else {
            m.x = (int32_t)n.u.primitivesOffset;
            m.y = (int32_t)(n.u.primitivesOffset + n.nPrimitives);
        }
    }
    meta[i] = m;
    if (i < internal_count) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            LT[3 * (size_t)i + k] = l[k];
            RT[3 * (size_t)i + k] = r[k];
        }
    }
}

// obvh_persistent_prepare: skipped nodes and metadata.w untouched.
__global__ __launch_bounds__(256) void aos_pack(const kr_bvh_node* __restrict__ nodes, const kr_mat4* __restrict__ mats,
                                                kr_bvh_node_packed* __restrict__ packed,
                                                kr_bvh_transformation_pair* __restrict__ pairs, uint32_t node_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= node_count) return;
    const kr_bvh_node n = nodes[i];
    if (n.nPrimitives == kCollapsedNode) return;
    kr_bvh_node_packed* p = &packed[i];
    if (n.nPrimitives == 0) {
        const uint32_t a = n.u.children.left, b = n.u.children.right;
        if (a >= node_count || b >= node_count || i >= (node_count - 1) / 2) return;
        const kr_bvh_node l = nodes[a], r = nodes[b];
        p->lbounds = l.bounds;
        p->rbounds = r.bounds;
        p->metadata.x = l.nPrimitives > 0 ? ~(int32_t)a : (int32_t)a;
        p->metadata.y = r.nPrimitives > 0 ? ~(int32_t)b : (int32_t)b;
        pairs[i].l = mats[a];
        pairs[i].r = mats[b];
    } else {
        p->lbounds = n.bounds;
        p->metadata.x = (int32_t)n.u.primitivesOffset;
        p->metadata.y = (int32_t)(n.u.primitivesOffset + n.nPrimitives);
    }
    p->metadata.z = n.axis;
}

// kr_cuda_soa_obvh's conversion loop.
__global__ __launch_bounds__(256) void aos_to_soa(const kr_bvh_node_packed* __restrict__ packed,
                                                  const kr_bvh_transformation_pair* __restrict__ pairs,
                                                  uint32_t internal_count, uint32_t node_count, float4* __restrict__ LT,
                                                  float4* __restrict__ RT, kr_ivec4* __restrict__ meta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < internal_count) {
        float4 l[3], r[3];
        m43_from_mat4(&pairs[i].l, l);
        m43_from_mat4(&pairs[i].r, r);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            LT[3 * (size_t)i + k] = l[k];
            RT[3 * (size_t)i + k] = r[k];
        }
    }
    if (i < node_count) meta[i] = packed[i].metadata;
}

// ----------------------------------------------------------- traversal --

struct M43 {
    float4 a, b, c;  // v[0..3], v[4..7], v[8..11]
};

__device__ __forceinline__ M43 load_m43(const float4* __restrict__ T, int32_t node) {
    const float4* p = T + 3 * (size_t)node;
    return M43{p[0], p[1], p[2]};
}

// kr_n43transform3p / kr_v43transform3p (vecmath.h:1262-1309), left to right.
__device__ __forceinline__ f3 m43_dir(const M43& m, f3 d) {
    return f3{m.a.x * d.x + m.a.w * d.y + m.b.z * d.z, m.a.y * d.x + m.b.x * d.y + m.b.w * d.z,
              m.a.z * d.x + m.b.y * d.y + m.c.x * d.z};
}
__device__ __forceinline__ f3 m43_point(const M43& m, f3 p) {
    return f3{m.a.x * p.x + m.a.w * p.y + m.b.z * p.z + m.c.y, m.a.y * p.x + m.b.x * p.y + m.b.w * p.z + m.c.z,
              m.a.z * p.x + m.b.y * p.y + m.c.x * p.z + m.c.w};
}

// kr_ray_unit_aabb_intersect (vecmath.h:1719-1749) fused with the caller's
// entry test (intersectors.cu:881-882): the box returns (1, -1) unless
// t0 <= t1 && t1 > 0, and then "t1 >= t0 && t1 >= 0 && t0 <= hitT" reduces
// to t0 <= hitT.  Returns whether the child is entered; t0 in *t_near.
__device__ __forceinline__ bool enter_unit_box(f3 rd, f3 ro, float hitT, float* t_near) {
    const float lx = (-0.5f - ro.x) / rd.x, ly = (-0.5f - ro.y) / rd.y, lz = (-0.5f - ro.z) / rd.z;
    const float hx = (0.5f - ro.x) / rd.x, hy = (0.5f - ro.y) / rd.y, hz = (0.5f - ro.z) / rd.z;
    const float t0 = kmax(kmax(kmin(lx, hx), kmin(ly, hy)), kmin(lz, hz));
    const float t1 = kmin(kmin(kmax(lx, hx), kmax(ly, hy)), kmax(lz, hz));
    *t_near = t0;
    return (t0 <= t1) && (t1 > 0.0f) && (t0 <= hitT);
}

// kr_ray_triangle_intersect (vecmath.h:1656-1676); returns t (t_max when
// rejected) and the barycentrics of an accepted hit.
__device__ __forceinline__ float ray_triangle(f3 dir, f3 org, f3 v1, f3 v2, f3 v3, float t_max, float* b1o,
                                              float* b2o) {
    const f3 e1 = sub3(v2, v1), e2 = sub3(v3, v1);
    const f3 s1 = vm::cross(dir, e2);
    const float invd = 1.0f / dot3(s1, e1);
    const f3 d = sub3(org, v1);
    const float b1 = dot3(d, s1) * invd;
    const f3 s2 = vm::cross(d, e1);
    const float b2 = dot3(dir, s2) * invd;
    const float t = dot3(e2, s2) * invd;
    if (b1 < 0.0f || b1 > 1.0f || b2 < 0.0f || b1 + b2 > 1.0f || t < 0.0f || t > t_max) return t_max;
    *b1o = b1;
    *b2o = b2;
    return t;
}

__device__ __forceinline__ f3 vload(const kr_vec3* V, uint32_t i) {
    const kr_vec3 v = V[i];
    return f3{v.x, v.y, v.z};
}

struct TravArgs {
    const float4* LT;
    const float4* RT;
    const int4* meta;
    const kr_vec3* V;
    const kr_uvec4* F;
    const uint32_t* prims;
    const kr_ray* rays;
    kr_intersection* out;
    uint32_t n;
    uint32_t max_steps;
    uint32_t internal_count;  // entries of LT / RT
    uint32_t node_count;      // entries of meta
    uint32_t face_count;      // entries of F (and of prims)
    uint32_t vertex_count;    // entries of V
    uint32_t* counter;
    int32_t* deep;
};

__global__ void trav_clear(uint32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

__global__ __launch_bounds__(kTravThreads) void soa_intersect(TravArgs a) {
    __shared__ int32_t lstack[kTravLdsSlots * kTravThreads];
    const uint32_t tid = threadIdx.x;
    const size_t stride = (size_t)gridDim.x * kTravThreads;
    int32_t* deep = a.deep + (size_t)blockIdx.x * kTravThreads + tid;
    const int lane = lane_id();

    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.counter, (uint32_t)kWave);
        base = (uint32_t)__shfl((int)base, 0, kWave);
        if (base >= a.n) break;  // wave-uniform
        const uint32_t ray = base + (uint32_t)lane;
        const bool live = ray < a.n;

        f3 org{0.0f, 0.0f, 0.0f}, dir{0.0f, 0.0f, 0.0f};
        float hitT = 0.0f;
        if (live) {
            const kr_ray r = a.rays[ray];
            org = f3{r.origin.x, r.origin.y, r.origin.z};
            dir = f3{r.direction.x, r.direction.y, r.direction.z};
            hitT = r.tmax;
        }
        uint32_t hit = 0xFFFFFFFFu, steps = 0;
        float bx = 0.0f, by = 0.0f;
        bool aborted = false;
        int sp = 0;
        int32_t cur = live ? 0 : kTravSentinel;
        int32_t leaf = 0;  // < 0: a leaf waiting to be intersected

        auto pop = [&]() -> int32_t {
            if (sp == 0) return kTravSentinel;
            const int s = sp--;
            return s <= kTravLdsSlots ? lstack[(s - 1) * kTravThreads + tid]
                                      : deep[(size_t)(s - kTravLdsSlots - 1) * stride];
        };

        while (cur != kTravSentinel) {
            // Descend until one leaf is found (postponed) or the stack is empty.
            while (cur >= 0 && cur != kTravSentinel) {
                if (++steps > a.max_steps || (uint32_t)cur >= a.internal_count) {
                    aborted = true;
                    cur = kTravSentinel;
                    break;
                }
                const int4 md = a.meta[cur];
                const M43 L = load_m43(a.LT, cur);
                const M43 R = load_m43(a.RT, cur);
                float tl, tr;
                const bool go_l = enter_unit_box(m43_dir(L, dir), m43_point(L, org), hitT, &tl);
                const bool go_r = enter_unit_box(m43_dir(R, dir), m43_point(R, org), hitT, &tr);
                int32_t next = md.x, other = md.y;
                if (go_l != go_r) {
                    if (go_r) next = other;
                } else if (!go_l) {
                    next = pop();
                } else {
                    if (tr < tl) {
                        const int32_t t = next;
                        next = other;
                        other = t;
                    }
                    if (sp == kTravSlots) {
                        aborted = true;
                        cur = kTravSentinel;
                        break;
                    }
                    const int s = ++sp;
                    if (s <= kTravLdsSlots) lstack[(s - 1) * kTravThreads + tid] = other;
                    else deep[(size_t)(s - kTravLdsSlots - 1) * stride] = other;
                }
                cur = next;
                if (cur < 0) {
                    leaf = cur;
                    cur = pop();
                    break;
                }
            }
            // Intersect the postponed leaf and the leaves popped right after it.
            while (leaf < 0) {
                const uint32_t lf = (uint32_t)~leaf;
                bool bad = ++steps > a.max_steps || lf >= a.node_count;
                const int4 md = bad ? make_int4(0, 0, 0, 0) : a.meta[lf];
                const int32_t beg = md.x, end = md.y, num = end - beg;
                bad = bad || beg < 0 || end < beg || (num > 1 && (!a.prims || (uint32_t)end > a.face_count));
                for (int32_t k = beg; k < end && !bad; ++k) {
                    const uint32_t pid = (num == 1) ? (uint32_t)k : a.prims[k];
                    if (pid >= a.face_count) {
                        bad = true;
                        break;
                    }
                    const kr_uvec4 f = a.F[pid];
                    if (f.x >= a.vertex_count || f.y >= a.vertex_count || f.z >= a.vertex_count) {
                        bad = true;
                        break;
                    }
                    float b1 = 0.0f, b2 = 0.0f;
                    const float t = ray_triangle(dir, org, vload(a.V, f.x), vload(a.V, f.y), vload(a.V, f.z), hitT,
                                                 &b1, &b2);
                    if (t < hitT) {
                        hitT = t;
                        hit = pid;
                        bx = b1;
                        by = b2;
                    }
                }
                if (bad) {
                    aborted = true;
                    cur = kTravSentinel;
                    break;
                }
                leaf = cur;
                if (cur < 0) cur = pop();
            }
        }

        if (live) {
            kr_intersection o;
            o.instance = aborted ? kTravAborted : 0u;
            o.primitive = hit;
            if (hit == 0xFFFFFFFFu) {
                o.geom_normal = kr_vec3{dir.x, dir.y, dir.z};
                o.barys = kr_vec2{0.5f, 0.5f};
            } else {
                const kr_uvec4 f = a.F[hit];
                const f3 va = vload(a.V, f.x), vb = vload(a.V, f.y), vc = vload(a.V, f.z);
                const f3 nrm = vm::normalize(vm::cross(sub3(vb, va), sub3(vc, va)));
                o.geom_normal = kr_vec3{nrm.x, nrm.y, nrm.z};
                o.barys = kr_vec2{bx, by};
            }
            a.out[ray] = o;
        }
    }
}

// ---------------------------------------------------------------- host --

static uint32_t trav_blocks(uint32_t ray_count) {
    const uint32_t b = (ray_count + kTravThreads - 1) / kTravThreads;
    return b < kTravMaxBlocks ? (b == 0 ? 1u : b) : kTravMaxBlocks;
}

size_t trav_carve(TravWs* ws, void* base, size_t cap, uint32_t ray_count) {
    Carver c{static_cast<char*>(base), 0, cap};
    const uint32_t blocks = trav_blocks(ray_count);
    ws->counter = c.take<uint32_t>(64);
    ws->deep = c.take<int32_t>((size_t)kTravDeepSlots * blocks * kTravThreads);
    ws->blocks = blocks;
    return c.used;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

kr_error soa_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, uint32_t node_count, uint32_t internal_count,
                         kr_mat43* left_T, kr_mat43* right_T, kr_ivec4* meta, hipStream_t stream) {
    if (node_count == 0) return nullptr;
    if (!aligned16(left_T) || !aligned16(right_T) || !aligned16(meta) || !aligned16(mats))
        return "obvh_soa_pack: left_T, right_T, metadata and transforms must be 16-byte aligned";
    hipLaunchKernelGGL(soa_pack, dim3((node_count + 255) / 256), dim3(256), 0, stream, nodes, mats, node_count,
                       internal_count, reinterpret_cast<float4*>(left_T), reinterpret_cast<float4*>(right_T),
                       reinterpret_cast<int4*>(meta));
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error aos_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, kr_bvh_node_packed* packed,
                         kr_bvh_transformation_pair* pairs, uint32_t node_count, hipStream_t stream) {
    if (node_count == 0) return nullptr;
    hipLaunchKernelGGL(aos_pack, dim3((node_count + 255) / 256), dim3(256), 0, stream, nodes, mats, packed, pairs,
                       node_count);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error aos_to_soa_launch(const kr_bvh_node_packed* packed, const kr_bvh_transformation_pair* pairs,
                           uint32_t internal_count, uint32_t node_count, kr_mat43* left_T, kr_mat43* right_T,
                           kr_ivec4* meta, hipStream_t stream) {
    const uint32_t n = internal_count > node_count ? internal_count : node_count;
    if (n == 0) return nullptr;
    if (!aligned16(left_T) || !aligned16(right_T) || !aligned16(pairs))
        return "kr_cuda_soa_obvh: arrays must be 16-byte aligned";
    hipLaunchKernelGGL(aos_to_soa, dim3((n + 255) / 256), dim3(256), 0, stream, packed, pairs, internal_count,
                       node_count, reinterpret_cast<float4*>(left_T), reinterpret_cast<float4*>(right_T), meta);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error intersect_launch(const kr_mat43* left_T, const kr_mat43* right_T, const kr_ivec4* meta, const kr_vec3* V,
                          const kr_uvec4* F, const uint32_t* prims, const kr_ray* rays, kr_intersection* out,
                          uint32_t ray_count, const TravCounts& counts, void* workspace, size_t ws_bytes,
                          hipStream_t stream) {
    if (ray_count == 0) return nullptr;
    if (!aligned16(left_T) || !aligned16(right_T) || !aligned16(meta))
        return "obvh_soa_intersect: left_T, right_T and metadata must be 16-byte aligned";
    if (ray_count > 0xFFFFFFFFu - kWave) return "obvh_soa_intersect: too many rays";
    TravWs ws;
    if (trav_carve(&ws, workspace, ws_bytes, ray_count) > ws_bytes || !workspace)
        return "obvh_soa_intersect: workspace too small";
    TravArgs a{reinterpret_cast<const float4*>(left_T),
               reinterpret_cast<const float4*>(right_T),
               reinterpret_cast<const int4*>(meta),
               V,
               F,
               prims,
               rays,
               out,
               ray_count,
               counts.max_steps,
               counts.internal_count,
               counts.node_count,
               counts.face_count,
               counts.vertex_count,
               ws.counter,
               ws.deep};
    hipLaunchKernelGGL(trav_clear, dim3(1), dim3(64), 0, stream, ws.counter);
    hipLaunchKernelGGL(soa_intersect, dim3(ws.blocks), dim3(kTravThreads), 0, stream, a);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
