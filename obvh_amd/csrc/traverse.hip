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
// ray per lane, a persistent grid; a lane whose ray is finished waits until
// a quarter of its wave is idle, then those lanes take consecutive rays from
// a global counter (one atomic per refill), so long rays do not hold idle
// lanes back.  Per ray: the reference's depth-first walk -- at an internal
// node both children are tested against the unit cube in their own frames
// (one entered: go, none: pop, both: enter the nearer by t0, left on ties,
// push the other); a leaf's triangles are tested in slice order, then the
// walk pops.  The reference runs the walk as two inner loops (descend to a
// leaf, then intersect it and any leaves popped right after it); that gives
// the same per-ray sequence of node visits and triangle tests as this
// single loop in which every lane does one step -- one node or one
// triangle -- per iteration, which keeps the 64 lanes of a wave in step
// (two inner loops each run as long as their slowest lane).  Per lane every
// floating-point operation and the strict "t < hitT" update equal the
// reference's, so the hit (primitive, barycentrics, normal) is
// bit-identical.
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

// Fast form of the same slab test: (+-0.5 - ro) * rcp(rd) instead of the
// correctly rounded quotients.  v_rcp_f32 is within 1 ulp, so with every
// |rd| in [2^-40, 2^40] and every product in [2^-100, 2^100] (no zero, no
// overflow, no denormal) each product is within 2^-21 relative of the
// correctly rounded quotient, and so are the entry / exit distances t0, t1
// (a max / min of them).  Returns false when those ranges do not hold.
__device__ __forceinline__ bool slab_fast(f3 rd, f3 ro, float& t0, float& t1) {
    const float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    const float rx = __builtin_amdgcn_rcpf(rd.x), ry = __builtin_amdgcn_rcpf(rd.y), rz = __builtin_amdgcn_rcpf(rd.z);
    const float lx = (-0.5f - ro.x) * rx, ly = (-0.5f - ro.y) * ry, lz = (-0.5f - ro.z) * rz;
    const float hx = (0.5f - ro.x) * rx, hy = (0.5f - ro.y) * ry, hz = (0.5f - ro.z) * rz;
    const float dmin = fminf(fminf(ax, ay), az), dmax = fmaxf(fmaxf(ax, ay), az);
    const float pmin = fminf(fminf(fminf(fabsf(lx), fabsf(ly)), fminf(fabsf(lz), fabsf(hx))), fminf(fabsf(hy), fabsf(hz)));
    const float pmax = fmaxf(fmaxf(fmaxf(fabsf(lx), fabsf(ly)), fmaxf(fabsf(lz), fabsf(hx))), fmaxf(fabsf(hy), fabsf(hz)));
    t0 = fmaxf(fmaxf(fminf(lx, hx), fminf(ly, hy)), fminf(lz, hz));
    t1 = fminf(fminf(fmaxf(lx, hx), fmaxf(ly, hy)), fmaxf(lz, hz));
    return dmin >= 0x1p-40f && dmax <= 0x1p40f && pmin >= 0x1p-100f && pmax <= 0x1p100f;
}

// x <= y decided from approximations within 2^-21 relative of the exact
// operands: certain only when they differ by more than 2^-19 (|x| + |y|).
__device__ __forceinline__ bool sure(float x, float y) { return fabsf(x - y) > 0x1p-19f * (fabsf(x) + fabsf(y)); }

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

// How a traversal step reads the tree.  Both views give the same per-ray
// sequence of node visits and triangle tests with the same values, so the
// same hits.
//
// SoaTree: the reference's kr_SoA_OBVH arrays (3x4 matrices per internal
// node, metadata per node, leaves' primitives through faces and vertices).
struct SoaTree {
    const float4* LT;
    const float4* RT;
    const int4* meta;
    const kr_vec3* V;
    const kr_uvec4* F;
    const uint32_t* prims;
    uint32_t internal_count;  // entries of LT / RT
    uint32_t node_count;      // entries of meta
    uint32_t face_count;      // entries of F (and of prims)
    uint32_t vertex_count;    // entries of V

    __device__ bool node(int32_t cur, M43& L, M43& R, int32_t& left, int32_t& right) const {
        if ((uint32_t)cur >= internal_count) return false;
        const int4 md = meta[cur];
        L = load_m43(LT, cur);
        R = load_m43(RT, cur);
        left = md.x;
        right = md.y;
        return true;
    }
    // Leaf code -> primitive cursor [k, end); one: a single primitive whose id is k.
    __device__ bool leaf(int32_t code, int32_t& k, int32_t& end, bool& one) const {
        const uint32_t lf = (uint32_t)~code;
        if (lf >= node_count) return false;
        const int4 md = meta[lf];
        k = md.x;
        end = md.y;
        one = end - k == 1;
        return !(k < 0 || end < k || (!one && end > k && (!prims || (uint32_t)end > face_count)));
    }
    __device__ bool tri(int32_t k, bool one, uint32_t& pid, f3& v0, f3& v1, f3& v2) const {
        pid = one ? (uint32_t)k : prims[k];
        if (pid >= face_count) return false;
        const kr_uvec4 f = F[pid];
        if (f.x >= vertex_count || f.y >= vertex_count || f.z >= vertex_count) return false;
        v0 = vload(V, f.x);
        v1 = vload(V, f.y);
        v2 = vload(V, f.z);
        return true;
    }
    __device__ void hit_vertices(uint32_t pid, int32_t, f3& va, f3& vb, f3& vc) const {
        const kr_uvec4 f = F[pid];
        va = vload(V, f.x);
        vb = vload(V, f.y);
        vc = vload(V, f.z);
    }
};

// RecTree: node records (128 B, one cache line: both children's matrices
// and child codes; a leaf child is ~(first triangle slot << 4 | count - 1))
// and triangle records (48 B: three vertices and the primitive id, a leaf's
// triangles in consecutive slots).
struct RecTree {
    const float4* rec;   // 8 float4 per internal node
    const float4* tris;  // 3 float4 per slot
    uint32_t internal_count, slot_count, face_count;

    __device__ bool node(int32_t cur, M43& L, M43& R, int32_t& left, int32_t& right) const {
        if ((uint32_t)cur >= internal_count) return false;
        const float4* p = rec + 8 * (size_t)cur;
        L = M43{p[0], p[1], p[2]};
        R = M43{p[3], p[4], p[5]};
        const float4 c = p[6];
        left = __float_as_int(c.x);
        right = __float_as_int(c.y);
        return true;
    }
    __device__ bool leaf(int32_t code, int32_t& k, int32_t& end, bool& one) const {
        const uint32_t x = (uint32_t)~code;
        k = (int32_t)(x >> 4);
        end = k + (int32_t)(x & 15u) + 1;
        one = false;
        return (uint32_t)end <= slot_count;
    }
    __device__ bool tri(int32_t k, bool, uint32_t& pid, f3& v0, f3& v1, f3& v2) const {
        const float4* p = tris + 3 * (size_t)k;
        const float4 a = p[0], b = p[1], c = p[2];
        v0 = f3{a.x, a.y, a.z};
        v1 = f3{a.w, b.x, b.y};
        v2 = f3{b.z, b.w, c.x};
        pid = __float_as_uint(c.y);
        return pid < face_count;
    }
    __device__ void hit_vertices(uint32_t, int32_t slot, f3& va, f3& vb, f3& vc) const {
        uint32_t pid;
        tri(slot, false, pid, va, vb, vc);
    }
};

struct TraceIO {
    const kr_ray* rays;
    kr_intersection* out;
    uint32_t n;
    uint32_t max_steps;
    uint32_t* counter;
    int32_t* deep;
};

__global__ void trav_clear(uint32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

template <class Tree>
__global__ __launch_bounds__(kTravThreads) void trace(Tree tree, TraceIO a) {
    __shared__ int32_t lstack[kTravLdsSlots * kTravThreads];
    const uint32_t tid = threadIdx.x;
    const size_t stride = (size_t)gridDim.x * kTravThreads;
    int32_t* deep = a.deep + (size_t)blockIdx.x * kTravThreads + tid;
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1;

    // Per-lane ray state.  A lane is busy while it holds an unfinished ray;
    // idle lanes are refilled from the counter once kTravRefill of them
    // wait (or none is busy).  cur >= 0: internal node to visit; cur < 0:
    // a leaf whose primitives [k, end) remain to be tested.
    uint32_t ray = 0, hit = 0, steps = 0;
    int32_t hit_slot = 0;
    bool busy = false, exhausted = false, aborted = false;
    f3 org{0.0f, 0.0f, 0.0f}, dir{0.0f, 0.0f, 0.0f};
    float hitT = 0.0f, bx = 0.0f, by = 0.0f;
    int sp = 0;
    int32_t cur = kTravSentinel, k = 0, end = 0;
    bool one = false;

    auto pop = [&]() -> int32_t {
        if (sp == 0) return kTravSentinel;
        const int s = sp--;
        return s <= kTravLdsSlots ? lstack[(s - 1) * kTravThreads + tid]
                                  : deep[(size_t)(s - kTravLdsSlots - 1) * stride];
    };

    for (;;) {
        const bool wants = !busy && !exhausted;
        const uint64_t idle = __ballot(wants);
        const int nidle = __popcll(idle);
        if (nidle >= kTravRefill || (nidle > 0 && __ballot(busy) == 0)) {  // wave-uniform
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(a.counter, (uint32_t)nidle);
            base = (uint32_t)__shfl((int)base, 0, kWave);
            if (wants) {
                const uint32_t r = base + (uint32_t)__popcll(idle & below);
                if (r < a.n) {
                    const kr_ray q = a.rays[r];
                    ray = r;
                    busy = true;
                    org = f3{q.origin.x, q.origin.y, q.origin.z};
                    dir = f3{q.direction.x, q.direction.y, q.direction.z};
                    hitT = q.tmax;
                    hit = 0xFFFFFFFFu;
                    steps = 0;
                    bx = by = 0.0f;
                    aborted = false;
                    sp = 0;
                    cur = 0;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(busy) == 0) break;  // every lane exhausted

        // One step per lane: an internal node or one primitive of a leaf.
        // (Holding triangle steps back until more lanes want one, or
        // refilling at 8 idle lanes instead of 16, measured slower.)
        bool enter_leaf = false;
        if (cur >= 0 && cur != kTravSentinel) {
            M43 L, R;
            int32_t next, other;
            if (++steps > a.max_steps || !tree.node(cur, L, R, next, other)) {
                aborted = true;
                cur = kTravSentinel;
            } else {
                const f3 rdl = m43_dir(L, dir), rol = m43_point(L, org);
                const f3 rdr = m43_dir(R, dir), ror = m43_point(R, org);
                // entry decisions from the fast slab test when each is certain
                // (t1 > 0 is exact: each product has its quotient's sign),
                // else from the exact test
                float tl, tr, t1l, t1r;
                bool fast = slab_fast(rdl, rol, tl, t1l) && slab_fast(rdr, ror, tr, t1r);
                bool go_l = false, go_r = false;
                if (fast) {
                    if (t1l > 0.0f) {
                        fast = sure(tl, t1l) && sure(tl, hitT);
                        go_l = tl <= t1l && tl <= hitT;
                    }
                    if (t1r > 0.0f) {
                        fast = fast && sure(tr, t1r) && sure(tr, hitT);
                        go_r = tr <= t1r && tr <= hitT;
                    }
                    if (go_l && go_r) fast = fast && sure(tr, tl);
                }
                if (!fast) {
                    go_l = enter_unit_box(rdl, rol, hitT, &tl);
                    go_r = enter_unit_box(rdr, ror, hitT, &tr);
                }
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
                        next = kTravSentinel;
                    } else {
                        const int s = ++sp;
                        if (s <= kTravLdsSlots) lstack[(s - 1) * kTravThreads + tid] = other;
                        else deep[(size_t)(s - kTravLdsSlots - 1) * stride] = other;
                    }
                }
                cur = next;
                enter_leaf = cur < 0;
            }
        } else if (cur < 0) {
            bool bad = false;
            if (k < end) {
                uint32_t pid;
                f3 v0, v1, v2;
                bad = !tree.tri(k, one, pid, v0, v1, v2);
                if (!bad) {
                    float b1 = 0.0f, b2 = 0.0f;
                    const float t = ray_triangle(dir, org, v0, v1, v2, hitT, &b1, &b2);
                    if (t < hitT) {
                        hitT = t;
                        hit = pid;
                        hit_slot = k;
                        bx = b1;
                        by = b2;
                    }
                }
                ++k;
            }
            if (bad) {
                aborted = true;
                cur = kTravSentinel;
            } else if (k >= end) {
                cur = pop();
                enter_leaf = cur < 0;
            }
        }
        if (enter_leaf) {  // the new current node is a leaf: its primitive cursor
            if (++steps > a.max_steps || !tree.leaf(cur, k, end, one)) {
                aborted = true;
                cur = kTravSentinel;
            }
        }

        if (busy && cur == kTravSentinel) {  // ray finished: write its record, free the lane
            kr_intersection o;
            o.instance = aborted ? kTravAborted : 0u;
            o.primitive = hit;
            if (hit == 0xFFFFFFFFu) {
                o.geom_normal = kr_vec3{dir.x, dir.y, dir.z};
                o.barys = kr_vec2{0.5f, 0.5f};
            } else {
                f3 va, vb, vc;
                tree.hit_vertices(hit, hit_slot, va, vb, vc);
                const f3 nrm = vm::normalize(vm::cross(sub3(vb, va), sub3(vc, va)));
                o.geom_normal = kr_vec3{nrm.x, nrm.y, nrm.z};
                o.barys = kr_vec2{bx, by};
            }
            a.out[ray] = o;
            busy = false;
        }
    }
}

// ----------------------------------------------------- record packing --
//
// One lane per node.  An internal node writes its record: the children's
// 3x4 matrices and codes (a leaf child: ~(slot << 4 | count - 1)); a leaf
// writes its triangles' records into its slots.  Slots: a leaf node v >= I
// (I = N - 1 internal nodes) owns slot N + (v - I); a collapsed leaf
// (v < I, 2..16 primitives) owns its slice of `prims`, slots
// [primitivesOffset, primitivesOffset + count).  A leaf that cannot be
// encoded (more than 16 primitives) gets an out-of-range code, a primitive
// id out of range gets 0xFFFFFFFF: the traversal stops such rays.

__device__ __forceinline__ int32_t rec_child(const kr_bvh_node* nodes, uint32_t c, uint32_t I, uint32_t N) {
    const kr_bvh_node n = nodes[c];
    if (n.nPrimitives == 0) return (int32_t)c;
    const uint32_t cnt = n.nPrimitives;
    if (cnt > 16 || (c < I && cnt < 2)) return 0x7FFFFFFF;
    const uint32_t slot = c >= I ? N + (c - I) : n.u.primitivesOffset;
    if (slot + cnt > 2 * N) return 0x7FFFFFFF;
    return ~(int32_t)((slot << 4) | (cnt - 1));
}

__global__ __launch_bounds__(256) void rec_pack(const kr_bvh_node* __restrict__ nodes, const kr_mat4* __restrict__ mats,
                                                const uint32_t* __restrict__ prims, const kr_vec3* __restrict__ V,
                                                const kr_uvec4* __restrict__ F, uint32_t N, uint32_t vertex_count,
                                                float4* __restrict__ rec, float4* __restrict__ tri) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t I = N - 1, total = 2 * N - 1;
    if (v >= total) return;
    const kr_bvh_node n = nodes[v];
    if (n.nPrimitives == kCollapsedNode) return;
    if (n.nPrimitives == 0) {
        if (v >= I) return;
        const uint32_t a = n.u.children.left, b = n.u.children.right;
        float4 l[3] = {}, r[3] = {};
        int32_t ca = 0x7FFFFFFF, cb = 0x7FFFFFFF;
        if (a < total && b < total) {
            m43_from_mat4(&mats[a], l);
            m43_from_mat4(&mats[b], r);
            ca = rec_child(nodes, a, I, N);
            cb = rec_child(nodes, b, I, N);
        }
        float4* p = rec + 8 * (size_t)v;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            p[j] = l[j];
            p[3 + j] = r[j];
        }
        p[6] = make_float4(__int_as_float(ca), __int_as_float(cb), 0.0f, 0.0f);
        p[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const uint32_t cnt = n.nPrimitives;
    if (cnt > 16 || (v < I && cnt < 2)) return;
    const uint32_t slot = v >= I ? N + (v - I) : n.u.primitivesOffset;
    if (slot + cnt > 2 * N) return;
    for (uint32_t j = 0; j < cnt; ++j) {
        uint32_t pid = cnt == 1 ? n.u.primitivesOffset : (prims ? prims[n.u.primitivesOffset + j] : 0xFFFFFFFFu);
        f3 v0{0, 0, 0}, v1 = v0, v2 = v0;
        if (pid < N) {
            const kr_uvec4 f = F[pid];
            if (f.x < vertex_count && f.y < vertex_count && f.z < vertex_count) {
                v0 = vload(V, f.x);
                v1 = vload(V, f.y);
                v2 = vload(V, f.z);
            } else {
                pid = 0xFFFFFFFFu;
            }
        } else {
            pid = 0xFFFFFFFFu;
        }
        float4* p = tri + 3 * (size_t)(slot + j);
        p[0] = make_float4(v0.x, v0.y, v0.z, v1.x);
        p[1] = make_float4(v1.y, v1.z, v2.x, v2.y);
        p[2] = make_float4(v2.z, __uint_as_float(pid), 0.0f, 0.0f);
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
    if (ray_count > kTravMaxRays) return "obvh_soa_intersect: too many rays";
    TravWs ws;
    if (trav_carve(&ws, workspace, ws_bytes, ray_count) > ws_bytes || !workspace)
        return "obvh_soa_intersect: workspace too small";
    SoaTree tree{reinterpret_cast<const float4*>(left_T),
                 reinterpret_cast<const float4*>(right_T),
                 reinterpret_cast<const int4*>(meta),
                 V,
                 F,
                 prims,
                 counts.internal_count,
                 counts.node_count,
                 counts.face_count,
                 counts.vertex_count};
    TraceIO io{rays, out, ray_count, counts.max_steps, ws.counter, ws.deep};
    hipLaunchKernelGGL(trav_clear, dim3(1), dim3(64), 0, stream, ws.counter);
    hipLaunchKernelGGL(trace<SoaTree>, dim3(ws.blocks), dim3(kTravThreads), 0, stream, tree, io);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error rec_pack_launch(const kr_bvh_node* nodes, const kr_mat4* mats, const uint32_t* prims, const kr_vec3* V,
                         const kr_uvec4* F, uint32_t face_count, uint32_t vertex_count, obvh_node_record* rec,
                         obvh_tri_record* tri, hipStream_t stream) {
    if (face_count < 2) return "obvh_records_pack: the tree needs at least two primitives";
    if (face_count > (1u << 26)) return "obvh_records_pack: more than 2^26 primitives";
    if (!aligned16(rec) || !aligned16(tri) || !aligned16(mats))
        return "obvh_records_pack: records and transforms must be 16-byte aligned";
    const uint32_t total = 2 * face_count - 1;
    hipLaunchKernelGGL(rec_pack, dim3((total + 255) / 256), dim3(256), 0, stream, nodes, mats, prims, V, F, face_count,
                       vertex_count, reinterpret_cast<float4*>(rec), reinterpret_cast<float4*>(tri));
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

kr_error rec_intersect_launch(const obvh_node_record* rec, const obvh_tri_record* tri, uint32_t face_count,
                              const kr_ray* rays, kr_intersection* out, uint32_t ray_count, void* workspace,
                              size_t ws_bytes, hipStream_t stream) {
    if (ray_count == 0) return nullptr;
    if (face_count < 2 || face_count > (1u << 26)) return "obvh_records_intersect: face_count outside [2, 2^26]";
    if (!aligned16(rec) || !aligned16(tri)) return "obvh_records_intersect: records must be 16-byte aligned";
    if (ray_count > kTravMaxRays) return "obvh_records_intersect: too many rays";
    TravWs ws;
    if (trav_carve(&ws, workspace, ws_bytes, ray_count) > ws_bytes || !workspace)
        return "obvh_records_intersect: workspace too small";
    RecTree tree{reinterpret_cast<const float4*>(rec), reinterpret_cast<const float4*>(tri), face_count - 1,
                 2 * face_count, face_count};
    TraceIO io{rays, out, ray_count, 2 * face_count - 1, ws.counter, ws.deep};
    hipLaunchKernelGGL(trav_clear, dim3(1), dim3(64), 0, stream, ws.counter);
    hipLaunchKernelGGL(trace<RecTree>, dim3(ws.blocks), dim3(kTravThreads), 0, stream, tree, io);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
