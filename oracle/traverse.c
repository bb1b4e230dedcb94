/*
 * traverse.c -- CPU restatement of the reference's OBVH packing and of its
 * closest-hit ray traversal over the packed OBB tree.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * oracle_obvh_pack follows obvh_persistent_prepare followed by
 * kr_cuda_soa_obvh (src/ads/common/cuda/transforms.cu:1298-1341,1366-1449):
 *   every node that is not marked 0xFFFF gets metadata (x, y, z, w):
 *     internal: x = left, y = right, each replaced by its bitwise NOT when
 *               that child has nPrimitives > 0 (a leaf); z = axis;
 *     leaf:     x = primitivesOffset, y = primitivesOffset + nPrimitives,
 *               z = axis;
 *   every internal node i gets left_T[i] / right_T[i] = kr_m43from4 of its
 *   children's transforms (vecmath.h:895-902: columns 0..2 of the upper
 *   3x4 block, i.e. v[0..2], v[4..6], v[8..10], v[12..14]).
 * The reference leaves w, the metadata of 0xFFFF nodes and the matrices of
 * non-internal slots unwritten; here they are zero.
 *
 * oracle_obvh_intersect follows soa_obvh_persistent_intersect
 * (src/ads/common/cuda/intersectors.cu:801-972).  Each ray is a depth-first
 * walk from node 0 with a 64-slot stack whose slot 0 holds the sentinel
 * 0x76543210.  At an internal node both children are tested with
 * kr_ray_unit_aabb_intersect (vecmath.h:1719-1749) in the child's frame
 * (direction by the 3x3 part, origin by the full 3x4 matrix,
 * vecmath.h:1262-1309); a child is entered when t1 >= t0, t1 >= 0 and
 * t0 <= hitT.  One child entered: go there.  None: pop.  Both: the nearer
 * by t0 is entered (the left one on ties) and the other is pushed.  A leaf
 * (negative index) tests its primitives in slice order with
 * kr_ray_triangle_intersect (vecmath.h:1656-1676), keeping a hit when its t
 * is strictly below hitT, then pops.  The reference's two-loop structure
 * (postponed leaf) performs the same node/leaf sequence per ray.
 * A one-primitive leaf's primitive id is its primitivesOffset; larger
 * leaves read primitives[].  Output per ray: primitive (0xFFFFFFFF = miss),
 * barys, geom_normal = normalize(cross(vb - va, vc - va)) for a hit or the
 * ray direction for a miss, instance = 0.  A ray that would push a 64th
 * entry (the reference writes past its array there) stops where it is,
 * keeps its hit so far and gets instance = 0xFFFFFFFF.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define SENTINEL 0x76543210
#define STACK_SLOTS 64
#define COLLAPSED 0xFFFFu

int oracle_obvh_pack(const onode* nodes, const omat4* mats, uint32_t node_count, uint32_t internal_count,
                     float* left_T, float* right_T, int32_t* meta) {
    memset(meta, 0, sizeof(int32_t) * 4 * (size_t)node_count);
    memset(left_T, 0, sizeof(float) * 12 * (size_t)internal_count);
    memset(right_T, 0, sizeof(float) * 12 * (size_t)internal_count);
    for (uint32_t i = 0; i < node_count; ++i) {
        const onode* n = &nodes[i];
        if (n->nprim == COLLAPSED) continue;
        int32_t* m = meta + 4 * (size_t)i;
        m[2] = n->axis;
        if (n->nprim == 0) {
            if (i >= internal_count || n->a >= node_count || n->b >= node_count) return -1;
            const onode* l = &nodes[n->a];
            const onode* r = &nodes[n->b];
            m[0] = l->nprim > 0 ? ~(int32_t)n->a : (int32_t)n->a;
            m[1] = r->nprim > 0 ? ~(int32_t)n->b : (int32_t)n->b;
            const float* ml = mats[n->a].v;
            const float* mr = mats[n->b].v;
            for (int c = 0; c < 4; ++c)
                for (int k = 0; k < 3; ++k) {
                    left_T[12 * (size_t)i + 3 * c + k] = ml[4 * c + k];
                    right_T[12 * (size_t)i + 3 * c + k] = mr[4 * c + k];
                }
        } else {
            m[0] = (int32_t)n->a;
            m[1] = (int32_t)(n->a + n->nprim);
        }
    }
    return 0;
}

/* kr_n43transform3p: x' = (m0 dx + m3 dy) + m6 dz, ... (no translation). */
static ov3 xform_dir(const float* m, ov3 d) {
    ov3 o;
    o.x = m[0] * d.x + m[3] * d.y + m[6] * d.z;
    o.y = m[1] * d.x + m[4] * d.y + m[7] * d.z;
    o.z = m[2] * d.x + m[5] * d.y + m[8] * d.z;
    return o;
}

/* kr_v43transform3p: x' = ((m0 px + m3 py) + m6 pz) + m9, ... */
static ov3 xform_pt(const float* m, ov3 p) {
    ov3 o;
    o.x = m[0] * p.x + m[3] * p.y + m[6] * p.z + m[9];
    o.y = m[1] * p.x + m[4] * p.y + m[7] * p.z + m[10];
    o.z = m[2] * p.x + m[5] * p.y + m[8] * p.z + m[11];
    return o;
}

static float kmin(float a, float b) { return (a < b) ? a : b; }
static float kmax(float a, float b) { return (a < b) ? b : a; }

/* Slab test against the cube [-0.5, 0.5]^3; (1, -1) on a miss. */
static void unit_box(ov3 rd, ov3 ro, float out[2]) {
    const float lo_x = (-0.5f - ro.x) / rd.x, lo_y = (-0.5f - ro.y) / rd.y, lo_z = (-0.5f - ro.z) / rd.z;
    const float hi_x = (0.5f - ro.x) / rd.x, hi_y = (0.5f - ro.y) / rd.y, hi_z = (0.5f - ro.z) / rd.z;
    const float near_x = kmin(lo_x, hi_x), near_y = kmin(lo_y, hi_y), near_z = kmin(lo_z, hi_z);
    const float far_x = kmax(lo_x, hi_x), far_y = kmax(lo_y, hi_y), far_z = kmax(lo_z, hi_z);
    const float t0 = kmax(kmax(near_x, near_y), near_z);
    const float t1 = kmin(kmin(far_x, far_y), far_z);
    if (!(t0 <= t1 && t1 > 0.0f)) {
        out[0] = 1.0f;
        out[1] = -1.0f;
        return;
    }
    out[0] = t0;
    out[1] = t1;
}

static ov3 vsub(ov3 a, ov3 b) {
    ov3 o = {a.x - b.x, a.y - b.y, a.z - b.z};
    return o;
}
static ov3 vcross(ov3 a, ov3 b) {
    ov3 o = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return o;
}
static float vdot(ov3 a, ov3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* Moeller-Trumbore as written in the reference: (b1, b2, t), or (0, 0, tmax)
 * when rejected. */
static void ray_triangle(ov3 dir, ov3 org, ov3 v1, ov3 v2, ov3 v3, float tmax, float out[3]) {
    const ov3 e1 = vsub(v2, v1), e2 = vsub(v3, v1);
    const ov3 s1 = vcross(dir, e2);
    const float invd = 1.0f / vdot(s1, e1);
    const ov3 d = vsub(org, v1);
    const float b1 = vdot(d, s1) * invd;
    const ov3 s2 = vcross(d, e1);
    const float b2 = vdot(dir, s2) * invd;
    const float t = vdot(e2, s2) * invd;
    if (b1 < 0.0f || b1 > 1.0f || b2 < 0.0f || b1 + b2 > 1.0f || t < 0.0f || t > tmax) {
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = tmax;
        return;
    }
    out[0] = b1;
    out[1] = b2;
    out[2] = t;
}

void oracle_ray_unit_box(const float* rd, const float* ro, float* out2) {
    ov3 d = {rd[0], rd[1], rd[2]}, o = {ro[0], ro[1], ro[2]};
    unit_box(d, o, out2);
}

void oracle_ray_triangle(const float* dir, const float* org, const float* v1, const float* v2, const float* v3,
                         float tmax, float* out3) {
    ov3 d = {dir[0], dir[1], dir[2]}, o = {org[0], org[1], org[2]};
    ov3 a = {v1[0], v1[1], v1[2]}, b = {v2[0], v2[1], v2[2]}, c = {v3[0], v3[1], v3[2]};
    ray_triangle(d, o, a, b, c, tmax, out3);
}

static ov3 vert(const float* V, uint32_t i) {
    ov3 p = {V[3 * (size_t)i], V[3 * (size_t)i + 1], V[3 * (size_t)i + 2]};
    return p;
}

typedef struct {
    const float *left_T, *right_T;
    const int32_t* meta;
    const float* V;
    const uint32_t *F, *prims;
    const oray* rays;
    oisect* out;
    uint32_t* stats;
} trace_ctx;

/* One ray (the body of oracle_obvh_intersect). */
static void trace_one(void* p, uint32_t r, int tid) {
    const trace_ctx* c = (const trace_ctx*)p;
    const float *left_T = c->left_T, *right_T = c->right_T, *V = c->V;
    const int32_t* meta = c->meta;
    const uint32_t *F = c->F, *prims = c->prims;
    const oray* rays = c->rays;
    oisect* out = c->out;
    uint32_t* stats = c->stats;
    (void)tid;
    {
        const ov3 org = rays[r].origin, dir = rays[r].direction;
        float hitT = rays[r].tmax;
        uint32_t hit = 0xFFFFFFFFu;
        float bx = 0.0f, by = 0.0f;
        int32_t stack[STACK_SLOTS];
        int sp = 0;
        stack[0] = SENTINEL;
        uint32_t visits = 0, tests = 0, deepest = 0, overflow = 0;
        int32_t cur = 0;
        while (cur != SENTINEL) {
            if (cur >= 0) {
                const float* L = left_T + 12 * (size_t)cur;
                const float* R = right_T + 12 * (size_t)cur;
                float li[2], ri[2];
                unit_box(xform_dir(L, dir), xform_pt(L, org), li);
                unit_box(xform_dir(R, dir), xform_pt(R, org), ri);
                ++visits;
                const int go_l = (li[1] >= li[0]) && (li[1] >= 0.0f) && (li[0] <= hitT);
                const int go_r = (ri[1] >= ri[0]) && (ri[1] >= 0.0f) && (ri[0] <= hitT);
                int32_t near_c = meta[4 * (size_t)cur], far_c = meta[4 * (size_t)cur + 1];
                if (go_l != go_r) {
                    cur = go_r ? far_c : near_c;
                } else if (!go_l) {
                    cur = stack[sp--];
                } else {
                    if (ri[0] < li[0]) {
                        const int32_t t = near_c;
                        near_c = far_c;
                        far_c = t;
                    }
                    if (sp + 1 >= STACK_SLOTS) {
                        overflow = 1;
                        break;
                    }
                    stack[++sp] = far_c;
                    if ((uint32_t)sp > deepest) deepest = (uint32_t)sp;
                    cur = near_c;
                }
            } else {
                const int32_t* m = meta + 4 * (size_t)(~cur);
                const int32_t beg = m[0], end = m[1], num = end - beg;
                for (int32_t k = beg; k < end; ++k) {
                    const uint32_t pid = (num == 1) ? (uint32_t)k : prims[k];
                    const uint32_t* f = F + 4 * (size_t)pid;
                    float tr[3];
                    ray_triangle(dir, org, vert(V, f[0]), vert(V, f[1]), vert(V, f[2]), hitT, tr);
                    ++tests;
                    if (tr[2] < hitT) {
                        hitT = tr[2];
                        hit = pid;
                        bx = tr[0];
                        by = tr[1];
                    }
                }
                cur = stack[sp--];
            }
        }
        oisect* o = &out[r];
        o->instance = overflow ? 0xFFFFFFFFu : 0u;
        o->primitive = hit;
        if (hit == 0xFFFFFFFFu) {
            o->bx = 0.5f;
            o->by = 0.5f;
            o->normal = dir;
        } else {
            const uint32_t* f = F + 4 * (size_t)hit;
            const ov3 va = vert(V, f[0]), vb = vert(V, f[1]), vc = vert(V, f[2]);
            const ov3 c = vcross(vsub(vb, va), vsub(vc, va));
            const float len = sqrtf(vdot(c, c));
            o->normal.x = c.x / len;
            o->normal.y = c.y / len;
            o->normal.z = c.z / len;
            o->bx = bx;
            o->by = by;
        }
        if (stats) {
            stats[3 * (size_t)r] = visits;
            stats[3 * (size_t)r + 1] = tests;
            stats[3 * (size_t)r + 2] = overflow ? 0xFFFFFFFFu : deepest;
        }
    }
}

int oracle_obvh_intersect(const float* left_T, const float* right_T, const int32_t* meta, const float* V,
                          const uint32_t* F, const uint32_t* prims, const oray* rays, uint32_t n, oisect* out,
                          uint32_t* stats) {
    return oracle_obvh_intersect_mt(left_T, right_T, meta, V, F, prims, rays, n, out, stats, 1);
}

int oracle_obvh_intersect_mt(const float* left_T, const float* right_T, const int32_t* meta, const float* V,
                             const uint32_t* F, const uint32_t* prims, const oray* rays, uint32_t n, oisect* out,
                             uint32_t* stats, int threads) {
    trace_ctx c = {left_T, right_T, meta, V, F, prims, rays, out, stats};
    oracle_parallel_for(n, 64, threads, trace_one, &c);
    int status = 0;
    for (uint32_t r = 0; r < n; ++r)
        if (out[r].instance == 0xFFFFFFFFu) status = -1;
    return status;
}
