/*
 * transform.c -- CPU restatement of the reference's parallel AABB-BVH ->
 * OBB-tree transform (src/ads/common/cuda/transforms.cu:213-964), run
 * sequentially or node-parallel on host threads (same results).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP transform).
 *
 * Semantics (each function names the reference lines it follows):
 *  K1  per node 7-slab k-DOP with extremal vertices.  Leaf: its 3 vertices
 *      in face order, strict comparisons.  Internal: merge(left, right),
 *      the right child's extreme is taken only if strictly better, so ties
 *      go to the left child (the reference keeps whichever child arrived
 *      second, transforms.cu:385,391 -- a scheduling-dependent choice that
 *      this build fixes to left-first so indices are reproducible).
 *  K2  per internal node DiTO candidate search over the 14 points in
 *      interleaved order [min0, max0, ..., min6, max6].
 *  K3  leaf OBB from the triangle frame; exact projection ranges of every
 *      subtree vertex on the node's candidate axes.
 *  K4  OBB-vs-AABB selection, extent floor 0.001, inverse(T*R*S).
 * vecmath semantics: kr_min(a,b) = a<b?a:b, kr_max(a,b) = a<b?b:a,
 * normalize divides by the length (vecmath.h:548-550,772-774,1396-1397).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define INVALID 0xFFFFFFFFu
static const float kEps = 0.000001f; /* transforms.cu:554,613 */

static ov3 v3(float x, float y, float z) { ov3 r = {x, y, z}; return r; }
static ov3 vsub(ov3 a, ov3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }        /* vecmath.h:804 */
static ov3 vadd(ov3 a, ov3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }        /* :780 */
static ov3 vmul1(ov3 a, float b) { return v3(a.x * b, a.y * b, a.z * b); }             /* :1347 */
static float vdot(ov3 a, ov3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }          /* :748 */
static ov3 vcross(ov3 a, ov3 b) {                                                       /* :812 */
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static float vlen2(ov3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }                /* :536 */
static ov3 vnormalize(ov3 a) {                                                          /* :548,772 */
    float l = sqrtf(vlen2(a));
    return v3(a.x / l, a.y / l, a.z / l);
}
static float vdist2(ov3 a, ov3 b) { return vlen2(vsub(b, a)); }                        /* :520 */
static float dist_inf_edge(ov3 q, ov3 p0, ov3 v) {                                      /* :524-530 */
    ov3 u0 = vsub(q, p0);
    float t = vdot(v, u0);
    float l2 = vlen2(v);
    return vlen2(u0) - t * t / l2;
}
static float kmin(float a, float b) { return (a < b) ? a : b; }                        /* :1396 */
static float kmax(float a, float b) { return (a < b) ? b : a; }                        /* :1397 */
static ov3 vmin3(ov3 a, ov3 b) { return v3(kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)); }
static ov3 vmax3(ov3 a, ov3 b) { return v3(kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)); }
static float quality(ov3 l) { return l.x * l.y + l.x * l.z + l.y * l.z; }             /* transforms.cu:414 */

/* ---- 4x4 matrices (column major), vecmath.h:816-834,946-1085,1214-1234,1286-1296 ---- */

static omat4 mscale(ov3 s) {
    omat4 m = {{s.x, 0, 0, 0, 0, s.y, 0, 0, 0, 0, s.z, 0, 0, 0, 0, 1}};
    return m;
}
static omat4 mtranslate(ov3 t) {
    omat4 m = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, t.x, t.y, t.z, 1}};
    return m;
}
static omat4 mrows(ov3 x, ov3 y, ov3 z) { /* kr_mobb3 -> kr_mrows3 with w = 0 */
    omat4 m = {{x.x, x.y, x.z, 0, y.x, y.y, y.z, 0, z.x, z.y, z.z, 0, 0, 0, 0, 1}};
    return m;
}

void oracle_mmul4(const omat4* pl, const omat4* pr, omat4* out) {
    const float* l = pl->v;
    const float* r = pr->v;
    omat4 o;
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i)
            o.v[4 * c + i] = l[i] * r[4 * c] + l[4 + i] * r[4 * c + 1] +
                             l[8 + i] * r[4 * c + 2] + l[12 + i] * r[4 * c + 3];
    *out = o;
}

/* kr_minverse4, vecmath.h:946-1075 (cofactor expansion, identity if det == 0). */
void oracle_minverse4(const omat4* pm, omat4* out) {
    const float* m = pm->v;
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
        omat4 id = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
        *out = id;
        return;
    }
    det = 1.0f / det;
    for (int i = 0; i < 16; ++i) out->v[i] = r[i] * det;
}

/* kr_vtransform3 = kr_mvmul43, vecmath.h:1286,1270-1276. */
static ov3 vtransform(const omat4* pm, ov3 r) {
    const float* l = pm->v;
    return v3(l[0] * r.x + l[4] * r.y + l[8] * r.z + l[12],
              l[1] * r.x + l[5] * r.y + l[9] * r.z + l[13],
              l[2] * r.x + l[6] * r.y + l[10] * r.z + l[14]);
}

/* inverse(T(mid) * (R(obb) * S(2 ext))), transforms.cu:832-834,939-942. */
static void obb_matrix(const oobb* o, omat4* out) {
    omat4 r = mrows(o->v0, o->v1, o->v2);
    omat4 s = mscale(v3(o->ext.x * 2.0f, o->ext.y * 2.0f, o->ext.z * 2.0f));
    omat4 t = mtranslate(o->mid);
    omat4 rs, trs;
    oracle_mmul4(&r, &s, &rs);
    oracle_mmul4(&t, &rs, &trs);
    oracle_minverse4(&trs, out);
}

/* ---- K1: k-DOP ---- */

typedef struct {
    float mn[7], mx[7];
    uint32_t imn[7], imx[7];
} kdop;

static ov3 vert(const float* V, uint32_t i) { return v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }

static void slab_proj(ov3 p, float pr[7]) { /* transforms.cu:224-251 */
    pr[0] = p.x;
    pr[1] = p.y;
    pr[2] = p.z;
    pr[3] = p.x + p.y + p.z;
    pr[4] = p.x + p.y - p.z;
    pr[5] = p.x - p.y + p.z;
    pr[6] = p.x - p.y - p.z;
}

/* transforms.cu:216-285 over (v0, v1, v2) of one face. */
static void leaf_kdop(const float* V, const uint32_t* face, kdop* k) {
    float pr[7];
    slab_proj(vert(V, face[0]), pr);
    for (int s = 0; s < 7; ++s) { k->mn[s] = k->mx[s] = pr[s]; k->imn[s] = k->imx[s] = face[0]; }
    for (int j = 1; j < 3; ++j) {
        slab_proj(vert(V, face[j]), pr);
        for (int s = 0; s < 7; ++s) {
            if (pr[s] < k->mn[s]) { k->mn[s] = pr[s]; k->imn[s] = face[j]; }
            if (pr[s] > k->mx[s]) { k->mx[s] = pr[s]; k->imx[s] = face[j]; }
        }
    }
}

/* transforms.cu:383-395 with ties to the left child. */
static void merge_kdop(const kdop* l, const kdop* r, kdop* o) {
    *o = *l;
    for (int s = 0; s < 7; ++s) {
        if (r->mn[s] < o->mn[s]) { o->mn[s] = r->mn[s]; o->imn[s] = r->imn[s]; }
        if (r->mx[s] > o->mx[s]) { o->mx[s] = r->mx[s]; o->imx[s] = r->imx[s]; }
    }
}

/* ---- K2: candidate search, transforms.cu:413-756 ---- */

typedef struct {
    ov3 b0, b1, b2;
    ov3 rmin, rmax;   /* candidate ranges over the 14 points (min_proj/max_proj[3..5]) */
    int degenerate;   /* 1: identity axes, the node's result is its AABB */
    int best_id;      /* winning candidate (DiTO's numbering: 0..2 base triangle,
                         3..11 upper, 12..20 lower tetrahedron), -1: the AABB's axes */
} candidate;

static void ext_proj(const ov3* P, ov3 d, float* mn, float* mx) { /* :418-435 */
    float p = vdot(P[0], d);
    float lo = p, hi = p;
    for (int i = 1; i < 14; ++i) {
        p = vdot(P[i], d);
        lo = kmin(lo, p);
        hi = kmax(hi, p);
    }
    *mn = lo; *mx = hi;
}

/* transforms.cu:484-544 */
static void from_normal_edges(const ov3* P, ov3 n, ov3 e0, ov3 e1, ov3 e2, candidate* c,
                              float* best, int id0) {
    ov3 e[3] = {e0, e1, e2};
    ov3 m[3] = {vcross(e0, n), vcross(e1, n), vcross(e2, n)};
    ov3 dmin, dmax;
    ext_proj(P, n, &dmin.y, &dmax.y);
    float leny = dmax.y - dmin.y;
    for (int k = 0; k < 3; ++k) {
        ext_proj(P, e[k], &dmin.x, &dmax.x);
        float lenx = dmax.x - dmin.x;
        ext_proj(P, m[k], &dmin.z, &dmax.z);
        float lenz = dmax.z - dmin.z;
        float q = quality(v3(lenx, leny, lenz));
        if (q < *best) {
            *best = q;
            c->rmax = dmax;
            c->rmin = dmin;
            c->b0 = e[k]; c->b1 = n; c->b2 = m[k];
            c->best_id = id0 + k;
        }
    }
}

static void candidate_search(const float* V, const kdop* k, candidate* c) {
    ov3 P[14]; /* interleaved [min0, max0, min1, max1, ...], :693-696,711 */
    for (int s = 0; s < 7; ++s) { P[2 * s] = vert(V, k->imn[s]); P[2 * s + 1] = vert(V, k->imx[s]); }
    ov3 alen = v3(k->mx[0] - k->mn[0], k->mx[1] - k->mn[1], k->mx[2] - k->mn[2]);
    float best = quality(alen);
    c->b0 = v3(1, 0, 0); c->b1 = v3(0, 1, 0); c->b2 = v3(0, 0, 1);
    c->rmin = v3(k->mn[0], k->mn[1], k->mn[2]);   /* :706-707 */
    c->rmax = v3(k->mx[0], k->mx[1], k->mx[2]);
    c->degenerate = 0;
    c->best_id = -1;

    /* Base triangle, :439-455,546-579. */
    int at = 0;
    float d2best = vdist2(P[0], P[1]);
    for (int s = 1; s < 7; ++s) {
        float d2 = vdist2(P[2 * s], P[2 * s + 1]);
        if (d2 > d2best) { d2best = d2; at = s; }
    }
    ov3 p0 = P[2 * at], p1 = P[2 * at + 1];
    if (vdist2(p0, p1) < kEps) { c->degenerate = 1; return; }
    ov3 e0 = vnormalize(vsub(p0, p1));
    int ip2 = 0;
    float dm = dist_inf_edge(P[0], p0, e0);
    for (int i = 1; i < 14; ++i) {
        float d = dist_inf_edge(P[i], p0, e0);
        if (d > dm) { dm = d; ip2 = i; }
    }
    if (dm < kEps) { c->degenerate = 1; return; }
    ov3 p2 = P[ip2];
    ov3 e1 = vnormalize(vsub(p1, p2));
    ov3 e2 = vnormalize(vsub(p2, p0));
    ov3 n = vnormalize(vcross(e1, e0));
    from_normal_edges(P, n, e0, e1, e2, c, &best, 0);

    /* Ditetrahedron, :582-674. */
    float pr = vdot(P[0], n);
    float lo = pr, hi = pr;
    int ilo = 0, ihi = 0;
    for (int i = 1; i < 14; ++i) {
        pr = vdot(P[i], n);
        if (pr < lo) { lo = pr; ilo = i; }
        if (pr > hi) { hi = pr; ihi = i; }
    }
    float tri = vdot(p0, n);
    int q0v = (hi - kEps > tri);
    int q1v = (lo + kEps < tri);
    if (!q0v && !q1v) { c->degenerate = 1; return; } /* :646-647,737-743 */
    if (q0v) {
        ov3 q0 = P[ihi];
        ov3 f0 = vnormalize(vsub(q0, p0));
        ov3 f1 = vnormalize(vsub(q0, p1));
        ov3 f2 = vnormalize(vsub(q0, p2));
        ov3 n0 = vnormalize(vcross(f1, e0));
        ov3 n1 = vnormalize(vcross(f2, e1));
        ov3 n2 = vnormalize(vcross(f0, e2));
        from_normal_edges(P, n0, e0, f1, f0, c, &best, 3);
        from_normal_edges(P, n1, e1, f2, f1, c, &best, 6);
        from_normal_edges(P, n2, e2, f0, f2, c, &best, 9);
    }
    if (q1v) {
        ov3 q1 = P[ilo];
        ov3 g0 = vnormalize(vsub(q1, p0));
        ov3 g1 = vnormalize(vsub(q1, p1));
        ov3 g2 = vnormalize(vsub(q1, p2));
        ov3 m0 = vnormalize(vcross(g1, e0));
        ov3 m1 = vnormalize(vcross(g2, e1));
        ov3 m2 = vnormalize(vcross(g0, e2));
        from_normal_edges(P, m0, e0, g1, g0, c, &best, 12);
        from_normal_edges(P, m1, e1, g2, g1, c, &best, 15);
        from_normal_edges(P, m2, e2, g0, g2, c, &best, 18);
    }
}

/* ---- K3 leaf part, transforms.cu:803-834 ---- */

static void leaf_obb(ov3 v0, ov3 v1, ov3 v2, oobb* o, omat4* M) {
    ov3 e0 = vsub(v1, v0);
    ov3 e1 = vsub(v2, v0);
    ov3 a2 = vcross(e0, e1);
    ov3 a0 = vnormalize(e0);
    ov3 a1 = vnormalize(vcross(a0, a2));
    o->v0 = a0; o->v1 = a1; o->v2 = a2;
    omat4 r = mrows(a0, a1, a2);
    omat4 ri;
    oracle_minverse4(&r, &ri);
    ov3 l0 = vtransform(&ri, v0), l1 = vtransform(&ri, v1), l2 = vtransform(&ri, v2);
    ov3 mx = vmax3(l0, vmax3(l1, l2));
    ov3 mn = vmin3(l0, vmin3(l1, l2));
    ov3 mid = vmul1(vadd(mx, mn), 0.5f);
    o->ext = vmax3(vmul1(vsub(mx, mn), 0.5f), v3(0.001f, 0.001f, 0.001f));
    o->mid = vmul1(o->v0, mid.x);
    o->mid = vadd(o->mid, vmul1(o->v1, mid.y));
    o->mid = vadd(o->mid, vmul1(o->v2, mid.z));
    obb_matrix(o, M);
}

/* ---- K4, transforms.cu:905-964 ---- */

static void finalize_node(const kdop* k, const candidate* c, ov3 rmin, ov3 rmax, oobb* o,
                          omat4* M) {
    ov3 alen = v3(k->mx[0] - k->mn[0], k->mx[1] - k->mn[1], k->mx[2] - k->mn[2]);
    ov3 olen = vsub(rmax, rmin);
    ov3 mn, mx;
    if (!c->degenerate && quality(olen) < quality(alen)) {
        o->v0 = c->b0; o->v1 = c->b1; o->v2 = c->b2;
        mn = rmin; mx = rmax;
    } else {
        o->v0 = v3(1, 0, 0); o->v1 = v3(0, 1, 0); o->v2 = v3(0, 0, 1);
        mn = v3(k->mn[0], k->mn[1], k->mn[2]);
        mx = v3(k->mx[0], k->mx[1], k->mx[2]);
    }
    ov3 mid = vmul1(vadd(mn, mx), 0.5f);
    o->ext = vmax3(vmul1(vsub(mx, mn), 0.5f), v3(0.001f, 0.001f, 0.001f));
    o->mid = vmul1(o->v0, mid.x);
    o->mid = vadd(o->mid, vmul1(o->v1, mid.y));
    o->mid = vadd(o->mid, vmul1(o->v2, mid.z));
    obb_matrix(o, M);
}

/* ---- driver ---- */

static int check_tree(const onode* nodes, const uint32_t* parents, uint32_t N) {
    uint32_t I = N - 1, total = 2 * N - 1;
    if (parents[0] != INVALID) return -1;
    for (uint32_t i = 0; i < I; ++i) {
        if (nodes[i].nprim != 0) return -1;
        uint32_t l = nodes[i].a, r = nodes[i].b;
        if (l >= total || r >= total || l == 0 || r == 0 || l == r) return -1;
        if (parents[l] != i || parents[r] != i) return -1;
    }
    for (uint32_t i = I; i < total; ++i)
        if (nodes[i].nprim == 0) return -1;
    return 0;
}

typedef struct {
    const onode* nodes;
    const float* V;
    const uint32_t* faces;
    uint32_t I;
    kdop* K;
    candidate* C;
    oobb* obbs;
    omat4* mats;
    float* refit_ranges;
    float* cand_axes;
    uint32_t* stacks; /* one DFS stack of `stack_len` entries per thread */
    uint32_t stack_len;
} tree_ctx;

static void candidates_one(void* p, uint32_t v, int tid) {
    const tree_ctx* t = (const tree_ctx*)p;
    (void)tid;
    candidate_search(t->V, &t->K[v], &t->C[v]);
}

/* Refit: exact ranges over every subtree vertex (transforms.cu:836-900),
 * computed per node over its own subtree; then the node's box (K4).
 * As in the reference kernel, each triangle first reduces its three
 * projections with kr_min / kr_max from vertex 0 (:855-885), and the result
 * enters the node's range only when strictly beyond it (:887-897, the
 * compare before the float CAS loop of atomics.h:9-31).  For finite input
 * that is the plain min / max over the vertices (up to the sign of a zero);
 * with a NaN projection it fixes which values survive: a NaN in the running
 * range of a triangle is replaced by the next value, a NaN triangle result
 * never enters the node's range. */
static void refit_one(void* p, uint32_t v, int tid) {
    const tree_ctx* t = (const tree_ctx*)p;
    const onode* nodes = t->nodes;
    const uint32_t I = t->I;
    const candidate* c = &t->C[v];
    uint32_t* stack = t->stacks + (size_t)tid * t->stack_len;
    ov3 rmin = c->rmin, rmax = c->rmax;
    if (!c->degenerate) {
        uint32_t sp = 0;
        stack[sp++] = v;
        while (sp) {
            uint32_t u = stack[--sp];
            if (u < I) { stack[sp++] = nodes[u].a; stack[sp++] = nodes[u].b; continue; }
            const uint32_t* f = &t->faces[4 * nodes[u].a];
            ov3 q = vert(t->V, f[0]);
            ov3 tmin = v3(vdot(q, c->b0), vdot(q, c->b1), vdot(q, c->b2));
            ov3 tmax = tmin;
            for (int j = 1; j < 3; ++j) {
                q = vert(t->V, f[j]);
                float d0 = vdot(q, c->b0), d1 = vdot(q, c->b1), d2 = vdot(q, c->b2);
                tmin.x = kmin(tmin.x, d0); tmax.x = kmax(tmax.x, d0);
                tmin.y = kmin(tmin.y, d1); tmax.y = kmax(tmax.y, d1);
                tmin.z = kmin(tmin.z, d2); tmax.z = kmax(tmax.z, d2);
            }
            if (tmin.x < rmin.x) rmin.x = tmin.x;
            if (tmin.y < rmin.y) rmin.y = tmin.y;
            if (tmin.z < rmin.z) rmin.z = tmin.z;
            if (tmax.x > rmax.x) rmax.x = tmax.x;
            if (tmax.y > rmax.y) rmax.y = tmax.y;
            if (tmax.z > rmax.z) rmax.z = tmax.z;
        }
    }
    if (t->cand_axes) {
        float* ca = &t->cand_axes[11 * v];
        ca[0] = c->b0.x; ca[1] = c->b0.y; ca[2] = c->b0.z;
        ca[3] = c->b1.x; ca[4] = c->b1.y; ca[5] = c->b1.z;
        ca[6] = c->b2.x; ca[7] = c->b2.y; ca[8] = c->b2.z;
        ca[9] = (float)c->degenerate;
        ca[10] = (float)c->best_id;
    }
    if (t->refit_ranges) {
        float* rr = &t->refit_ranges[6 * v];
        rr[0] = rmin.x; rr[1] = rmin.y; rr[2] = rmin.z;
        rr[3] = rmax.x; rr[4] = rmax.y; rr[5] = rmax.z;
    }
    finalize_node(&t->K[v], c, rmin, rmax, &t->obbs[v], &t->mats[v]);
}

static void leaf_one(void* p, uint32_t i, int tid) {
    const tree_ctx* t = (const tree_ctx*)p;
    (void)tid;
    const uint32_t v = t->I + i;
    const uint32_t* f = &t->faces[4 * t->nodes[v].a];
    leaf_obb(vert(t->V, f[0]), vert(t->V, f[1]), vert(t->V, f[2]), &t->obbs[v], &t->mats[v]);
}

int oracle_obb_tree(const onode* nodes, const uint32_t* parents, const float* V,
                    const uint32_t* faces, uint32_t N, oobb* obbs, omat4* mats,
                    uint32_t* kdop_ids, float* refit_ranges, float* cand_axes) {
    return oracle_obb_tree_mt(nodes, parents, V, faces, N, obbs, mats, kdop_ids, refit_ranges, cand_axes, 1);
}

int oracle_obb_tree_mt(const onode* nodes, const uint32_t* parents, const float* V,
                       const uint32_t* faces, uint32_t N, oobb* obbs, omat4* mats,
                       uint32_t* kdop_ids, float* refit_ranges, float* cand_axes, int threads) {
    if (N == 0) return 0;
    uint32_t I = N - 1, total = 2 * N - 1;
    if (check_tree(nodes, parents, N) != 0) return -1;
    if (threads < 1) threads = 1;
    if (threads > OPAR_MAX_THREADS) threads = OPAR_MAX_THREADS;

    kdop* K = (kdop*)malloc(sizeof(kdop) * total);
    candidate* C = (candidate*)malloc(sizeof(candidate) * (I ? I : 1));
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * total);
    uint32_t* depth = (uint32_t*)malloc(sizeof(uint32_t) * total);
    if (!K || !C || !order || !depth) { free(K); free(C); free(order); free(depth); return -1; }

    /* Post-order of the tree (children before parents), iterative DFS; the
     * deepest node bounds the refit's DFS stacks. */
    uint32_t max_depth = 0;
    {
        uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (total + 1));
        uint8_t* seen = (uint8_t*)calloc(total, 1);
        uint32_t sp = 0, cnt = 0;
        stack[sp++] = 0;
        depth[0] = 0;
        while (sp) {
            uint32_t v = stack[sp - 1];
            if (v >= I || seen[v]) { order[cnt++] = v; --sp; continue; }
            seen[v] = 1;
            const uint32_t d = depth[v] + 1;
            if (d > max_depth) max_depth = d;
            depth[nodes[v].a] = d;
            depth[nodes[v].b] = d;
            stack[sp++] = nodes[v].b;
            stack[sp++] = nodes[v].a;
        }
        free(stack); free(seen);
        if (cnt != total) { free(K); free(C); free(order); free(depth); return -1; }
    }
    free(depth);

    /* K1 in post-order (sequential: each merge needs its children). */
    for (uint32_t t = 0; t < total; ++t) {
        uint32_t v = order[t];
        if (v >= I) leaf_kdop(V, &faces[4 * nodes[v].a], &K[v]);
        else merge_kdop(&K[nodes[v].a], &K[nodes[v].b], &K[v]);
    }
    free(order);

    /* A DFS from a node at depth d holds at most (max_depth - d) + 2 entries. */
    tree_ctx ctx = {nodes, V, faces, I, K, C, obbs, mats, refit_ranges, cand_axes, NULL, max_depth + 2};
    ctx.stacks = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)ctx.stack_len * (size_t)threads);
    if (!ctx.stacks) { free(K); free(C); return -1; }
    oracle_parallel_for(I, 64, threads, candidates_one, &ctx);  /* K2 */
    oracle_parallel_for(I, 16, threads, refit_one, &ctx);       /* K3 ranges + K4 */
    oracle_parallel_for(N, 1024, threads, leaf_one, &ctx);      /* K3 leaf boxes */
    free(ctx.stacks);

    if (kdop_ids) {
        for (uint32_t v = 0; v < total; ++v) {
            memcpy(&kdop_ids[14 * v], K[v].imn, 7 * sizeof(uint32_t));
            memcpy(&kdop_ids[14 * v + 7], K[v].imx, 7 * sizeof(uint32_t));
        }
    }
    free(K); free(C);
    return 0;
}
