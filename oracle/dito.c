/*
 * dito.c -- CPU restatement of DiTO-14 (Larsson & Kallberg, "Fast
 * Computation of Tight-Fitting Oriented Bounding Boxes", Game Engine Gems 2)
 * as the reference implements it in src/common/algorithm/dito/dito.cpp.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP DiTO kernels).
 *
 * Floating point: float, no contraction, evaluation order of each
 * expression as written in the reference (left to right).  Every function
 * names the reference lines it follows.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static const float kEps = 0.000001f; /* dito.cpp:298,352,464 */

static ov3 v3(float x, float y, float z) { ov3 r = {x, y, z}; return r; }
static ov3 vsub(ov3 u, ov3 v) { return v3(u.x - v.x, u.y - v.y, u.z - v.z); }   /* :69-72 */
static ov3 vadd(ov3 u, ov3 v) { return v3(u.x + v.x, u.y + v.y, u.z + v.z); }   /* :63-66 */
static ov3 vscale(float s, ov3 u) { return v3(s * u.x, s * u.y, s * u.z); }     /* :75-78 */
static float vdot(ov3 u, ov3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }  /* :81 */
/* dito.cpp:84-87 -- the y component is the negated difference. */
static ov3 vcross(ov3 u, ov3 v) {
    return v3(u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x);
}
static float vsqlen(ov3 u) { return u.x * u.x + u.y * u.y + u.z * u.z; }       /* :90 */
static float vlen(ov3 u) { return sqrtf(vsqlen(u)); }                            /* :93 */
static float sqdist(ov3 u, ov3 v) { return vsqlen(vsub(v, u)); }                 /* :96 */
/* dito.cpp:99-103: multiply by the reciprocal length. */
static ov3 vnormalize(ov3 u) {
    float inv = 1.0f / vlen(u);
    return v3(u.x * inv, u.y * inv, u.z * inv);
}
/* dito.cpp:274-278: half box area. */
static float quality(ov3 len) { return len.x * len.y + len.x * len.z + len.y * len.z; }

static ov3 load(const float* xyz, int i) { return v3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]); }

/* The 7 slab projections in the reference's order (dito.cpp:115-171). */
static void slab_proj(ov3 p, float pr[7]) {
    pr[0] = p.x;
    pr[1] = p.y;
    pr[2] = p.z;
    pr[3] = p.x + p.y + p.z;
    pr[4] = p.x + p.y - p.z;
    pr[5] = p.x - p.y + p.z;
    pr[6] = p.x - p.y - p.z;
}

/* dito.cpp:105-191: initialise with point 0, strict comparisons, so ties
 * keep the lowest index. */
static void extremal_points_7(const float* xyz, int n, float mn[7], float mx[7],
                              int imn[7], int imx[7]) {
    float pr[7];
    slab_proj(load(xyz, 0), pr);
    for (int k = 0; k < 7; ++k) { mn[k] = mx[k] = pr[k]; imn[k] = imx[k] = 0; }
    for (int i = 1; i < n; ++i) {
        slab_proj(load(xyz, i), pr);
        for (int k = 0; k < 7; ++k) {
            if (pr[k] < mn[k]) { mn[k] = pr[k]; imn[k] = i; }
            if (pr[k] > mx[k]) { mx[k] = pr[k]; imx[k] = i; }
        }
    }
}

/* dito.cpp:193-206 */
static void extremal_projs(ov3 d, const ov3* pts, int np, float* mn, float* mx) {
    float p = vdot(pts[0], d);
    float lo = p, hi = p;
    for (int i = 1; i < np; ++i) {
        p = vdot(pts[i], d);
        if (p < lo) lo = p;
        if (p > hi) hi = p;
    }
    *mn = lo; *mx = hi;
}

/* Same over the raw input (used for the final dimensions, dito.cpp:421-426). */
static void extremal_projs_all(ov3 d, const float* xyz, int n, float* mn, float* mx) {
    float p = vdot(load(xyz, 0), d);
    float lo = p, hi = p;
    for (int i = 1; i < n; ++i) {
        p = vdot(load(xyz, i), d);
        if (p < lo) lo = p;
        if (p > hi) hi = p;
    }
    *mn = lo; *mx = hi;
}

typedef struct {
    const ov3* sel;
    int np;
    float best;
    ov3 b0, b1, b2;
    int best_id;      /* candidate counter value of the winner, -1 = initial */
    int counter;      /* candidates evaluated so far */
} search_state;

/* dito.cpp:309-345: evaluate (e_k, n, m_k), k = 0..2, m_k = e_k x n. */
static void eval_triangle(search_state* s, ov3 n, ov3 e0, ov3 e1, ov3 e2) {
    ov3 e[3] = {e0, e1, e2};
    ov3 m[3] = {vcross(e0, n), vcross(e1, n), vcross(e2, n)};
    float nmin, nmax;
    extremal_projs(n, s->sel, s->np, &nmin, &nmax);
    float leny = nmax - nmin;
    for (int k = 0; k < 3; ++k) {
        float emin, emax, mmin, mmax;
        extremal_projs(e[k], s->sel, s->np, &emin, &emax);
        extremal_projs(m[k], s->sel, s->np, &mmin, &mmax);
        ov3 len = v3(emax - emin, leny, mmax - mmin);
        float q = quality(len);
        if (q < s->best) {
            s->best = q; s->b0 = e[k]; s->b1 = n; s->b2 = m[k];
            s->best_id = s->counter;
        }
        s->counter++;
    }
}

/* dito.cpp:428-440 */
static void finalize_obb(ov3 v0, ov3 v1, ov3 v2, ov3 mn, ov3 mx, ov3 len, oobb* o) {
    o->v0 = v0; o->v1 = v1; o->v2 = v2;
    o->ext = vscale(0.5f, len);
    ov3 q = vscale(0.5f, vadd(mn, mx));
    o->mid = vscale(q.x, v0);
    o->mid = vadd(o->mid, vscale(q.y, v1));
    o->mid = vadd(o->mid, vscale(q.z, v2));
}

/* dito.cpp:442-449 */
static void finalize_aabb(ov3 mid, ov3 len, oobb* o) {
    o->mid = mid;
    o->ext = vscale(0.5f, len);
    o->v0 = v3(1, 0, 0); o->v1 = v3(0, 1, 0); o->v2 = v3(0, 0, 1);
}

/* dito.cpp:451-476 */
static void finalize_line(ov3 u, const float* xyz, int n, oobb* o) {
    ov3 r = u;
    if (fabsf(u.x) > fabsf(u.y) && fabsf(u.x) > fabsf(u.z)) r.x = 0;
    else if (fabsf(u.y) > fabsf(u.z)) r.y = 0;
    else r.z = 0;
    if (vsqlen(r) < kEps) r = v3(1, 1, 1);
    ov3 v = vnormalize(vcross(u, r));
    ov3 w = vnormalize(vcross(u, v));
    ov3 mn, mx;
    extremal_projs_all(u, xyz, n, &mn.x, &mx.x);
    extremal_projs_all(v, xyz, n, &mn.y, &mx.y);
    extremal_projs_all(w, xyz, n, &mn.z, &mx.z);
    finalize_obb(u, v, w, mn, mx, vsub(mx, mn), o);
}

static void trace_reset(odito_trace* t) {
    memset(t, 0, sizeof(*t));
    t->pair_k = t->p2_sel = t->q0_sel = t->q1_sel = -1;
    t->best = -1;
}

/* DiTO_14<float>, dito.cpp:478-538. */
void oracle_dito14(const float* xyz, int n, oobb* out, odito_trace* trace) {
    odito_trace tr;
    trace_reset(&tr);
    if (n <= 0) { /* :498-502 */
        finalize_aabb(v3(0, 0, 0), v3(0, 0, 0), out);
        tr.degenerate = 3; tr.result_aabb = 1;
        if (trace) *trace = tr;
        return;
    }
    float mn[7], mx[7];
    int imn[7], imx[7];
    extremal_points_7(xyz, n, mn, mx, imn, imx);
    for (int k = 0; k < 7; ++k) { tr.slab_min_idx[k] = (uint32_t)imn[k]; tr.slab_max_idx[k] = (uint32_t)imx[k]; }

    /* :508-516 */
    ov3 al_mid = v3((mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f, (mn[2] + mx[2]) * 0.5f);
    ov3 al_len = v3(mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]);
    float al_val = quality(al_len);

    /* :519-520: selected points [min0..min6, max0..max6] or all points. */
    ov3 selbuf[14];
    ov3 minv[7], maxv[7];
    for (int k = 0; k < 7; ++k) { minv[k] = load(xyz, imn[k]); maxv[k] = load(xyz, imx[k]); }
    int np;
    if (n > 14) {
        for (int k = 0; k < 7; ++k) { selbuf[k] = minv[k]; selbuf[7 + k] = maxv[k]; }
        np = 14;
    } else {
        for (int i = 0; i < n; ++i) selbuf[i] = load(xyz, i);
        np = n;
    }
    search_state s;
    s.sel = selbuf; s.np = np; s.best = al_val;
    s.b0 = v3(1, 0, 0); s.b1 = v3(0, 1, 0); s.b2 = v3(0, 0, 1);
    s.best_id = -1; s.counter = 0;

    /* Base triangle, dito.cpp:347-380; furthest pair :280-292. */
    int kstar = 0;
    float best_d = sqdist(maxv[0], minv[0]);
    for (int k = 1; k < 7; ++k) {
        float d = sqdist(maxv[k], minv[k]);
        if (d > best_d) { best_d = d; kstar = k; }
    }
    tr.pair_k = kstar;
    ov3 p0 = minv[kstar], p1 = maxv[kstar];
    if (sqdist(p0, p1) < kEps) { /* degenerate case 1, :524 */
        finalize_aabb(al_mid, al_len, out);
        tr.degenerate = 1; tr.result_aabb = 1;
        if (trace) *trace = tr;
        return;
    }
    ov3 e0 = vnormalize(vsub(p0, p1));
    /* :230-258: furthest point from the infinite line (p0, e0). */
    int ip2 = 0;
    float dmax;
    {
        ov3 u0 = vsub(s.sel[0], p0);
        float t = vdot(e0, u0);
        dmax = vsqlen(u0) - t * t / vsqlen(e0);
        for (int i = 1; i < np; ++i) {
            u0 = vsub(s.sel[i], p0);
            t = vdot(e0, u0);
            float d = vsqlen(u0) - t * t / vsqlen(e0);
            if (d > dmax) { dmax = d; ip2 = i; }
        }
    }
    tr.p2_sel = ip2;
    if (dmax < kEps) { /* degenerate case 2, :525 */
        finalize_line(e0, xyz, n, out);
        tr.degenerate = 2; tr.result_aabb = 0;
        if (trace) *trace = tr;
        return;
    }
    ov3 p2 = s.sel[ip2];
    ov3 e1 = vnormalize(vsub(p1, p2));
    ov3 e2 = vnormalize(vsub(p2, p0));
    ov3 nrm = vnormalize(vcross(e1, e0));
    eval_triangle(&s, nrm, e0, e1, e2);

    /* Ditetrahedron, dito.cpp:294-307,382-418. */
    {
        float p = vdot(s.sel[0], nrm);
        float lo = p, hi = p;
        int ilo = 0, ihi = 0;
        for (int i = 1; i < np; ++i) {
            p = vdot(s.sel[i], nrm);
            if (p < lo) { lo = p; ilo = i; }
            if (p > hi) { hi = p; ihi = i; }
        }
        float tri = vdot(p0, nrm);
        int q0v = (hi - kEps > tri) ? 1 : 0;
        int q1v = (lo + kEps < tri) ? 1 : 0;
        tr.q0_sel = ihi; tr.q1_sel = ilo; tr.q0_valid = q0v; tr.q1_valid = q1v;
        s.counter = 3;
        if (q0v) {
            ov3 q0 = s.sel[ihi];
            ov3 f0 = vnormalize(vsub(q0, p0));
            ov3 f1 = vnormalize(vsub(q0, p1));
            ov3 f2 = vnormalize(vsub(q0, p2));
            ov3 n0 = vnormalize(vcross(f1, e0));
            ov3 n1 = vnormalize(vcross(f2, e1));
            ov3 n2 = vnormalize(vcross(f0, e2));
            eval_triangle(&s, n0, e0, f1, f0);
            eval_triangle(&s, n1, e1, f2, f1);
            eval_triangle(&s, n2, e2, f0, f2);
        }
        s.counter = 12;
        if (q1v) {
            ov3 q1 = s.sel[ilo];
            ov3 g0 = vnormalize(vsub(q1, p0));
            ov3 g1 = vnormalize(vsub(q1, p1));
            ov3 g2 = vnormalize(vsub(q1, p2));
            ov3 m0 = vnormalize(vcross(g1, e0));
            ov3 m1 = vnormalize(vcross(g2, e1));
            ov3 m2 = vnormalize(vcross(g0, e2));
            eval_triangle(&s, m0, e0, g1, g0);
            eval_triangle(&s, m1, e1, g2, g1);
            eval_triangle(&s, m2, e2, g0, g2);
        }
    }
    tr.best = s.best_id;

    /* Final dimensions over all points, dito.cpp:530-537. */
    ov3 bmn, bmx;
    extremal_projs_all(s.b0, xyz, n, &bmn.x, &bmx.x);
    extremal_projs_all(s.b1, xyz, n, &bmn.y, &bmx.y);
    extremal_projs_all(s.b2, xyz, n, &bmn.z, &bmx.z);
    ov3 blen = vsub(bmx, bmn);
    float bval = quality(blen);
    if (bval < al_val) {
        finalize_obb(s.b0, s.b1, s.b2, bmn, bmx, blen, out);
        tr.result_aabb = 0;
    } else {
        finalize_aabb(al_mid, al_len, out);
        tr.result_aabb = 1;
    }
    if (trace) *trace = tr;
}

/* kr_points_obb, geometry.cpp:232-252: ext = kr_vmax3(ext, 0.001),
 * kr_max(a, b) = (a < b) ? b : a (vecmath.h:1397). */
static float kmax(float a, float b) { return (a < b) ? b : a; }

void oracle_points_obb(const float* xyz, int n, oobb* out, odito_trace* trace) {
    oracle_dito14(xyz, n, out, trace);
    out->ext.x = kmax(out->ext.x, 0.001f);
    out->ext.y = kmax(out->ext.y, 0.001f);
    out->ext.z = kmax(out->ext.z, 0.001f);
}

void oracle_points_obb_batch(const float* xyz, const uint64_t* off, int nsets,
                             oobb* out, odito_trace* traces) {
    oracle_points_obb_batch_mt(xyz, off, nsets, out, traces, 1);
}

typedef struct {
    const float* xyz;
    const uint64_t* off;
    oobb* out;
    odito_trace* traces;
} batch_ctx;

static void batch_one(void* p, uint32_t s, int tid) {
    const batch_ctx* c = (const batch_ctx*)p;
    (void)tid;
    oracle_points_obb(c->xyz + 3 * c->off[s], (int)(c->off[s + 1] - c->off[s]), &c->out[s],
                      c->traces ? &c->traces[s] : 0);
}

void oracle_points_obb_batch_mt(const float* xyz, const uint64_t* off, int nsets,
                                oobb* out, odito_trace* traces, int threads) {
    if (nsets <= 0) return;
    batch_ctx c = {xyz, off, out, traces};
    oracle_parallel_for((uint32_t)nsets, 1, threads, batch_one, &c);
}
