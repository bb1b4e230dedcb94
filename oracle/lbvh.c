/*
 * lbvh.c -- CPU restatement of the reference LBVH build
 * (src/ads/lbvh_gpu/lbvh.cu:252-457 and the build sequence :1547-1692).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP LBVH builder, and the
 * tree source of the CPU-only transform tests).
 *
 * Steps:
 *  1. centroid AABB over triangle-box centres.  The reference kernel
 *     (src/ads/common/cuda/transforms.cu:1127-1190) seeds every LDS slot of
 *     a block with vertices[0] and lets threads past face_count return
 *     before the block reduction, so when face_count is not a multiple of
 *     its 256-thread block the box also contains vertices[0].  That is
 *     restated here (the built tree depends on it).
 *  2. 30-bit Morton code of the triangle-box centre, normalised by the
 *     centroid box with a safe divide (lbvh.cu:252-282, vecmath.h:698-716,760).
 *  3. stable sort of leaves by code (thrust::stable_sort_by_key, :1670).
 *  4. Karras radix tree with the duplicate-code fallback 32 + clz(i ^ j)
 *     (:284-396).
 *  5. bottom-up bounds, primitive counts and SAH costs with ct = 1.0,
 *     ci = 1.2 (:399-457).
 */
#include "oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

#define INVALID 0xFFFFFFFFu

static ov3 v3(float x, float y, float z) { ov3 r = {x, y, z}; return r; }
static float kmin(float a, float b) { return (a < b) ? a : b; }
static float kmax(float a, float b) { return (a < b) ? b : a; }
static ov3 vmin3(ov3 a, ov3 b) { return v3(kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)); }
static ov3 vmax3(ov3 a, ov3 b) { return v3(kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)); }
static ov3 vert(const float* V, uint32_t i) { return v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }

typedef struct { ov3 mn, mx; } box;

static box expand3(box b, ov3 a) { box r = {vmin3(b.mn, a), vmax3(b.mx, a)}; return r; } /* vecmath.h:1640 */
static box expand(box b, box a) { return expand3(expand3(b, a.mn), a.mx); }             /* vecmath.h:1646 */
static float area(box b) {                                                              /* vecmath.h:1580 */
    ov3 d = v3(b.mx.x - b.mn.x, b.mx.y - b.mn.y, b.mx.z - b.mn.z);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
static ov3 center(box b) {                                                              /* vecmath.h:1559 */
    return v3((b.mn.x + b.mx.x) * 0.5f, (b.mn.y + b.mx.y) * 0.5f, (b.mn.z + b.mx.z) * 0.5f);
}

static box tri_box(const float* V, const uint32_t* f) { /* lbvh.cu:264-269 */
    ov3 a = vert(V, f[0]);
    box b = {a, a};
    b = expand3(b, vert(V, f[1]));
    b = expand3(b, vert(V, f[2]));
    return b;
}

static uint32_t stretch3(uint32_t x) { /* vecmath.h:698-707 */
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
}
static uint32_t quant10(float v) { return (uint32_t)kmin(kmax(v * 1024.0f, 0.0f), 1023.0f); }
static uint32_t morton30(ov3 a) { /* vecmath.h:710-716 */
    return stretch3(quant10(a.x)) << 2 | stretch3(quant10(a.y)) << 1 | stretch3(quant10(a.z));
}
static float safe_div(float a, float b) { return (b != 0.0f) ? a / b : 0.0f; } /* vecmath.h:760-762 */

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static int delta(const uint32_t* c, int i, int j, int n) { /* lbvh.cu:289-308 */
    if (j > n - 1 || j < 0) return -1;
    if (c[i] == c[j]) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(c[i] ^ c[j]);
}
static int lid(int num) { /* lbvh.cu:310-315 */
    if (num == 1) return 0;
    if (num % 2 == 1) return num / 2 + 1;
    return num / 2;
}

typedef struct { uint32_t code, prim; } keyed;
static int cmp_keyed(const void* a, const void* b) {
    const keyed* x = (const keyed*)a;
    const keyed* y = (const keyed*)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->prim < y->prim ? -1 : (x->prim > y->prim); /* stability */
}

int oracle_lbvh_build(const float* V, const uint32_t* F, uint32_t nf, onode* nodes,
                      uint32_t* parents, uint32_t* counts, float* costs,
                      uint32_t* mortons_sorted) {
    if (nf == 0) return -1;
    uint32_t I = nf - 1;
    onode* leaves = nodes + I;

    /* 1. centroid box (+ vertices[0] when the last 256-block is partial). */
    box cb = {v3(FLT_MAX, FLT_MAX, FLT_MAX), v3(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
    for (uint32_t f = 0; f < nf; ++f) cb = expand3(cb, center(tri_box(V, &F[4 * f])));
    if (nf % 256u != 0) cb = expand3(cb, vert(V, 0));

    /* 2. codes */
    keyed* kv = (keyed*)malloc(sizeof(keyed) * nf);
    box* tb = (box*)malloc(sizeof(box) * nf);
    if (!kv || !tb) { free(kv); free(tb); return -1; }
    ov3 ext = v3(cb.mx.x - cb.mn.x, cb.mx.y - cb.mn.y, cb.mx.z - cb.mn.z);
    for (uint32_t f = 0; f < nf; ++f) {
        box b = tri_box(V, &F[4 * f]);
        tb[f] = b;
        box o = {v3(b.mn.x + -cb.mn.x, b.mn.y + -cb.mn.y, b.mn.z + -cb.mn.z),
                 v3(b.mx.x + -cb.mn.x, b.mx.y + -cb.mn.y, b.mx.z + -cb.mn.z)};
        ov3 c = center(o);
        c = v3(safe_div(c.x, ext.x), safe_div(c.y, ext.y), safe_div(c.z, ext.z));
        kv[f].code = morton30(c);
        kv[f].prim = f;
    }
    /* 3. stable sort */
    qsort(kv, nf, sizeof(keyed), cmp_keyed);
    uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * nf);
    for (uint32_t j = 0; j < nf; ++j) {
        codes[j] = kv[j].code;
        onode* L = &leaves[j];
        memset(L, 0, sizeof(*L));
        L->bmin = tb[kv[j].prim].mn;
        L->bmax = tb[kv[j].prim].mx;
        L->a = kv[j].prim;
        L->nprim = 1;
    }
    if (mortons_sorted) memcpy(mortons_sorted, codes, sizeof(uint32_t) * nf);
    free(kv); free(tb);

    /* 4. Karras radix tree */
    int n = (int)nf;
    if (nf == 1) parents[0] = INVALID;
    for (int i = 0; i < (int)I; ++i) {
        int da = delta(codes, i, i + 1, n), db = delta(codes, i, i - 1, n);
        int d = (da - db > 0) - (da - db < 0);
        int dmin = delta(codes, i, i - d, n);
        int lmax = 2;
        while (delta(codes, i, i + lmax * d, n) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(codes, i, i + (l + t) * d, n) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(codes, i, j, n);
        int s = 0;
        for (int t = lid(l); t >= 1; t = lid(t))
            if (delta(codes, i, i + (s + t) * d, n) > dnode) s += t;
        int split = i + s * d + (d < 0 ? d : 0);
        int lleaf = ((i < j ? i : j) == split);
        int rleaf = ((i > j ? i : j) == split + 1);
        uint32_t li = (uint32_t)split + (lleaf ? I : 0);
        uint32_t ri = (uint32_t)split + 1 + (rleaf ? I : 0);
        parents[li] = (uint32_t)i;
        parents[ri] = (uint32_t)i;
        onode* N = &nodes[i];
        memset(N, 0, sizeof(*N));
        N->a = li; N->b = ri;
        N->bmin = v3(FLT_MAX, FLT_MAX, FLT_MAX);
        N->bmax = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        if (i == 0) parents[0] = INVALID;
    }
    free(codes);

    /* 5. bounds / counts / costs bottom-up (children before parents: a node's
     * children are resolved by a post-order walk from the root). */
    const float ct = 1.0f, ci = (float)1.2;
    for (uint32_t j = 0; j < nf; ++j) {
        box b = {leaves[j].bmin, leaves[j].bmax};
        costs[I + j] = ct * area(b);
        counts[I + j] = 1;
    }
    if (I > 0) {
        uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (2 * nf));
        uint8_t* seen = (uint8_t*)calloc(I, 1);
        uint32_t sp = 0;
        stack[sp++] = 0;
        while (sp) {
            uint32_t v = stack[sp - 1];
            if (v >= I) { --sp; continue; }
            if (!seen[v]) {
                seen[v] = 1;
                stack[sp++] = nodes[v].b;
                stack[sp++] = nodes[v].a;
                continue;
            }
            --sp;
            const onode* L = &nodes[nodes[v].a];
            const onode* R = &nodes[nodes[v].b];
            box bl = {L->bmin, L->bmax}, br = {R->bmin, R->bmax};
            box bn = expand(bl, br);
            counts[v] = counts[nodes[v].a] + counts[nodes[v].b];
            costs[v] = ci * area(bn) + costs[nodes[v].b] + costs[nodes[v].a];
            nodes[v].bmin = bn.mn;
            nodes[v].bmax = bn.mx;
        }
        free(stack); free(seen);
    }
    return 0;
}
