/*
 * atrbvh.c -- CPU restatement of the reference's ATRBVH builder.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/ads/atrbvh_gpu (paths relative to the reference root):
 *   scene box      atrbvh.cu:262-290 (union of the triangle boxes)
 *   Morton codes   LBVH.cu:9-93, Commons.cuh:422-430,602-645: triangle-box
 *                  centre (max + min) * 0.5, normalised by the scene box,
 *                  21 bits per axis, x << 2 | y << 1 | z
 *   sort           CubWrapper.cu:8-46: stable radix sort of (code, index)
 *   radix tree     LBVH.cu:156-302 (64-bit keys, duplicates by 64 + clz(i^j))
 *   boxes          LBVH.cu:414-446,551-665
 *   optimizer      AgglomerativeTreeletOptimizer.cpp:26-100 (iterations,
 *                  gamma = treelet size, doubled each iteration),
 *                  AgglomerativeTreelet.cu:14-222,405-561 (small-treelet
 *                  kernel), Treelet.cuh:21-127 (treelet formation),
 *                  TriangularMatrix.cuh:37-52
 *   output         atrbvh.cu:95-179 (node copy, bounds / counts / costs)
 *
 * The reference optimises each treelet with one 32-lane warp; its
 * reductions are __shfl_down tournaments whose tie-breaks depend on the
 * lane pattern.  They are simulated here lane by lane (warp_* below), so the
 * chosen pairs and leaves are those of the warp code.  Box min/max follow
 * fminf/fmaxf with -0 ordered below +0 (zmin/zmax); the final node bounds
 * use kr_min/kr_max (a < b ? a : b) as kr_aabb_expand does.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE 0xFFFFFFFFu
#define LANES 32

static float zmin(float a, float b) { return a < b ? a : (b < a ? b : (signbit(a) ? a : b)); }
static float zmax(float a, float b) { return a > b ? a : (b > a ? b : (signbit(a) ? b : a)); }
static float kmin(float a, float b) { return (a < b) ? a : b; }
static float kmax(float a, float b) { return (a < b) ? b : a; }

typedef struct { float mn[3], mx[3]; } box3;

/* calculateBoundingBoxSurfaceArea, Commons.cuh:96-103 */
static float box_area(const box3* b) {
    float x = b->mx[0] - b->mn[0], y = b->mx[1] - b->mn[1], z = b->mx[2] - b->mn[2];
    return 2 * (x * y + x * z + y * z);
}
/* calculateBoundingBoxAndSurfaceArea, Commons.cuh:173-181 */
static float union_area(const box3* a, const box3* b) {
    float x = zmax(a->mx[0], b->mx[0]) - zmin(a->mn[0], b->mn[0]);
    float y = zmax(a->mx[1], b->mx[1]) - zmin(a->mn[1], b->mn[1]);
    float z = zmax(a->mx[2], b->mx[2]) - zmin(a->mn[2], b->mn[2]);
    return 2 * (x * y + x * z + y * z);
}

/* expandBits64, Commons.cuh:602-613 */
static uint64_t expand_bits64(uint64_t v) {
    v &= 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
static uint64_t discretize21(float x) {
    float s = fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f); /* NaN -> 0 like fmaxf */
    return (uint64_t)s;
}

typedef struct {
    uint64_t key;
    uint32_t idx;
} kv;
static int kv_cmp(const void* a, const void* b) {
    const kv* x = (const kv*)a;
    const kv* y = (const kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

/* longestCommonPrefix, LBVH.cu:173-187 */
static int lcp(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 64 + clz32((uint32_t)(i ^ j));
    return clz64(k[i] ^ k[j]);
}

typedef struct {
    uint32_t* left;
    uint32_t* right;
    uint32_t* parent;
    box3* box;
    float* area;
    int32_t* tri;
    float* sah;
    uint32_t n; /* triangles */
} atree;

static int is_leaf(const atree* t, uint32_t v) { return v >= t->n - 1; }

/* findLeafToExpand / findMinimum: __shfl_down tournament over 32 lanes with
 * the given participant count; better(a, b) = b strictly better than a. */
static void warp_argbest(int count, int32_t idx[LANES], float val[LANES], int want_max) {
    int amount = count / 2;
    while (count > 1) {
        int32_t ni[LANES];
        float nv[LANES];
        for (int l = 0; l < LANES; ++l) {
            int src = l + amount < LANES ? l + amount : l;
            int32_t oi = idx[src];
            float ov = val[src];
            int better = want_max ? (ov > val[l]) : (ov < val[l]);
            ni[l] = better ? oi : idx[l];
            nv[l] = better ? ov : val[l];
        }
        memcpy(idx, ni, sizeof(ni));
        memcpy(val, nv, sizeof(nv));
        count = (count + 1) / 2;
        amount = count / 2;
    }
}

/* LOWER_TRM_ROW / LOWER_TRM_COL, TriangularMatrix.cuh:37-52 */
static int trm_row(int index) { return (int)((-1 + sqrtf(8.0f * index + 1)) / 2) + 1; }
static int trm_col(int index) {
    int y = trm_row(index);
    return index - (y * (y - 1)) / 2;
}
#define TRM(row, col) ((((row) - 1) * (row)) / 2 + (col))

static void treelet(atree* t, uint32_t root, int ts, float ci) {
    uint32_t tin[32], tl[32];
    float ta[32];
    /* formTreelet, Treelet.cuh:55-127 */
    tin[0] = root;
    tl[0] = t->left[root];
    tl[1] = t->right[root];
    ta[0] = t->area[tl[0]];
    ta[1] = t->area[tl[1]];
    for (int it = 0; it < ts - 2; ++it) {
        int32_t idx[LANES];
        float val[LANES];
        for (int l = 0; l < LANES; ++l) {
            idx[l] = -1;
            val[l] = -FLT_MAX;
            if (l < 2 + it) {
                idx[l] = l;
                val[l] = is_leaf(t, tl[l]) ? -FLT_MAX : ta[l];
            }
        }
        warp_argbest(2 + it, idx, val, 1);
        int big = idx[0];
        uint32_t rep = tl[big];
        uint32_t l2 = t->left[rep], r2 = t->right[rep];
        tin[it + 1] = rep;
        tl[big] = l2;
        tl[it + 2] = r2;
        ta[big] = t->area[l2];
        ta[it + 2] = t->area[r2];
    }
    /* lane state */
    box3 bb[LANES];
    uint32_t node[LANES];
    float sah[LANES];
    float tla[32];
    for (int l = 0; l < ts; ++l) {
        bb[l] = t->box[tl[l]];
        node[l] = tl[l];
        sah[l] = t->sah[tl[l]];
        tla[l] = ta[l];
    }
    float dist[32 * 31 / 2];
    for (int a = 1; a < ts; ++a)
        for (int b = 0; b < a; ++b) dist[TRM(a, b)] = union_area(&bb[a], &bb[b]);
    for (int last = ts - 1; last > 0; --last) {
        /* findMinimumDistance, AgglomerativeTreelet.cu:78-93 */
        int size = last * (1 + last) / 2;
        int32_t idx[LANES];
        float val[LANES];
        for (int l = 0; l < LANES; ++l) {
            idx[l] = 0;
            val[l] = FLT_MAX;
            for (int j = l; j < size; j += LANES)
                if (dist[j] < val[l]) { val[l] = dist[j]; idx[l] = j; }
        }
        warp_argbest(LANES, idx, val, 0);
        int row = trm_row(idx[0]), col = trm_col(idx[0]);
        /* updateState, AgglomerativeTreelet.cu:95-146 */
        float s = sah[row] + sah[col];
        uint32_t li = node[row], ri = node[col];
        box3 A = bb[row], B = bb[col];
        for (int k = 0; k < 3; ++k) { /* expandBoundingBox(A, B) */
            A.mn[k] = zmin(B.mn[k], A.mn[k]);
            A.mx[k] = zmax(B.mx[k], A.mx[k]);
        }
        node[col] = tin[last - 1];
        bb[col] = A;
        tla[last] = s + ci * box_area(&A);
        sah[col] = tla[last];
        uint32_t last_node = node[last];
        float last_sah = sah[last];
        box3 lb = bb[last];
        node[row] = last_node;
        sah[row] = last_sah;
        bb[row] = lb;
        node[last] = li;
        tl[last] = ri;
        bb[last] = A;
        /* updateDistancesMatrix, AgglomerativeTreelet.cu:148-185 */
        if (last > 1) {
            for (int l = 0; l < last; ++l) {
                if (l == row) continue;
                int dr = row > l ? row : l, dc = row < l ? row : l;
                dist[TRM(dr, dc)] = dist[TRM(last, l)];
            }
            for (int l = 0; l < last; ++l) {
                if (l == col) continue;
                int dr = col > l ? col : l, dc = col < l ? col : l;
                dist[TRM(dr, dc)] = union_area(&bb[dr], &bb[dc]);
            }
        }
    }
    /* updateTreelet, AgglomerativeTreelet.cu:187-208 */
    if (tla[1] < t->sah[tin[0]]) {
        for (int l = 1; l < ts; ++l) {
            uint32_t v = tin[l - 1];
            t->left[v] = node[l];
            t->right[v] = tl[l];
            t->parent[node[l]] = v;
            t->parent[tl[l]] = v;
            t->sah[v] = tla[l];
            t->box[v] = bb[l];
            t->area[v] = box_area(&bb[l]);
        }
    }
}

/* One optimiser iteration (agglomerativeSmallTreeletKernel): post-order over
 * the current tree; a node is optimised as a treelet root right after its
 * subtree when it holds at least gamma triangles. */
static void optimize_pass(atree* t, int ts, int gamma, float ci, float ct) {
    const uint32_t I = t->n - 1, total = 2 * t->n - 1;
    for (uint32_t v = I; v < total; ++v) {
        t->tri[v] = 1;
        t->sah[v] = ct * t->area[v];
    }
    if (t->n < 2) return;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * 2 * total);
    uint8_t* expanded = (uint8_t*)calloc(total, 1);
    int64_t sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        uint32_t v = stack[sp - 1];
        if (is_leaf(t, v)) { --sp; continue; }
        if (!expanded[v]) {
            expanded[v] = 1;
            stack[sp++] = t->right[v];
            stack[sp++] = t->left[v];
            continue;
        }
        --sp;
        uint32_t l = t->left[v], r = t->right[v];
        int32_t tri = t->tri[l] + t->tri[r];
        t->tri[v] = tri;
        t->sah[v] = ci * t->area[v] + t->sah[l] + t->sah[r];
        if (tri >= gamma) treelet(t, v, ts, ci);
    }
    free(stack);
    free(expanded);
}

int oracle_atrbvh_build(const float* vertices, const uint32_t* faces, uint32_t nfaces, int treelet_size,
                        int iterations, float ci, float ct, onode* nodes, uint32_t* parents,
                        uint32_t* counts, float* costs) {
    if (nfaces < 1 || treelet_size < 3 || treelet_size > 20) return -1;
    const uint32_t N = nfaces, I = N - 1, total = 2 * N - 1;
    /* scene box over the triangle boxes (kr_aabb_expand3 / kr_aabb_expand) */
    float smin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, smax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t f = 0; f < N; ++f) {
        float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                float c = vertices[3 * faces[4 * f + j] + k];
                bmin[k] = kmin(bmin[k], c);
                bmax[k] = kmax(bmax[k], c);
            }
        for (int k = 0; k < 3; ++k) {
            smin[k] = kmin(kmin(smin[k], bmin[k]), bmax[k]);
            smax[k] = kmax(kmax(smax[k], bmin[k]), bmax[k]);
        }
    }
    /* Morton codes of the triangle-box centres */
    kv* keys = (kv*)malloc(sizeof(kv) * N);
    box3* leafbox = (box3*)malloc(sizeof(box3) * N);
    for (uint32_t f = 0; f < N; ++f) {
        const float* a = &vertices[3 * faces[4 * f]];
        const float* b = &vertices[3 * faces[4 * f + 1]];
        const float* c = &vertices[3 * faces[4 * f + 2]];
        box3 bx;
        for (int k = 0; k < 3; ++k) {
            bx.mn[k] = zmin(zmin(a[k], b[k]), c[k]);
            bx.mx[k] = zmax(zmax(a[k], b[k]), c[k]);
        }
        leafbox[f] = bx;
        uint64_t d[3];
        for (int k = 0; k < 3; ++k) {
            float cen = (bx.mx[k] + bx.mn[k]) * 0.5f;
            float nrm = (cen - smin[k]) / (smax[k] - smin[k]);
            d[k] = expand_bits64(discretize21(nrm));
        }
        keys[f].key = (d[0] << 2) | (d[1] << 1) | d[2];
        keys[f].idx = f;
    }
    qsort(keys, N, sizeof(kv), kv_cmp);
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint32_t i = 0; i < N; ++i) k[i] = keys[i].key;

    atree t;
    t.n = N;
    t.left = (uint32_t*)calloc(total, sizeof(uint32_t));
    t.right = (uint32_t*)calloc(total, sizeof(uint32_t));
    t.parent = (uint32_t*)malloc(total * sizeof(uint32_t));
    t.box = (box3*)calloc(total, sizeof(box3));
    t.area = (float*)calloc(total, sizeof(float));
    t.tri = (int32_t*)calloc(total, sizeof(int32_t));
    t.sah = (float*)calloc(total, sizeof(float));
    uint32_t* data = (uint32_t*)malloc(total * sizeof(uint32_t));
    t.parent[0] = NONE;
    data[I] = keys[0].idx;
    /* buildTreeKernel, LBVH.cu:193-302 */
    for (int i = 0; i < (int)I; ++i) {
        int n = (int)N;
        int lcp1 = lcp(k, n, i, i + 1), lcp2 = lcp(k, n, i, i - 1);
        int dir = (0 < lcp1 - lcp2) - (lcp1 - lcp2 < 0);
        int min_lcp = lcp(k, n, i, i - dir);
        int lmax = 2;
        while (lcp(k, n, i, i + lmax * dir) > min_lcp) lmax <<= 1;
        int l = 0;
        for (int s = lmax >> 1; s >= 1; s >>= 1)
            if (lcp(k, n, i, i + (l + s) * dir) > min_lcp) l += s;
        int j = i + l * dir;
        int node_lcp = lcp(k, n, i, j);
        int s = 0, divisor = 2;
        int max_div = 1 << (32 - clz32((uint32_t)l));
        while (divisor <= max_div) {
            int step = (l + divisor - 1) / divisor;
            if (lcp(k, n, i, i + (s + step) * dir) > node_lcp) s += step;
            divisor <<= 1;
        }
        int split = i + s * dir + (dir < 0 ? dir : 0);
        uint32_t li = (uint32_t)split, ri = (uint32_t)split + 1;
        int lo = i < j ? i : j, hi = i < j ? j : i;
        if (lo == split) {
            li = (uint32_t)split + I;
            data[li] = keys[split].idx;
        }
        if (hi == split + 1) {
            ri = (uint32_t)split + 1 + I;
            data[ri] = keys[split + 1].idx;
        }
        t.left[i] = li;
        t.right[i] = ri;
        t.parent[li] = (uint32_t)i;
        t.parent[ri] = (uint32_t)i;
    }
    if (N == 1) data[0] = keys[0].idx;
    /* leaf boxes and bottom-up unions, areas */
    for (uint32_t v = I; v < total; ++v) {
        t.box[v] = leafbox[data[v]];
        t.area[v] = box_area(&t.box[v]);
    }
    if (N > 1) {
        uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * total);
        uint32_t head = 0, tail = 0;
        order[tail++] = 0;
        while (head < tail) {
            uint32_t v = order[head++];
            if (v < I) {
                order[tail++] = t.left[v];
                order[tail++] = t.right[v];
            }
        }
        for (uint32_t q = total; q-- > 0;) {
            uint32_t v = order[q];
            if (v >= I) continue;
            const box3* a = &t.box[t.left[v]];
            const box3* b = &t.box[t.right[v]];
            for (int c = 0; c < 3; ++c) {
                t.box[v].mn[c] = zmin(a->mn[c], b->mn[c]);
                t.box[v].mx[c] = zmax(a->mx[c], b->mx[c]);
            }
            t.area[v] = box_area(&t.box[v]);
        }
        free(order);
        int gamma = treelet_size;
        for (int it = 0; it < iterations; ++it) {
            optimize_pass(&t, treelet_size, gamma, ci, ct);
            gamma *= 2;
        }
    }
    /* copyTree_kernel (atrbvh.cu:95-121) and calculate_aabbs_kernel
     * (atrbvh.cu:124-179, ct = 1, ci = 1.2 as written there) */
    memset(nodes, 0, sizeof(onode) * total);
    for (uint32_t v = 0; v < total; ++v) {
        parents[v] = t.parent[v];
        if (v < I) {
            nodes[v].a = t.left[v];
            nodes[v].b = t.right[v];
        } else {
            const box3* b = &t.box[v];
            nodes[v].bmin = (ov3){b->mn[0], b->mn[1], b->mn[2]};
            nodes[v].bmax = (ov3){b->mx[0], b->mx[1], b->mx[2]};
            nodes[v].nprim = 1;
            nodes[v].a = data[v];
            counts[v] = 1;
            float dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
            costs[v] = 1.0f * (2.0f * (dx * dy + dx * dz + dy * dz));
        }
    }
    if (N > 1) {
        uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * total);
        uint32_t head = 0, tail = 0;
        order[tail++] = 0;
        while (head < tail) {
            uint32_t v = order[head++];
            if (v < I) {
                order[tail++] = nodes[v].a;
                order[tail++] = nodes[v].b;
            }
        }
        for (uint32_t q = total; q-- > 0;) {
            uint32_t v = order[q];
            if (v >= I) continue;
            const onode* A = &nodes[nodes[v].a];
            const onode* B = &nodes[nodes[v].b];
            ov3 mn = A->bmin, mx = A->bmax;
            /* kr_aabb_expand(A, B) = expand3(expand3(A, B.min), B.max) */
            mn.x = kmin(mn.x, B->bmin.x); mn.y = kmin(mn.y, B->bmin.y); mn.z = kmin(mn.z, B->bmin.z);
            mx.x = kmax(mx.x, B->bmin.x); mx.y = kmax(mx.y, B->bmin.y); mx.z = kmax(mx.z, B->bmin.z);
            mn.x = kmin(mn.x, B->bmax.x); mn.y = kmin(mn.y, B->bmax.y); mn.z = kmin(mn.z, B->bmax.z);
            mx.x = kmax(mx.x, B->bmax.x); mx.y = kmax(mx.y, B->bmax.y); mx.z = kmax(mx.z, B->bmax.z);
            nodes[v].bmin = mn;
            nodes[v].bmax = mx;
            counts[v] = counts[nodes[v].a] + counts[nodes[v].b];
            float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
            costs[v] = (float)1.2 * (2.0f * (dx * dy + dx * dz + dy * dz)) + costs[nodes[v].b] + costs[nodes[v].a];
        }
        free(order);
    }
    (void)ct;
    free(keys);
    free(leafbox);
    free(k);
    free(t.left);
    free(t.right);
    free(t.parent);
    free(t.box);
    free(t.area);
    free(t.tri);
    free(t.sah);
    free(data);
    return 0;
}
