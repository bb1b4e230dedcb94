/*
 * cpu_transform.c -- restatement of the reference CPU transform
 * kr_bvh_obb_tree (src/ads/common/transforms.cpp:349-481).
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline for the transform.
 *
 * Method (a different algorithm from the GPU path): every leaf gets the
 * DiTO box of its 3 vertices; climbing from each leaf in order, the second
 * visit of an internal node gathers all leaves below it in breadth-first
 * order (transforms.cpp:349-369, 443-452) and runs kr_points_obb over their
 * vertices.  Matrix = inverse(T * (R * S)).  node.axis records whether the
 * OBB beats the AABB by surface area (obb_cost = 1, :473).  The BFS queue is
 * an array here where the reference uses std::list; the leaf order, and so
 * the result, is the same.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define INVALID 0xFFFFFFFFu

static float kmax(float a, float b) { return (a < b) ? b : a; }

static void obb_matrix(const oobb* o, omat4* out) {
    omat4 r = {{o->v0.x, o->v0.y, o->v0.z, 0, o->v1.x, o->v1.y, o->v1.z, 0,
                o->v2.x, o->v2.y, o->v2.z, 0, 0, 0, 0, 1}};
    omat4 s = {{o->ext.x * 2.0f, 0, 0, 0, 0, o->ext.y * 2.0f, 0, 0, 0, 0, o->ext.z * 2.0f, 0,
                0, 0, 0, 1}};
    omat4 t = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, o->mid.x, o->mid.y, o->mid.z, 1}};
    omat4 rs, trs;
    oracle_mmul4(&r, &s, &rs);
    oracle_mmul4(&t, &rs, &trs);
    oracle_minverse4(&trs, out);
}

static float box_area(ov3 mn, ov3 mx) { /* vecmath.h:1580-1584 */
    float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}
static float obb_area(const oobb* o) { /* vecmath.h:1574-1578 */
    float dx = 2.0f * o->ext.x, dy = 2.0f * o->ext.y, dz = 2.0f * o->ext.z;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

static void floor_ext(oobb* o) {
    o->ext.x = kmax(o->ext.x, 0.001f);
    o->ext.y = kmax(o->ext.y, 0.001f);
    o->ext.z = kmax(o->ext.z, 0.001f);
}

int64_t oracle_cpu_obb_tree(onode* nodes, const uint32_t* parents, const float* V,
                            const uint32_t* F, uint32_t N, oobb* obbs, omat4* mats,
                            uint32_t leaf_begin, uint32_t leaf_end) {
    if (N == 0) return 0;
    uint32_t I = N - 1;
    uint32_t* visit = (uint32_t*)calloc(I ? I : 1, sizeof(uint32_t));
    float* pts = (float*)malloc(sizeof(float) * 9 * (size_t)N);
    uint32_t* queue = (uint32_t*)malloc(sizeof(uint32_t) * (2 * (size_t)N));
    if (!visit || !pts || !queue) { free(visit); free(pts); free(queue); return -1; }
    int64_t fitted = 0;
    if (leaf_end > N) leaf_end = N;
    for (uint32_t i = leaf_begin; i < leaf_end; ++i) {
        onode* leaf = &nodes[I + i];
        const uint32_t* f = &F[4 * leaf->a];
        for (int j = 0; j < 3; ++j) memcpy(&pts[3 * j], &V[3 * f[j]], 3 * sizeof(float));
        oobb o;
        oracle_points_obb(pts, 3, &o, 0);
        floor_ext(&o);
        obbs[I + i] = o;
        obb_matrix(&o, &mats[I + i]);
        leaf->axis = 1;

        uint32_t p = parents[I + i];
        while (p != INVALID) {
            if (visit[p]++ == 0) break; /* first visit stops (transforms.cpp:436) */
            /* BFS leaves of p, transforms.cpp:349-369 */
            uint32_t head = 0, tail = 0, nv = 0;
            queue[tail++] = p;
            while (head < tail) {
                uint32_t u = queue[head++];
                if (nodes[u].nprim > 0) {
                    const uint32_t* g = &F[4 * nodes[u].a];
                    for (int j = 0; j < 3; ++j) memcpy(&pts[3 * (nv + j)], &V[3 * g[j]], 3 * sizeof(float));
                    nv += 3;
                    continue;
                }
                queue[tail++] = nodes[u].a;
                queue[tail++] = nodes[u].b;
            }
            oracle_points_obb(pts, (int)nv, &o, 0);
            floor_ext(&o);
            obbs[p] = o;
            obb_matrix(&o, &mats[p]);
            nodes[p].axis = (box_area(nodes[p].bmin, nodes[p].bmax) > 1.0f * obb_area(&o)) ? 1 : 0;
            ++fitted;
            p = parents[p];
        }
    }
    free(visit); free(pts); free(queue);
    return fitted;
}
