/*
 * collapse.c -- CPU restatement of the reference's BVH collapse and SAH.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * oracle_collapse follows kr_cuda_bvh_collapse
 * (src/ads/common/cuda/transforms.cu:18-211,1061-1125):
 *   mark    (bvh_mark_collapse_kernel, :64-137) bottom-up: a leaf costs
 *           ct * area(bounds); an internal node p with count = counts[l] +
 *           counts[r], area = area(p.bounds), node_cost = ci*area + costs[l]
 *           + costs[r], leaf_cost = ct*area*count is collapsed when
 *           leaf_cost < node_cost and count <= 8: costs[p] = leaf_cost, p and
 *           both children are marked, and both children get nPrimitives =
 *           0xFFFF in the collapsed copy; otherwise costs[p] = node_cost.
 *   emit    (bvh_collapse_kernel, :139-209) for every internal node that is
 *           not marked, each marked child with more than one primitive
 *           becomes a leaf of the collapsed tree holding the distinct
 *           primitive ids of its subtree, gathered by the reference's
 *           right-first stack walk (bvh_node_collapse, :18-61).
 * The reference gives each such leaf its slice of `primitives` by an atomic
 * counter (arrival order).  Here the slices follow increasing node index, a
 * deterministic order; each slice holds the same ids in the same order.
 *
 * oracle_bvh_sah follows kr_bvh_sah (src/ads/common/transforms.cpp:250-303):
 * a LIFO stack walk from the root (push left, then right), summing ci*area
 * for internal nodes and ct*n*area for leaves (AABB area for one-primitive
 * leaves even when OBBs are given), divided by the root area.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define COLLAPSED 0xFFFFu
#define MAX_LEAF_SIZE 8u
#define NONE 0xFFFFFFFFu

/* kr_aabb_surface_area3, vecmath.h:1583-1586 (extents = max - min). */
static float aabb_area(const onode* n) {
    float dx = n->bmax.x - n->bmin.x, dy = n->bmax.y - n->bmin.y, dz = n->bmax.z - n->bmin.z;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

/* kr_obb_surface_area3, vecmath.h:1572-1580 (extents = 2 * ext). */
static float obb_area(const oobb* o) {
    float dx = 2.0f * o->ext.x, dy = 2.0f * o->ext.y, dz = 2.0f * o->ext.z;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

int64_t oracle_collapse(const onode* nodes, const uint32_t* parents, uint32_t* counts, float* costs,
                        onode* collapsed, uint32_t* primitives, uint8_t* table, float ci, float ct,
                        uint32_t nprim) {
    if (nprim == 0) return 0;
    const uint32_t I = nprim - 1, total = 2 * nprim - 1;
    uint32_t root = NONE;
    for (uint32_t v = 0; v < total; ++v) {
        if (parents[v] == NONE) {
            if (root != NONE) return -1;
            root = v;
        }
    }
    if (root == NONE) return -1;
    memcpy(collapsed, nodes, sizeof(onode) * total);
    uint8_t* mark = table ? table : (uint8_t*)malloc(total);
    memset(mark, 0, total);

    /* breadth-first order from the root, processed backwards: children first */
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * total);
    uint32_t head = 0, tail = 0;
    order[tail++] = root;
    while (head < tail) {
        uint32_t v = order[head++];
        if (v < I) {
            if (nodes[v].a >= total || nodes[v].b >= total || tail + 2 > total) {
                free(order);
                if (!table) free(mark);
                return -1;
            }
            order[tail++] = nodes[v].a;
            order[tail++] = nodes[v].b;
        }
    }
    if (tail != total) {
        free(order);
        if (!table) free(mark);
        return -1;
    }
    for (uint32_t k = total; k-- > 0;) {
        uint32_t v = order[k];
        if (v >= I) {
            costs[v] = ct * aabb_area(&nodes[v]);
            continue;
        }
        uint32_t l = nodes[v].a, r = nodes[v].b;
        uint32_t count = counts[l] + counts[r];
        counts[v] = count;
        float area = aabb_area(&nodes[v]);
        float node_cost = ci * area + costs[l] + costs[r];
        float leaf_cost = ct * area * (float)count;
        if (leaf_cost < node_cost && count <= MAX_LEAF_SIZE) {
            costs[v] = leaf_cost;
            mark[v] = mark[l] = mark[r] = 1;
            collapsed[l].nprim = COLLAPSED;
            collapsed[r].nprim = COLLAPSED;
        } else {
            costs[v] = node_cost;
        }
    }

    /* emit: leaves of the collapsed tree, slices in node-index order */
    int64_t offset = 0;
    uint32_t stack[64];
    for (uint32_t c = 0; c < total; ++c) {
        uint32_t p = parents[c];
        if (p == NONE || mark[p] || !mark[c] || counts[c] <= 1) continue;
        uint32_t* buf = primitives + offset;
        int32_t position = 0;
        /* right-first stack walk of bvh_node_collapse */
        int sp = 0;
        stack[sp++] = nodes[c].a;
        stack[sp++] = nodes[c].b;
        while (sp > 0) {
            uint32_t u = stack[--sp];
            if (nodes[u].nprim > 0) {
                uint32_t data = nodes[u].a;
                int unique = 1;
                for (int32_t i = 0; i < position; ++i)
                    if (buf[i] == data) { unique = 0; break; }
                if (unique) buf[position++] = data;
            } else {
                if (sp + 2 > 64) { free(order); if (!table) free(mark); return -1; }
                stack[sp++] = nodes[u].a;
                stack[sp++] = nodes[u].b;
            }
        }
        counts[c] = (uint32_t)position;
        collapsed[c].nprim = (uint16_t)position;
        collapsed[c].a = (uint32_t)offset;
        offset += counts[c] > 0 ? counts[c] : 0;
    }
    free(order);
    if (!table) free(mark);
    return offset;
}

float oracle_bvh_sah(const onode* nodes, int32_t root_idx, float ci, float ct, const oobb* obbs,
                     uint32_t total) {
    if (total == 0) return 0.0f;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (total + 2));
    int64_t sp = 0;
    const onode* root = &nodes[root_idx];
    float sah = 0.0f;
    float root_area = obbs ? obb_area(&obbs[root_idx]) : aabb_area(root);
    if (root->nprim > 0) sah += ct * root_area * root->nprim;
    else stack[sp++] = (uint32_t)root_idx;
    while (sp > 0) {
        uint32_t idx = stack[--sp];
        const onode* node = &nodes[idx];
        float area = obbs ? obb_area(&obbs[idx]) : aabb_area(node);
        if (node->nprim > 0) {
            if (node->nprim == COLLAPSED) continue;
            if (node->nprim == 1) area = aabb_area(node);
            sah += ct * node->nprim * area;
        } else {
            sah += ci * area;
            if (sp + 2 > (int64_t)total + 2) break;
            stack[sp++] = node->a;
            stack[sp++] = node->b;
        }
    }
    free(stack);
    return sah / root_area;
}
