// cpu_dito.cpp -- DiTO-14 on the host: kr_points_obb, the reference's CPU
// fit (src/common/geometry.cpp:232-252 over DiTO::DiTO_14<float>,
// src/common/algorithm/dito/dito.cpp:478-538).  The bench integrator plugin
// times it beside kr_cuda_points_obb, as the reference's bench does.
//
// Floating point as the reference evaluates it: float throughout, no
// contraction (built with -ffp-contract=off), every expression left to
// right: dot = (ax bx + ay by) + az bz, normalise = multiply by 1 / length,
// cross with the negated y difference, quality = (lx ly + lx lz) + ly lz.
#include <math.h>

#include <array>
#include <chrono>

#include "obvh/obvh.h"

namespace {

struct V3 {
    float x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scaled(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x}; }
inline V3 unit(V3 a) {
    const float r = 1.0f / sqrtf(len2(a));
    return {a.x * r, a.y * r, a.z * r};
}
inline float half_area(V3 l) { return l.x * l.y + l.x * l.z + l.y * l.z; }

constexpr float kEps = 0.000001f;

struct Box {
    V3 mid, v0, v1, v2, ext;
};

// The fit of one point set; the point list `sel` (at most 14 points) is
// what the axis search runs over.
class Dito14 {
  public:
    Dito14(const kr_vec3* p, int n) : p_(p), n_(n) {}

    Box fit() {
        if (n_ <= 0) return aligned({0, 0, 0}, {0, 0, 0});
        slabs();
        const V3 amid = {(lo_[0] + hi_[0]) * 0.5f, (lo_[1] + hi_[1]) * 0.5f, (lo_[2] + hi_[2]) * 0.5f};
        const V3 alen = {hi_[0] - lo_[0], hi_[1] - lo_[1], hi_[2] - lo_[2]};
        best_ = half_area(alen);
        axes_ = {V3{1, 0, 0}, V3{0, 1, 0}, V3{0, 0, 1}};
        if (n_ > 14) {
            for (int k = 0; k < 7; ++k) {
                sel_[k] = at(ilo_[k]);
                sel_[7 + k] = at(ihi_[k]);
            }
            ns_ = 14;
        } else {
            for (int i = 0; i < n_; ++i) sel_[i] = at(i);
            ns_ = n_;
        }
        // the farthest pair of slab extremes (first on ties)
        int pk = 0;
        float dk = len2(at(ilo_[0]) - at(ihi_[0]));
        for (int k = 1; k < 7; ++k) {
            const float d = len2(at(ilo_[k]) - at(ihi_[k]));
            if (d > dk) {
                dk = d;
                pk = k;
            }
        }
        const V3 p0 = at(ilo_[pk]), p1 = at(ihi_[pk]);
        if (len2(p1 - p0) < kEps) return aligned(amid, alen);
        const V3 e0 = unit(p0 - p1);
        // the point farthest from the line through p0 along e0
        int ip = 0;
        float dp = line_dist2(sel_[0], p0, e0);
        for (int i = 1; i < ns_; ++i) {
            const float d = line_dist2(sel_[i], p0, e0);
            if (d > dp) {
                dp = d;
                ip = i;
            }
        }
        if (dp < kEps) return line_box(e0);
        const V3 p2 = sel_[ip];
        const V3 e1 = unit(p1 - p2), e2 = unit(p2 - p0);
        const V3 nrm = unit(cross(e1, e0));
        try_triangle(nrm, e0, e1, e2);
        // the ditetrahedron: apexes at the extremes along the normal
        float plo = dot(sel_[0], nrm), phi = plo;
        int ilo = 0, ihi = 0;
        for (int i = 1; i < ns_; ++i) {
            const float q = dot(sel_[i], nrm);
            if (q < plo) {
                plo = q;
                ilo = i;
            }
            if (q > phi) {
                phi = q;
                ihi = i;
            }
        }
        const float base = dot(p0, nrm);
        if (phi - kEps > base) apex(sel_[ihi], p0, p1, p2, e0, e1, e2);
        if (plo + kEps < base) apex(sel_[ilo], p0, p1, p2, e0, e1, e2);
        // the chosen axes over every point; the box only if strictly smaller
        V3 mn, mx;
        range_all(axes_[0], mn.x, mx.x);
        range_all(axes_[1], mn.y, mx.y);
        range_all(axes_[2], mn.z, mx.z);
        const V3 blen = mx - mn;
        if (half_area(blen) < half_area(alen)) return oriented(axes_[0], axes_[1], axes_[2], mn, mx, blen);
        return aligned(amid, alen);
    }

  private:
    V3 at(int i) const { return {p_[i].x, p_[i].y, p_[i].z}; }

    static void proj7(V3 p, float (&s)[7]) {
        const float a = p.x + p.y, b = p.x - p.y;
        s[0] = p.x;
        s[1] = p.y;
        s[2] = p.z;
        s[3] = a + p.z;
        s[4] = a - p.z;
        s[5] = b + p.z;
        s[6] = b - p.z;
    }

    // Extremes of the 7 slab directions, first index on ties.
    void slabs() {
        float s[7];
        proj7(at(0), s);
        for (int k = 0; k < 7; ++k) {
            lo_[k] = hi_[k] = s[k];
            ilo_[k] = ihi_[k] = 0;
        }
        for (int i = 1; i < n_; ++i) {
            proj7(at(i), s);
            for (int k = 0; k < 7; ++k) {
                if (s[k] < lo_[k]) {
                    lo_[k] = s[k];
                    ilo_[k] = i;
                }
                if (s[k] > hi_[k]) {
                    hi_[k] = s[k];
                    ihi_[k] = i;
                }
            }
        }
    }

    static float line_dist2(V3 q, V3 p0, V3 dir) {
        const V3 u = q - p0;
        const float t = dot(dir, u);
        return len2(u) - t * t / len2(dir);
    }

    void range_sel(V3 d, float& mn, float& mx) const {
        mn = mx = dot(sel_[0], d);
        for (int i = 1; i < ns_; ++i) {
            const float q = dot(sel_[i], d);
            if (q < mn) mn = q;
            if (q > mx) mx = q;
        }
    }
    void range_all(V3 d, float& mn, float& mx) const {
        mn = mx = dot(at(0), d);
        for (int i = 1; i < n_; ++i) {
            const float q = dot(at(i), d);
            if (q < mn) mn = q;
            if (q > mx) mx = q;
        }
    }

    // The three axis triples (e_k, n, e_k x n) of one triangle.
    void try_triangle(V3 n, V3 a, V3 b, V3 c) {
        const std::array<V3, 3> e = {a, b, c};
        float nmn, nmx;
        range_sel(n, nmn, nmx);
        for (int k = 0; k < 3; ++k) {
            const V3 m = cross(e[k], n);
            float emn, emx, mmn, mmx;
            range_sel(e[k], emn, emx);
            range_sel(m, mmn, mmx);
            const float q = half_area({emx - emn, nmx - nmn, mmx - mmn});
            if (q < best_) {
                best_ = q;
                axes_ = {e[k], n, m};
            }
        }
    }

    // The three faces of the tetrahedron on apex q over the base triangle.
    void apex(V3 q, V3 p0, V3 p1, V3 p2, V3 e0, V3 e1, V3 e2) {
        const V3 f0 = unit(q - p0), f1 = unit(q - p1), f2 = unit(q - p2);
        const V3 n0 = unit(cross(f1, e0)), n1 = unit(cross(f2, e1)), n2 = unit(cross(f0, e2));
        try_triangle(n0, e0, f1, f0);
        try_triangle(n1, e1, f2, f1);
        try_triangle(n2, e2, f0, f2);
    }

    // All points on a line: a box around the line direction.
    Box line_box(V3 u) const {
        V3 r = u;
        if (fabsf(u.x) > fabsf(u.y) && fabsf(u.x) > fabsf(u.z)) r.x = 0;
        else if (fabsf(u.y) > fabsf(u.z)) r.y = 0;
        else r.z = 0;
        if (len2(r) < kEps) r = {1, 1, 1};
        const V3 v = unit(cross(u, r));
        const V3 w = unit(cross(u, v));
        V3 mn, mx;
        range_all(u, mn.x, mx.x);
        range_all(v, mn.y, mx.y);
        range_all(w, mn.z, mx.z);
        return oriented(u, v, w, mn, mx, mx - mn);
    }

    static Box oriented(V3 a, V3 b, V3 c, V3 mn, V3 mx, V3 len) {
        const V3 h = scaled(0.5f, mn + mx);
        V3 mid = scaled(h.x, a);
        mid = mid + scaled(h.y, b);
        mid = mid + scaled(h.z, c);
        return {mid, a, b, c, scaled(0.5f, len)};
    }
    static Box aligned(V3 mid, V3 len) { return {mid, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, scaled(0.5f, len)}; }

    const kr_vec3* p_;
    int n_;
    float lo_[7], hi_[7];
    int ilo_[7], ihi_[7];
    std::array<V3, 14> sel_{};
    int ns_ = 0;
    float best_ = 0;
    std::array<V3, 3> axes_{};
};

inline float floor_ext(float e) { return (e < 0.001f) ? 0.001f : e; }
inline kr_vec3 kv(V3 a) { return kr_vec3{a.x, a.y, a.z}; }

}  // namespace

extern "C" kr_obb3 kr_points_obb(const kr_vec3* points, kr_size count, kr_scalar* elapsed_time) {
    const auto t0 = std::chrono::steady_clock::now();
    Box b = Dito14(points, (int)count).fit();
    const auto t1 = std::chrono::steady_clock::now();
    if (elapsed_time) *elapsed_time = std::chrono::duration<float, std::milli>(t1 - t0).count();
    kr_obb3 o;
    o.mid = kv(b.mid);
    o.v0 = kv(b.v0);
    o.v1 = kv(b.v1);
    o.v2 = kv(b.v2);
    o.ext = kr_vec3{floor_ext(b.ext.x), floor_ext(b.ext.y), floor_ext(b.ext.z)};
    return o;
}
