// devmath.hpp -- float3 arithmetic with a fixed evaluation order.
//
// The library is compiled with -ffp-contract=off and correctly rounded
// division and square root, so every expression below rounds after each
// operation, left to right, exactly as the reference sources are written.
// Two flavours exist because the reference has two:
//   dm::  src/common/algorithm/dito/dito.cpp:56-103   (CPU DiTO-14)
//         normalize = multiply by 1/length; cross y = -(u.x v.z - u.z v.x)
//   vm::  src/common/vecmath.h:520-812,1396-1425      (GPU transform)
//         normalize = divide by length; cross y = a.z b.x - a.x b.z
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace obvh {

struct f3 {
    float x, y, z;
};

__host__ __device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
// a.x*b.x + a.y*b.y + a.z*b.z, i.e. ((ax*bx) + (ay*by)) + (az*bz)
__host__ __device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// Half box area, dito.cpp:274-278 and transforms.cu:413-416.
__host__ __device__ __forceinline__ float quality(f3 l) { return l.x * l.y + l.x * l.z + l.y * l.z; }
// kr_min / kr_max, vecmath.h:1396-1397.
__host__ __device__ __forceinline__ float kmin(float a, float b) { return (a < b) ? a : b; }
__host__ __device__ __forceinline__ float kmax(float a, float b) { return (a < b) ? b : a; }

namespace dm {  // dito.cpp flavour
__host__ __device__ __forceinline__ f3 scale(float s, f3 u) { return f3{s * u.x, s * u.y, s * u.z}; }
__host__ __device__ __forceinline__ f3 cross(f3 u, f3 v) {
    return f3{u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x};
}
__host__ __device__ __forceinline__ f3 normalize(f3 u) {
    float inv = 1.0f / sqrtf(len2(u));
    return f3{u.x * inv, u.y * inv, u.z * inv};
}
// sqPtPtDist(u, v) = |v - u|^2
__host__ __device__ __forceinline__ float sqdist(f3 u, f3 v) { return len2(sub3(v, u)); }
}  // namespace dm

namespace vm {  // vecmath.h flavour
__host__ __device__ __forceinline__ f3 mul1(f3 a, float b) { return f3{a.x * b, a.y * b, a.z * b}; }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ f3 normalize(f3 a) {
    float l = sqrtf(len2(a));
    return f3{a.x / l, a.y / l, a.z / l};
}
// kr_vdistance3sqr(a, b) = |b - a|^2
__host__ __device__ __forceinline__ float dist2(f3 a, f3 b) { return len2(sub3(b, a)); }
// kr_vdistance_to_inf_edge3, vecmath.h:524-530
__host__ __device__ __forceinline__ float dist_inf_edge(f3 q, f3 p0, f3 v) {
    f3 u0 = sub3(q, p0);
    float t = dot3(v, u0);
    float l2 = len2(v);
    return len2(u0) - t * t / l2;
}
__host__ __device__ __forceinline__ f3 vmin3(f3 a, f3 b) { return f3{kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)}; }
__host__ __device__ __forceinline__ f3 vmax3(f3 a, f3 b) { return f3{kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)}; }
}  // namespace vm

// The 7 slab projections, in the reference order (dito.cpp:115-171,
// transforms.cu:224-251): x, y, z, x+y+z, x+y-z, x-y+z, x-y-z.
__host__ __device__ __forceinline__ void slab7(f3 p, float pr[7]) {
    float s = p.x + p.y;
    float t = p.x - p.y;
    pr[0] = p.x;
    pr[1] = p.y;
    pr[2] = p.z;
    pr[3] = s + p.z;
    pr[4] = s - p.z;
    pr[5] = t + p.z;
    pr[6] = t - p.z;
}

// Order-preserving map float -> int32 (for integer atomic min/max).
__host__ __device__ __forceinline__ int32_t f2ord(float f) {
    int32_t i = __builtin_bit_cast(int32_t, f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__host__ __device__ __forceinline__ float ord2f(int32_t i) {
    return __builtin_bit_cast(float, i >= 0 ? i : (i ^ 0x7FFFFFFF));
}

}  // namespace obvh
