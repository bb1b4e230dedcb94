// common.hpp -- shared device helpers: wave64 reductions, scoped fences and
// atomics, workspace carving, host error plumbing.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "devmath.hpp"

namespace obvh {

constexpr int kWave = 64;

// Agent (device) scope release/acquire.  On gfx950 these lower to an L2
// write-back / invalidate, which makes data written by one XCD visible to
// another; used only at set/tile completion points, never per element.
__device__ __forceinline__ void fence_release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ uint32_t atomic_add_acqrel(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / kWave); }

// Butterfly reductions over the 64 lanes (every lane gets the result).
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v = fminf(v, __shfl_xor(v, m, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        uint32_t o = (uint32_t)__shfl_xor((int)v, m, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Integer min / max over the 64 lanes, result uniform (in every lane).  Four
// DPP steps leave each 16-lane row's result in all of its lanes (xor 1, xor 2,
// half-row mirror, row mirror); the four rows are combined on the scalar
// unit.  Every lane of the wave must be active.
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
    const int32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
    const int32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}

// (value, key) lexicographic reductions: smaller value wins (MIN) or larger
// value wins (MAX); equal values (== compare, so -0 == +0) keep the smaller
// key.  Associative and commutative, so the result does not depend on the
// reduction order.
template <bool MIN>
__device__ __forceinline__ bool pair_better(float v, uint32_t k, float ov, uint32_t ok) {
    if (MIN) return (ov < v) || (ov == v && ok < k);
    return (ov > v) || (ov == v && ok < k);
}
template <bool MIN>
__device__ __forceinline__ void wave_pair(float& v, uint32_t& k) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        float ov = __shfl_xor(v, m, kWave);
        uint32_t ok = (uint32_t)__shfl_xor((int)k, m, kWave);
        if (pair_better<MIN>(v, k, ov, ok)) { v = ov; k = ok; }
    }
}

// Bump allocator over a caller-provided workspace (256 B aligned pieces).
struct Carver {
    char* base;
    size_t used;
    size_t cap;
    template <class T>
    T* take(size_t count) {
        size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T);
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
    bool ok() const { return used <= cap; }
};

}  // namespace obvh

// Host-side error helpers (defined in abi.hip).
void obvh_set_error(const char* msg);
const char* obvh_hip_error(hipError_t e, const char* where);

#define OBVH_HIP_TRY(expr)                                           \
    do {                                                             \
        hipError_t e__ = (expr);                                     \
        if (e__ != hipSuccess) return obvh_hip_error(e__, #expr);    \
    } while (0)
