// upload.hip -- gather of many host point sets into one SoA batch.
//
// A serving batch holds thousands of sets of a few thousand to a few
// hundred thousand points, each in its own pinned host buffer.  One copy per
// set costs a fixed setup per copy, which dominates at these sizes; here one
// kernel reads every set straight from pinned host memory over PCIe (the
// buffers are device-accessible) and writes the batch's (3, row) SoA array:
// one lane per 4 points, consecutive lanes read consecutive 16-byte pieces
// of a set's rows.
#include "common.hpp"
#include "upload.hpp"

namespace obvh {

// 16-byte vectors at 4-byte alignment: a set's rows and the batch's rows
// start at any point index (global_load/store_dwordx4 accept that).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ uint32_t set_of(const uint64_t* __restrict__ offs, uint32_t nsets, uint64_t i) {
    // the largest s with offs[s] <= i (empty sets are skipped)
    uint32_t lo = 0, hi = nsets;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One lane per 4 consecutive points of the batch: one 16-byte read per row
// of the set when the 4 points lie in one set, point by point otherwise.
__global__ __launch_bounds__(256) void gather_sets(const float* const* __restrict__ sets,
                                                   const uint64_t* __restrict__ offs, uint32_t nsets,
                                                   float* __restrict__ soa, uint64_t row, uint64_t total) {
    const uint64_t nq = (total + 3) / 4;
    for (uint64_t q = (uint64_t)blockIdx.x * 256u + threadIdx.x; q < nq; q += (uint64_t)gridDim.x * 256u) {
        const uint64_t i = 4 * q;
        const uint32_t s = set_of(offs, nsets, i);
        const uint64_t s0 = offs[s], n = offs[s + 1] - s0, j = i - s0;
        if (j + 4 <= n) {
            const float* p = sets[s];
            const f4u x = *reinterpret_cast<const f4u*>(p + j);
            const f4u y = *reinterpret_cast<const f4u*>(p + n + j);
            const f4u z = *reinterpret_cast<const f4u*>(p + 2 * n + j);
            *reinterpret_cast<f4u*>(soa + i) = x;
            *reinterpret_cast<f4u*>(soa + row + i) = y;
            *reinterpret_cast<f4u*>(soa + 2 * row + i) = z;
        } else {
            for (uint64_t e = i; e < i + 4 && e < total; ++e) {
                const uint32_t t = set_of(offs, nsets, e);
                const uint64_t t0 = offs[t], m = offs[t + 1] - t0, k = e - t0;
                const float* p = sets[t];
                soa[e] = p[k];
                soa[row + e] = p[m + k];
                soa[2 * row + e] = p[2 * m + k];
            }
        }
    }
}

kr_error gather_sets_launch(const float* const* sets, const uint64_t* offs, uint32_t nsets, float* soa, uint64_t row,
                            uint64_t total, hipStream_t stream) {
    if (nsets == 0 || total == 0) return nullptr;
    uint64_t blocks = ((total + 3) / 4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(gather_sets, dim3((uint32_t)blocks), dim3(256), 0, stream, sets, offs, nsets, soa, row, total);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
