// upload.hip -- gather of many host point sets into one SoA batch.
//
// A serving batch holds thousands of sets of a few thousand to a few
// hundred thousand points, each in its own pinned host buffer.  One copy per
// set costs a fixed setup per copy, which dominates at these sizes; here one
// kernel reads every set straight from pinned host memory over PCIe (the
// buffers are device-accessible) and writes the batch's (3, row) SoA array:
// one lane per point, consecutive lanes read consecutive floats of a set.
#include "common.hpp"
#include "upload.hpp"

namespace obvh {

__global__ __launch_bounds__(256) void gather_sets(const float* const* __restrict__ sets,
                                                   const uint64_t* __restrict__ offs, uint32_t nsets,
                                                   float* __restrict__ soa, uint64_t row, uint64_t total) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256u) {
        // the set holding point i: the largest s with offs[s] <= i (empty sets are skipped)
        uint32_t lo = 0, hi = nsets;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offs[mid] <= i) lo = mid;
            else hi = mid;
        }
        const uint64_t n = offs[lo + 1] - offs[lo], j = i - offs[lo];
        const float* p = sets[lo];
        soa[i] = p[j];
        soa[row + i] = p[n + j];
        soa[2 * row + i] = p[2 * n + j];
    }
}

kr_error gather_sets_launch(const float* const* sets, const uint64_t* offs, uint32_t nsets, float* soa, uint64_t row,
                            uint64_t total, hipStream_t stream) {
    if (nsets == 0 || total == 0) return nullptr;
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(gather_sets, dim3((uint32_t)blocks), dim3(256), 0, stream, sets, offs, nsets, soa, row, total);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
