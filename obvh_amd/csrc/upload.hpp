// upload.hpp -- internal interface of the batch gather (upload.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "obvh/obvh.h"

namespace obvh {

kr_error gather_sets_launch(const float* const* sets, const uint64_t* offs, uint32_t nsets, float* soa, uint64_t row,
                            uint64_t total, hipStream_t stream);

}  // namespace obvh
