// chen.hpp -- entry points of chen.hip.
#pragma once

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

int stream_signature_batch(void* out, const void* increments, const int64_t* offsets,
                           int64_t n_streams, int32_t width, int32_t depth, rpy_dtype dtype,
                           void* stream);

// Prefix length P the specialised kernel uses for (width, depth, dtype), or
// 0 when the shape runs on the generic kernel.
int chen_prefix_length(int32_t width, int32_t depth, int dtype_is_f64);

// Host run of the lane decomposition for one double-precision stream.
int signature_host_emulation(double* out, const double* incr, int64_t len, int32_t width,
                             int32_t depth, int32_t prefix);

}  // namespace rpy_amd
