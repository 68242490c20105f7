// chen.hpp -- entry points of chen.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

// Stream-ordered temporary allocation from the device's default memory
// pool.  On first use per device the pool's release threshold is raised so
// freed blocks stay cached and later calls do not go back to the driver
// (rpy_release_workspace trims the pool).
hipError_t workspace_alloc(void** ptr, size_t bytes, hipStream_t stream);

int stream_signature_batch(void* out, const void* increments, const int64_t* offsets,
                           int64_t n_streams, int32_t width, int32_t depth, rpy_dtype dtype,
                           void* stream);

// The same with every temporary in a fixed per-device scratch array; no
// allocation.  RPY_ERR_SHAPE when the batch needs more scratch than it holds.
int stream_signature_batch_noalloc(void* out, const void* increments, const int64_t* offsets,
                                   int64_t n_streams, int32_t width, int32_t depth,
                                   rpy_dtype dtype, void* stream);
// Bytes of that scratch array.
int64_t stream_signature_scratch_bytes();

// Same with explicit row ranges: stream s owns rows [begins[s], ends[s]).
int stream_signature_ranges(void* out, const void* increments, const int64_t* begins,
                            const int64_t* ends, int64_t n_streams, int32_t width, int32_t depth,
                            rpy_dtype dtype, void* stream);

// Prefix length P of the variant a call of (width, depth, dtype) runs, or 0
// when the shape runs on the generic kernel.
int chen_prefix_length(int32_t width, int32_t depth, int dtype_is_f64);

// Flags of a signature launch.  kSkipEmpty: streams of no rows write
// nothing (their output rows are left as they are); kSortedDesc: the order
// is by length, longest first, so the first empty stream ends the queue.
constexpr int kSkipEmpty = 1;
constexpr int kSortedDesc = 2;

// Launches one specialised signature kernel (chen_kernels.hpp) on a sorted
// stream order; `counter` is the queue counter, cleared by the caller.
using SigLauncher = int (*)(void* out, const void* incr, const int64_t* begins,
                            const int64_t* ends, const int32_t* order, int64_t n_streams,
                            const DevBasis& b, uint32_t* counter, int flags, hipStream_t st);

// One compiled signature kernel variant of one shape and dtype.
struct SigVariant {
    int32_t width, depth;
    rpy_dtype dtype;
    rpy_kernel_variant v;
    SigLauncher launch;
};

struct SigVariantList {
    const SigVariant* items;
    size_t n;
};

// Extra variants compiled for the tuned shapes (chen_tune_W_D.hip).
SigVariantList sig_tune_variants_4_6();
SigVariantList sig_tune_variants_5_5();
SigVariantList sig_tune_variants_6_5();
SigVariantList sig_tune_variants_8_4();

// rpy_tune_* of the signature forward (op RPY_OP_SIGNATURE).
int32_t sig_variants(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
                     int32_t max);
int sig_select(int32_t width, int32_t depth, rpy_dtype dtype, const rpy_kernel_variant* v);
int sig_config(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
               int32_t* source);

// Block-compensated float32 signatures (chen_block.hip): float32 segment
// signatures of at most block_segment_slots() segments per stream, merged
// per stream with compensated float32 sums.
struct BlockPlan {
    int64_t n_sub;     // streams per round (the segment buffer holds n_sub streams)
    int64_t slots;     // n_sub * block_segment_slots()
    size_t scan_bytes, c_bytes, o_bytes, r_bytes, seg_bytes;
    size_t total;      // workspace bytes of one call
};
bool block_merge_supported(int32_t width, int32_t depth);
int64_t block_segment_slots();
int block_plan(int64_t n_sub, int64_t tsize, BlockPlan& p, hipStream_t st);
// seg_launch: the float32 signature variant run on the segments; T: nominal
// rows per segment.
int block_run(float* out, const float* increments, const int64_t* begins, const int64_t* ends,
              int64_t n_streams, const DevBasis& b, int64_t T, SigLauncher seg_launch,
              const BlockPlan& p, char* ws, hipStream_t st);

// Host run of the lane decomposition for one double-precision stream.
int signature_host_emulation(double* out, const double* incr, int64_t len, int32_t width,
                             int32_t depth, int32_t prefix);

}  // namespace rpy_amd

namespace rpy_amd {

// dyadic.hip
int64_t dyadic_cache_elements(int64_t n_streams, int32_t width, int32_t depth, int32_t levels);
int dyadic_cache_build(void* cache, const void* increments, const double* params,
                       const int64_t* offsets, int64_t n_streams, int32_t width, int32_t depth,
                       int32_t resolution, int32_t levels, rpy_dtype dtype, void* stream);
int dyadic_cache_query(void* out, const void* cache, const void* increments, const double* params,
                       const int64_t* offsets, int64_t n_streams, int32_t width, int32_t depth,
                       int32_t resolution, int32_t levels, const int64_t* q_stream,
                       const double* q_inf, const double* q_sup, int64_t n_queries,
                       rpy_dtype dtype, void* stream);
int dyadic_cover_host(int64_t A, int64_t B, int r, int64_t* ks, int* ns, int max_pieces);

// backward.hip
int vjp_host_emulation(double* grad, const double* incr, int64_t len, int32_t width,
                       int32_t depth, int32_t prefix, const double* sig, const double* gsig);
int stream_signature_vjp(void* grad, const void* increments, const int64_t* offsets,
                         int64_t n_streams, int32_t width, int32_t depth, const void* signatures,
                         const void* grad_signatures, rpy_dtype dtype, void* stream);
// rpy_tune_* of the signature backward (op RPY_OP_SIGNATURE_VJP).
int32_t vjp_variants(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
                     int32_t max);
int vjp_select(int32_t width, int32_t depth, rpy_dtype dtype, const rpy_kernel_variant* v);
int vjp_config(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
               int32_t* source);
int dense_ft_adjoint_right_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                               const rpy_tensor_basis* basis, int64_t batch, int64_t os,
                               int64_t ps, int64_t as, void* stream);

// log_vjp.hip
int dense_ft_log_vjp(void* out, const void* arg, const void* cotangent, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t as,
                     int64_t cs, void* stream);

// gram.hip
int gram_bf16x3(float* G, const float* S, int64_t n, int64_t k, int64_t lds, int64_t ldg,
                void* stream);

}  // namespace rpy_amd

namespace rpy_amd {

// shuffle.hip
int dense_st_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                 const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ls,
                 int64_t rs, void* stream);
int dense_st_adj_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ps,
                     int64_t as, void* stream);

}  // namespace rpy_amd
