// capi.hip -- the extern "C" boundary of libroughpy_amd.so.
//
// Thin: argument checks, error reporting (status code + thread-local
// message, the C analogue of the TypeError/ValueError the reference
// bindings raise in roughpy/compute/_src/py_ternary_array_fn.hpp:83-160 and
// call_config.cpp:11-31), and dispatch to the kernels.

#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/roughpy_amd.h"
#include "chen.hpp"
#include "dense_ops.hpp"
#include "strided.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {
thread_local std::string g_last_error;

__global__ void widen_kernel(double* __restrict__ out, const float* __restrict__ in, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        out[i] = double(in[i]);
}

// Coefficients of a tensor operand (0 for an invalid basis: the entry
// point's own checks report it).
int64_t basis_size(const rpy_tensor_basis* b)
{
    if (!b || b->width < 1 || b->depth < 0 || b->depth > kMaxDepth) return 0;
    const int64_t n = tensor_size_checked(b->width, b->depth);
    return n < 0 ? 0 : n;
}

size_t elem_size(rpy_dtype dt) { return dt == RPY_F64 ? 8 : 4; }

uint32_t blocks_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return uint32_t(b);
}
}  // namespace

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

hipError_t workspace_alloc(void** ptr, size_t bytes, hipStream_t stream)
{
    constexpr int kMaxDevices = 64;
    static std::atomic<bool> kept[kMaxDevices];
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev >= 0 && dev < kMaxDevices && !kept[dev].load(std::memory_order_acquire)) {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t threshold = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
        kept[dev].store(true, std::memory_order_release);
    }
    return hipMallocAsync(ptr, bytes, stream);
}

int make_basis(const rpy_tensor_basis* tb, DevBasis& out)
{
    if (!tb) return fail(RPY_ERR_INVALID_ARGUMENT, "null basis");
    if (tb->width < 1) return fail(RPY_ERR_BASIS, "width must be positive, got %d", tb->width);
    if (tb->depth < 0 || tb->depth > kMaxDepth)
        return fail(RPY_ERR_BASIS, "depth must be in [0, %d], got %d", kMaxDepth, tb->depth);
    out.width = tb->width;
    out.depth = tb->depth;
    out.db[0] = 0;
    for (int k = 0; k <= tb->depth; ++k) {
        if (out.db[k] > (INT64_C(1) << 40) / tb->width)
            return fail(RPY_ERR_BASIS, "tensor algebra (width %d, depth %d) is too large",
                        tb->width, tb->depth);
        out.db[k + 1] = 1 + int64_t(tb->width) * out.db[k];
    }
    // the kernels index levels with 32-bit offsets
    if (out.level_size(tb->depth) >= (INT64_C(1) << 31))
        return fail(RPY_ERR_BASIS, "top level of (width %d, depth %d) exceeds 2^31 entries",
                    tb->width, tb->depth);
    if (tb->degree_begin) {
        for (int k = 0; k <= tb->depth + 1; ++k)
            if (tb->degree_begin[k] != out.db[k])
                return fail(RPY_ERR_BASIS, "degree_begin[%d] = %lld does not match the dense layout",
                            k, (long long)tb->degree_begin[k]);
    }
    return RPY_OK;
}

}  // namespace rpy_amd

using namespace rpy_amd;

extern "C" {

const char* rpy_last_error(void) { return g_last_error.c_str(); }

int32_t rpy_version(void) { return 1; }

int32_t rpy_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rpy_release_workspace(void)
{
    // Temporaries are stream-ordered allocations returned at the end of
    // each call; trim the device pool back to the system.
    int dev = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    hipMemPool_t pool;
    RPY_HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, dev));
    RPY_HIP_CHECK(hipMemPoolTrimTo(pool, 0));
    return RPY_OK;
}

int64_t rpy_tensor_size(int32_t width, int32_t depth) { return tensor_size_checked(width, depth); }

int64_t rpy_lie_basis_size(int32_t width, int32_t depth)
{
    if (width < 1 || depth < 0) return -1;
    return lie_size(width, depth);
}

int rpy_lie_basis_build(int32_t width, int32_t depth, int64_t* degree_begin, int64_t* data)
{
    if (width < 1 || depth < 0) return fail(RPY_ERR_BASIS, "invalid width/depth");
    if (!degree_begin || !data) return fail(RPY_ERR_INVALID_ARGUMENT, "null output array");
    return build_hall(width, depth, degree_begin, data);
}

int rpy_check_lie_basis_data(const rpy_lie_basis* basis)
{
    if (!basis || !basis->degree_begin || !basis->data)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null basis data");
    return check_hall(basis);
}

int rpy_l2t_matrix(const rpy_lie_basis* basis, int64_t* nnz, int64_t* offsets, int64_t* indices,
                   double* data)
{
    if (!basis || !basis->degree_begin || !basis->data || !nnz)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_hall(basis);
    if (rc) return rc;
    return build_l2t(basis, nnz, offsets, indices, data);
}

int rpy_t2l_matrix(const rpy_lie_basis* basis, int64_t* nnz, int64_t* offsets, int64_t* indices,
                   double* data)
{
    if (!basis || !basis->degree_begin || !basis->data || !nnz)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_hall(basis);
    if (rc) return rc;
    return build_t2l(basis, nnz, offsets, indices, data);
}

int rpy_dense_ft_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, rpy_degree_bounds ob, rpy_degree_bounds lb,
                     rpy_degree_bounds rb, int64_t batch, int64_t os, int64_t ls, int64_t rs,
                     int64_t oi, int64_t li, int64_t ri, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)lhs, ls, li, n, false},
                                 {(void*)rhs, rs, ri, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_fma(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, ob, lb, rb, batch,
                            v.stride[0], v.stride[1], v.stride[2], stream);
    });
}

int rpy_dense_ft_inplace_mul(void* lhs, const void* rhs, rpy_dtype dtype,
                             const rpy_tensor_basis* basis, rpy_degree_bounds lb,
                             rpy_degree_bounds rb, int64_t batch, int64_t ls, int64_t rs,
                             int64_t li, int64_t ri, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[2] = {{lhs, ls, li, n, true}, {(void*)rhs, rs, ri, n, false}};
    return with_unit_inc(ops, 2, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_inplace_mul(v.ptr[0], v.ptr[1], dtype, basis, lb, rb, batch, v.stride[0],
                                    v.stride[1], stream);
    });
}

int rpy_dense_ft_antipode(void* out, const void* arg, rpy_dtype dtype,
                          const rpy_tensor_basis* basis, rpy_degree_bounds ob,
                          rpy_degree_bounds ab, int64_t batch, int64_t os, int64_t as,
                          int64_t oi, int64_t ai, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[2] = {{out, os, oi, n, true}, {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 2, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_antipode(v.ptr[0], v.ptr[1], dtype, basis, ob, ab, batch, v.stride[0],
                                 v.stride[1], stream);
    });
}

int rpy_dense_ft_adjoint_left_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                                  const rpy_tensor_basis* basis, rpy_degree_bounds ob,
                                  rpy_degree_bounds pb, rpy_degree_bounds ab, int64_t batch,
                                  int64_t os, int64_t ps, int64_t as, int64_t oi, int64_t pi,
                                  int64_t ai, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)op, ps, pi, n, false},
                                 {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_adjoint_left_mul(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, ob, pb, ab,
                                         batch, v.stride[0], v.stride[1], v.stride[2], stream);
    });
}

int rpy_dense_st_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ls,
                     int64_t rs, int64_t oi, int64_t li, int64_t ri, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)lhs, ls, li, n, false},
                                 {(void*)rhs, rs, ri, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_st_fma(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, batch, v.stride[0],
                            v.stride[1], v.stride[2], stream);
    });
}

int rpy_dense_st_adj_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                         const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ps,
                         int64_t as, int64_t oi, int64_t pi, int64_t ai, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)op, ps, pi, n, false},
                                 {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_st_adj_mul(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, batch, v.stride[0],
                                v.stride[1], v.stride[2], stream);
    });
}

int rpy_dense_ft_exp(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                     rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                     int64_t as, int64_t oi, int64_t ai, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[2] = {{out, os, oi, n, true}, {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 2, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_exp(v.ptr[0], v.ptr[1], dtype, basis, ob, ab, batch, v.stride[0],
                            v.stride[1], stream);
    });
}

int rpy_dense_ft_fmexp(void* out, const void* mul, const void* expo, rpy_dtype dtype,
                       const rpy_tensor_basis* basis, rpy_degree_bounds ob,
                       rpy_degree_bounds mb, rpy_degree_bounds eb, int64_t batch, int64_t os,
                       int64_t ms, int64_t es, int64_t oi, int64_t mi, int64_t ei, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)mul, ms, mi, n, false},
                                 {(void*)expo, es, ei, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_fmexp(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, ob, mb, eb, batch,
                              v.stride[0], v.stride[1], v.stride[2], stream);
    });
}

int rpy_dense_ft_log(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                     rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                     int64_t as, int64_t oi, int64_t ai, void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[2] = {{out, os, oi, n, true}, {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 2, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_log(v.ptr[0], v.ptr[1], dtype, basis, ob, ab, batch, v.stride[0],
                            v.stride[1], stream);
    });
}

int rpy_dense_ft_log_vjp(void* out, const void* arg, const void* cotangent, rpy_dtype dtype,
                         const rpy_tensor_basis* basis, int64_t batch, int64_t out_stride,
                         int64_t arg_stride, int64_t cotangent_stride, void* stream)
{
    return dense_ft_log_vjp(out, arg, cotangent, dtype, basis, batch, out_stride, arg_stride,
                            cotangent_stride, stream);
}

int rpy_dense_sparse_map(void* out, const void* arg, const rpy_sparse_matrix* matrix,
                         rpy_dtype dtype, double scale, int64_t batch, int64_t os, int64_t as,
                         int64_t oi, int64_t ai, void* stream)
{
    if (!matrix) return fail(RPY_ERR_INVALID_ARGUMENT, "null matrix");
    const DenseOperand ops[2] = {{out, os, oi, matrix->rows, true},
                                 {(void*)arg, as, ai, matrix->cols, false}};
    return with_unit_inc(ops, 2, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_sparse_map(v.ptr[0], v.ptr[1], matrix, dtype, scale, batch, v.stride[0],
                                v.stride[1], stream);
    });
}

int rpy_stream_signature_batch(void* out, const void* increments, const int64_t* offsets,
                               int64_t n_streams, int32_t width, int32_t depth, rpy_dtype dtype,
                               void* stream)
{
    return stream_signature_batch(out, increments, offsets, n_streams, width, depth, dtype,
                                  stream);
}

int rpy_stream_signature_batch_noalloc(void* out, const void* increments,
                                       const int64_t* offsets, int64_t n_streams, int32_t width,
                                       int32_t depth, rpy_dtype dtype, void* stream)
{
    return stream_signature_batch_noalloc(out, increments, offsets, n_streams, width, depth,
                                          dtype, stream);
}

namespace {

// out = t2l(log(sig)) for a batch of signatures (float32 or float64): one
// fused kernel for the fixed shapes, else the float64 log (widening float32
// first where no fixed-shape log exists) and the projection.
int log_to_lie(double* out, const void* sig, rpy_dtype dtype, int32_t width, int32_t depth,
               const rpy_sparse_matrix* t2l, int64_t batch, int64_t lsize, int64_t tsize,
               int64_t os, int64_t ss, hipStream_t st)
{
    void* stream = reinterpret_cast<void*>(st);
    // compensated float32 signatures are float32 data
    const rpy_dtype store = dtype == RPY_F64 ? RPY_F64 : RPY_F32;
    const int frc =
            dense_log_t2l_fixed(out, sig, store, width, depth, t2l, batch, os, ss, stream);
    if (frc != RPY_ERR_UNSUPPORTED) return frc;
    const size_t tel = size_t(batch) * size_t(tsize);
    // sig widened to f64 (if needed) | log (f64), each segment starting on
    // a 256-byte boundary
    auto seg = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    const size_t wide_bytes = dtype == RPY_F64 ? 0 : seg(tel * 8);
    char* ws = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&ws, wide_bytes + seg(tel * 8), st));
    int rc = RPY_OK;
    double* lg = (double*)(ws + wide_bytes);
    rpy_tensor_basis tb{width, depth, nullptr};
    bool logged = false;
    const double* sig64 = (const double*)sig;
    int64_t s64 = ss;
    if (dtype != RPY_F64) {
        // float32 signatures: the float64 log widens them as it loads them;
        // shapes without a fixed-shape log widen first
        rc = dense_ft_log_from_f32(lg, (const float*)sig, &tb, batch, tsize, ss, stream);
        logged = rc == RPY_OK;
        if (rc == RPY_ERR_UNSUPPORTED) {
            rc = RPY_OK;
            double* wide = (double*)ws;
            if (ss == tsize) {
                hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(int64_t(tel))), dim3(256), 0, st,
                                   wide, (const float*)sig, int64_t(tel));
            } else {
                for (int64_t e = 0; e < batch && rc == RPY_OK; ++e)
                    hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(tsize)), dim3(256), 0, st,
                                       wide + e * tsize, (const float*)sig + e * ss, tsize);
            }
            if (hipGetLastError() != hipSuccess) rc = fail(RPY_ERR_HIP, "widen launch failed");
            sig64 = wide;
            s64 = tsize;
        }
    }
    if (!rc && !logged)
        rc = dense_ft_log(lg, sig64, RPY_F64, &tb, {0, -1}, {0, -1}, batch, tsize, s64, stream);
    // the projection writes the output (no clearing pass, no read of it)
    if (!rc) rc = dense_sparse_map(out, lg, t2l, RPY_F64, 1.0, batch, os, tsize, stream, false);
    hipError_t ferr = hipFreeAsync(ws, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

}  // namespace

int rpy_stream_log_signature_batch(double* out, const void* increments, const int64_t* offsets,
                                   int64_t n_streams, int32_t width, int32_t depth,
                                   rpy_dtype dtype, const rpy_sparse_matrix* t2l, void* stream)
{
    if (!t2l) return fail(RPY_ERR_INVALID_ARGUMENT, "null t2l matrix");
    const int64_t tsize = tensor_size_checked(width, depth);
    const int64_t lsize = (width >= 1 && depth >= 0) ? lie_size(width, depth) : -1;
    if (tsize < 0 || lsize < 0) return fail(RPY_ERR_BASIS, "invalid width/depth");
    if (t2l->rows != lsize || t2l->cols != tsize)
        return fail(RPY_ERR_SHAPE, "t2l matrix must be %lld x %lld", (long long)lsize,
                    (long long)tsize);
    if (n_streams < 0) return fail(RPY_ERR_SHAPE, "negative number of streams");
    if (n_streams == 0) return RPY_OK;
    if (!out) return fail(RPY_ERR_INVALID_ARGUMENT, "null output");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t tel = size_t(n_streams) * size_t(tsize);
    const size_t elt = (dtype == RPY_F64) ? 8 : 4;
    void* sig = nullptr;
    RPY_HIP_CHECK(workspace_alloc(&sig, tel * elt, st));
    int rc = stream_signature_batch(sig, increments, offsets, n_streams, width, depth, dtype,
                                    stream);
    if (!rc)
        rc = log_to_lie(out, sig, dtype, width, depth, t2l, n_streams, lsize, tsize, lsize,
                        tsize, st);
    hipError_t ferr = hipFreeAsync(sig, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

int rpy_dense_log_to_lie(double* out, const void* sig, rpy_dtype dtype, int32_t width,
                         int32_t depth, const rpy_sparse_matrix* t2l, int64_t batch,
                         int64_t out_stride, int64_t sig_stride, void* stream)
{
    if (!t2l) return fail(RPY_ERR_INVALID_ARGUMENT, "null t2l matrix");
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    const int64_t tsize = tensor_size_checked(width, depth);
    const int64_t lsize = (width >= 1 && depth >= 0) ? lie_size(width, depth) : -1;
    if (tsize < 0 || lsize < 0) return fail(RPY_ERR_BASIS, "invalid width/depth");
    if (t2l->rows != lsize || t2l->cols != tsize)
        return fail(RPY_ERR_SHAPE, "t2l matrix must be %lld x %lld", (long long)lsize,
                    (long long)tsize);
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch");
    if (batch == 0) return RPY_OK;
    if (!out || !sig) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    if (batch > 1 && (out_stride < lsize || sig_stride < tsize))
        return fail(RPY_ERR_SHAPE, "batch strides smaller than the Lie and tensor sizes");
    return log_to_lie(out, sig, dtype, width, depth, t2l, batch, lsize, tsize, out_stride,
                      sig_stride, reinterpret_cast<hipStream_t>(stream));
}

// Test hook (not part of the compute surface): runs the lane decomposition
// of the signature kernel on the host for one float64 stream.
int rpy_debug_signature_host_emulation(double* out, const double* increments, int64_t length,
                                       int32_t width, int32_t depth, int32_t prefix)
{
    if (!out || (!increments && length > 0)) return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    return signature_host_emulation(out, increments, length, width, depth, prefix);
}

int rpy_debug_vjp_host_emulation(double* grad, const double* increments, int64_t length,
                                 int32_t width, int32_t depth, int32_t prefix,
                                 const double* signature, const double* grad_signature)
{
    if (!grad || !signature || !grad_signature || (!increments && length > 0))
        return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    return vjp_host_emulation(grad, increments, length, width, depth, prefix, signature,
                              grad_signature);
}

// Prefix length of the signature kernel variant a call runs, 0 = generic.
int32_t rpy_debug_signature_prefix(int32_t width, int32_t depth, int32_t is_f64)
{
    return chen_prefix_length(width, depth, is_f64);
}

int32_t rpy_tune_variants(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                          rpy_kernel_variant* out, int32_t max)
{
    if (max < 0 || (!out && max > 0)) {
        fail(RPY_ERR_INVALID_ARGUMENT, "bad output array");
        return 0;
    }
    if (op == RPY_OP_SIGNATURE) return sig_variants(width, depth, dtype, out, max);
    if (op == RPY_OP_SIGNATURE_VJP) return vjp_variants(width, depth, dtype, out, max);
    fail(RPY_ERR_INVALID_ARGUMENT, "unknown kernel op %d", op);
    return 0;
}

int rpy_tune_select(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                    const rpy_kernel_variant* variant)
{
    if (op == RPY_OP_SIGNATURE) return sig_select(width, depth, dtype, variant);
    if (op == RPY_OP_SIGNATURE_VJP) return vjp_select(width, depth, dtype, variant);
    return fail(RPY_ERR_INVALID_ARGUMENT, "unknown kernel op %d", op);
}

int rpy_tune_config(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                    rpy_kernel_variant* out, int32_t* source)
{
    if (op == RPY_OP_SIGNATURE) return sig_config(width, depth, dtype, out, source);
    if (op == RPY_OP_SIGNATURE_VJP) return vjp_config(width, depth, dtype, out, source);
    return fail(RPY_ERR_INVALID_ARGUMENT, "unknown kernel op %d", op);
}

}  // extern "C"

extern "C" {

int64_t rpy_dyadic_cache_elements(int64_t n_streams, int32_t width, int32_t depth, int32_t levels)
{
    return dyadic_cache_elements(n_streams, width, depth, levels);
}

int rpy_dyadic_cache_build(void* cache, const void* increments, const double* params,
                           const int64_t* offsets, int64_t n_streams, int32_t width,
                           int32_t depth, int32_t resolution, int32_t levels, rpy_dtype dtype,
                           void* stream)
{
    return dyadic_cache_build(cache, increments, params, offsets, n_streams, width, depth,
                              resolution, levels, dtype, stream);
}

int rpy_dyadic_cache_query(void* out, const void* cache, const void* increments,
                           const double* params, const int64_t* offsets, int64_t n_streams,
                           int32_t width, int32_t depth, int32_t resolution, int32_t levels,
                           const int64_t* query_stream, const double* query_inf,
                           const double* query_sup, int64_t n_queries, rpy_dtype dtype,
                           void* stream)
{
    return dyadic_cache_query(out, cache, increments, params, offsets, n_streams, width, depth,
                              resolution, levels, query_stream, query_inf, query_sup, n_queries,
                              dtype, stream);
}

int rpy_stream_signature_vjp(void* grad_increments, const void* increments,
                             const int64_t* offsets, int64_t n_streams, int32_t width,
                             int32_t depth, const void* signatures, const void* grad_signatures,
                             rpy_dtype dtype, void* stream)
{
    return stream_signature_vjp(grad_increments, increments, offsets, n_streams, width, depth,
                                signatures, grad_signatures, dtype, stream);
}

int rpy_dense_ft_adjoint_right_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                                   const rpy_tensor_basis* basis, int64_t batch, int64_t os,
                                   int64_t ps, int64_t as, int64_t oi, int64_t pi, int64_t ai,
                                   void* stream)
{
    const int64_t n = basis_size(basis);
    const DenseOperand ops[3] = {{out, os, oi, n, true}, {(void*)op, ps, pi, n, false},
                                 {(void*)arg, as, ai, n, false}};
    return with_unit_inc(ops, 3, batch, elem_size(dtype), stream, [&](const UnitIncViews& v) {
        return dense_ft_adjoint_right_mul(v.ptr[0], v.ptr[1], v.ptr[2], dtype, basis, batch,
                                          v.stride[0], v.stride[1], v.stride[2], stream);
    });
}

int rpy_signature_gram(float* gram, const float* signatures, int64_t n, int64_t k,
                       int64_t ld_signatures, int64_t ld_gram, void* stream)
{
    return gram_bf16x3(gram, signatures, n, k, ld_signatures, ld_gram, stream);
}

int32_t rpy_debug_dyadic_cover(int64_t begin, int64_t end, int32_t resolution, int64_t* k_out,
                               int32_t* n_out, int32_t max_pieces)
{
    return dyadic_cover_host(begin, end, resolution, k_out, n_out, max_pieces);
}

}  // extern "C"
