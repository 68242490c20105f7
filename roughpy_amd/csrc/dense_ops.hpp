// dense_ops.hpp -- entry points of dense_ops.hip.
#pragma once

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

int dense_ft_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                 const rpy_tensor_basis* basis, rpy_degree_bounds ob, rpy_degree_bounds lb,
                 rpy_degree_bounds rb, int64_t batch, int64_t os, int64_t ls, int64_t rs,
                 void* stream);
int dense_ft_inplace_mul(void* lhs, const void* rhs, rpy_dtype dtype,
                         const rpy_tensor_basis* basis, rpy_degree_bounds lb,
                         rpy_degree_bounds rb, int64_t batch, int64_t ls, int64_t rs,
                         void* stream);
int dense_ft_antipode(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                      rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                      int64_t as, void* stream);
int dense_ft_adjoint_left_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                              const rpy_tensor_basis* basis, rpy_degree_bounds ob,
                              rpy_degree_bounds pb, rpy_degree_bounds ab, int64_t batch,
                              int64_t os, int64_t ps, int64_t as, void* stream);
int dense_ft_exp(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                 rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                 int64_t as, void* stream);
int dense_ft_fmexp(void* out, const void* mul, const void* expo, rpy_dtype dtype,
                   const rpy_tensor_basis* basis, rpy_degree_bounds ob, rpy_degree_bounds mb,
                   rpy_degree_bounds eb, int64_t batch, int64_t os, int64_t ms, int64_t es,
                   void* stream);
int dense_ft_log(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                 rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                 int64_t as, void* stream);
// Whether dense_ft_mul_new has a fixed-shape product for (width, depth, dtype).
bool has_fixed_mul(int32_t width, int32_t depth, rpy_dtype dtype);
// out = lhs (x) rhs at full depth into a buffer of its own (out is written,
// not read); RPY_ERR_UNSUPPORTED for shapes without a fixed-shape kernel.
int dense_ft_mul_new(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ls,
                     int64_t rs, void* stream);

// The float64 tensor log of float32 tensors (the widening fused into the
// kernel's loads); RPY_ERR_UNSUPPORTED, with no error text set, for shapes
// without a fixed-shape kernel.
int dense_ft_log_from_f32(double* out, const float* arg, const rpy_tensor_basis* basis,
                          int64_t batch, int64_t os, int64_t as, void* stream);
// The float64 log of float32 or float64 signatures projected by t2l (rows =
// the Lie dimension, columns = the tensor size) in one kernel, the log never
// leaving LDS; RPY_ERR_UNSUPPORTED, with no error text set, for shapes without
// a fused kernel.
int dense_log_t2l_fixed(double* out, const void* sig, rpy_dtype sig_dtype, int width, int depth,
                        const rpy_sparse_matrix* t2l, int64_t batch, int64_t os, int64_t as,
                        void* stream);
// out += scale * M arg (accumulate), or out = scale * M arg.
int dense_sparse_map(void* out, const void* arg, const rpy_sparse_matrix* m, rpy_dtype dtype,
                     double scale, int64_t batch, int64_t os, int64_t as, void* stream,
                     bool accumulate = true);

}  // namespace rpy_amd
