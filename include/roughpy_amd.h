/*
 * roughpy_amd.h -- C-ABI of the MI355X (gfx950) dense free-tensor library.
 *
 * Every entry point takes plain pointers and sizes.  Tensor/Lie data
 * pointers are DEVICE pointers (HBM) unless the comment says "host".  The
 * caller owns every buffer; the library allocates only its internal
 * workspace (a grow-only per-device pool, freed by rpy_release_workspace).
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *
 * Return value: RPY_OK (0) or an rpy_status error code; rpy_last_error()
 * returns a thread-local message for the last failure on this thread.
 *
 * Dense layout (roughpy_compute/README.md:110-134): level k of a tensor
 * occupies [db[k], db[k+1]) with db[0]=0, db[k+1]=1+width*db[k]; the word
 * (i_1..i_k) sits at db[k] + sum_j i_j width^(k-j).  A batch is a leading
 * dimension: element e of operand X starts at X + e*X_stride (in elements).
 */
#ifndef ROUGHPY_AMD_H
#define ROUGHPY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rpy_status {
    RPY_OK = 0,
    RPY_ERR_INVALID_ARGUMENT = 1, /* ValueError in the Python layer */
    RPY_ERR_SHAPE = 2,            /* ValueError */
    RPY_ERR_DTYPE = 3,            /* TypeError */
    RPY_ERR_BASIS = 4,            /* ValueError */
    RPY_ERR_HIP = 5,              /* RuntimeError */
    RPY_ERR_OUT_OF_MEMORY = 6,    /* MemoryError */
    RPY_ERR_UNSUPPORTED = 7       /* NotImplementedError */
} rpy_status;

/* Arithmetic type of a call.  RPY_F32_COMPENSATED reads and writes float
 * data and carries sums beyond float precision; each entry point that
 * accepts it states how. */
typedef enum rpy_dtype {
    RPY_F32 = 0,
    RPY_F64 = 1,
    RPY_F32_COMPENSATED = 2
} rpy_dtype;

/* Mirror of rpy::compute::BasisBase (roughpy_compute/common/basis.hpp:8-18)
 * for a tensor basis.  degree_begin is a HOST array of depth+2 entries; it
 * may be NULL, in which case the standard one is used (and checked when
 * given). */
typedef struct rpy_tensor_basis {
    int32_t width;
    int32_t depth;
    const int64_t *degree_begin;
} rpy_tensor_basis;

/* Mirror of the per-operand DegreeBounds of the reference bindings
 * (roughpy/compute/_src/call_config.hpp:11-15): max_degree -1 = basis depth. */
typedef struct rpy_degree_bounds {
    int32_t min_degree;
    int32_t max_degree;
} rpy_degree_bounds;

/* Mirror of rpy::compute::LieBasis (basis.hpp:64-81): degree_begin (depth+2)
 * and data ((size+1) x 2 parent pairs, row 0 = (0,0), letters (0,l)); both
 * HOST arrays. */
typedef struct rpy_lie_basis {
    int32_t width;
    int32_t depth;
    const int64_t *degree_begin;
    const int64_t *data;
} rpy_lie_basis;

/* Compressed sparse matrix, mirror of CompressedMatrixData
 * (roughpy/compute/_src/dense_basic.cpp:554-562).  All arrays are DEVICE
 * arrays when passed to a compute entry point.  compressed_rows != 0: CSR
 * (n_offsets = rows), otherwise CSC (n_offsets = cols).  data has the call's
 * dtype. */
typedef struct rpy_sparse_matrix {
    int64_t rows;
    int64_t cols;
    int64_t nnz;
    const int64_t *offsets; /* n_offsets + 1 */
    const int64_t *indices; /* nnz */
    const void *data;       /* nnz */
    int32_t compressed_rows;
} rpy_sparse_matrix;

/* ---- library ---------------------------------------------------------- */

const char *rpy_last_error(void);
int32_t rpy_version(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int32_t rpy_device_count(void);
/* Frees the internal workspace pool of the current device. */
int rpy_release_workspace(void);

/* ---- bases (host) ------------------------------------------------------- */

/* Size of a truncated tensor algebra; -1 on overflow. */
int64_t rpy_tensor_size(int32_t width, int32_t depth);

/* Lie basis size (sum of Witt dimensions), mirror of compute_lie_dim
 * (roughpy/compute/_src/lie_basis.c:1804-1813). */
int64_t rpy_lie_basis_size(int32_t width, int32_t depth);

/* Builds the greedy Hall set of LieBasis(width, depth)
 * (roughpy/compute/_src/lie_basis.c:201-286).  degree_begin: host, depth+2;
 * data: host, (size+1)*2.  */
int rpy_lie_basis_build(int32_t width, int32_t depth, int64_t *degree_begin,
                        int64_t *data);

/* Validity check of user Hall-set data, Bourbaki major, mirror of
 * check_lie_basis_data (lie_basis.c:1333-1419, 1651-1682).  Returns RPY_OK
 * when valid, RPY_ERR_BASIS with the reason in rpy_last_error otherwise. */
int rpy_check_lie_basis_data(const rpy_lie_basis *basis);

/* l2t map as CSR (tensor_dim rows x lie_dim cols), float64 values, host
 * arrays; mirror of LieBasis.get_l2t_matrix (lie_basis.c:902-994).  Call
 * with NULL arrays to query nnz (written to *nnz). */
int rpy_l2t_matrix(const rpy_lie_basis *basis, int64_t *nnz, int64_t *offsets,
                   int64_t *indices, double *data);

/* t2l map as CSR (lie_dim rows x tensor_dim cols), float64, host arrays;
 * mirror of LieBasis.get_t2l_matrix (lie_basis.c:1150-1278): Dynkin right
 * bracketing expanded in the Hall basis and scaled by 1/degree. */
int rpy_t2l_matrix(const rpy_lie_basis *basis, int64_t *nnz, int64_t *offsets,
                   int64_t *indices, double *data);

/* ---- dense free-tensor operations (device) ------------------------------ */

/* Operand layout of the dense operations.  Element e of operand X is read
 * and written at X + e*X_stride, and its coefficient i at
 * X + e*X_stride + i*X_inc (all in elements).  X_inc is the element stride
 * of the last axis: 1 for contiguous tensors, anything nonzero otherwise
 * (the reference walks such arrays with StridedDenseIterator,
 * roughpy_compute/dense/views.hpp:23-170, chosen in
 * roughpy/compute/_src/py_ternary_array_fn.hpp:41-62).  An output's
 * elements must not overlap.  Operands with X_inc != 1 are gathered into a
 * contiguous workspace copy and scattered back; X_inc == 1 takes no copy.
 *
 * Degree bounds (rpy_degree_bounds) follow update_algebra_params
 * (roughpy/compute/_src/call_config.cpp:11-31): max_degree -1 or above the
 * basis depth means the depth; min_degree above the resolved max_degree is
 * RPY_ERR_INVALID_ARGUMENT.  An input's coefficients outside its bounds are
 * treated as zero; an output's coefficients outside its bounds are left
 * unchanged.  Every tensor operand holds basis size coefficients. */

/* out += lhs (x) rhs, truncated; replaces dense_ft_fma(out, lhs, rhs,
 * basis, out_depth, lhs_depth, rhs_depth)
 * (roughpy/compute/_src/dense_basic.cpp:85-151, kernel
 * roughpy_compute/dense/basic/free_tensor_fma.hpp:13-53).  dtype RPY_F32,
 * RPY_F64, or RPY_F32_COMPENSATED (float data, each output coefficient
 * accumulated in double and rounded to float once). */
int rpy_dense_ft_fma(void *out, const void *lhs, const void *rhs,
                     rpy_dtype dtype, const rpy_tensor_basis *basis,
                     rpy_degree_bounds out_bounds, rpy_degree_bounds lhs_bounds,
                     rpy_degree_bounds rhs_bounds, int64_t batch,
                     int64_t out_stride, int64_t lhs_stride,
                     int64_t rhs_stride, int64_t out_inc, int64_t lhs_inc,
                     int64_t rhs_inc, void *stream);

/* lhs <- lhs (x) rhs; replaces dense_ft_inplace_mul(lhs, rhs, basis,
 * out_depth, rhs_depth) (dense_basic.cpp:195-260, kernel
 * free_tensor_inplace_mul.hpp:13-87). */
int rpy_dense_ft_inplace_mul(void *lhs, const void *rhs, rpy_dtype dtype,
                             const rpy_tensor_basis *basis,
                             rpy_degree_bounds lhs_bounds,
                             rpy_degree_bounds rhs_bounds, int64_t batch,
                             int64_t lhs_stride, int64_t rhs_stride,
                             int64_t lhs_inc, int64_t rhs_inc, void *stream);

/* out = antipode(arg) on the levels max(mins)..min(maxes) of the two
 * bounds (and level 1 whenever that range starts at or below 1, as
 * free_tensor_antipode.hpp:60-80 does); replaces dense_ft_antipode(out, arg,
 * basis, out_depth) (dense_basic.cpp:295-360, kernel
 * free_tensor_antipode.hpp:45-100), whose one depth argument bounds the
 * argument (dense_basic.cpp:310-318). */
int rpy_dense_ft_antipode(void *out, const void *arg, rpy_dtype dtype,
                          const rpy_tensor_basis *basis,
                          rpy_degree_bounds out_bounds,
                          rpy_degree_bounds arg_bounds, int64_t batch,
                          int64_t out_stride, int64_t arg_stride,
                          int64_t out_inc, int64_t arg_inc, void *stream);

/* out += L_op^*(arg) (adjoint of left multiplication); replaces
 * dense_ft_adjoint_left_mul(out, op, arg, basis, out_depth, lhs_depth,
 * rhs_depth) (roughpy/compute/_src/dense_basic.cpp:387-490, kernel
 * free_tensor_adjoint_left_mul.hpp:10-77), with that kernel's degree ranges:
 * argument degrees from max(arg min - op max, out min) to
 * min(arg max - op min, out max). */
int rpy_dense_ft_adjoint_left_mul(void *out, const void *op, const void *arg,
                                  rpy_dtype dtype,
                                  const rpy_tensor_basis *basis,
                                  rpy_degree_bounds out_bounds,
                                  rpy_degree_bounds op_bounds,
                                  rpy_degree_bounds arg_bounds,
                                  int64_t batch, int64_t out_stride,
                                  int64_t op_stride, int64_t arg_stride,
                                  int64_t out_inc, int64_t op_inc,
                                  int64_t arg_inc, void *stream);

/* out += lhs (shuffle) rhs at full depth (1..12); replaces dense_st_fma(out,
 * lhs, rhs, basis) (roughpy/compute/_src/dense_basic.cpp:449-534, kernel
 * roughpy_compute/dense/basic/shuffle_tensor_product.hpp:24-85).  Same
 * mask order as the reference, so float results round the same way. */
int rpy_dense_st_fma(void *out, const void *lhs, const void *rhs,
                     rpy_dtype dtype, const rpy_tensor_basis *basis,
                     int64_t batch, int64_t out_stride, int64_t lhs_stride,
                     int64_t rhs_stride, int64_t out_inc, int64_t lhs_inc,
                     int64_t rhs_inc, void *stream);

/* out += adjoint of shuffle multiplication by op, applied to arg:
 * <out, x> = <arg, op (shuffle) x>; the kernel of
 * roughpy_compute/dense/basic/shuffle_tensor_adjoint_mul.hpp:13-70 (the
 * reference exposes it to C++ only). */
int rpy_dense_st_adj_mul(void *out, const void *op, const void *arg,
                         rpy_dtype dtype, const rpy_tensor_basis *basis,
                         int64_t batch, int64_t out_stride, int64_t op_stride,
                         int64_t arg_stride, int64_t out_inc, int64_t op_inc,
                         int64_t arg_inc, void *stream);

/* out = exp(arg) (arg[0] ignored) truncated at out_bounds.max_degree m:
 * coefficients 0 .. db[m+1]-1 of out are written; replaces
 * dense_ft_exp(out, arg, basis, out_depth, arg_depth)
 * (roughpy/compute/_src/dense_intermediate.cpp:62-102, kernel
 * free_tensor_exp.hpp:13-37).  out_bounds.min_degree must be 0.  The
 * reference reads both views with out_depth's bounds (:38-52), so a drop-in
 * binding passes out_depth for arg_bounds too. */
int rpy_dense_ft_exp(void *out, const void *arg, rpy_dtype dtype,
                     const rpy_tensor_basis *basis,
                     rpy_degree_bounds out_bounds, rpy_degree_bounds arg_bounds,
                     int64_t batch, int64_t out_stride, int64_t arg_stride,
                     int64_t out_inc, int64_t arg_inc, void *stream);

/* out = multiplier * exp(exponent) (exponent[0] ignored) truncated at
 * out_bounds.max_degree; replaces dense_ft_fmexp(out, multiplier, exponent,
 * basis, out_depth, multiplier_depth, exponent_depth)
 * (dense_intermediate.cpp:153-208, kernel free_tensor_fmexp.hpp:15-40),
 * which reads all three views with out_depth's bounds (:124-144): a drop-in
 * binding passes out_depth for all three. */
int rpy_dense_ft_fmexp(void *out, const void *multiplier, const void *exponent,
                       rpy_dtype dtype, const rpy_tensor_basis *basis,
                       rpy_degree_bounds out_bounds,
                       rpy_degree_bounds multiplier_bounds,
                       rpy_degree_bounds exponent_bounds, int64_t batch,
                       int64_t out_stride, int64_t mul_stride,
                       int64_t exp_stride, int64_t out_inc, int64_t mul_inc,
                       int64_t exp_inc, void *stream);

/* out = log(arg), assuming arg[0] == 1, truncated at out_bounds.max_degree,
 * with arg read within arg_bounds; replaces dense_ft_log(out, arg, basis,
 * out_depth, arg_depth) (dense_intermediate.cpp:254-295, kernel
 * free_tensor_log.hpp:12-40).  out is write-only (the reference's Python
 * wrapper zero-fills it, roughpy/compute/__init__.py:279). */
int rpy_dense_ft_log(void *out, const void *arg, rpy_dtype dtype,
                     const rpy_tensor_basis *basis,
                     rpy_degree_bounds out_bounds, rpy_degree_bounds arg_bounds,
                     int64_t batch, int64_t out_stride, int64_t arg_stride,
                     int64_t out_inc, int64_t arg_inc, void *stream);

/* Cotangent of the truncated log (the log of rpy_dense_ft_log at full
 * depth; arg's degree 0 is not read): out[e] = d<cotangent[e], log(arg[e])>
 * / d arg[e], degree 0 of out set to 0.  No counterpart in the reference,
 * which has no derivatives; it is the log step of the log-signature
 * backward (Horner chain of the log run in reverse, float64 arithmetic for
 * either dtype).  Batch strides in elements; each element's last axis is
 * contiguous. */
int rpy_dense_ft_log_vjp(void *out, const void *arg, const void *cotangent,
                         rpy_dtype dtype, const rpy_tensor_basis *basis,
                         int64_t batch, int64_t out_stride, int64_t arg_stride,
                         int64_t cotangent_stride, void *stream);

/* out[e] += scale * M arg[e] for a device CSR/CSC matrix; replaces
 * dense_lie_to_tensor / dense_tensor_to_lie (dense_basic.cpp:862-1047,
 * 1116-1298, kernel apply_sparse_linear_map.hpp:16-55).  out has
 * matrix.rows entries per element, arg matrix.cols.  scale 1 = none. */
int rpy_dense_sparse_map(void *out, const void *arg,
                         const rpy_sparse_matrix *matrix, rpy_dtype dtype,
                         double scale, int64_t batch, int64_t out_stride,
                         int64_t arg_stride, int64_t out_inc, int64_t arg_inc,
                         void *stream);

/* ---- streams (device) --------------------------------------------------- */

/* Signatures of n_streams piecewise-linear streams of uneven length.
 * increments: device, rows of `width` values, stream s owning rows
 * [offsets[s], offsets[s+1]); offsets: device int64, n_streams+1.
 * out: device, n_streams x tensor_size(width, depth).  Result of stream s is
 * the Chen product of exp(row) over its rows, i.e. Context::compute_signature
 * (algebra/src/lite_context.h:545-563) applied per stream, the value
 * LieIncrementStream.signature() returns for the stream's full support
 * (streams/src/stream.cpp:156-171).  dtype: RPY_F64 (double data), RPY_F32
 * (float data and arithmetic) or RPY_F32_COMPENSATED (float data).  The
 * compensated mode runs the float32 kernels with the coefficients of levels
 * 1 and 2 accumulated by Kahan's compensated summation (their rounding
 * errors are the ones every later step multiplies into the higher levels)
 * at the shapes with specialised kernels (widths 2..8); its block-merge
 * variant (RPY_KIND_BLOCK_MERGE, at (3,4), (4,6), (5,5), (6,5) and (8,4))
 * cuts each stream into at most 16 segments of `block` rows, computes them
 * on those lanes and multiplies them per stream with Kahan sums; at other
 * shapes the mode runs float64 arithmetic on the float data, each
 * coefficient rounded to float once at the end.  Work is one persistent
 * launch over a length-sorted queue. */
int rpy_stream_signature_batch(void *out, const void *increments,
                               const int64_t *offsets, int64_t n_streams,
                               int32_t width, int32_t depth, rpy_dtype dtype,
                               void *stream);

/* rpy_stream_signature_batch with every temporary (length keys, the sort's
 * scratch, the work-queue counter) in a fixed scratch array of 8 MiB per GPU
 * instead of allocated memory: the call allocates nothing.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Same arguments, results and errors, plus RPY_ERR_SHAPE when
 * the batch needs more scratch than the array holds (about 300,000 streams;
 * the message states the bytes needed).  The block-compensated mode runs
 * its segment buffer in rounds of the streams that fit the array (about 30
 * streams per round at width 5, depth 5).  The array keeps the last call's
 * contents and is shared by every call on that GPU: calls that overlap on
 * one GPU (other streams, other host threads) race and must be ordered by
 * the caller. */
int rpy_stream_signature_batch_noalloc(void *out, const void *increments,
                                       const int64_t *offsets,
                                       int64_t n_streams, int32_t width,
                                       int32_t depth, rpy_dtype dtype,
                                       void *stream);

/* Log-signatures in the Hall basis: t2l(log(signature)) per stream
 * (lite_context.h:899-902; Stream.log_signature, stream.cpp:107-123).
 * The log and the t2l projection run in float64 whatever `dtype` is; out is
 * float64, n_streams x lie_size.  t2l: device CSR matrix from
 * rpy_t2l_matrix (float64 values). */
int rpy_stream_log_signature_batch(double *out, const void *increments,
                                   const int64_t *offsets, int64_t n_streams,
                                   int32_t width, int32_t depth,
                                   rpy_dtype dtype,
                                   const rpy_sparse_matrix *t2l,
                                   void *stream);

/* The last two steps of rpy_stream_log_signature_batch on signatures the
 * caller already holds (float32 or float64, batch x tensor size, rows
 * `sig_stride` apart): out (float64, batch x lie_size, rows `out_stride`
 * apart) = t2l(log(sig)), the log in float64.  For the fixed shapes the log
 * stays in LDS (one kernel); other shapes run the log and the projection as
 * two launches.  No reference counterpart as one call: it is
 * dense_ft_log (roughpy/compute/_src/dense_intermediate.cpp:260-280)
 * followed by the t2l map (lite_context.h:899-902), for callers that keep
 * the signatures, such as the log-signature backward. */
int rpy_dense_log_to_lie(double *out, const void *sig, rpy_dtype dtype,
                         int32_t width, int32_t depth, const rpy_sparse_matrix *t2l,
                         int64_t batch, int64_t out_stride, int64_t sig_stride,
                         void *stream);

/* ---- dyadic signature cache (device) ------------------------------------ */

/* Elements (of the cache dtype) of a dyadic cache storing node signatures of
 * powers 0..levels over [0, 1) for n_streams streams; -1 on bad arguments. */
int64_t rpy_dyadic_cache_elements(int64_t n_streams, int32_t width,
                                  int32_t depth, int32_t levels);

/* Builds the cache: node (j, s, k) = signature of the rows of stream s whose
 * param lies in [k 2^-j, (k+1) 2^-j), for j = 0..levels, k < 2^j.  Rows of
 * stream s are [offsets[s], offsets[s+1]) with strictly increasing params
 * (device, float64; rows outside [0, 1) are never covered).  The reference
 * memoises resolution leaves and rebuilds coarser nodes with CBH on every
 * query (streams/src/dyadic_caching_layer.cpp:86-130); this cache builds
 * every stored level once, in one fused pass: one signature launch for level
 * `levels`, then one batched product per coarser level.  cache: device,
 * rpy_dyadic_cache_elements(...) elements of dtype (RPY_F32 / RPY_F64). */
int rpy_dyadic_cache_build(void *cache, const void *increments,
                           const double *params, const int64_t *offsets,
                           int64_t n_streams, int32_t width, int32_t depth,
                           int32_t resolution, int32_t levels, rpy_dtype dtype,
                           void *stream);

/* Answers n_queries clopen interval queries [query_inf[q], query_sup[q]) on
 * stream query_stream[q] at `resolution` (all device arrays): out[q] is the
 * signature Stream.signature(RealInterval(inf, sup), resolution) returns for
 * a LieIncrementStream of that resolution (streams/src/stream.cpp:107-171):
 * the ordered Chen product over the maximal dyadic pieces of
 * [floor(inf 2^r), floor(sup 2^r)) 2^-r, read from the cache for pieces of
 * power <= levels and computed from the rows for finer ones.  An empty or
 * out-of-range query gives the unit tensor. */
int rpy_dyadic_cache_query(void *out, const void *cache, const void *increments,
                           const double *params, const int64_t *offsets,
                           int64_t n_streams, int32_t width, int32_t depth,
                           int32_t resolution, int32_t levels,
                           const int64_t *query_stream, const double *query_inf,
                           const double *query_sup, int64_t n_queries,
                           rpy_dtype dtype, void *stream);

/* ---- derivatives (device) ----------------------------------------------- */

/* Vector-Jacobian product of rpy_stream_signature_batch: grad_increments
 * (same shape as increments) receives d<G_s, S_s>/dx for every row, where
 * S_s = signatures[s] (the forward outputs) and G_s = grad_signatures[s].
 * The reference only has the forward-mode derivative
 * (algebra/src/lite_context.h:580-636); this is its reverse mode, by reverse
 * Chen recomputation (prefixes rebuilt by multiplying with exp(-x), nothing
 * stored per step).  dtype RPY_F32 or RPY_F64; width <= 32. */
int rpy_stream_signature_vjp(void *grad_increments, const void *increments,
                             const int64_t *offsets, int64_t n_streams,
                             int32_t width, int32_t depth,
                             const void *signatures,
                             const void *grad_signatures, rpy_dtype dtype,
                             void *stream);

/* out += R_op^*(arg), the adjoint of right multiplication by op:
 * out[u] += sum_v arg[u v] op[v]; right-hand counterpart of
 * ft_adj_lmul (roughpy_compute/dense/basic/free_tensor_adjoint_left_mul.hpp:
 * 10-77), used by the backward of the tensor product. */
int rpy_dense_ft_adjoint_right_mul(void *out, const void *op, const void *arg,
                                   rpy_dtype dtype,
                                   const rpy_tensor_basis *basis,
                                   int64_t batch, int64_t out_stride,
                                   int64_t op_stride, int64_t arg_stride,
                                   int64_t out_inc, int64_t op_inc,
                                   int64_t arg_inc, void *stream);

/* ---- signature kernel by signature dot (device) ------------------------- */

/* gram[i, j] = <S_i, S_j> for the n dense signatures S (n x k float32, row
 * stride ld_signatures), the signature kernel of
 * examples/signature-kernel-by-signature-dot.py:15-23 (np.inner of dense
 * signatures) for a whole batch: one GEMM on MFMA with each float32 split
 * into bf16 hi + lo parts (hi.hi + hi.lo + lo.hi, float32 accumulate).
 * gram: n x n float32, row stride ld_gram.  All pointers device. */
int rpy_signature_gram(float *gram, const float *signatures, int64_t n,
                       int64_t k, int64_t ld_signatures, int64_t ld_gram,
                       void *stream);

/* ---- stream queries (host) ---------------------------------------------- */

/* -min(0, e-1) where arg = m 2^e, 0.5 <= |m| < 1; mirror of
 * param_to_resolution (streams/include/roughpy/streams/stream_base.h:73-83). */
int32_t rpy_param_to_resolution(double arg);

/* Maximal aligned dyadic pieces k 2^-n of the clopen [inf, sup) at tolerance
 * tol, left to right; mirror of to_dyadic_intervals
 * (intervals/src/dyadic_interval.cpp:191-228).  *count receives the number of
 * pieces; k_out/n_out (host, max_pieces entries) may be NULL to query it. */
int rpy_to_dyadic_intervals(double inf, double sup, int32_t tol,
                            int64_t *k_out, int32_t *n_out, int32_t max_pieces,
                            int32_t *count);

/* Rows of a LieIncrementStream that the query [inf, sup) at
 * query_resolution covers, as [begin, end) ranges over the strictly
 * increasing params (host).  Leaves at stream_resolution are whole;
 * a dissection piece finer than the stream contributes its leaf only when it
 * starts where the leaf starts and holds a row -- the semantics of
 * DyadicCachingLayer::log_signature (streams/src/dyadic_caching_layer.cpp:
 * 86-130) over LieIncrementStream leaves (streams/src/
 * lie_increment_stream.cpp:174-199).  ranges: host, 2*max_ranges entries. */
int rpy_stream_select_rows(const double *params, int64_t n_params,
                           int32_t stream_resolution, double inf, double sup,
                           int32_t query_resolution, int64_t *ranges,
                           int64_t max_ranges, int64_t *n_ranges);

/* The rows of many queries at once, for Stream::simplify
 * (streams/src/stream.cpp:232-258), which queries the stream's
 * log-signature on every interval of a partition: query q is
 * [infs[q], sups[q]) trimmed to the support [support_inf, support_sup)
 * (Stream::check_support_and_trim, stream.cpp:37-52; no rows when it misses
 * the support or is empty), dissected at query_resolution and selected as in
 * rpy_stream_select_rows.  offsets (host, n_queries + 1) receives the CSR
 * offsets of the selected row indices; rows (host, max_rows entries) receives
 * the indices, or may be NULL to size the output first.  The params are
 * checked once per call.  Large batches run in blocks of queries on up to 16
 * host threads, joined before the call returns; the result equals a single
 * pass in query order. */
int rpy_stream_select_rows_batch(const double *params, int64_t n_params,
                                 int32_t stream_resolution, double support_inf,
                                 double support_sup, const double *infs,
                                 const double *sups, int64_t n_queries,
                                 int32_t query_resolution, int64_t *offsets,
                                 int64_t *rows, int64_t max_rows);

/* ---- test hooks (host only; no counterpart in the reference) ----------- */

/* Runs the lane decomposition of the specialised signature kernel on the
 * host for one float64 stream of `length` increment rows (host arrays), with
 * prefix length `prefix`.  Lets CPU-only tests check the kernel's index map. */
int rpy_debug_signature_host_emulation(double *out, const double *increments,
                                       int64_t length, int32_t width,
                                       int32_t depth, int32_t prefix);

/* Runs the lane algebra of the register-resident signature backward on the
 * host for one float64 stream (host arrays) with prefix length `prefix`:
 * grad (length x width) from the increments, the stream's signature and the
 * cotangent of the signature.  (width, depth, prefix): (8, 2..3, 2),
 * (8, 4, 2..3), (4, 3..5, 3), (4, 6, 3..4), (2, 6, 6), (2, 7, 6..7). */
int rpy_debug_vjp_host_emulation(double *grad, const double *increments,
                                 int64_t length, int32_t width, int32_t depth,
                                 int32_t prefix, const double *signature,
                                 const double *grad_signature);

/* Prefix length of the signature kernel variant a call of (width, depth,
 * RPY_F64 if is_f64 else RPY_F32) runs (rpy_tune_config), or 0 when it runs
 * the generic kernel. */
int32_t rpy_debug_signature_prefix(int32_t width, int32_t depth,
                                   int32_t is_f64);

/* Host run of the query kernel's dissection of [begin, end) (units of
 * 2^-resolution) into maximal aligned pieces; returns the piece count. */
int32_t rpy_debug_dyadic_cover(int64_t begin, int64_t end, int32_t resolution,
                               int64_t *k_out, int32_t *n_out,
                               int32_t max_pieces);

/* ---- kernel variants and the tuned table (host only; no counterpart in
 * the reference, whose CPU kernels have no launch shape to choose) -------
 *
 * The signature forward and backward each have several compiled kernel
 * variants per (width, depth, dtype).  A call runs, in order of precedence:
 * the variant chosen with rpy_tune_select in this process; the row of the
 * tuned table compiled into the library (written by the autotuner,
 * roughpy_amd/tune.py, from timings on MI355X); else the built-in default.
 * All variants of one shape compute the same signature (up to rounding). */

typedef enum rpy_kernel_op {
    RPY_OP_SIGNATURE = 0,      /* rpy_stream_signature_batch and friends */
    RPY_OP_SIGNATURE_VJP = 1   /* rpy_stream_signature_vjp */
} rpy_kernel_op;

typedef enum rpy_kernel_kind {
    RPY_KIND_GENERIC = 0,      /* one workgroup per stream, state in memory */
    RPY_KIND_LANE_CHUNK = 1,   /* 64 (stream, prefix) lanes per wave */
    RPY_KIND_STREAM_GROUP = 2, /* W^P lanes of one stream per workgroup */
    RPY_KIND_WAVE_VJP = 3,     /* backward: W^P lanes per stream, registers */
    RPY_KIND_BLOCK_MERGE = 4,  /* RPY_F32_COMPENSATED: segments of `block`
                                * rows on the compensated lanes, merged per
                                * stream with Kahan sums */
    RPY_KIND_STREAM_TABLE = 5  /* the stream group with the path partials of
                                * a block's steps computed once per prefix
                                * group into LDS */
} rpy_kernel_kind;

typedef struct rpy_kernel_variant {
    int32_t kind;          /* rpy_kernel_kind */
    int32_t prefix;        /* prefix length P (W^P lanes per stream); 0 generic */
    int32_t block;         /* increment rows per LDS block (stream group) or
                            * per segment (block merge), else 0 */
    int32_t waves_per_eu;  /* occupancy hint of the lane-chunk kernel, else 0 */
} rpy_kernel_variant;

/* Writes up to `max` of the variants compiled for (op, width, depth, dtype)
 * to `out` (may be NULL with max 0); returns how many there are. */
int32_t rpy_tune_variants(int32_t op, int32_t width, int32_t depth,
                          rpy_dtype dtype, rpy_kernel_variant *out,
                          int32_t max);

/* Makes later calls of (op, width, depth, dtype) in this process run
 * `variant` (one of rpy_tune_variants'), or restores the tuned / default
 * choice when `variant` is NULL.  RPY_ERR_UNSUPPORTED for a variant that is
 * not compiled.  Not synchronised with calls in flight on other threads. */
int rpy_tune_select(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                    const rpy_kernel_variant *variant);

/* The variant a call of (op, width, depth, dtype) runs now; *source (may be
 * NULL) is 0 for the built-in default, 1 for the tuned table, 2 for a
 * rpy_tune_select choice. */
int rpy_tune_config(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                    rpy_kernel_variant *out, int32_t *source);

#ifdef __cplusplus
}
#endif

#endif /* ROUGHPY_AMD_H */
