// strided.hpp -- dense operands whose coefficients are not adjacent.
//
// The reference bindings walk the last axis of a NumPy array with its own
// stride (StridedDenseIterator, roughpy_compute/dense/views.hpp:23-170, chosen
// in roughpy/compute/_src/py_ternary_array_fn.hpp:41-62 whenever a stride is
// not 1).  Here an operand with element stride `inc` != 1 is gathered into a
// contiguous workspace copy, the contiguous kernels run on the copies, and
// written operands are scattered back.  Calls whose operands all have
// inc == 1 take no copy.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

// One operand of a batched call: element e, coefficient i at
// ptr[e * stride + i * inc] (in elements), n coefficients per element.
// A written operand is also read: coefficients the call leaves alone keep
// their values through the copy.
struct DenseOperand {
    void* ptr;
    int64_t stride;
    int64_t inc;
    int64_t n;
    bool written;
};

constexpr int kMaxDenseOperands = 4;

// Contiguous views of the operands, filled by prepare_unit_inc.
struct UnitIncViews {
    void* ptr[kMaxDenseOperands];
    int64_t stride[kMaxDenseOperands];
    char* workspace;
};

// Validates the element strides (nonzero; a written operand's elements must
// not overlap) and gathers every operand with inc != 1 into workspace.
int prepare_unit_inc(const DenseOperand* ops, int n_ops, int64_t batch, size_t elem_size,
                     hipStream_t st, UnitIncViews& v);
// Scatters the written gathered operands back and releases the workspace;
// `rc` is the status of the call in between (nothing is scattered after a
// failure).
int finish_unit_inc(const DenseOperand* ops, int n_ops, int64_t batch, size_t elem_size,
                    hipStream_t st, UnitIncViews& v, int rc);

// Runs op(views) with every operand in a unit-stride layout.
template <typename Op>
int with_unit_inc(const DenseOperand* ops, int n_ops, int64_t batch, size_t elem_size,
                  void* stream, Op&& op)
{
    bool strided = false;
    for (int i = 0; i < n_ops; ++i) strided |= ops[i].inc != 1;
    UnitIncViews v{};
    if (!strided || batch <= 0) {
        for (int i = 0; i < n_ops; ++i) {
            v.ptr[i] = ops[i].ptr;
            v.stride[i] = ops[i].stride;
        }
        return op(v);
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int rc = prepare_unit_inc(ops, n_ops, batch, elem_size, st, v);
    if (rc) return rc;
    rc = op(v);
    return finish_unit_inc(ops, n_ops, batch, elem_size, st, v, rc);
}

}  // namespace rpy_amd
