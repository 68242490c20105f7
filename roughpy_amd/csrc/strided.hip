// strided.hip -- gathers and scatters of dense operands with an element
// stride other than 1 (strided.hpp).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "chen.hpp"
#include "strided.hpp"

namespace rpy_amd {

namespace {

constexpr int kBlock = 256;

// dst[e * n + i] = src[e * stride + i * inc]
template <typename U>
__global__ __launch_bounds__(kBlock) void gather_kernel(U* __restrict__ dst,
                                                       const U* __restrict__ src, int64_t batch,
                                                       int64_t n, int64_t stride, int64_t inc)
{
    const int64_t total = batch * n;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * kBlock) {
        const int64_t e = t / n;
        const int64_t i = t - e * n;
        dst[t] = src[e * stride + i * inc];
    }
}

// dst[e * stride + i * inc] = src[e * n + i]
template <typename U>
__global__ __launch_bounds__(kBlock) void scatter_kernel(U* __restrict__ dst,
                                                        const U* __restrict__ src,
                                                        int64_t batch, int64_t n, int64_t stride,
                                                        int64_t inc)
{
    const int64_t total = batch * n;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * kBlock) {
        const int64_t e = t / n;
        const int64_t i = t - e * n;
        dst[e * stride + i * inc] = src[t];
    }
}

uint32_t grid_for_total(int64_t total)
{
    return uint32_t(std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 65536)));
}

int64_t aligned(int64_t bytes) { return (bytes + 255) & ~int64_t(255); }

// The map (e, i) -> e * stride + i * inc over e < batch, i < n is one to one
// when one axis nests inside the other's step.
bool disjoint(int64_t batch, int64_t n, int64_t stride, int64_t inc)
{
    const int64_t s = std::llabs(stride), c = std::llabs(inc);
    if (batch <= 1) return true;  // inc != 0 is checked by the caller
    if (n <= 1) return s >= 1;
    return s >= c * (n - 1) + 1 || c >= s * (batch - 1) + 1;
}

template <typename U>
void launch_copy(bool gather, void* dst, const void* src, int64_t batch, int64_t n,
                 int64_t stride, int64_t inc, hipStream_t st)
{
    const uint32_t g = grid_for_total(batch * n);
    if (gather)
        hipLaunchKernelGGL(gather_kernel<U>, dim3(g), dim3(kBlock), 0, st, (U*)dst,
                           (const U*)src, batch, n, stride, inc);
    else
        hipLaunchKernelGGL(scatter_kernel<U>, dim3(g), dim3(kBlock), 0, st, (U*)dst,
                           (const U*)src, batch, n, stride, inc);
}

void copy(bool gather, size_t es, void* dst, const void* src, int64_t batch, int64_t n,
          int64_t stride, int64_t inc, hipStream_t st)
{
    if (es == 8)
        launch_copy<uint64_t>(gather, dst, src, batch, n, stride, inc, st);
    else
        launch_copy<uint32_t>(gather, dst, src, batch, n, stride, inc, st);
}

}  // namespace

int prepare_unit_inc(const DenseOperand* ops, int n_ops, int64_t batch, size_t es,
                     hipStream_t st, UnitIncViews& v)
{
    if (n_ops > kMaxDenseOperands) return fail(RPY_ERR_INVALID_ARGUMENT, "too many operands");
    if (es != 4 && es != 8) return fail(RPY_ERR_DTYPE, "unsupported element size");
    int64_t total = 0;
    for (int i = 0; i < n_ops; ++i) {
        const DenseOperand& o = ops[i];
        if (o.inc == 0) return fail(RPY_ERR_SHAPE, "element stride of operand %d is zero", i);
        if (o.n < 0) return fail(RPY_ERR_SHAPE, "negative coefficient count");
        if (o.inc != 1) {
            if (o.written && !disjoint(batch, o.n, o.stride, o.inc))
                return fail(RPY_ERR_SHAPE,
                            "output operand %d: batch stride %lld and element stride %lld "
                            "make its elements overlap", i, (long long)o.stride,
                            (long long)o.inc);
            if (!o.ptr && o.n > 0) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
            total += aligned(batch * o.n * int64_t(es));
        }
    }
    v.workspace = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&v.workspace, size_t(total), st));
    int64_t off = 0;
    for (int i = 0; i < n_ops; ++i) {
        const DenseOperand& o = ops[i];
        if (o.inc == 1) {
            v.ptr[i] = o.ptr;
            v.stride[i] = o.stride;
            continue;
        }
        v.ptr[i] = v.workspace + off;
        v.stride[i] = o.n;
        off += aligned(batch * o.n * int64_t(es));
        copy(true, es, v.ptr[i], o.ptr, batch, o.n, o.stride, o.inc, st);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        (void)hipFreeAsync(v.workspace, st);
        return fail(RPY_ERR_HIP, "gather launch failed: %s", hipGetErrorString(err));
    }
    return RPY_OK;
}

int finish_unit_inc(const DenseOperand* ops, int n_ops, int64_t batch, size_t es,
                    hipStream_t st, UnitIncViews& v, int rc)
{
    if (!rc) {
        for (int i = 0; i < n_ops; ++i)
            if (ops[i].inc != 1 && ops[i].written)
                copy(false, es, ops[i].ptr, v.ptr[i], batch, ops[i].n, ops[i].stride, ops[i].inc,
                     st);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "scatter launch failed: %s", hipGetErrorString(err));
    }
    hipError_t ferr = hipFreeAsync(v.workspace, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

}  // namespace rpy_amd
