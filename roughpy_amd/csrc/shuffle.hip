// shuffle.hip -- the shuffle product of dense tensors and its adjoint.
//
// (x sh y)(w) = sum over subsets S of the letter positions of w of
// x(w_S) y(w_S^c), the v1 kernel of
// roughpy_compute/dense/basic/shuffle_tensor_product.hpp:24-85 (bit j of the
// mask sends letter j, counted from the last letter, to the lhs word).  One
// thread owns one output coefficient; the masks are walked in the
// reference's order (the last letter's bit fastest, rhs before lhs), so the
// float rounding follows the reference loop.
//
// The adjoint of shuffle multiplication by `op`
// (shuffle_tensor_adjoint_mul.hpp:13-70) scatters arg(w) op(w_S) into
// out(w_S^c).  Here it is a gather: one thread owns one out word u and sums,
// over op words v and over the interleavings of u and v (Gosper's walk over
// the k-subsets of the positions), op(v) arg(interleaving) -- the same
// terms, with no atomics.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "chen.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxShuffleDepth = 12;

// letters of the word `idx` of length n, letters[0] the last letter
template <int D>
RPY_FI void unpack(int (&letters)[D], uint32_t idx, int n, const Pow& pw)
{
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (d < n) {
            const uint32_t q = fdiv(idx, pw.pw[1]);
            letters[d] = int(idx - q * pw.pw[1].d);
            idx = q;
        }
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kBlock) void st_fma_kernel(T* __restrict__ out,
                                                       const T* __restrict__ lhs,
                                                       const T* __restrict__ rhs, DevBasis b,
                                                       Pow pw, int64_t batch, int64_t os,
                                                       int64_t ls, int64_t rs)
{
    const uint32_t W = uint32_t(b.width);
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* l = lhs + e * ls;
        const T* r = rhs + e * rs;
        if (threadIdx.x == 0) o[0] += l[0] * r[0];
        for (int n = 1; n <= D; ++n) {
            const uint32_t size = uint32_t(b.level_size(n));
            for (uint32_t idx = threadIdx.x; idx < size; idx += kBlock) {
                int letters[D];
                unpack<D>(letters, idx, n, pw);
                const int c = letters[0];
                T acc = T(0);
                // prefix masks: bit j (j >= 1) of the full mask is bit j-1 here
                const uint32_t npm = 1u << (n - 1);
                for (uint32_t pm = 0; pm < npm; ++pm) {
                    uint32_t li = 0, ri = 0;
                    int ld = 0, rd = 0;
#pragma unroll
                    for (int dd = D - 1; dd >= 1; --dd) {
                        if (dd < n) {
                            if ((pm >> (dd - 1)) & 1u) {
                                ++ld;
                                li = li * W + uint32_t(letters[dd]);
                            } else {
                                ++rd;
                                ri = ri * W + uint32_t(letters[dd]);
                            }
                        }
                    }
                    // last letter to the rhs word, then to the lhs word
                    acc += l[b.db[ld] + li] * r[b.db[rd + 1] + ri * W + c];
                    acc += l[b.db[ld + 1] + li * W + c] * r[b.db[rd] + ri];
                }
                o[b.db[n] + idx] += acc;
            }
        }
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kBlock) void st_adj_mul_kernel(T* __restrict__ out,
                                                           const T* __restrict__ op,
                                                           const T* __restrict__ arg,
                                                           DevBasis b, Pow pw, int64_t batch,
                                                           int64_t os, int64_t ps, int64_t as)
{
    const uint32_t W = uint32_t(b.width);
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* p = op + e * ps;
        const T* a = arg + e * as;
        for (int m = 0; m <= D; ++m) {
            const uint32_t usize = uint32_t(b.level_size(m));
            for (uint32_t uidx = threadIdx.x; uidx < usize; uidx += kBlock) {
                int ul[D > 0 ? D : 1];
                unpack<D>(ul, uidx, m, pw);
                T acc = T(0);
                for (int k = 0; k + m <= D; ++k) {
                    const int n = m + k;
                    const uint32_t vsize = uint32_t(b.level_size(k));
                    for (uint32_t vidx = 0; vidx < vsize; ++vidx) {
                        int vl[D > 0 ? D : 1];
                        unpack<D>(vl, vidx, k, pw);
                        const T pv = p[b.db[k] + vidx];
                        T sum = T(0);
                        // k-subsets of the n positions (bit j = position j from
                        // the last letter) in increasing order
                        uint32_t S = (k == 0) ? 0u : ((1u << k) - 1u);
                        const uint32_t limit = 1u << n;
                        while (S < limit) {
                            // merge: positions n-1 .. 0, first letter first
                            uint32_t widx = 0;
                            int iu = m - 1, iv = k - 1;
#pragma unroll
                            for (int pos = D - 1; pos >= 0; --pos) {
                                if (pos < n) {
                                    int letter;
                                    if ((S >> pos) & 1u) {
                                        letter = 0;
#pragma unroll
                                        for (int t = 0; t < D; ++t)
                                            if (t == iv) letter = vl[t];
                                        --iv;
                                    } else {
                                        letter = 0;
#pragma unroll
                                        for (int t = 0; t < D; ++t)
                                            if (t == iu) letter = ul[t];
                                        --iu;
                                    }
                                    widx = widx * W + uint32_t(letter);
                                }
                            }
                            sum += a[b.db[n] + widx];
                            if (k == 0) break;
                            // next k-subset (Gosper)
                            const uint32_t lowest = S & (0u - S);
                            const uint32_t ripple = S + lowest;
                            S = ripple | (((S ^ ripple) >> 2) / lowest);
                        }
                        acc += pv * sum;
                    }
                }
                o[b.db[m] + uidx] += acc;
            }
        }
    }
}


template <typename T, int D>
void launch_st_fma(void* out, const void* lhs, const void* rhs, const DevBasis& b, const Pow& pw,
                   int64_t batch, int64_t os, int64_t ls, int64_t rs, hipStream_t st)
{
    const uint32_t grid = uint32_t(std::min<int64_t>(batch, 8192));
    hipLaunchKernelGGL((st_fma_kernel<T, D>), dim3(grid), dim3(kBlock), 0, st, (T*)out,
                       (const T*)lhs, (const T*)rhs, b, pw, batch, os, ls, rs);
}

template <typename T, int D>
void launch_st_adj(void* out, const void* op, const void* arg, const DevBasis& b, const Pow& pw,
                   int64_t batch, int64_t os, int64_t ps, int64_t as, hipStream_t st)
{
    const uint32_t grid = uint32_t(std::min<int64_t>(batch, 8192));
    hipLaunchKernelGGL((st_adj_mul_kernel<T, D>), dim3(grid), dim3(kBlock), 0, st, (T*)out,
                       (const T*)op, (const T*)arg, b, pw, batch, os, ps, as);
}

template <typename T>
void dispatch_st(bool adjoint, void* out, const void* x, const void* y, const DevBasis& b,
                 const Pow& pw, int64_t batch, int64_t s0, int64_t s1, int64_t s2,
                 hipStream_t st)
{
    switch (b.depth) {
#define RPY_D(D)                                                              \
    case D:                                                                   \
        if (adjoint)                                                          \
            launch_st_adj<T, D>(out, x, y, b, pw, batch, s0, s1, s2, st);     \
        else                                                                  \
            launch_st_fma<T, D>(out, x, y, b, pw, batch, s0, s1, s2, st);     \
        break;
        RPY_D(1) RPY_D(2) RPY_D(3) RPY_D(4) RPY_D(5) RPY_D(6)
        RPY_D(7) RPY_D(8) RPY_D(9) RPY_D(10) RPY_D(11) RPY_D(12)
#undef RPY_D
        default: break;
    }
}

int shuffle_common(bool adjoint, void* out, const void* x, const void* y, rpy_dtype dtype,
                   const rpy_tensor_basis* basis, int64_t batch, int64_t s0, int64_t s1,
                   int64_t s2, void* stream)
{
    DevBasis b;
    int rc = make_basis(basis, b);
    if (rc) return rc;
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (b.depth < 1 || b.depth > kMaxShuffleDepth)
        return fail(RPY_ERR_UNSUPPORTED, "shuffle products support depth 1..%d",
                    kMaxShuffleDepth);
    if (b.size() > int64_t(0xffffffff)) return fail(RPY_ERR_UNSUPPORTED, "tensor too large");
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch");
    if (batch > 1 && (s0 < b.size() || s1 < b.size() || s2 < b.size()))
        return fail(RPY_ERR_SHAPE, "batch stride smaller than the tensor size");
    if (batch == 0) return RPY_OK;
    if (!out || !x || !y) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const Pow pw = make_pow(b.width, b.depth);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (dtype == RPY_F64)
        dispatch_st<double>(adjoint, out, x, y, b, pw, batch, s0, s1, s2, st);
    else
        dispatch_st<float>(adjoint, out, x, y, b, pw, batch, s0, s1, s2, st);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

}  // namespace

int dense_st_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                 const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ls,
                 int64_t rs, void* stream)
{
    return shuffle_common(false, out, lhs, rhs, dtype, basis, batch, os, ls, rs, stream);
}

int dense_st_adj_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ps,
                     int64_t as, void* stream)
{
    return shuffle_common(true, out, op, arg, dtype, basis, batch, os, ps, as, stream);
}

}  // namespace rpy_amd
