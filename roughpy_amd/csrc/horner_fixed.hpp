// horner_fixed.hpp -- compile-time-shape helpers of the tensor Horner
// kernels (the exp / fmexp / log of dense_ops.hip and the fused log and
// projection of log_map.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "chen_core.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

// number of words shorter than k letters (the offset of level k)
template <int W>
constexpr int fdb(int k)
{
    int s = 0;
    for (int j = 0; j < k; ++j) s += int(ipow(W, j));
    return s;
}

// a <- a (x) x with x's unit dropped, levels ml..1 top-down.  a and xs hold
// levels 0..D-1 in LDS; when ml == D the top level's one x term (lhs degree
// 0) reads the global xg and the top level's outputs go straight to the
// global o (plus m when given): no product ever reads a's top level.  xg
// may be of a narrower type (float32 input of a float64 log).
template <typename T, int W, int D, int ml, typename Tin>
RPY_FI void wg_mul_by_nounit_fixed(T* a, const T* xs, T f, const Tin* __restrict__ xg,
                                   T* __restrict__ o, const T* __restrict__ mg)
{
    chen::static_for<ml>([&](auto mi) {
        constexpr int m = ml - int(decltype(mi)::value);  // level, top-down
        constexpr bool top = m == D;
        constexpr uint32_t tasks = uint32_t(ipow(W, m - 1));
        constexpr uint32_t dbm = uint32_t(fdb<W>(m));
        for (uint32_t u = threadIdx.x; u < tasks; u += blockDim.x) {
            T acc[W];
            chen::static_for<W>([&](auto c) { acc[c] = T(0); });
            chen::static_for<m>([&](auto li) {
                constexpr int ld = m - 1 - int(decltype(li)::value);  // lhs degree, high to low
                constexpr uint32_t k = uint32_t(ipow(W, m - 1 - ld));  // rhs letters from u
                constexpr uint32_t dbl = uint32_t(fdb<W>(ld));
                constexpr uint32_t dbr = uint32_t(fdb<W>(m - ld));
                const uint32_t pi = u / k;
                const uint32_t si = u - pi * k;
                const T av = a[dbl + pi];
                if constexpr (top && ld == 0) {
                    const Tin* xb = xg + dbr + si * W;
                    chen::static_for<W>([&](auto c) { acc[c] += (av * T(xb[c])) * f; });
                } else {
                    const T* xb = xs + dbr + si * W;
                    chen::static_for<W>([&](auto c) { acc[c] += (av * xb[c]) * f; });
                }
            });
            if constexpr (top) {
                T* oo = o + dbm + u * W;
                if (mg)
                    chen::static_for<W>([&](auto c) { oo[c] = acc[c] + mg[dbm + u * W + c]; });
                else
                    chen::static_for<W>([&](auto c) { oo[c] = acc[c]; });
            } else {
                T* ao = a + dbm + u * W;
                chen::static_for<W>([&](auto c) { ao[c] = acc[c]; });
            }
        }
        __syncthreads();
    });
    if (threadIdx.x == 0) a[0] = T(0);
    __syncthreads();
}

}  // namespace

}  // namespace rpy_amd
