// chen_core.hpp -- per-lane Chen-step signature update (host + device).
//
// Decomposition.  A stream's running signature S (width W, depth D) is split
// across W^P lanes.  Lane q owns the prefix word u = (u_0..u_{P-1}) (digits of
// q, first letter most significant) and holds in registers
//   path[k] = S_k(u_0..u_{k-1})          for k < P   (P values, path[0] = S_0)
//   sub     = S_k(u v), |v| = k - P      for k >= P  (sum_k W^(k-P) values)
// One Chen step S <- S (x) exp(x) for a degree-one increment x only ever
// reads, for a word with prefix u, coefficients of that word's own prefixes.
// So each lane updates its subtree with no communication at all; the path
// levels are replicated (recomputed identically) by every lane.
//
// Update (Horner form of free_tensor_fmexp.hpp:15-40 specialised to a
// degree-one exponent): for every target level M,
//   P_0 = S_0,  P_K = S_K + P_{K-1} x_{letter K} / (M - K + 1),  S_M' = P_M,
// which equals sum_j S_{M-j} x^(j) / j!.  Targets run from M = D down to 1,
// so lower levels are still old when higher targets read them.  The partial
// P_K of a prefix is shared by all W^(M-K) words below it, which makes the
// total work sum_K (D - K + 1) W^K fused multiply-adds per step -- the same
// count as the reference's truncated in-place Horner.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "rpy_common.hpp"

namespace rpy_amd {
namespace chen {

template <typename F, int... I>
RPY_HD inline __attribute__((always_inline)) void static_for_impl(
        F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

// Calls f(integral_constant<int, i>) for i = 0 .. N-1, fully unrolled.
template <int N, typename F>
RPY_HD inline __attribute__((always_inline)) void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int W, int D, int P>
struct Shape {
    static_assert(W >= 1 && D >= 1 && P >= 1 && P <= D, "bad shape");
    static constexpr int NP = int(ipow(W, P));  // lanes per stream
    // offset of subtree level k (P <= k <= D+1) inside `sub`
    static constexpr int off(int k)
    {
        int s = 0;
        for (int j = P; j < k; ++j) s += int(ipow(W, j - P));
        return s;
    }
    static constexpr int SUB = off(D + 1);
};

template <typename T, int W, int D, int P>
struct Lane {
    using S = Shape<W, D, P>;
    T path[P];
    T sub[S::SUB];

    RPY_HD inline __attribute__((always_inline)) void init_identity()
    {
        path[0] = T(1);
        static_for<P - 1>([&](auto k) { path[k + 1] = T(0); });
        static_for<S::SUB>([&](auto i) { sub[i] = T(0); });
    }

    // x[a]: increment; xu[j] = x[u_j] (the lane's path letters).
    RPY_HD inline __attribute__((always_inline)) void step(const T (&x)[W], const T (&xu)[P])
    {
        // subtree targets, top-down
        static_for<D - P + 1>([&](auto i) {
            constexpr int M = D - int(decltype(i)::value);
            target<M>(x, xu);
        });
        // path targets M = P-1 .. 1
        static_for<P - 1>([&](auto i) {
            constexpr int M = P - 1 - int(decltype(i)::value);
            path_target<M>(xu);
        });
    }

  private:
    template <int d>
    static RPY_HD inline __attribute__((always_inline)) T scaled(T v)
    {
        if constexpr (d == 1)
            return v;
        else
            return v * (T(1) / T(d));
    }

    static RPY_HD inline __attribute__((always_inline)) T fmaT(T a, T b, T c)
    {
#ifdef __HIP_DEVICE_COMPILE__
        return __builtin_fma(a, b, c);
#else
        return std::fma(a, b, c);
#endif
    }

    template <int M>
    RPY_HD inline __attribute__((always_inline)) T path_partial(const T (&xu)[P]) const
    {
        // P_{P-1} along the lane's prefix, for target M
        T p = path[0];
        static_for<P - 1>([&](auto k0) {
            constexpr int K = int(decltype(k0)::value) + 1;
            p = fmaT(p, scaled<M - K + 1>(xu[K - 1]), path[K]);
        });
        return p;
    }

    template <int M>
    RPY_HD inline __attribute__((always_inline)) void path_target(const T (&xu)[P])
    {
        // S_M(u_0..u_{M-1}) for M < P
        T p = path[0];
        static_for<M - 1>([&](auto k0) {
            constexpr int K = int(decltype(k0)::value) + 1;
            p = fmaT(p, scaled<M - K + 1>(xu[K - 1]), path[K]);
        });
        path[M] = fmaT(p, xu[M - 1], path[M]);
    }

    template <int M>
    RPY_HD inline __attribute__((always_inline)) void target(const T (&x)[W], const T (&xu)[P])
    {
        const T p = path_partial<M>(xu);
        // K = P: the single word u
        const T q0 = fmaT(p, scaled<M - P + 1>(xu[P - 1]), sub[0]);
        if constexpr (M == P) {
            sub[0] = q0;
        } else {
            T q[1] = {q0};
            stage<P + 1, M, 1>(q, x);
        }
    }

    // q_K[v] = sub_K[v] + q_{K-1}[v / W] * x[v % W] / (M - K + 1)
    template <int K, int M, int N>
    RPY_HD inline __attribute__((always_inline)) void stage(const T (&qprev)[N], const T (&x)[W])
    {
        constexpr int NK = int(ipow(W, K - P));
        constexpr int O = S::off(K);
        T xs[W];
        static_for<W>([&](auto a) { xs[a] = scaled<M - K + 1>(x[a]); });
        if constexpr (K == M) {
            static_for<NK>([&](auto v) {
                constexpr int vi = int(decltype(v)::value);
                sub[O + vi] = fmaT(qprev[vi / W], xs[vi % W], sub[O + vi]);
            });
        } else {
            T q[NK];
            static_for<NK>([&](auto v) {
                constexpr int vi = int(decltype(v)::value);
                q[vi] = fmaT(qprev[vi / W], xs[vi % W], sub[O + vi]);
            });
            stage<K + 1, M, NK>(q, x);
        }
    }
};

}  // namespace chen
}  // namespace rpy_amd
