// chen_bwd.hpp -- per-lane reverse-Chen step for signature gradients
// (host + device).
//
// Loss l = <G, S>, S = exp(x_1) ... exp(x_L).  Walking t = L .. 1 the lane
// keeps, in the prefix decomposition of chen_core.hpp (lane q owns the words
// with prefix u = digits of q, plus its path words u_0..u_{k-1}, k < P):
//   A = S_{<t}   recovered from S_{<=t} by one Chen step with -x_t,
//   H = cotangent of S_{<=t}  (starts at G),
// and produces dl/dx_t.  Differentiating the Horner chain of the forward
// step (target level M: P_0 = A_0, P_K = A_K + P_{K-1} x_{letter} / (M-K+1),
// S'_M = P_M) gives, edge by edge,
//   adjP_M = H_M,
//   adjP_{K-1}(V) += adjP_K(Va) x_a / (M-K+1),
//   dl/dx_a        += adjP_K(Va) P_{K-1}(V) / (M-K+1),
//   H_new_K        = H_K + sum_{M > K} adjP_K^{(M)}.
// Targets are taken in ascending M so H_M is still old when target M reads
// it and the update of the lower levels can be done in place.
//
// Below level P everything stays in the lane.  At the path levels the
// adjoint of a path word is a sum over all lanes below it; by linearity each
// lane propagates its own share down its path (`dpath`, `gp`) and the shares
// are summed across lanes afterwards: the path-level cotangent update
// (aligned groups of W^(P-k) lanes) and the W gradient values.  Path targets
// (M < P) are evaluated by one designated lane per path word (the lane whose
// remaining digits are all zero) so nothing is counted twice.
#pragma once

#include "chen_core.hpp"

namespace rpy_amd {
namespace chen {

// Increment accessor for the backward step: the scaled table x_a / d lives
// in registers (the whole wave shares one stream, so constant indices), the
// lane's path letters x_{u_j} are loaded per lane.  Neg gives -x (the
// recovery step A <- A (x) exp(-x)).
template <typename T, int W, int D, int P, bool Neg>
struct RegRow {
    static constexpr bool kFold = false;
    const T (&tab)[D * W];  // tab[(d-1) W + a] = x_a * (1/d)
    const T (&pu)[P];       // x_{u_j}
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T x(int a) const
    {
        const T v = tab[(d - 1) * W + a];
        return Neg ? -v : v;
    }
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T xu(int j) const
    {
        const T v = (d == 1) ? pu[j] : pu[j] * (T(1) / T(d));
        return Neg ? -v : v;
    }
};

template <typename T, int W, int D>
RPY_HD inline __attribute__((always_inline)) void build_table(T (&tab)[D * W], const T* x)
{
    static_for<D>([&](auto d0) {
        constexpr int d = int(decltype(d0)::value) + 1;
        static_for<W>([&](auto a0) {
            constexpr int a = int(decltype(a0)::value);
            tab[(d - 1) * W + a] = (d == 1) ? x[a] : x[a] * (T(1) / T(d));
        });
    });
}

template <typename T, int W, int D, int P>
struct BwdLane {
    using S = Shape<W, D, P>;
    Lane<T, W, D, P> A;  // prefix signature S_{<t}
    T hpath[P];          // hpath[k] = H_k(u_0..u_{k-1}), k >= 1
    T hsub[S::SUB];      // H on the lane's subtree, same layout as A.sub
    // per-step shares (valid after contributions())
    T g[W];              // dl/dx_a from subtree edges
    T gp[P];             // dl/dx_{u_j} from the lane's path edge j
    T dpath[P];          // share of the path cotangent update, k >= 1
    int zeros_from;      // designated for path words of level >= zeros_from

    RPY_HD inline void set_prefix(const int (&u)[P])
    {
        zeros_from = P;
        for (int j = P - 1; j >= 0 && u[j] == 0; --j) zeros_from = j;
    }

    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void contributions(const Acc& acc)
    {
        static_for<W>([&](auto a) { g[decltype(a)::value] = T(0); });
        static_for<P>([&](auto j) {
            gp[decltype(j)::value] = T(0);
            dpath[decltype(j)::value] = T(0);
        });
        static_for<P - 1>([&](auto i) { path_target<int(decltype(i)::value) + 1>(acc); });
        static_for<D - P + 1>([&](auto i) { target<P + int(decltype(i)::value)>(acc); });
    }

    // contributions() in two parts for callers that update hpath in
    // between: the subtree targets (which do not read hpath; they also clear
    // the step's accumulators), then the path targets.  The shares are
    // accumulated in a different order than by contributions().
    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void contributions_sub(const Acc& acc)
    {
        static_for<W>([&](auto a) { g[decltype(a)::value] = T(0); });
        static_for<P>([&](auto j) {
            gp[decltype(j)::value] = T(0);
            dpath[decltype(j)::value] = T(0);
        });
        static_for<D - P + 1>([&](auto i) { target<P + int(decltype(i)::value)>(acc); });
    }
    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void contributions_path(const Acc& acc)
    {
        static_for<P - 1>([&](auto i) { path_target<int(decltype(i)::value) + 1>(acc); });
    }

    // this lane's share of dl/dx_a (before the cross-lane sum)
    RPY_HD inline __attribute__((always_inline)) T grad_share(int a, const int (&u)[P]) const
    {
        T v = T(0);
        static_for<W>([&](auto b) {
            if (int(decltype(b)::value) == a) v = g[decltype(b)::value];
        });
        static_for<P>([&](auto j) {
            if (u[decltype(j)::value] == a) v += gp[decltype(j)::value];
        });
        return v;
    }

  private:
    static RPY_HD inline __attribute__((always_inline)) T fmaT(T a, T b, T c)
    {
        if constexpr (std::is_same<T, float>::value)
            return __builtin_fmaf(a, b, c);
        else
            return __builtin_fma(a, b, c);
    }

    template <int d>
    static RPY_HD inline __attribute__((always_inline)) T scale(T v)
    {
        if constexpr (d == 1)
            return v;
        else
            return v * (T(1) / T(d));
    }

    // forward partials P_0..P_{N-1} along the lane's path for target M
    template <int M, int N, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void path_partials(T (&pp)[P], const Acc& acc) const
    {
        pp[0] = A.path[0];
        static_for<N - 1>([&](auto k0) {
            constexpr int K = int(decltype(k0)::value) + 1;
            pp[K] = fmaT(pp[K - 1], acc.template xu<M - K + 1>(K - 1), A.path[K]);
        });
    }

    // the lane's share c of adjP_N at its path word of level N, pushed down
    template <int M, int N, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void path_chain(T c, const T (&pp)[P],
                                                                 const Acc& acc)
    {
        static_for<N>([&](auto i) {
            constexpr int K = N - int(decltype(i)::value);  // N .. 1
            gp[K - 1] = fmaT(scale<M - K + 1>(c), pp[K - 1], gp[K - 1]);
            c = c * acc.template xu<M - K + 1>(K - 1);
            if constexpr (K - 1 >= 1) dpath[K - 1] += c;
        });
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void path_target(const Acc& acc)
    {
        T pp[P];
        path_partials<M, M>(pp, acc);
        const T c = (M >= zeros_from) ? hpath[M] : T(0);
        path_chain<M, M>(c, pp, acc);
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void target(const Acc& acc)
    {
        T pp[P];
        path_partials<M, P>(pp, acc);
        T adj;
        if constexpr (M == P) {
            adj = hsub[0];
        } else {
            const T qP = fmaT(pp[P - 1], acc.template xu<M - P + 1>(P - 1), A.sub[0]);
            adj = stage<P + 1, M, 0>(qP, acc);
            hsub[0] += adj;
        }
        path_chain<M, P>(adj, pp, acc);
    }

    // Node V at subtree level K-1 with forward partial qprev; returns
    // adjP_{K-1}(V) and updates H below V and the gradient.
    template <int K, int M, int V, typename Acc>
    RPY_HD inline __attribute__((always_inline)) T stage(T qprev, const Acc& acc)
    {
        constexpr int O = S::off(K);
        const T qf = scale<M - K + 1>(qprev);
        T adjacc = T(0);
        static_for<W>([&](auto a) {
            constexpr int ai = int(decltype(a)::value);
            constexpr int v = V * W + ai;
            T adjc;
            if constexpr (K == M) {
                adjc = hsub[O + v];
            } else {
                const T q = fmaT(qprev, acc.template x<M - K + 1>(ai), A.sub[O + v]);
                adjc = stage<K + 1, M, v>(q, acc);
                hsub[O + v] += adjc;
            }
            g[ai] = fmaT(adjc, qf, g[ai]);
            adjacc = fmaT(adjc, acc.template x<M - K + 1>(ai), adjacc);
        });
        return adjacc;
    }
};

}  // namespace chen
}  // namespace rpy_amd
