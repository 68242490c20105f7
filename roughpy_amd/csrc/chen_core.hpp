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
//
// The scaled increments x_a / d come through a row accessor: ScaledRow reads
// them from a table of D*W values built once per increment and shared by all
// lanes of a stream (row[(d-1)*W + a] = x_a * (1/d)), so a step issues only
// fused multiply-adds; RawRow reads the W raw values (and the lane's path
// letters) once and scales in registers, which moves 3x fewer bytes from
// LDS for wide float64 lanes.
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

// Prefix length per shape: the smallest P whose per-lane state fits the
// register budget (about 160 VGPRs of state for float and for double).  A
// smaller P means fewer lanes per stream recomputing the path levels: at width
// 8, depth 4 float64, P = 2 (75 values) runs 2.2x faster than P = 3.
constexpr int choose_p(int W, int D, int budget)
{
    int p = 1;
    while (p < D) {
        int sub = 0;
        for (int j = p; j <= D; ++j) sub += int(ipow(W, j - p));
        if (sub + p <= budget) break;
        ++p;
    }
    return p;
}

constexpr int kBudgetF32 = 160;  // state values per lane (1 VGPR each)
constexpr int kBudgetF64 = 80;   // state values per lane (2 VGPRs each)

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
    static constexpr int ROW = D * W;  // scaled-increment table per step
};

// FMAs one lane issues per Chen step: for each subtree target M the path
// partial (P-1), then the subtree levels P..M; plus the path targets.
template <int W, int D, int P>
constexpr int lane_fmas()
{
    int n = 0;
    for (int M = P; M <= D; ++M) {
        n += P - 1;
        for (int K = P; K <= M; ++K) n += int(ipow(W, K - P));
    }
    for (int M = 1; M < P; ++M) n += M;
    return n;
}

// Fills row[(d-1)*W + a] = x[a] * (1/d) for d = 1..D.
template <typename T, int W, int D>
RPY_HD inline void build_row(T* row, const T* x)
{
    for (int d = 1; d <= D; ++d)
        for (int a = 0; a < W; ++a) row[(d - 1) * W + a] = (d == 1) ? x[a] : x[a] * (T(1) / T(d));
}

// x_a / d read from the scaled table.
template <typename T, int W, int D, int P>
struct ScaledRow {
    static constexpr bool kFold = false;
    const T* row;
    const int (&u)[P];
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T x(int a) const { return row[(d - 1) * W + a]; }
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T xu(int j) const { return row[(d - 1) * W + u[j]]; }
};

// x_a / d computed from the raw increment held in registers.
template <typename T, int W, int D, int P>
struct RawRow {
    static constexpr bool kFold = false;
    T raw[W];
    T rawu[P];
    RPY_HD inline __attribute__((always_inline)) RawRow(const T* row, const int (&u)[P])
    {
        for (int a = 0; a < W; ++a) raw[a] = row[a];
        for (int j = 0; j < P; ++j) rawu[j] = row[u[j]];
    }
    template <int d>
    static RPY_HD inline __attribute__((always_inline)) T sc(T v)
    {
        if constexpr (d == 1)
            return v;
        else
            return v * (T(1) / T(d));
    }
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T x(int a) const { return sc<d>(raw[a]); }
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T xu(int j) const { return sc<d>(rawu[j]); }
};

// The row of one step as the stream-per-workgroup kernel holds it: the raw
// increment x_a (the same for every lane of a stream) in registers, and the
// lane's path letters x_{u_j} / d from a per-step table of scaled values,
// letter-major with `LS` slots per letter (tab[a * LS + d - 1] = x_a / d).
// A subtree stage whose divisor d > 1 scales its partial by 1/d once
// (kFold) instead of reading scaled subtree values.
template <typename T, int W, int D, int P, int LS>
struct UniformRow {
    static constexpr bool kFold = true;
    const T (&raw)[W];
    const T* const (&tab)[P];  // tab[j] = the table row of letter u_j
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T x(int a) const
    {
        static_assert(d == 1, "folded rows hold the raw increment only");
        return raw[a];
    }
    template <int d>
    RPY_HD inline __attribute__((always_inline)) T xu(int j) const { return tab[j][d - 1]; }
};

// Levels whose coefficients the compensated lanes (Comp) accumulate with
// Kahan's compensated summation.  The rounding errors of the lowest levels
// are the ones that grow most: every later step multiplies them into all
// higher levels.  Levels 1 and 2 are one path value and one subtree value
// of a lane at P = 2, so compensating them costs a few operations per step.
constexpr int kCompLevels = 2;

template <typename T, int W, int D, int P, bool Comp = false>
struct Lane {
    using S = Shape<W, D, P>;
    // compensated path levels 1..CP and subtree positions 0..CS-1 (levels
    // P..kCompLevels)
    static constexpr int CP = !Comp ? 0 : (P - 1 < kCompLevels ? P - 1 : kCompLevels);
    static constexpr int CS = (Comp && kCompLevels >= P) ? S::off(kCompLevels + 1) : 0;
    T path[P];
    T sub[S::SUB];
    // Kahan terms: the value of a compensated coefficient is hi - c
    T cpath[CP + 1];
    T csub[CS > 0 ? CS : 1];

    RPY_HD inline __attribute__((always_inline)) void init_identity()
    {
        path[0] = T(1);
        static_for<P - 1>([&](auto k) { path[k + 1] = T(0); });
        static_for<S::SUB>([&](auto i) { sub[i] = T(0); });
        static_for<CP + 1>([&](auto k) { cpath[k] = T(0); });
        static_for<(CS > 0 ? CS : 1)>([&](auto i) { csub[i] = T(0); });
    }

    // Coefficients as written out: subtree position i, path level k.
    template <int i>
    RPY_HD inline __attribute__((always_inline)) T sub_out() const
    {
        if constexpr (i < CS)
            return sub[i] - csub[i];
        else
            return sub[i];
    }
    template <int k>
    RPY_HD inline __attribute__((always_inline)) T path_out() const
    {
        if constexpr (k >= 1 && k <= CP)
            return path[k] - cpath[k];
        else
            return path[k];
    }

    // One Chen step.  acc: row accessor (ScaledRow or RawRow) giving x_a / d
    // and x_{u_j} / d for the lane's path letters u_j.
    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void step(const Acc& acc)
    {
        // subtree targets, top-down
        static_for<D - P + 1>([&](auto i) {
            constexpr int M = D - int(decltype(i)::value);
            target<M>(acc);
        });
        // path targets M = P-1 .. 1
        static_for<P - 1>([&](auto i) {
            constexpr int M = P - 1 - int(decltype(i)::value);
            path_target<M>(acc);
        });
    }

    // The path work of one step, for the lanes whose prefix shares
    // u_0..u_{P-2}: writes the path partial of every subtree target M to
    // pp[(M - P) * stride] (the value target M would compute) and advances
    // the path levels.  Lanes that take these partials (step_pp) do no path
    // work, and the result is the same.
    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void path_step(const Acc& acc, T* pp,
                                                                 int stride)
    {
        static_for<D - P + 1>([&](auto i) {
            constexpr int M = D - int(decltype(i)::value);
            pp[(M - P) * stride] = path_partial<M>(acc);
        });
        static_for<P - 1>([&](auto i) {
            constexpr int M = P - 1 - int(decltype(i)::value);
            path_target<M>(acc);
        });
    }

    // One Chen step of the subtree levels with the path partials given
    // (pp[(M - P) * stride] for target M); the path levels are not advanced.
    template <typename Acc>
    RPY_HD inline __attribute__((always_inline)) void step_pp(const Acc& acc, const T* pp,
                                                               int stride)
    {
        static_for<D - P + 1>([&](auto i) {
            constexpr int M = D - int(decltype(i)::value);
            target_p<M>(acc, pp[(M - P) * stride]);
        });
    }

  private:

    // hi += t with Kahan's compensation c (the value is hi - c).  No term
    // is a product, so contraction cannot change the error term.
    static RPY_HD inline __attribute__((always_inline)) void kahan(T& hi, T& c, T t)
    {
        const T y = t - c;
        const T s = hi + y;
        c = (s - hi) - y;
        hi = s;
    }

    static RPY_HD inline __attribute__((always_inline)) T fmaT(T a, T b, T c)
    {
        // type-exact fused multiply-add (no promotion of float to double)
        if constexpr (std::is_same<T, float>::value)
            return __builtin_fmaf(a, b, c);
        else
            return __builtin_fma(a, b, c);
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) T path_partial(const Acc& acc) const
    {
        // P_{P-1} along the lane's prefix, for target M
        T p = path[0];
        static_for<P - 1>([&](auto k0) {
            constexpr int K = int(decltype(k0)::value) + 1;
            p = fmaT(p, acc.template xu<M - K + 1>(K - 1), path[K]);
        });
        return p;
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void path_target(const Acc& acc)
    {
        // S_M(u_0..u_{M-1}) for M < P
        T p = path[0];
        static_for<M - 1>([&](auto k0) {
            constexpr int K = int(decltype(k0)::value) + 1;
            p = fmaT(p, acc.template xu<M - K + 1>(K - 1), path[K]);
        });
        if constexpr (M <= CP)
            kahan(path[M], cpath[M], p * acc.template xu<1>(M - 1));
        else
            path[M] = fmaT(p, acc.template xu<1>(M - 1), path[M]);
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void target(const Acc& acc)
    {
        target_p<M>(acc, path_partial<M>(acc));
    }

    template <int M, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void target_p(const Acc& acc, const T p)
    {
        // K = P: the single word u
        if constexpr (M == P && CS > 0) {
            kahan(sub[0], csub[0], p * acc.template xu<1>(P - 1));
        } else {
            const T q0 = fmaT(p, acc.template xu<M - P + 1>(P - 1), sub[0]);
            if constexpr (M == P)
                sub[0] = q0;
            else
                stage<P + 1, M, 0>(q0, acc);
        }
    }

    // Depth-first over the subtree below word V (level K-1 of the subtree):
    //   q_K[V W + a] = sub_K[V W + a] + q_{K-1}[V] x[a] / (M - K + 1).
    // Only one partial per level is live at a time, which keeps the register
    // footprint at the lane state plus O(W) temporaries.
    template <int K, int M, int V, typename Acc>
    RPY_HD inline __attribute__((always_inline)) void stage(T qprev, const Acc& acc)
    {
        constexpr int O = S::off(K);
        constexpr int d = M - K + 1;
        // folded rows: q_{K-1} / d once, times the raw x_a for each child
        // (exact for d a power of two)
        const T qs = (Acc::kFold && d > 1) ? qprev * (T(1) / T(d)) : qprev;
        static_for<W>([&](auto a) {
            constexpr int ai = int(decltype(a)::value);
            constexpr int v = V * W + ai;
            T xa;
            if constexpr (Acc::kFold)
                xa = acc.template x<1>(ai);
            else
                xa = acc.template x<d>(ai);
            if constexpr (K == M && O + v < CS) {
                kahan(sub[O + v], csub[O + v], qs * xa);
            } else {
                const T q = fmaT(qs, xa, sub[O + v]);
                if constexpr (K == M)
                    sub[O + v] = q;
                else
                    stage<K + 1, M, v>(q, acc);
            }
        });
    }
};

}  // namespace chen
}  // namespace rpy_amd
