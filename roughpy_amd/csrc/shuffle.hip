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
#include <atomic>
#include <type_traits>

#include "chen.hpp"
#include "chen_core.hpp"
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


// ---------------------------------------------------------------------------
// Tiled shuffle product for fixed width and depth.
//
// One workgroup per tensor pair.  Every output word w = u c of level n (c the
// last letter) is owned by the thread of the prefix u, which keeps the W
// outputs of the tile c = 0..W-1 in registers (the letter tile of the
// reference's v2 kernel, shuffle_tensor_product.hpp:101-179).  For one mask
// of u's letters the lhs and rhs words of u are shared by the whole tile, so
// a mask costs 2 scalar and 2 W-long reads for 2W fused multiply-adds.  (A
// two-letter tile halves the reads per multiply-add but measured slower on
// gfx950: W^2 accumulators and blocks exhaust the registers.)
//
// Levels 0..D-1 of both operands are staged in LDS when they fit.  The top
// level is read only by the top output level under the all-lhs and all-rhs
// masks, once per coefficient, so it is read from global memory there and
// takes no LDS.
//
// Masks are walked in the v1 order (the bit of the letter before the last
// fastest; rhs-last term before lhs-last term), so each output is summed in
// the same order as the reference loop.  The mask of u's letters is split
// into an outer part whose lhs/rhs partial words are rebuilt once per 2^h
// masks and an inner part of the h <= 3 lowest letters unrolled at compile
// time, where the full index is the partial index shifted by a constant
// power of W plus a constant combination of the low letters:
//   db[ld + k] + li = db[ld] + W^ld db[k] + li_hi W^k + li_lo.
// ---------------------------------------------------------------------------

template <int W>
constexpr uint32_t cpow(int k)
{
    return uint32_t(ipow(W, k));
}

// number of words shorter than k letters (the offset of level k)
template <int W>
constexpr uint32_t cdb(int k)
{
    uint32_t s = 0;
    for (int j = 0; j < k; ++j) s += cpow<W>(j);
    return s;
}

template <typename T>
constexpr uint32_t padded(uint32_t n)
{
    constexpr uint32_t v = 16 / sizeof(T);
    return (n + v - 1) / v * v;
}

// The adjoint's arg in LDS: one pad word after every 32, so that reads at a
// stride of a power of the width (a letter of u or v at one position) fall on
// different banks of the 32 four-byte LDS banks rather than on one.
RPY_FI uint32_t adj_pad(uint32_t i) { return i + (i >> 5); }

template <typename T>
constexpr uint32_t adj_arg_words(uint32_t n)
{
    return padded<T>(n + (n >> 5) + 1);
}

// LDS bytes for the adjoint's staged pair (op, padded arg), 0 if it does not
// fit the budget
template <typename T, int W, int D>
constexpr size_t st_lds_bytes()
{
    const size_t b = (size_t(padded<T>(cdb<W>(D + 1))) + adj_arg_words<T>(cdb<W>(D + 1))) *
                     sizeof(T);
    return b <= 160 * 1024 ? b : 0;
}

// the same for levels 0..D-1 only (the product's staging)
template <typename T, int W, int D>
constexpr size_t st_prod_lds_bytes()
{
    const size_t b = 2 * size_t(padded<T>(cdb<W>(D))) * sizeof(T);
    return b <= 160 * 1024 ? b : 0;
}

constexpr int kStInner = 3;  // mask letters unrolled at compile time

template <typename T>
RPY_FI T fmaT(T a, T b, T c)
{
    if constexpr (std::is_same<T, float>::value)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

template <typename T>
RPY_FI void stage_lds(T* __restrict__ dst, const T* __restrict__ src, uint32_t n)
{
    constexpr uint32_t V = 16 / sizeof(T);
    uint32_t done = 0;
    if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
        const uint32_t nv = n / V;
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        for (uint32_t i = threadIdx.x; i < nv; i += blockDim.x) d4[i] = s4[i];
        done = nv * V;
    }
    for (uint32_t i = done + threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

template <int Nv, typename T>
RPY_FI void load_block(T (&dst)[Nv], const T* src, uint32_t off)
{
    chen::static_for<Nv>([&](auto i) { dst[i] = src[off + uint32_t(decltype(i)::value)]; });
}

// lhs/rhs partial words of the outer (high) prefix letters for outer mask om
struct StHigh {
    uint32_t liH, riH, AL, AR, dbL, dbR;
};

template <int W, int m, int h>
RPY_FI StHigh st_high(const uint32_t* dig, uint32_t om)
{
    StHigh s{0, 0, 1, 1, 0, 0};
    chen::static_for<m - h>([&](auto j) {
        constexpr int p = m - int(decltype(j)::value);
        const bool left = (om >> (p - 1 - h)) & 1u;
        const uint32_t d = dig[p - 1];
        s.liH = left ? s.liH * W + d : s.liH;
        s.dbL = left ? s.dbL * W + 1 : s.dbL;
        s.AL = left ? s.AL * W : s.AL;
        s.riH = left ? s.riH : s.riH * W + d;
        s.dbR = left ? s.dbR : s.dbR * W + 1;
        s.AR = left ? s.AR : s.AR * W;
    });
    return s;
}

// One mask (outer part `hs`, inner part im) for the tile of W outputs u c:
// acc[c] += lhs(u_S) rhs(u_S' c) + lhs(u_S c) rhs(u_S').  GlobR / GlobL: the
// rhs / lhs block of level n comes from the global operand (the all-rhs and
// all-lhs masks of the top level).
template <typename T, int W, int h, int im, bool GlobR, bool GlobL>
RPY_FI void st_mask(T* acc, const uint32_t* dig, const StHigh& hs, const T* L, const T* R,
                    const T* GL, const T* GR)
{
    constexpr int k = __builtin_popcount(im);
    uint32_t liL = 0, riL = 0;
    chen::static_for<h>([&](auto j) {
        constexpr int p = h - int(decltype(j)::value);
        if constexpr ((im >> (p - 1)) & 1)
            liL = liL * W + dig[p - 1];
        else
            riL = riL * W + dig[p - 1];
    });
    // block of level ld + j starting at li W^j (lhs), likewise rhs; the
    // powers and level offsets are compile-time constants
    constexpr uint32_t dl0 = cdb<W>(k), dl1 = cdb<W>(k + 1);
    constexpr uint32_t pl0 = cpow<W>(k), pl1 = cpow<W>(k + 1);
    constexpr uint32_t dr0 = cdb<W>(h - k), dr1 = cdb<W>(h - k + 1);
    constexpr uint32_t pr0 = cpow<W>(h - k), pr1 = cpow<W>(h - k + 1);
    const uint32_t l0 = hs.dbL + hs.AL * dl0 + hs.liH * pl0 + liL;
    const uint32_t l1 = hs.dbL + hs.AL * dl1 + hs.liH * pl1 + liL * W;
    const uint32_t r0 = hs.dbR + hs.AR * dr0 + hs.riH * pr0 + riL;
    const uint32_t r1 = hs.dbR + hs.AR * dr1 + hs.riH * pr1 + riL * W;
    const T a0 = L[l0];
    const T b0 = R[r0];
    T rb1[W], lb1[W];
    load_block(rb1, GlobR ? GR : R, r1);
    load_block(lb1, GlobL ? GL : L, l1);
    chen::static_for<W>([&](auto c) {
        acc[c] = fmaT(a0, rb1[c], acc[c]);
        acc[c] = fmaT(lb1[c], b0, acc[c]);
    });
}

// Output level n.  L, R hold levels 0..D-1 (LDS or
// global); GL, GR are the global operands, read for the level-n blocks of the
// top level (Top), whose two masks that reach them -- the first (all letters
// to rhs) and the last (all to lhs) -- are peeled off the loop so that every
// read inside it has one address space.
template <typename T, int W, int n, bool Top>
RPY_FI void st_level(T* __restrict__ o, const T* L, const T* R, const T* GL, const T* GR)
{
    constexpr int m = n - 1;  // prefix letters
    constexpr int h = m < kStInner ? m : kStInner;  // inner (lowest) prefix letters
    constexpr uint32_t tasks = cpow<W>(m);
    constexpr uint32_t outer = 1u << (m - h);
    constexpr int TW = W;
    constexpr int full = (1 << h) - 1;
    for (uint32_t u = threadIdx.x; u < tasks; u += blockDim.x) {
        uint32_t dig[m > 0 ? m : 1];  // dig[p-1] = p-th prefix letter from the end
        {
            uint32_t x = u;
            chen::static_for<m>([&](auto p) {
                dig[p] = x % uint32_t(W);
                x /= uint32_t(W);
            });
        }
        T acc[TW];
        chen::static_for<TW>([&](auto c) { acc[c] = T(0); });
        if constexpr (Top) {
            // first mask; with no prefix letters it is also the last
            const StHigh hs = st_high<W, m, h>(dig, 0);
            st_mask<T, W, h, 0, true, m == 0>(acc, dig, hs, L, R, GL, GR);
        }
        for (uint32_t om = 0; om < outer; ++om) {
            const StHigh hs = st_high<W, m, h>(dig, om);
            chen::static_for<(1 << h)>([&](auto imc) {
                constexpr int im = int(decltype(imc)::value);
                if constexpr (Top) {
                    if (im == 0 && om == 0) return;
                    if (im == full && om == outer - 1) return;
                }
                st_mask<T, W, h, im, false, false>(acc, dig, hs, L, R, GL, GR);
            });
        }
        if constexpr (Top && m > 0) {
            const StHigh hs = st_high<W, m, h>(dig, outer - 1);
            st_mask<T, W, h, full, false, true>(acc, dig, hs, L, R, GL, GR);
        }
        constexpr uint32_t obase = cdb<W>(n);
        T* oo = o + obase + u * uint32_t(TW);
        chen::static_for<TW>([&](auto c) { oo[c] += acc[c]; });
    }
}

template <typename T, int W, int D, bool InLds>
__global__ __launch_bounds__(256) void st_fma_tiled_kernel(T* __restrict__ out,
                                                          const T* __restrict__ lhs,
                                                          const T* __restrict__ rhs,
                                                          int64_t batch, int64_t os, int64_t ls,
                                                          int64_t rs)
{
    constexpr uint32_t NS = cdb<W>(D);  // staged levels 0..D-1
    extern __shared__ __align__(16) unsigned char st_smem[];
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* GL = lhs + e * ls;
        const T* GR = rhs + e * rs;
        const T* L = GL;
        const T* R = GR;
        if constexpr (InLds) {
            T* Ls = reinterpret_cast<T*>(st_smem);
            T* Rs = Ls + padded<T>(NS);
            __syncthreads();  // the previous pair's readers are done
            stage_lds(Ls, GL, NS);
            stage_lds(Rs, GR, NS);
            __syncthreads();
            L = Ls;
            R = Rs;
        }
        if (threadIdx.x == 0) o[0] += L[0] * R[0];
        chen::static_for<D>([&](auto i) {
            constexpr int n = int(decltype(i)::value) + 1;
            st_level<T, W, n, InLds && n == D>(o, L, R, GL, GR);
        });
    }
}

// ---------------------------------------------------------------------------
// Adjoint of shuffle multiplication, fixed width and depth.
//
// out(u) = sum_k sum_{k-subsets S of the m+k positions} sum_{|v| = k}
//          op(v) arg(w), w the interleaving with v on S and u elsewhere
// (position j = j-th letter from the end, weight W^j in the level index).
// The work for an output of level m is sum_k C(m+k, k) W^k terms: thousands
// for m <= 2, one for m = D.  A thread takes the W outputs u' c (c < W) of a
// level m >= 1 together -- one op value per W multiply-adds, as the product's
// letter tile -- and level 0 alone.  Units whose work exceeds about twice the
// per-task grain are "heavy": each (unit, k) job is split over blocks of masks
// and chunks of the v range into sub-jobs whose partial sums land in LDS and
// are added in a fixed order afterwards; "light" units get one thread each.
// Every sum has a fixed order, so results are deterministic.
// ---------------------------------------------------------------------------

constexpr int kAdjMaxParts = 4096;

struct AdjPlan {
    int m0;                                              // heavy levels are m < m0
    int n_heavy;                                         // sub-jobs
    int n_heavy_u;                                       // heavy units
    int parts[kMaxShuffleDepth + 1][kMaxShuffleDepth + 1];  // parts[m][k]
    int chunk[kMaxShuffleDepth + 1][kMaxShuffleDepth + 1];  // v-high values per part
    int scount[kMaxShuffleDepth + 1][kMaxShuffleDepth + 1]; // masks per part
    int per_u[kMaxShuffleDepth + 1];                     // sub-jobs per heavy unit of level m
    int base[kMaxShuffleDepth + 2];                      // first sub-job of level m
    int ubase[kMaxShuffleDepth + 2];                     // first heavy unit of level m
};

template <int W, int D>
RPY_FI uint32_t wpow_rt(int j)
{
    uint32_t r = 1;
#pragma unroll
    for (int q = 0; q < D; ++q)
        if (q < j) r *= uint32_t(W);
    return r;
}

// offset of level j
template <int W, int D>
RPY_FI uint32_t db_rt(int j)
{
    return (wpow_rt<W, D + 1>(j) - 1u) / uint32_t(W - 1);
}

// The j-th k-subset of {0, .., n-1} in increasing mask order (the order of
// Gosper's hack): the combinatorial number system, j = sum_i C(c_i, i) over
// the members c_k > .. > c_1.
RPY_FI uint32_t unrank_subset(uint32_t j, int k, int n)
{
    uint32_t S = 0;
    for (int i = k; i >= 1; --i) {
        int c = i - 1;       // C(i - 1, i) = 0 <= j
        uint32_t bc = 0;     // C(c, i)
        uint32_t bn = 1;     // C(c + 1, i)
        while (c + 1 < n && bn <= j) {
            ++c;
            bc = bn;
            bn = bn * uint32_t(c + 1) / uint32_t(c + 1 - i);
        }
        S |= 1u << c;
        j -= bc;
        n = c;
    }
    return S;
}

// Sum over the k-subsets S of the m+k positions with ranks [sb, sb + sc)
// (in increasing mask order) and v-high values vh in [vb, ve) of
// op(v) arg(interleaving), v = vh W + a, for the TW outputs u0 + c, c < TW:
// acc[c] += the terms of output u0 + c.  With TW = W, u0's last letter is 0
// and the outputs share everything but that letter: for every S it sits at
// the lowest position outside S (weight wc), so output u0 + c reads arg at
// the offset of u0 plus c wc, and one op value serves W multiply-adds.  Each
// acc[c] receives its output's terms in the order of the one-output job.
template <typename T, int W, int D, bool Pad, int TW>
RPY_FI void adj_job(T (&accs)[TW], const T* P, const T* A, int m, uint32_t u, int k, uint32_t vb,
                    uint32_t ve, uint32_t sb = 0, uint32_t sc = ~0u)
{
    auto ar = [&](uint32_t i) { return A[Pad ? adj_pad(i) : i]; };
    const int n = m + k;
    if (k == 0) {
        const T p0 = P[0];
        chen::static_for<TW>([&](auto c) {
            accs[c] = fmaT(p0, ar(db_rt<W, D>(m) + u + uint32_t(decltype(c)::value)), accs[c]);
        });
        return;
    }
    T acc[TW];
    chen::static_for<TW>([&](auto c) { acc[c] = T(0); });
    const uint32_t dbk = db_rt<W, D>(k);
    const uint32_t dbn = db_rt<W, D>(n);
    const uint32_t limit = 1u << n;
    uint32_t S = sb == 0 ? (1u << k) - 1u : unrank_subset(sb, k, n);
    for (uint32_t cnt = 0; cnt < sc && S < limit; ++cnt) {
        // U: the u letters on the complement of S (lowest position = last
        // letter); wp[i] = W^(position of the i-th lowest bit of S); for
        // W = 2^b, M: the bit fields of S's positions above its lowest;
        // wc: W^(lowest position outside S), u's last letter
        uint32_t U = 0;
        uint32_t wp[D];
        uint32_t M = 0;
        uint32_t wc = 0;
        if constexpr ((W & (W - 1)) == 0) {
            // a level index holds letter j in bits [j b, j b + b): U is u
            // with an empty field opened at each position of S, in one pass
            // over S's k bits
            constexpr uint32_t b = uint32_t(__builtin_ctz(uint32_t(W)));
            U = u;
            uint32_t t = S;
            chen::static_for<D>([&](auto ic) {
                constexpr int i = int(decltype(ic)::value);
                wp[i] = 0;
                if (i < k) {
                    const uint32_t sh = uint32_t(__builtin_ctz(t)) * b;
                    t &= t - 1u;
                    const uint32_t low = (1u << sh) - 1u;
                    U = ((U & ~low) << b) | (U & low);
                    wp[i] = 1u << sh;
                    if (i > 0) M |= uint32_t(W - 1) << sh;
                }
            });
            if constexpr (TW > 1) wc = 1u << (uint32_t(__builtin_ctz(~S)) * b);
        } else {
            uint32_t ur = u;
            chen::static_for<D>([&](auto jc) {
                constexpr int j = int(decltype(jc)::value);
                if (j < n && !((S >> j) & 1u)) {
                    U += (ur % uint32_t(W)) * cpow<W>(j);
                    ur /= uint32_t(W);
                }
            });
            uint32_t t = S;
            chen::static_for<D>([&](auto ic) {
                constexpr int i = int(decltype(ic)::value);
                wp[i] = 0;
                if (i < k) {
                    wp[i] = wpow_rt<W, D>(__builtin_ctz(t));
                    t &= t - 1u;
                }
            });
            if constexpr (TW > 1) wc = wpow_rt<W, D>(__builtin_ctz(~S));
        }
        // v-high digits of vb and the matching offset
        uint32_t dv[D];
        uint32_t V = 0;
        {
            uint32_t x = vb;
            chen::static_for<D>([&](auto ic) {
                constexpr int i = int(decltype(ic)::value);
                dv[i] = 0;
                if (i >= 1 && i < k) {
                    dv[i] = x % uint32_t(W);
                    x /= uint32_t(W);
                    V += dv[i] * wp[i];
                }
            });
        }
        // one op value times the tile's TW arg values
        auto terms = [&](T pv, uint32_t ai) {
            chen::static_for<TW>([&](auto c) {
                constexpr uint32_t cv = uint32_t(decltype(c)::value);
                acc[cv] = fmaT(pv, ar(ai + cv * wc), acc[cv]);
            });
        };
        if constexpr ((W & (W - 1)) == 0) {
            // W = 2^b: a level index holds letter j in bits [j b, j b + b), so
            // the v-high offset V is vh's digits deposited into the bit fields
            // of S's positions above its lowest, and vh + 1 is the masked
            // increment ((V | ~M) + 1) & M over those fields
            constexpr uint32_t b = uint32_t(__builtin_ctz(uint32_t(W)));
            uint32_t ob = dbk + vb * W;
            // (float64 at W >= 4 measured slower unrolled: (4,6) 6.87 -> 7.15 ms;
            // so did the float32 tiles, 2.49 -> 2.95 ms, out of registers)
            constexpr bool kTwo = (sizeof(T) == 4 || W * W <= 4) && D >= 2 && TW == 1;
            if (kTwo && k >= 2 && vb % uint32_t(W) == 0 && ve % uint32_t(W) == 0) {
                // v's last two letters unrolled (W^2 terms per step, in the
                // same order as one letter at a time): the walk runs over
                // the fields above the second-lowest position of S
                const uint32_t S1 = S & (S - 1u);
                const uint32_t M2 = M & ~(uint32_t(W - 1) << (uint32_t(__builtin_ctz(S1)) * b));
                for (uint32_t vh = vb; vh < ve; vh += uint32_t(W)) {
                    const uint32_t ab = dbn + U + V;
                    chen::static_for<W>([&](auto a1) {
                        constexpr uint32_t a1v = uint32_t(decltype(a1)::value);
                        chen::static_for<W>([&](auto a0) {
                            constexpr uint32_t a0v = uint32_t(decltype(a0)::value);
                            terms(P[ob + a1v * W + a0v], ab + a1v * wp[D >= 2 ? 1 : 0] + a0v * wp[0]);
                        });
                    });
                    V = ((V | ~M2) + 1u) & M2;
                    ob += uint32_t(W * W);
                }
            } else {
                for (uint32_t vh = vb; vh < ve; ++vh) {
                    const uint32_t ab = dbn + U + V;
                    chen::static_for<W>([&](auto a) {
                        terms(P[ob + a], ab + uint32_t(decltype(a)::value) * wp[0]);
                    });
                    V = ((V | ~M) + 1u) & M;
                    ob += uint32_t(W);
                }
            }
        } else {
            for (uint32_t vh = vb; vh < ve; ++vh) {
                const uint32_t ob = dbk + vh * W;
                const uint32_t ab = dbn + U + V;
                chen::static_for<W>([&](auto a) {
                    terms(P[ob + a], ab + uint32_t(decltype(a)::value) * wp[0]);
                });
                // odometer over the v-high digits
                bool carry = true;
                chen::static_for<D>([&](auto ic) {
                    constexpr int i = int(decltype(ic)::value);
                    if (i >= 1 && i < k && carry) {
                        V += wp[i];
                        if (++dv[i] == uint32_t(W)) {
                            dv[i] = 0;
                            V -= uint32_t(W) * wp[i];
                        } else {
                            carry = false;
                        }
                    }
                });
            }
        }
        // next k-subset (Gosper)
        const uint32_t lowest = S & (0u - S);
        const uint32_t ripple = S + lowest;
        S = ripple | (((S ^ ripple) >> 2) / lowest);
    }
    chen::static_for<TW>([&](auto c) { accs[c] += acc[c]; });
}

// Outputs of level m >= 1 go in units of W (the tile u' c, c < W); level 0
// is one unit of one output.
template <int W>
RPY_FI uint32_t adj_units(int m)
{
    uint32_t r = 1;
    for (int j = 1; j < m; ++j) r *= uint32_t(W);
    return r;
}

// Threads per workgroup: a float64 pair that takes more than half the LDS
// runs alone on its CU, and 16 waves then hide the gather's LDS latency
// ((4,6): 6.00 -> 4.67 ms against 8 waves); other pairs share the CU with 4
// waves per workgroup.
template <typename T, int W, int D>
constexpr int adj_block()
{
    return (sizeof(T) == 8 && st_lds_bytes<T, W, D>() > 80 * 1024) ? 1024 : 256;
}

template <typename T, int W, int D, bool InLds>
__global__ __launch_bounds__((adj_block<T, W, D>())) void st_adj_tiled_kernel(T* __restrict__ out,
                                                           const T* __restrict__ op,
                                                           const T* __restrict__ arg,
                                                           int64_t batch, int64_t os, int64_t ps,
                                                           int64_t as, AdjPlan plan)
{
    constexpr uint32_t N = cdb<W>(D + 1);
    extern __shared__ __align__(16) unsigned char st_smem[];
    T* part = reinterpret_cast<T*>(st_smem) + (InLds ? padded<T>(N) + adj_arg_words<T>(N) : 0);
    // light units, level m0 and up
    uint32_t n_light = 0;
    for (int m = plan.m0; m <= D; ++m) n_light += adj_units<W>(m);
    const uint32_t total = uint32_t(plan.n_heavy) + n_light;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* P;
        const T* A;
        __syncthreads();  // the previous pair's readers are done
        if constexpr (InLds) {
            T* Ps = reinterpret_cast<T*>(st_smem);
            T* As = Ps + padded<T>(N);
            stage_lds(Ps, op + e * ps, N);
            const T* a = arg + e * as;
            for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) As[adj_pad(i)] = a[i];
            __syncthreads();
            P = Ps;
            A = As;
        } else {
            P = op + e * ps;
            A = arg + e * as;
        }
        for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
            if (t < uint32_t(plan.n_heavy)) {
                int m = 0;
                while (m + 1 < plan.m0 && t >= uint32_t(plan.base[m + 1])) ++m;
                const uint32_t r0 = t - uint32_t(plan.base[m]);
                const uint32_t ui = r0 / uint32_t(plan.per_u[m]);
                uint32_t r = r0 - ui * uint32_t(plan.per_u[m]);
                int k = 0;
                while (r >= uint32_t(plan.parts[m][k])) {
                    r -= uint32_t(plan.parts[m][k]);
                    ++k;
                }
                T s[W];
                chen::static_for<W>([&](auto c) { s[c] = T(0); });
                uint32_t vb = 0, ve = 1, sb = 0, sc = ~0u;
                if (k > 0) {
                    // part r: mask block r / vparts, v-high block r % vparts
                    const uint32_t vh_n = wpow_rt<W, D>(k - 1);
                    const uint32_t c = uint32_t(plan.chunk[m][k]);
                    const uint32_t vparts = (vh_n + c - 1) / c;
                    const uint32_t pb = r / vparts;
                    vb = (r - pb * vparts) * c;
                    ve = vb + c < vh_n ? vb + c : vh_n;
                    sc = uint32_t(plan.scount[m][k]);
                    sb = pb * sc;
                }
                if (vb < ve) {
                    if (m == 0) {
                        T s1[1] = {T(0)};
                        adj_job<T, W, D, InLds, 1>(s1, P, A, 0, 0, k, vb, ve, sb, sc);
                        s[0] = s1[0];
                    } else {
                        adj_job<T, W, D, InLds, W>(s, P, A, m, ui * uint32_t(W), k, vb, ve, sb,
                                                   sc);
                    }
                }
                chen::static_for<W>([&](auto c) { part[t * W + uint32_t(decltype(c)::value)] = s[c]; });
            } else {
                uint32_t r = t - uint32_t(plan.n_heavy);
                int m = plan.m0;
                while (r >= adj_units<W>(m)) {
                    r -= adj_units<W>(m);
                    ++m;
                }
                if (m == 0) {
                    T s[1] = {T(0)};
                    for (int k = 0; k <= D; ++k) {
                        T sk[1] = {T(0)};
                        adj_job<T, W, D, InLds, 1>(sk, P, A, 0, 0, k, 0,
                                                   k == 0 ? 1u : wpow_rt<W, D>(k - 1));
                        s[0] += sk[0];
                    }
                    o[0] += s[0];
                } else {
                    T s[W];
                    chen::static_for<W>([&](auto c) { s[c] = T(0); });
                    for (int k = 0; m + k <= D; ++k) {
                        T sk[W];
                        chen::static_for<W>([&](auto c) { sk[c] = T(0); });
                        adj_job<T, W, D, InLds, W>(sk, P, A, m, r * uint32_t(W), k, 0,
                                                   k == 0 ? 1u : wpow_rt<W, D>(k - 1));
                        chen::static_for<W>([&](auto c) { s[c] += sk[c]; });
                    }
                    T* oo = o + db_rt<W, D>(m) + r * uint32_t(W);
                    chen::static_for<W>([&](auto c) { oo[uint32_t(decltype(c)::value)] += s[c]; });
                }
            }
        }
        __syncthreads();
        // heavy outputs: each (unit, letter) sums its sub-jobs in order
        for (uint32_t hc = threadIdx.x; hc < uint32_t(plan.n_heavy_u) * W; hc += blockDim.x) {
            const uint32_t hu = hc / uint32_t(W);
            const uint32_t c = hc - hu * uint32_t(W);
            int m = 0;
            while (m + 1 < plan.m0 && hu >= uint32_t(plan.ubase[m + 1])) ++m;
            if (m == 0 && c > 0) continue;
            const uint32_t ui = hu - uint32_t(plan.ubase[m]);
            const uint32_t p0 = uint32_t(plan.base[m]) + ui * uint32_t(plan.per_u[m]);
            T s = T(0);
            for (int q = 0; q < plan.per_u[m]; ++q) s += part[(p0 + uint32_t(q)) * W + c];
            o[m == 0 ? 0u : db_rt<W, D>(m) + ui * uint32_t(W) + c] += s;
        }
    }
}

// Splits the adjoint work of a (W, D) tensor for `block` threads.  A unit is
// the W outputs u' c of a level m >= 1 (one output at level 0); units whose
// work exceeds about twice the grain are heavy and split into sub-jobs, each
// with W partial sums in LDS (at most kAdjMaxParts values).
inline AdjPlan make_adj_plan(int W, int D, int block)
{
    auto binom = [](int a, int b) {
        double r = 1;
        for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
        return r;
    };
    auto wp = [&](int e) { return double(ipow(W, e)); };
    auto units = [&](int m) { return m == 0 ? 1.0 : wp(m - 1); };
    auto width = [&](int m) { return m == 0 ? 1.0 : double(W); };
    double total = 0;
    for (int n = 0; n <= D; ++n) total += wp(n) * double(int64_t(1) << n);
    double grain = std::max(64.0, total / (4.0 * block));
    AdjPlan p{};
    for (;;) {
        p = AdjPlan{};
        // heavy levels: every level up to the last one whose unit work
        // exceeds twice the grain (the work is not monotone in m)
        p.m0 = 0;
        for (int m = 0; m <= D; ++m) {
            double work = 0;
            for (int k = 0; m + k <= D; ++k) work += binom(m + k, k) * wp(k);
            if (work * width(m) > 2 * grain) p.m0 = m + 1;
        }
        int64_t nh = 0, nu = 0;
        for (int m = 0; m < p.m0; ++m) {
            p.base[m] = int(nh);
            p.ubase[m] = int(nu);
            int per = 0;
            for (int k = 0; m + k <= D; ++k) {
                int parts = 1;
                int chunk = 1;
                int scount = 1;
                if (k >= 1) {
                    // split over blocks of masks first (each mask's setup is
                    // then done once), then over blocks of v-high values
                    const double vh = wp(k - 1);
                    const double masks = binom(m + k, k);
                    const double target = std::min(
                            masks * vh,
                            std::max(1.0, std::ceil(masks * wp(k) * width(m) / grain)));
                    scount = int(std::ceil(masks / std::min(masks, target)));
                    const int sparts = int(std::ceil(masks / scount));
                    const double vp = std::min(vh, std::max(1.0, std::ceil(target / sparts)));
                    chunk = int(std::ceil(vh / vp));
                    parts = sparts * int(std::ceil(vh / chunk));
                }
                p.parts[m][k] = parts;
                p.chunk[m][k] = chunk;
                p.scount[m][k] = scount;
                per += parts;
            }
            p.per_u[m] = per;
            nh += int64_t(per) * int64_t(units(m));
            nu += int64_t(units(m));
        }
        p.base[p.m0] = int(nh);
        p.ubase[p.m0] = int(nu);
        p.n_heavy = int(nh);
        p.n_heavy_u = int(nu);
        if (nh * W <= kAdjMaxParts) break;
        grain *= 2;
    }
    return p;
}

size_t st_device_lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 64 * 1024;
    return size_t(v);
}

template <typename T, int W, int D>
bool launch_st_tiled(void* out, const void* lhs, const void* rhs, int64_t batch, int64_t os,
                     int64_t ls, int64_t rs, hipStream_t st)
{
    constexpr size_t lds = st_prod_lds_bytes<T, W, D>();
    constexpr bool in_lds = lds > 0;
    const uint32_t grid = uint32_t(std::min<int64_t>(batch, 16384));
    // threads: the top level's tasks, rounded to whole waves, at most 256
    constexpr int top_tasks = int(cpow<W>(D - 1));
    constexpr int block = top_tasks >= 256 ? 256 : (top_tasks <= 64 ? 64 : (top_tasks + 63) / 64 * 64);
    if constexpr (in_lds) {
        if (lds > st_device_lds_limit()) return false;
        static std::atomic<bool> reserved{false};
        if (lds > 64 * 1024 && !reserved.load(std::memory_order_acquire)) {
            if (hipFuncSetAttribute((const void*)st_fma_tiled_kernel<T, W, D, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)) != hipSuccess)
                return false;
            reserved.store(true, std::memory_order_release);
        }
        hipLaunchKernelGGL((st_fma_tiled_kernel<T, W, D, true>), dim3(grid), dim3(block), lds,
                           st, (T*)out, (const T*)lhs, (const T*)rhs, batch, os, ls, rs);
    } else {
        hipLaunchKernelGGL((st_fma_tiled_kernel<T, W, D, false>), dim3(grid), dim3(block), 0, st,
                           (T*)out, (const T*)lhs, (const T*)rhs, batch, os, ls, rs);
    }
    return true;
}

template <typename T, int W, int D>
bool launch_st_adj_tiled(void* out, const void* op, const void* arg, int64_t batch, int64_t os,
                         int64_t ps, int64_t as, hipStream_t st)
{
    constexpr size_t staged = st_lds_bytes<T, W, D>();
    constexpr bool in_lds = staged > 0;
    constexpr int block = adj_block<T, W, D>();
    static const AdjPlan plan = make_adj_plan(W, D, block);
    const size_t lds = staged + (size_t(plan.n_heavy) * W * sizeof(T) + 15) / 16 * 16;
    if (lds > st_device_lds_limit()) return false;
    const uint32_t grid = uint32_t(std::min<int64_t>(batch, 16384));
    static std::atomic<bool> reserved{false};
    if (lds > 64 * 1024 && !reserved.load(std::memory_order_acquire)) {
        if (hipFuncSetAttribute((const void*)st_adj_tiled_kernel<T, W, D, in_lds>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                int(lds)) != hipSuccess)
            return false;
        reserved.store(true, std::memory_order_release);
    }
    hipLaunchKernelGGL((st_adj_tiled_kernel<T, W, D, in_lds>), dim3(grid), dim3(block), lds, st,
                       (T*)out, (const T*)op, (const T*)arg, batch, os, ps, as, plan);
    return true;
}

template <typename T>
using StTiledFn = bool (*)(void*, const void*, const void*, int64_t, int64_t, int64_t, int64_t,
                           hipStream_t);

// shapes with a fixed-shape kernel: width 2..8, depth up to the largest one
// whose tensor stays near 2^21 coefficients
template <typename T>
StTiledFn<T> find_st_tiled(int w, int d, bool adjoint)
{
#define RPY_S(W, D) \
    if (w == W && d == D) return adjoint ? &launch_st_adj_tiled<T, W, D> : &launch_st_tiled<T, W, D>;
    RPY_S(2, 1) RPY_S(2, 2) RPY_S(2, 3) RPY_S(2, 4) RPY_S(2, 5) RPY_S(2, 6)
    RPY_S(2, 7) RPY_S(2, 8) RPY_S(2, 9) RPY_S(2, 10) RPY_S(2, 11) RPY_S(2, 12)
    RPY_S(3, 1) RPY_S(3, 2) RPY_S(3, 3) RPY_S(3, 4) RPY_S(3, 5) RPY_S(3, 6)
    RPY_S(3, 7) RPY_S(3, 8) RPY_S(3, 9) RPY_S(3, 10)
    RPY_S(4, 1) RPY_S(4, 2) RPY_S(4, 3) RPY_S(4, 4) RPY_S(4, 5) RPY_S(4, 6)
    RPY_S(4, 7) RPY_S(4, 8)
    RPY_S(5, 1) RPY_S(5, 2) RPY_S(5, 3) RPY_S(5, 4) RPY_S(5, 5) RPY_S(5, 6) RPY_S(5, 7)
    RPY_S(6, 1) RPY_S(6, 2) RPY_S(6, 3) RPY_S(6, 4) RPY_S(6, 5) RPY_S(6, 6)
    RPY_S(7, 1) RPY_S(7, 2) RPY_S(7, 3) RPY_S(7, 4) RPY_S(7, 5) RPY_S(7, 6)
    RPY_S(8, 1) RPY_S(8, 2) RPY_S(8, 3) RPY_S(8, 4) RPY_S(8, 5) RPY_S(8, 6)
#undef RPY_S
    return nullptr;
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
    {
        bool done = false;
        if (dtype == RPY_F64) {
            if (auto fn = find_st_tiled<double>(b.width, b.depth, adjoint))
                done = fn(out, x, y, batch, s0, s1, s2, st);
        } else {
            if (auto fn = find_st_tiled<float>(b.width, b.depth, adjoint))
                done = fn(out, x, y, batch, s0, s1, s2, st);
        }
        if (done) {
            RPY_HIP_CHECK(hipGetLastError());
            return RPY_OK;
        }
    }
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
