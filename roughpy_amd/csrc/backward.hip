// backward.hip -- reverse-mode derivatives of batched signatures.
//
// The reference has only a forward-mode directional derivative
// (LiteContext::sig_derivative_impl, algebra/src/lite_context.h:580-636).
// Training needs the vector-Jacobian product of
//     S = exp(x_1) exp(x_2) ... exp(x_L)
// with an upstream cotangent G, i.e. dl/dx_t for l = <G, S>.
//
// Reverse Chen recomputation: walking t = L .. 1, keep
//   A   = S_{<=t}, recovered from S_{<=t+1} by A <- A (x) exp(-x_{t+1}),
//   H_t = R*_{exp(x_{t+1}) ... exp(x_L)} G  (cotangent of S_{<=t}),
// so no prefix signature is stored.  One step back-propagates through the
// Horner chain of the forward Chen step (for every target level M,
// P_K = A_K + P_{K-1} x / (M-K+1), S'_M = P_M):
//   adjP_M = H_M,   adjP_{K-1}(w') = sum_a adjP_K(w'a) x_a / (M-K+1),
//   dl/dx_a += sum_{w'} adjP_K(w'a) P_{K-1}(w') / (M-K+1),
//   H_{t-1,K} = sum_{M >= K} adjP^{(M)}_K   (= R*_{exp(x_t)} H_t).
// Two kernels:
// * sig_vjp_wave_kernel: one workgroup of W^P lanes (whole waves) per
//   stream, the lanes splitting the words by prefix (chen_bwd.hpp), A and H
//   in registers, cross-lane sums by shuffles and, across waves, through LDS;
//   widths 2, 4 and 8 (width 8 depth <= 4, width 4 depth <= 6, width 2
//   depth 6..10).
// * sig_vjp_kernel: any shape; one workgroup per stream with its four
//   working tensors in LDS when they fit, else in a global scratch slot.
//
// Also the adjoint of right multiplication, R*_B(g)(u) = sum_v g(uv) B(v),
// the right-hand counterpart of free_tensor_adjoint_left_mul.hpp:10-77.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <type_traits>
#include <vector>

#include "chen.hpp"
#include "chen_bwd.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"
#include "tuning.hpp"

namespace rpy_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxW = 32;

template <typename T>
RPY_FI T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// A <- A (x) exp(x) (x given, degree one), per-coefficient Horner, top-down.
template <typename T>
RPY_FI void wg_chen_step_global(T* A, const T* x, const T* inv, const DevBasis& b,
                                const Pow& pw)
{
    const uint32_t W = uint32_t(b.width);
    for (int m = b.depth; m >= 1; --m) {
        const uint32_t n = uint32_t(b.level_size(m));
        for (uint32_t idx = threadIdx.x; idx < n; idx += kBlock) {
            T p = A[0];
            for (int K = 1; K <= m; ++K) {
                const uint32_t pre = fdiv(idx, pw.pw[m - K]);
                const uint32_t letter = pre - fdiv(pre, pw.pw[1]) * W;
                p = p * (x[letter] * inv[m - K + 1]) + A[b.db[K] + pre];
            }
            A[b.db[m] + idx] = p;
        }
        __syncthreads();
    }
}

// Forward partial P_{K}(w) of target M for the word w (|w| = K), from A.
template <typename T>
RPY_FI T forward_partial(const T* A, const T* x, const T* inv, const DevBasis& b, const Pow& pw,
                         int M, int K, uint32_t w)
{
    const uint32_t W = uint32_t(b.width);
    T p = A[0];
    for (int j = 1; j <= K; ++j) {
        const uint32_t pre = fdiv(w, pw.pw[K - j]);
        const uint32_t letter = pre - fdiv(pre, pw.pw[1]) * W;
        p = p * (x[letter] * inv[M - j + 1]) + A[b.db[j] + pre];
    }
    return p;
}

template <typename T, bool InLds>
__global__ __launch_bounds__(kBlock) void sig_vjp_kernel(
        T* __restrict__ grad, const T* __restrict__ incr, const int64_t* __restrict__ offsets,
        int64_t n_streams, const T* __restrict__ sig, const T* __restrict__ gsig,
        T* __restrict__ scratch, DevBasis b, Pow pw)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int W = b.width;
    const int D = b.depth;
    const int64_t tsize = b.size();
    T* A = InLds ? reinterpret_cast<T*>(smem) : scratch + int64_t(blockIdx.x) * 4 * tsize;
    T* Hc = A + tsize;
    T* Hn = Hc + tsize;
    T* Q = Hn + tsize;
    __shared__ T xs[kMaxW];
    __shared__ T xneg[kMaxW];
    __shared__ T inv[kMaxDepth + 2];
    __shared__ T red[kBlock / 64][kMaxW];
    if (threadIdx.x <= kMaxDepth + 1) inv[threadIdx.x] = threadIdx.x ? T(1) / T(threadIdx.x) : T(0);

    for (int64_t s = blockIdx.x; s < n_streams; s += gridDim.x) {
        const int64_t r0 = offsets[s], r1 = offsets[s + 1];
        for (int64_t i = threadIdx.x; i < tsize; i += kBlock) {
            A[i] = sig[s * tsize + i];
            Hc[i] = gsig[s * tsize + i];
        }
        __syncthreads();
        for (int64_t row = r1 - 1; row >= r0; --row) {
            if (threadIdx.x < W) {
                const T v = incr[row * W + threadIdx.x];
                xs[threadIdx.x] = v;
                xneg[threadIdx.x] = -v;
            }
            for (int64_t i = threadIdx.x; i < tsize; i += kBlock) Hn[i] = T(0);
            __syncthreads();
            // A <- S_{< row}
            wg_chen_step_global(A, xneg, inv, b, pw);
            T g[kMaxW];
            for (int a = 0; a < W; ++a) g[a] = T(0);
            if (threadIdx.x == 0) Hn[0] = Hc[0];
            for (int M = D; M >= 1; --M) {
                // adjP_M = H_M contributes to the new cotangent at level M
                const uint32_t nM = uint32_t(b.level_size(M));
                for (uint32_t idx = threadIdx.x; idx < nM; idx += kBlock)
                    Hn[b.db[M] + idx] += Hc[b.db[M] + idx];
                const T* src = Hc;  // holds adjP_K at level K
                for (int K = M; K >= 1; --K) {
                    const T f = inv[M - K + 1];
                    const uint32_t np = uint32_t(b.level_size(K - 1));
                    for (uint32_t wp = threadIdx.x; wp < np; wp += kBlock) {
                        const T pprev = forward_partial(A, xs, inv, b, pw, M, K - 1, wp);
                        const T* ch = src + b.db[K] + int64_t(wp) * W;
                        T acc = T(0);
                        for (int a = 0; a < W; ++a) {
                            const T v = ch[a];
                            g[a] += v * pprev * f;
                            acc += v * xs[a];
                        }
                        const T adj = acc * f;
                        Q[b.db[K - 1] + wp] = adj;
                        Hn[b.db[K - 1] + wp] += adj;
                    }
                    __syncthreads();
                    src = Q;
                }
            }
            // reduce the increment gradient over the workgroup
            for (int a = 0; a < W; ++a) {
                const T v = wave_sum(g[a]);
                if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][a] = v;
            }
            __syncthreads();
            if (threadIdx.x < W) {
                T v = T(0);
                for (int wv = 0; wv < kBlock / 64; ++wv) v += red[wv][threadIdx.x];
                grad[row * W + threadIdx.x] = v;
            }
            // H <- Hn
            T* tmp = Hc;
            Hc = Hn;
            Hn = tmp;
            __syncthreads();
        }
    }
}

// out += R*_B(g): out_k[u] += sum_{j} sum_{|v| = j} g_{k+j}[u v] B_j[v]
template <typename T>
__global__ __launch_bounds__(kBlock) void adj_rmul_kernel(T* __restrict__ out,
                                                         const T* __restrict__ B,
                                                         const T* __restrict__ g, DevBasis b,
                                                         int64_t batch, int64_t os, int64_t bs,
                                                         int64_t gs)
{
    const int D = b.depth;
    const int64_t n = b.size();
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const T* Be = B + e * bs;
        const T* ge = g + e * gs;
        T* o = out + e * os;
        for (int64_t idx = threadIdx.x; idx < n; idx += kBlock) {
            int k = 0;
            while (idx >= b.db[k + 1]) ++k;
            const int64_t u = idx - b.db[k];
            T acc = o[idx];
            for (int j = D - k; j >= 0; --j) {
                const int64_t nj = b.level_size(j);
                const T* gk = ge + b.db[k + j] + u * nj;
                const T* bj = Be + b.db[j];
                for (int64_t v = 0; v < nj; ++v) acc += gk[v] * bj[v];
            }
            o[idx] = acc;
        }
    }
}

template <int W, int P>
RPY_HD inline void prefix_digits(int q, int (&u)[P])
{
    for (int j = P - 1; j >= 0; --j) {
        u[j] = q % W;
        q /= W;
    }
}

// Loads lane q's share of a dense tensor (path words and subtree).
template <typename T, int W, int D, int P>
RPY_HD inline __attribute__((always_inline)) void load_share(T* path, T* sub, const T* t,
                                                             const DevBasis& b, int q)
{
    using Sh = chen::Shape<W, D, P>;
    chen::static_for<P>([&](auto k0) {
        constexpr int k = int(decltype(k0)::value);
        path[k] = t[b.db[k] + q / int(ipow(W, P - k))];
    });
    chen::static_for<D - P + 1>([&](auto j0) {
        constexpr int k = P + int(decltype(j0)::value);
        constexpr int nk = int(ipow(W, k - P));
        chen::static_for<nk>([&](auto v0) {
            constexpr int v = int(decltype(v0)::value);
            sub[Sh::off(k) + v] = t[b.db[k] + int64_t(q) * nk + v];
        });
    });
}

// One workgroup of NP = W^P threads per stream (one or more whole waves);
// every path-level lane group (W^(P-k) lanes, k >= 1) lies inside one wave,
// so only the W gradient totals cross waves (through LDS, once per step).
template <typename T, int W, int D, int P>
__global__ __launch_bounds__((chen::Shape<W, D, P>::NP)) void sig_vjp_wave_kernel(
        T* __restrict__ grad, const T* __restrict__ incr, const int64_t* __restrict__ offsets,
        const T* __restrict__ sig, const T* __restrict__ gsig, DevBasis b)
{
    using Sh = chen::Shape<W, D, P>;
    constexpr int kWaves = Sh::NP / 64;
    static_assert(Sh::NP % 64 == 0 && kWaves <= 16, "whole waves per stream");
    static_assert(ipow(W, P - 1) <= 64, "path groups inside one wave");
    constexpr int LW = (W == 2) ? 1 : (W == 4) ? 2 : 3;
    static_assert((1 << LW) == W, "width must be 2, 4 or 8");
    __shared__ T red[2][kWaves > 1 ? kWaves : 1][W];
    const int64_t s = blockIdx.x;
    const int q_lane = threadIdx.x;          // prefix index
    const int lane = threadIdx.x % 64;       // lane in the wave
    const int wave = threadIdx.x / 64;
    const int64_t tsize = b.size();
    int u[P];
    prefix_digits<W, P>(q_lane, u);
    chen::BwdLane<T, W, D, P> st;
    st.set_prefix(u);
    load_share<T, W, D, P>(st.A.path, st.A.sub, sig + s * tsize, b, q_lane);
    load_share<T, W, D, P>(st.hpath, st.hsub, gsig + s * tsize, b, q_lane);
    const int64_t r0 = offsets[s], r1 = offsets[s + 1];
    int parity = 0;
    for (int64_t row = r1 - 1; row >= r0; --row) {
        const T* xr = incr + row * W;
        T tab[D * W];
        chen::build_table<T, W, D>(tab, xr);
        T pu[P];
        for (int j = 0; j < P; ++j) pu[j] = xr[u[j]];
        st.A.step(chen::RegRow<T, W, D, P, true>{tab, pu});
        st.contributions(chen::RegRow<T, W, D, P, false>{tab, pu});
        // path cotangents: sum the shares over the W^(P-k) lanes below each
        // path word (aligned lane groups)
        chen::static_for<P - 1>([&](auto i) {
            constexpr int k = int(decltype(i)::value) + 1;
            constexpr int grp = int(ipow(W, P - k));
            T v = st.dpath[k];
            for (int off = 1; off < grp; off <<= 1) v += __shfl_xor(v, off);
            st.hpath[k] += v;
        });
        // gradient: W values summed over the wave by a reduce-scatter; lane
        // ends with the total of letter lane >> (6 - LW)
        T v[W];
        chen::static_for<W>([&](auto a) { v[decltype(a)::value] = st.g[decltype(a)::value]; });
        chen::static_for<P>([&](auto j) {
            chen::static_for<W>([&](auto a) {
                constexpr int ai = int(decltype(a)::value);
                if (u[decltype(j)::value] == ai) v[ai] += st.gp[decltype(j)::value];
            });
        });
        chen::static_for<LW>([&](auto i0) {
            constexpr int i = int(decltype(i0)::value);
            constexpr int half = (W >> i) / 2;
            constexpr int sft = 32 >> i;
            const bool upper = (lane & sft) != 0;
            chen::static_for<half>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const T send = upper ? v[k] : v[k + half];
                const T keep = upper ? v[k + half] : v[k];
                v[k] = keep + __shfl_xor(send, sft);
            });
        });
        T tot = v[0];
        for (int sft = 32 >> LW; sft >= 1; sft >>= 1) tot += __shfl_xor(tot, sft);
        const bool writer = (lane & ((64 >> LW) - 1)) == 0;
        if constexpr (kWaves == 1) {
            if (writer) grad[row * W + (lane >> (6 - LW))] = tot;
        } else {
            // wave totals to LDS (two buffers, so one barrier per step)
            if (writer) red[parity][wave][lane >> (6 - LW)] = tot;
            __syncthreads();
            if (threadIdx.x < W) {
                T g = T(0);
                for (int w = 0; w < kWaves; ++w) g += red[parity][w][threadIdx.x];
                grad[row * W + threadIdx.x] = g;
            }
            parity ^= 1;
        }
    }
}

// Stream-group backward for splits whose lane groups do not align with
// power-of-two lane boundaries (widths 3, 5, 6, 7): a workgroup of W^P lanes
// (ceil(W^P / 64) waves) per stream, the same lane algebra as the wave
// kernel, and the cross-lane sums through LDS with two barriers per step,
// each in a fixed order (deterministic):
//   * path cotangents: the lanes write their shares; one thread per group of
//     W lanes sharing u_0..u_{P-2} sums the group's W shares of every path
//     level; each lane then adds the W^(P-1-k) group sums below its level-k
//     path word;
//   * gradient: the W values are summed over each wave by a reduce-scatter
//     (W padded to 8), then over the waves by W threads.
// The path cotangents are only read by the path targets, so each step runs
// the prefix recovery and the subtree targets before it waits for the
// previous step's group sums: the group summing overlaps that work.
template <int W, int P>
struct VjpGroupCfg {
    static constexpr int NP = int(ipow(W, P));
    static constexpr int NW = (NP + 63) / 64;
    static constexpr int NT = NW * 64;
    static constexpr int NG = int(ipow(W, P - 1));  // groups of W lanes
};

template <typename T, int W, int D, int P>
__global__ __launch_bounds__((VjpGroupCfg<W, P>::NT)) void sig_vjp_group_kernel(
        T* __restrict__ grad, const T* __restrict__ incr, const int64_t* __restrict__ offsets,
        const T* __restrict__ sig, const T* __restrict__ gsig, DevBasis b)
{
    using C = VjpGroupCfg<W, P>;
    constexpr int NP = C::NP, NW = C::NW, NT = C::NT, NG = C::NG;
    static_assert(W <= 8 && P >= 2 && NG <= NT, "group backward shape");
    __shared__ T s_dp[2][P][NT];  // path cotangent shares (levels 1..P-1)
    __shared__ T s_hs[2][P][NG];  // their sums over each group of W lanes
    __shared__ T s_g[2][NW][8];   // wave totals of the gradient
    const int64_t s = blockIdx.x;
    const int tid = threadIdx.x;
    const bool active = tid < NP;
    const int q = active ? tid : 0;
    const int lane = tid % 64;
    const int wave = tid / 64;
    const int64_t tsize = b.size();
    int u[P];
    prefix_digits<W, P>(q, u);
    chen::BwdLane<T, W, D, P> st;
    st.set_prefix(u);
    load_share<T, W, D, P>(st.A.path, st.A.sub, sig + s * tsize, b, q);
    load_share<T, W, D, P>(st.hpath, st.hsub, gsig + s * tsize, b, q);
    const int64_t r0 = offsets[s], r1 = offsets[s + 1];
    int par = 0;
    for (int64_t row = r1 - 1; row >= r0; --row) {
        const T* xr = incr + row * W;
        T tab[D * W];
        chen::build_table<T, W, D>(tab, xr);
        T pu[P];
        for (int j = 0; j < P; ++j) pu[j] = xr[u[j]];
        st.A.step(chen::RegRow<T, W, D, P, true>{tab, pu});
        st.contributions_sub(chen::RegRow<T, W, D, P, false>{tab, pu});
        if (row != r1 - 1) {
            __syncthreads();  // the previous step's group sums are written
            // level-k path word of this lane: groups g0 .. g0 + W^(P-1-k) - 1
            chen::static_for<P - 1>([&](auto i) {
                constexpr int k = int(decltype(i)::value) + 1;
                constexpr int span = int(ipow(W, P - 1 - k));
                const int g0 = (q / W / span) * span;
                T acc = T(0);
                for (int c = 0; c < span; ++c) acc += s_hs[par ^ 1][k][g0 + c];
                st.hpath[k] += acc;
            });
        }
        st.contributions_path(chen::RegRow<T, W, D, P, false>{tab, pu});
        // shares of the path cotangents
        chen::static_for<P - 1>([&](auto i) {
            constexpr int k = int(decltype(i)::value) + 1;
            s_dp[par][k][tid] = active ? st.dpath[k] : T(0);
        });
        // gradient: wave reduce-scatter of the W values (padded to 8); lane
        // ends with the wave total of letter lane >> 3 when lane % 8 == 0
        T v[8];
        chen::static_for<8>([&](auto a) {
            constexpr int ai = int(decltype(a)::value);
            if constexpr (ai < W) {
                T x = st.g[ai];
                chen::static_for<P>([&](auto j) {
                    if (u[decltype(j)::value] == ai) x += st.gp[decltype(j)::value];
                });
                v[ai] = active ? x : T(0);
            } else {
                v[ai] = T(0);
            }
        });
        chen::static_for<3>([&](auto i0) {
            constexpr int i = int(decltype(i0)::value);
            constexpr int half = 4 >> i;
            constexpr int sft = 32 >> i;
            const bool upper = (lane & sft) != 0;
            chen::static_for<half>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const T send = upper ? v[k] : v[k + half];
                const T keep = upper ? v[k + half] : v[k];
                v[k] = keep + __shfl_xor(send, sft);
            });
        });
        T tot = v[0];
        for (int sft = 4; sft >= 1; sft >>= 1) tot += __shfl_xor(tot, sft);
        if ((lane & 7) == 0) s_g[par][wave][lane >> 3] = tot;
        __syncthreads();
        if (tid < NG) {
            chen::static_for<P - 1>([&](auto i) {
                constexpr int k = int(decltype(i)::value) + 1;
                T acc = T(0);
                for (int c = 0; c < W; ++c) acc += s_dp[par][k][tid * W + c];
                s_hs[par][k][tid] = acc;
            });
        }
        if (tid < W) {
            T g = T(0);
            for (int w = 0; w < NW; ++w) g += s_g[par][w][tid];
            grad[row * W + tid] = g;
        }
        par ^= 1;
    }
}

// Host run of the wave kernel's lane algebra for one float64 stream (the
// cross-lane sums in plain loops).
template <int W, int D, int P>
void vjp_emulate(double* grad, const double* incr, int64_t len, const double* sig,
                 const double* gsig, const DevBasis& b)
{
    using Sh = chen::Shape<W, D, P>;
    std::vector<chen::BwdLane<double, W, D, P>> lanes(Sh::NP);
    std::vector<std::array<int, P>> digits(Sh::NP);
    for (int q = 0; q < Sh::NP; ++q) {
        int u[P];
        prefix_digits<W, P>(q, u);
        for (int j = 0; j < P; ++j) digits[q][j] = u[j];
        lanes[q].set_prefix(u);
        load_share<double, W, D, P>(lanes[q].A.path, lanes[q].A.sub, sig, b, q);
        load_share<double, W, D, P>(lanes[q].hpath, lanes[q].hsub, gsig, b, q);
    }
    for (int64_t row = len - 1; row >= 0; --row) {
        const double* xr = incr + row * W;
        double tab[D * W];
        chen::build_table<double, W, D>(tab, xr);
        for (int q = 0; q < Sh::NP; ++q) {
            int u[P];
            double pu[P];
            for (int j = 0; j < P; ++j) {
                u[j] = digits[q][j];
                pu[j] = xr[u[j]];
            }
            lanes[q].A.step(chen::RegRow<double, W, D, P, true>{tab, pu});
            lanes[q].contributions(chen::RegRow<double, W, D, P, false>{tab, pu});
        }
        for (int k = 1; k < P; ++k) {
            const int grp = int(ipow(W, P - k));
            for (int g0 = 0; g0 < Sh::NP; g0 += grp) {
                double sum = 0.0;
                for (int q = g0; q < g0 + grp; ++q) sum += lanes[q].dpath[k];
                for (int q = g0; q < g0 + grp; ++q) lanes[q].hpath[k] += sum;
            }
        }
        for (int a = 0; a < W; ++a) {
            double sum = 0.0;
            for (int q = 0; q < Sh::NP; ++q) {
                int u[P];
                for (int j = 0; j < P; ++j) u[j] = digits[q][j];
                sum += lanes[q].grad_share(a, u);
            }
            grad[row * W + a] = sum;
        }
    }
}

template <typename T, int W, int D, int P>
void launch_wave(void* grad, const void* incr, const int64_t* offsets, int64_t n,
                 const void* sig, const void* gsig, const DevBasis& b, hipStream_t st)
{
    hipLaunchKernelGGL((sig_vjp_wave_kernel<T, W, D, P>), dim3(uint32_t(n)),
                       dim3(chen::Shape<W, D, P>::NP), 0, st,
                       (T*)grad, (const T*)incr, offsets, (const T*)sig, (const T*)gsig, b);
}

template <typename T, int W, int D, int P>
void launch_group(void* grad, const void* incr, const int64_t* offsets, int64_t n,
                  const void* sig, const void* gsig, const DevBasis& b, hipStream_t st)
{
    hipLaunchKernelGGL((sig_vjp_group_kernel<T, W, D, P>), dim3(uint32_t(n)),
                       dim3(VjpGroupCfg<W, P>::NT), 0, st, (T*)grad, (const T*)incr, offsets,
                       (const T*)sig, (const T*)gsig, b);
}

// Register-resident backward variants (kind RPY_KIND_WAVE_VJP): a
// workgroup of W^P lanes per stream.  The defaults are 64-lane splits (one
// wave per stream): for float64 at (8,4), (4,6) and (2,10) the lane state
// takes all 256 VGPRs (at (4,6) the compiler also keeps part of it in
// AGPRs), which still beats splitting the stream over several waves that
// must meet at a barrier every step.  (8,4) with P = 3 and (4,6) with P = 4
// (512 and 256 lanes, smaller states) are compiled for the autotuner.
using VjpLauncher = void (*)(void* grad, const void* incr, const int64_t* offsets, int64_t n,
                             const void* sig, const void* gsig, const DevBasis& b,
                             hipStream_t st);

struct VjpVariant {
    int32_t width, depth;
    rpy_dtype dtype;
    int32_t kind;  // RPY_KIND_WAVE_VJP or RPY_KIND_STREAM_GROUP
    int32_t prefix;
    bool is_default;
    VjpLauncher launch;
};

template <int W, int D, int P, bool Default>
void add_wave(std::vector<VjpVariant>& v)
{
    v.push_back({W, D, RPY_F32, RPY_KIND_WAVE_VJP, P, Default, &launch_wave<float, W, D, P>});
    v.push_back({W, D, RPY_F64, RPY_KIND_WAVE_VJP, P, Default, &launch_wave<double, W, D, P>});
}

// Stream-group backward (sig_vjp_group_kernel), kind RPY_KIND_STREAM_GROUP.
template <int W, int D, int P, bool Default>
void add_group(std::vector<VjpVariant>& v)
{
    v.push_back({W, D, RPY_F32, RPY_KIND_STREAM_GROUP, P, Default,
                 &launch_group<float, W, D, P>});
    v.push_back({W, D, RPY_F64, RPY_KIND_STREAM_GROUP, P, Default,
                 &launch_group<double, W, D, P>});
}

const std::vector<VjpVariant>& vjp_registry()
{
    static const std::vector<VjpVariant> r = [] {
        std::vector<VjpVariant> v;
        add_wave<8, 2, 2, true>(v);
        add_wave<8, 3, 2, true>(v);
        add_wave<8, 4, 2, true>(v);
        add_wave<4, 3, 3, true>(v);
        add_wave<4, 4, 3, true>(v);
        add_wave<4, 5, 3, true>(v);
        add_wave<4, 6, 3, true>(v);
        add_wave<2, 6, 6, true>(v);
        add_wave<2, 7, 6, true>(v);
        add_wave<2, 8, 6, true>(v);
        add_wave<2, 9, 6, true>(v);
        add_wave<2, 10, 6, true>(v);
        add_wave<8, 4, 3, false>(v);
        add_wave<4, 6, 4, false>(v);
        add_group<3, 4, 3, true>(v);
        add_group<5, 5, 3, true>(v);
        add_group<6, 5, 3, true>(v);
        return v;
    }();
    return r;
}

rpy_kernel_variant vjp_kind(const VjpVariant& x)
{
    return rpy_kernel_variant{x.kind, x.prefix, 0, 0};
}

const VjpVariant* find_vjp(int32_t w, int32_t d, rpy_dtype dt, const rpy_kernel_variant& k)
{
    if (k.block != 0 || k.waves_per_eu != 0) return nullptr;
    for (const VjpVariant& x : vjp_registry())
        if (x.width == w && x.depth == d && x.dtype == dt && x.kind == k.kind &&
            x.prefix == k.prefix)
            return &x;
    return nullptr;
}

bool vjp_exists(int32_t w, int32_t d, rpy_dtype dt, const rpy_kernel_variant& k)
{
    return same_variant(k, rpy_kernel_variant{RPY_KIND_GENERIC, 0, 0, 0}) ||
           find_vjp(w, d, dt, k) != nullptr;
}

// The variant a call runs: nullptr for the generic (LDS / workspace) kernel.
const VjpVariant* choose_vjp(int32_t w, int32_t d, rpy_dtype dt, int32_t* source)
{
    rpy_kernel_variant def{RPY_KIND_GENERIC, 0, 0, 0};
    for (const VjpVariant& x : vjp_registry())
        if (x.width == w && x.depth == d && x.dtype == dt && x.is_default) def = vjp_kind(x);
    const rpy_kernel_variant k = resolve_variant(
            RPY_OP_SIGNATURE_VJP, w, d, dt, def,
            [&](const rpy_kernel_variant& v) { return vjp_exists(w, d, dt, v); }, source);
    return find_vjp(w, d, dt, k);
}

size_t lds_limit(hipStream_t)
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 64 * 1024;
    return size_t(v);
}

}  // namespace

int vjp_host_emulation(double* grad, const double* incr, int64_t len, int32_t width,
                       int32_t depth, int32_t prefix, const double* sig, const double* gsig)
{
    rpy_tensor_basis tb{width, depth, nullptr};
    DevBasis b;
    int rc = make_basis(&tb, b);
    if (rc) return rc;
#define RPY_EMU(W, D, P)                                           \
    if (width == W && depth == D && prefix == P) {                 \
        vjp_emulate<W, D, P>(grad, incr, len, sig, gsig, b);       \
        return RPY_OK;                                             \
    }
    RPY_EMU(8, 2, 2) RPY_EMU(8, 3, 2) RPY_EMU(8, 4, 2) RPY_EMU(8, 4, 3)
    RPY_EMU(4, 3, 3) RPY_EMU(4, 4, 3) RPY_EMU(4, 5, 3) RPY_EMU(4, 6, 3) RPY_EMU(4, 6, 4)
    RPY_EMU(2, 6, 6) RPY_EMU(2, 7, 6) RPY_EMU(2, 7, 7)
#undef RPY_EMU
    return fail(RPY_ERR_UNSUPPORTED, "no host emulation of the wave backward for this shape");
}

int stream_signature_vjp(void* grad, const void* increments, const int64_t* offsets,
                         int64_t n_streams, int32_t width, int32_t depth, const void* signatures,
                         const void* grad_signatures, rpy_dtype dtype, void* stream)
{
    rpy_tensor_basis tb{width, depth, nullptr};
    DevBasis b;
    int rc = make_basis(&tb, b);
    if (rc) return rc;
    if (depth < 1) return fail(RPY_ERR_INVALID_ARGUMENT, "signatures need depth >= 1");
    if (width > kMaxW) return fail(RPY_ERR_UNSUPPORTED, "backward supports width <= %d", kMaxW);
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (n_streams < 0) return fail(RPY_ERR_SHAPE, "negative number of streams");
    if (n_streams == 0) return RPY_OK;
    if (!grad || !increments || !offsets || !signatures || !grad_signatures)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    if (n_streams > int64_t(0x7fffffff))
        return fail(RPY_ERR_SHAPE, "too many streams for one launch");
    const Pow pw = make_pow(width, depth);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (const VjpVariant* v = choose_vjp(width, depth, dtype, nullptr)) {
        v->launch(grad, increments, offsets, n_streams, signatures, grad_signatures, b, st);
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    const int64_t grid = std::min<int64_t>(n_streams, 2048);
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    const size_t lds = 4 * size_t(b.size()) * es;
    void* scratch = nullptr;
    // the kernel's static LDS (increment, reciprocals, reduction) is < 4 KiB
    const bool in_lds = lds + 4096 <= lds_limit(st);
    if (!in_lds)
        RPY_HIP_CHECK(workspace_alloc(&scratch, size_t(grid) * 4 * size_t(b.size()) * es, st));
    const dim3 g = dim3(uint32_t(grid)), blk = dim3(kBlock);
    const size_t sh = in_lds ? lds : 0;
    if (in_lds) {
        hipError_t e = dtype == RPY_F64
                               ? hipFuncSetAttribute(
                                         (const void*)sig_vjp_kernel<double, true>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(sh))
                               : hipFuncSetAttribute(
                                         (const void*)sig_vjp_kernel<float, true>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(sh));
        if (e != hipSuccess)
            return fail(RPY_ERR_HIP, "cannot reserve %zu bytes of LDS: %s", sh,
                        hipGetErrorString(e));
    }
    if (dtype == RPY_F64) {
        if (in_lds)
            hipLaunchKernelGGL((sig_vjp_kernel<double, true>), g, blk, sh, st, (double*)grad,
                               (const double*)increments, offsets, n_streams,
                               (const double*)signatures, (const double*)grad_signatures,
                               (double*)nullptr, b, pw);
        else
            hipLaunchKernelGGL((sig_vjp_kernel<double, false>), g, blk, 0, st, (double*)grad,
                               (const double*)increments, offsets, n_streams,
                               (const double*)signatures, (const double*)grad_signatures,
                               (double*)scratch, b, pw);
    } else {
        if (in_lds)
            hipLaunchKernelGGL((sig_vjp_kernel<float, true>), g, blk, sh, st, (float*)grad,
                               (const float*)increments, offsets, n_streams,
                               (const float*)signatures, (const float*)grad_signatures,
                               (float*)nullptr, b, pw);
        else
            hipLaunchKernelGGL((sig_vjp_kernel<float, false>), g, blk, 0, st, (float*)grad,
                               (const float*)increments, offsets, n_streams,
                               (const float*)signatures, (const float*)grad_signatures,
                               (float*)scratch, b, pw);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "vjp launch failed: %s", hipGetErrorString(err));
    if (scratch) {
        hipError_t ferr = hipFreeAsync(scratch, st);
        if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    }
    return rc;
}

int dense_ft_adjoint_right_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                               const rpy_tensor_basis* basis, int64_t batch, int64_t os,
                               int64_t ps, int64_t as, void* stream)
{
    DevBasis b;
    int rc = make_basis(basis, b);
    if (rc) return rc;
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch");
    if (batch > 1 && (os < b.size() || ps < b.size() || as < b.size()))
        return fail(RPY_ERR_SHAPE, "batch stride smaller than the tensor size");
    if (batch == 0) return RPY_OK;
    if (!out || !op || !arg) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const uint32_t grid = uint32_t(std::min<int64_t>(batch, 8192));
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(adj_rmul_kernel<double>, dim3(grid), dim3(kBlock), 0, st, (double*)out,
                           (const double*)op, (const double*)arg, b, batch, os, ps, as);
    else
        hipLaunchKernelGGL(adj_rmul_kernel<float>, dim3(grid), dim3(kBlock), 0, st, (float*)out,
                           (const float*)op, (const float*)arg, b, batch, os, ps, as);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

int32_t vjp_variants(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
                     int32_t max)
{
    if (dtype != RPY_F32 && dtype != RPY_F64) return 0;
    int32_t n = 0;
    for (const VjpVariant& x : vjp_registry()) {
        if (x.width != width || x.depth != depth || x.dtype != dtype) continue;
        if (out && n < max) out[n] = vjp_kind(x);
        ++n;
    }
    if (out && n < max) out[n] = rpy_kernel_variant{RPY_KIND_GENERIC, 0, 0, 0};
    return n + 1;
}

int vjp_select(int32_t width, int32_t depth, rpy_dtype dtype, const rpy_kernel_variant* v)
{
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (v && !vjp_exists(width, depth, dtype, *v))
        return fail(RPY_ERR_UNSUPPORTED,
                    "no signature backward variant (kind %d, prefix %d, block %d, waves/EU %d) "
                    "for width %d depth %d",
                    v->kind, v->prefix, v->block, v->waves_per_eu, width, depth);
    set_selected_variant(RPY_OP_SIGNATURE_VJP, width, depth, dtype, v);
    return RPY_OK;
}

int vjp_config(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
               int32_t* source)
{
    if (!out) return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    const VjpVariant* v = choose_vjp(width, depth, dtype, source);
    *out = v ? vjp_kind(*v) : rpy_kernel_variant{RPY_KIND_GENERIC, 0, 0, 0};
    return RPY_OK;
}

}  // namespace rpy_amd
