// dyadic.hip -- dyadic signature cache over [0, 1) and interval queries.
//
// Semantics (the reference's, checked against the oracle): a query [inf, sup)
// at resolution r is dissected into maximal aligned dyadic pieces
// (intervals/src/dyadic_interval.cpp:191-228) and its signature is the Chen
// product, left to right, of the rows whose parameter lies in each piece --
// what Stream::signature returns for a LieIncrementStream of resolution r
// (streams/src/stream.cpp:107-171, streams/src/dyadic_caching_layer.cpp:
// 86-130).  The reference memoises only resolution-r leaves and recomputes
// every coarser node per query; here every node of power 0..J (a node of
// power j covers 2^(r-j) leaves) is built once, so a query costs at most
// 2(J+1) node products plus at most 2(2^(r-J) - 1) Chen steps from raw rows.
//
// Cache layout (level-major, so level J is one contiguous batch the
// signature kernel writes directly, and level j-1 is one strided batch of
// products of level j):
//   node(j, s, k) at element offset ((2^j - 1) * n_streams + s * 2^j + k) * tsize.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "chen.hpp"
#include "chen_bwd.hpp"
#include "dense_ops.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRes = 30;
constexpr int kMaxPieces = 2 * kMaxRes + 2;

RPY_HD inline int64_t node_offset(int j, int64_t n_streams, int64_t s, int64_t k)
{
    return ((int64_t(1) << j) - 1) * n_streams + s * (int64_t(1) << j) + k;
}

// first index in [lo, hi) with params[i] >= v
RPY_FI int64_t lower_bound_dev(const double* params, int64_t lo, int64_t hi, double v)
{
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (params[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// begins/ends of the level-J segments: rows of stream s with params in
// [k 2^-J, (k+1) 2^-J).
__global__ void segment_bounds_kernel(const double* __restrict__ params,
                                      const int64_t* __restrict__ offsets, int64_t n_streams,
                                      int J, int64_t* __restrict__ begins,
                                      int64_t* __restrict__ ends)
{
    const int64_t nseg = n_streams << J;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nseg;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = t >> J;
        const int64_t k = t - (s << J);
        const int64_t lo = offsets[s], hi = offsets[s + 1];
        begins[t] = lower_bound_dev(params, lo, hi, std::ldexp(double(k), -J));
        ends[t] = lower_bound_dev(params, lo, hi, std::ldexp(double(k + 1), -J));
    }
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = T(0);
}

// Maximal aligned dyadic pieces of [A, B) (units of 2^-r, 0 <= A <= B <=
// 2^r), left to right: at each step the largest aligned block that starts at
// A and ends by B.  Returns the count; piece i is (ks[i], ns[i]).
RPY_HD inline int dyadic_cover(int64_t A, int64_t B, int r, int64_t* ks, int* ns)
{
    int n = 0;
    while (A < B) {
        int lg = r;
        if (A != 0) {
            lg = 0;
            while (((A >> lg) & 1) == 0) ++lg;
        }
        while (A + (int64_t(1) << lg) > B) --lg;
        ns[n] = r - lg;
        ks[n] = A >> lg;
        ++n;
        A += int64_t(1) << lg;
    }
    return n;
}

// S <- S (x) N, both in LDS, top-down with a barrier per level.
template <typename T>
RPY_FI void wg_mul_inplace(T* S, const T* N, const DevBasis& b, const Pow& pw)
{
    for (int m = b.depth; m >= 1; --m) {
        const uint32_t n = uint32_t(b.level_size(m));
        for (uint32_t idx = threadIdx.x; idx < n; idx += kBlock) {
            T acc = T(0);
            for (int i = 0; i <= m; ++i) {  // lhs degree m-i, high to low
                const uint32_t pre = fdiv(idx, pw.pw[i]);
                const uint32_t suf = idx - pre * pw.pw[i].d;
                acc += S[b.db[m - i] + pre] * N[b.db[i] + suf];
            }
            S[b.db[m] + idx] = acc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) S[0] = S[0] * N[0];
    __syncthreads();
}

// S <- S (x) exp(x) for a degree-one increment x, the Horner of the
// signature kernel evaluated per coefficient (top-down, barrier per level).
template <typename T>
RPY_FI void wg_chen_step(T* S, const T* x, const T* inv, const DevBasis& b, const Pow& pw)
{
    const uint32_t W = uint32_t(b.width);
    for (int m = b.depth; m >= 1; --m) {
        const uint32_t n = uint32_t(b.level_size(m));
        for (uint32_t idx = threadIdx.x; idx < n; idx += kBlock) {
            T p = S[0];
            for (int K = 1; K <= m; ++K) {
                const uint32_t pre = fdiv(idx, pw.pw[m - K]);       // prefix of length K
                const uint32_t letter = pre - fdiv(pre, pw.pw[1]) * W;
                p = p * (x[letter] * inv[m - K + 1]) + S[b.db[K] + pre];
            }
            S[b.db[m] + idx] = p;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dyadic_query_kernel(
        T* __restrict__ out, const T* __restrict__ cache, const T* __restrict__ incr,
        const double* __restrict__ params, const int64_t* __restrict__ offsets, int64_t n_streams,
        DevBasis b, Pow pw, int r, int J, const int64_t* __restrict__ q_stream,
        const double* __restrict__ q_inf, const double* __restrict__ q_sup, int64_t n_queries)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t tsize = b.size();
    T* S = reinterpret_cast<T*>(smem);
    T* N = S + tsize;
    __shared__ int64_t pk[kMaxPieces];
    __shared__ int pn[kMaxPieces];
    __shared__ int npieces;
    __shared__ int64_t row_lo, row_hi;
    __shared__ T xrow[32];
    __shared__ T inv[kMaxDepth + 2];
    const int W = b.width;
    if (threadIdx.x <= kMaxDepth + 1) inv[threadIdx.x] = threadIdx.x ? T(1) / T(threadIdx.x) : T(0);

    for (int64_t q = blockIdx.x; q < n_queries; q += gridDim.x) {
        const int64_t s = q_stream[q];
        const bool valid = s >= 0 && s < n_streams;
        if (threadIdx.x == 0) {
            int cnt = 0;
            if (valid) {
                const double scale = std::ldexp(1.0, r);
                const double full = scale;
                double a = std::floor(q_inf[q] * scale), c = std::floor(q_sup[q] * scale);
                a = std::min(std::max(a, 0.0), full);
                c = std::min(std::max(c, 0.0), full);
                if (a < c) cnt = dyadic_cover(int64_t(a), int64_t(c), r, pk, pn);
            }
            npieces = cnt;
        }
        for (int64_t i = threadIdx.x; i < tsize; i += kBlock) S[i] = (i == 0) ? T(1) : T(0);
        __syncthreads();
        const int np = npieces;
        for (int p = 0; p < np; ++p) {
            const int j = pn[p];
            const int64_t k = pk[p];
            if (j <= J) {
                const T* node = cache + node_offset(j, n_streams, s, k) * tsize;
                for (int64_t i = threadIdx.x; i < tsize; i += kBlock) N[i] = node[i];
                __syncthreads();
                wg_mul_inplace(S, N, b, pw);
            } else {
                if (threadIdx.x == 0) {
                    const int64_t lo = offsets[s], hi = offsets[s + 1];
                    row_lo = lower_bound_dev(params, lo, hi, std::ldexp(double(k), -j));
                    row_hi = lower_bound_dev(params, lo, hi, std::ldexp(double(k + 1), -j));
                }
                __syncthreads();
                const int64_t rlo = row_lo, rhi = row_hi;
                __syncthreads();  // every wave has read the bounds before they change
                for (int64_t row = rlo; row < rhi; ++row) {
                    if (threadIdx.x < W) xrow[threadIdx.x] = incr[row * W + threadIdx.x];
                    __syncthreads();
                    wg_chen_step(S, xrow, inv, b, pw);
                }
            }
        }
        T* o = out + q * tsize;
        for (int64_t i = threadIdx.x; i < tsize; i += kBlock) o[i] = S[i];
        __syncthreads();
    }
}

constexpr int kQueryChunk = 48;  // fine-run rows staged per pass

// offset of level k in a dense tensor of width W
template <int W>
constexpr int dbeg(int k)
{
    int s = 0;
    for (int j = 0; j < k; ++j) s += int(ipow(W, j));
    return s;
}

// Unrolled S <- S (x) N for a compile-time shape, both tensors in LDS: for
// each output level (top-down, barrier per level) every thread takes outputs
// idx = tid + 256 q and sums lhs degree m-i times rhs degree i, i = 0..m
// (the reference order, lhs degree high to low), with constant divisors.
template <typename T, int W, int D>
RPY_FI void wg_mul_inplace_t(T* S, const T* N)
{
    // the thread of a prefix u computes the W outputs u c of level m: for
    // rhs degree i >= 1 the lhs word is a prefix of u (one read) and the rhs
    // words are the rest of u followed by c (W contiguous reads)
    chen::static_for<D>([&](auto i0) {
        constexpr int m = D - int(decltype(i0)::value);
        constexpr int tasks = int(ipow(W, m - 1));
        constexpr int dbm = dbeg<W>(m);
        for (int u = threadIdx.x; u < tasks; u += kBlock) {
            T acc[W];
            const T n0 = N[0];
            chen::static_for<W>([&](auto c) { acc[c] = S[dbm + u * W + int(decltype(c)::value)] * n0; });
            chen::static_for<m>([&](auto k0) {
                constexpr int i = int(decltype(k0)::value) + 1;  // rhs degree
                constexpr int wi = int(ipow(W, i - 1));
                constexpr int dbl = dbeg<W>(m - i);
                constexpr int dbr = dbeg<W>(i);
                const int pre = u / wi;
                const int suf = u - pre * wi;
                const T lv = S[dbl + pre];
                const T* rb = N + dbr + suf * W;
                chen::static_for<W>([&](auto c) { acc[c] += lv * rb[int(decltype(c)::value)]; });
            });
            chen::static_for<W>([&](auto c) { S[dbm + u * W + int(decltype(c)::value)] = acc[c]; });
        }
        __syncthreads();
    });
    if (threadIdx.x == 0) S[0] = S[0] * N[0];
    __syncthreads();
}

// Query plans, one thread per query, computed ahead of the query kernel so
// that a query's workgroup starts from one load.  The cover of [A, B) is
// [fine run][stored nodes][fine run] (piece sizes grow, then shrink), so one
// greedy pass over it yields the run bounds and the stored nodes' cache
// offsets.  Record (int64 words): stream (-1 if invalid), the four run row
// bounds, the node count, then the node offsets (at most 2 (J + 1)).
constexpr int kPlanHead = 6;

inline int plan_record_words(int J) { return kPlanHead + 2 * (J + 1); }

__global__ __launch_bounds__(256) void dyadic_query_plan_kernel(
        int64_t* __restrict__ plans, int rec, const double* __restrict__ params,
        const int64_t* __restrict__ offsets, int64_t n_streams, int r, int J,
        const int64_t* __restrict__ q_stream, const double* __restrict__ q_inf,
        const double* __restrict__ q_sup, int64_t n_queries)
{
    const int64_t q = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (q >= n_queries) return;
    int64_t* P = plans + q * rec;
    const int64_t s = q_stream[q];
    const bool valid = s >= 0 && s < n_streams;
    int64_t runs[4] = {0, 0, 0, 0};
    int64_t cnt = 0;
    if (valid) {
        const int64_t lo = offsets[s], hi = offsets[s + 1];
        runs[0] = runs[1] = runs[2] = runs[3] = lo;
        const double scale = std::ldexp(1.0, r);
        double a = std::floor(q_inf[q] * scale), e = std::floor(q_sup[q] * scale);
        a = std::min(std::max(a, 0.0), scale);
        e = std::min(std::max(e, 0.0), scale);
        if (a < e) {
            int64_t A = int64_t(a);
            const int64_t B = int64_t(e);
            bool stored = false, left = false, right = false;
            double l0 = 0.0, l1 = 0.0, r0 = 0.0, r1 = 0.0;
            while (A < B) {
                int lg = r;
                if (A != 0) lg = __builtin_ctzll(uint64_t(A));
                while (A + (int64_t(1) << lg) > B) --lg;
                const int n = r - lg;
                const int64_t k = A >> lg;
                const double ps = std::ldexp(double(k), -n), pe = std::ldexp(double(k + 1), -n);
                if (n > J) {
                    if (!stored) {
                        if (!left) l0 = ps;
                        left = true;
                        l1 = pe;
                    } else {
                        if (!right) r0 = ps;
                        right = true;
                        r1 = pe;
                    }
                } else if (!right && cnt < 2 * (J + 1)) {
                    stored = true;
                    P[kPlanHead + cnt] = node_offset(n, n_streams, s, k);
                    ++cnt;
                }
                A += int64_t(1) << lg;
            }
            if (left) {
                runs[0] = lower_bound_dev(params, lo, hi, l0);
                runs[1] = lower_bound_dev(params, lo, hi, l1);
            }
            if (right) {
                runs[2] = lower_bound_dev(params, lo, hi, r0);
                runs[3] = lower_bound_dev(params, lo, hi, r1);
            }
        }
    }
    P[0] = valid ? s : -1;
    P[1] = runs[0];
    P[2] = runs[1];
    P[3] = runs[2];
    P[4] = runs[3];
    P[5] = cnt;
}

// Queries for shapes whose prefix split has at most 128 lanes.  Per query:
// 1. the query's plan record (below) gives [left fine run][stored nodes]
//    [right run];
// 2. the left run's rows are staged chunk by chunk in LDS as scaled tables
//    (x_a / d) by the whole workgroup, and lanes 0..NP-1 step through them,
//    each holding its prefix share of the run's signature in registers
//    (chen_core.hpp), then write it to S;
// 3. S is multiplied by each stored node (the next node's coefficients are
//    loaded into registers while the current product runs);
// 4. the lanes load S back into registers, step through the right run the
//    same way (S (x) exp(x_1) (x) ... equals S times the run's signature) and
//    write the result to the output.
// Two tensors of LDS per workgroup, so four workgroups share a CU at width 5
// depth 5 in float32.
template <typename T, int W, int D, int P>
__global__ __launch_bounds__(kBlock) void dyadic_query_fast_kernel(
        T* __restrict__ out, const T* __restrict__ cache, const T* __restrict__ incr,
        const int64_t* __restrict__ plans, int rec, int64_t n_queries)
{
    using Sh = chen::Shape<W, D, P>;
    static_assert(Sh::NP <= kBlock, "a run's lanes fit the workgroup");
    constexpr int kT = dbeg<W>(D + 1);                  // tensor size
    constexpr int kPer = (kT + kBlock - 1) / kBlock;    // node values per thread
    constexpr int kDW = D * W;
    extern __shared__ __align__(16) unsigned char smem[];
    T* S = reinterpret_cast<T*>(smem);
    T* N = S + kT;
    __shared__ int64_t prec[kPlanHead + 2 * 25];
    __shared__ T XT[kQueryChunk * kDW];
    const int lane = threadIdx.x;
    const bool active = lane < Sh::NP;
    int u[P];
    {
        int t = lane;
        for (int j = P - 1; j >= 0; --j) {
            u[j] = t % W;
            t /= W;
        }
    }

    // one run of rows [r0, r1) stepped into the lanes' registers
    auto run = [&](chen::Lane<T, W, D, P>& st, int64_t r0, int64_t r1) {
        for (int64_t cs = r0; cs < r1; cs += kQueryChunk) {
            const int n = int(min<int64_t>(kQueryChunk, r1 - cs));
            if constexpr (sizeof(T) == 8) {
                // float64: all of a thread's loads of the chunk go out before
                // its stores (7.28 -> 5.19 ms for 100k queries; the same in
                // float32 measured slower, 2.52 -> 2.85 ms)
                constexpr int kIt = (kQueryChunk * kDW + kBlock - 1) / kBlock;
                T xv[kIt];
                chen::static_for<kIt>([&](auto k0) {
                    constexpr int k = int(decltype(k0)::value);
                    const int e = int(threadIdx.x) + k * kBlock;
                    if (e < n * kDW) {
                        const int row = e / kDW;
                        const int da = e - row * kDW;
                        const int a = da - (da / W) * W;
                        xv[k] = incr[(cs + row) * W + a];
                    }
                });
                chen::static_for<kIt>([&](auto k0) {
                    constexpr int k = int(decltype(k0)::value);
                    const int e = int(threadIdx.x) + k * kBlock;
                    if (e < n * kDW) {
                        const int da = e - (e / kDW) * kDW;
                        const int d = da / W + 1;
                        XT[e] = (d == 1) ? xv[k] : xv[k] * (T(1) / T(d));
                    }
                });
            } else {
                for (int e = threadIdx.x; e < n * kDW; e += kBlock) {
                    const int row = e / kDW;
                    const int da = e - row * kDW;
                    const int d = da / W + 1;
                    const int a = da - (d - 1) * W;
                    const T x = incr[(cs + row) * W + a];
                    XT[e] = (d == 1) ? x : x * (T(1) / T(d));
                }
            }
            __syncthreads();
            if (active)
                for (int i = 0; i < n; ++i)
                    st.step(chen::ScaledRow<T, W, D, P>{XT + i * kDW, u});
            __syncthreads();
        }
    };
    // lane state <-> dense tensor: the lane's subtree, and the path values
    // from the lanes that start each prefix group
    auto store = [&](const chen::Lane<T, W, D, P>& st, T* dst) {
        chen::static_for<D - P + 1>([&](auto j0) {
            constexpr int k = P + int(decltype(j0)::value);
            constexpr int nk = int(ipow(W, k - P));
            chen::static_for<nk>([&](auto v0) {
                constexpr int v = int(decltype(v0)::value);
                dst[dbeg<W>(k) + lane * nk + v] = st.sub[Sh::off(k) + v];
            });
        });
        chen::static_for<P>([&](auto k0) {
            constexpr int k = int(decltype(k0)::value);
            constexpr int grp = int(ipow(W, P - k));
            if (lane % grp == 0) dst[dbeg<W>(k) + lane / grp] = st.path[k];
        });
    };
    auto fetch = [&](chen::Lane<T, W, D, P>& st, const T* src) {
        chen::static_for<D - P + 1>([&](auto j0) {
            constexpr int k = P + int(decltype(j0)::value);
            constexpr int nk = int(ipow(W, k - P));
            chen::static_for<nk>([&](auto v0) {
                constexpr int v = int(decltype(v0)::value);
                st.sub[Sh::off(k) + v] = src[dbeg<W>(k) + lane * nk + v];
            });
        });
        chen::static_for<P>([&](auto k0) {
            constexpr int k = int(decltype(k0)::value);
            constexpr int grp = int(ipow(W, P - k));
            st.path[k] = src[dbeg<W>(k) + lane / grp];
        });
    };

    // each thread below rec holds one word of the next query's plan
    int64_t nxt = 0;
    if (int(threadIdx.x) < rec && int64_t(blockIdx.x) < n_queries)
        nxt = plans[int64_t(blockIdx.x) * rec + threadIdx.x];
    for (int64_t q = blockIdx.x; q < n_queries; q += gridDim.x) {
        if (int(threadIdx.x) < rec) {
            prec[threadIdx.x] = nxt;
            const int64_t qn = q + gridDim.x;
            if (qn < n_queries) nxt = plans[qn * rec + threadIdx.x];
        }
        __syncthreads();
        const int64_t l0 = prec[1], l1 = prec[2], rr0 = prec[3], rr1 = prec[4];
        const int nn = int(prec[5]);
        // first stored node: its loads overlap the left run
        T pre[kPer];
        auto load_node = [&](int p) {
            const T* node = cache + prec[kPlanHead + p] * kT;
            chen::static_for<kPer>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + k * kBlock;
                if (i < kT) pre[k] = node[i];
            });
        };
        if (nn > 0) load_node(0);

        chen::Lane<T, W, D, P> st;
        st.init_identity();
        run(st, l0, l1);
        if (active) store(st, S);
        for (int p = 0; p < nn; ++p) {
            chen::static_for<kPer>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + k * kBlock;
                if (i < kT) N[i] = pre[k];
            });
            __syncthreads();  // N staged; S written by the lanes
            if (p + 1 < nn) load_node(p + 1);
            wg_mul_inplace_t<T, W, D>(S, N);
        }
        if (nn == 0) __syncthreads();  // S written by the lanes
        T* o = out + q * kT;
        if (rr1 > rr0) {
            if (active) fetch(st, S);
            run(st, rr0, rr1);
            if (active) store(st, o);
        } else {
            for (int i = threadIdx.x; i < kT; i += kBlock) o[i] = S[i];
        }
        __syncthreads();
    }
}

size_t device_lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 64 * 1024;
    return size_t(v);
}

// Launches the fast query kernel if its LDS (two tensors plus the staged
// run table) fits the device; returns false to fall back.
template <typename T, int W, int D, int P>
bool launch_query_fast(void* out, const void* cache, const void* incr, const double* params,
                       const int64_t* offsets, int64_t n_streams, int r, int J,
                       const int64_t* q_stream, const double* q_inf, const double* q_sup,
                       int64_t n_queries, uint32_t grid, hipStream_t st)
{
    const size_t dyn = 2 * size_t(dbeg<W>(D + 1)) * sizeof(T);
    const size_t stat = size_t(kQueryChunk) * D * W * sizeof(T) + 1024;
    if (dyn + stat > device_lds_limit()) return false;
    if (dyn + stat > 64 * 1024) {
        // extended LDS, reserved once per process and shape
        static const hipError_t err =
                hipFuncSetAttribute((const void*)dyadic_query_fast_kernel<T, W, D, P>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(dyn));
        if (err != hipSuccess) return false;
    }
    const int rec = plan_record_words(J);
    int64_t* plans = nullptr;
    if (workspace_alloc((void**)&plans, size_t(n_queries) * size_t(rec) * sizeof(int64_t), st) !=
        hipSuccess)
        return false;
    const uint32_t pgrid = uint32_t((n_queries + 255) / 256);
    hipLaunchKernelGGL(dyadic_query_plan_kernel, dim3(pgrid), dim3(256), 0, st, plans, rec,
                       params, offsets, n_streams, r, J, q_stream, q_inf, q_sup, n_queries);
    hipLaunchKernelGGL((dyadic_query_fast_kernel<T, W, D, P>), dim3(grid), dim3(kBlock), dyn, st,
                       (T*)out, (const T*)cache, (const T*)incr, plans, rec, n_queries);
    (void)hipFreeAsync(plans, st);
    return true;
}

using QueryLauncher = bool (*)(void*, const void*, const void*, const double*, const int64_t*,
                               int64_t, int, int, const int64_t*, const double*, const double*,
                               int64_t, uint32_t, hipStream_t);

// Prefix length of a query run: the smallest P with at most 128 lanes whose
// lane state fits `budget` values, or 0.
constexpr int query_p(int W, int D, int budget)
{
    for (int p = 1; p <= D; ++p) {
        if (ipow(W, p) > 128) return 0;
        int sub = 0;
        for (int j = p; j <= D; ++j) sub += int(ipow(W, j - p));
        if (sub + p <= budget) return p;
    }
    return 0;
}

template <typename T, int W, int D>
constexpr QueryLauncher query_entry()
{
    constexpr int P = query_p(W, D, sizeof(T) == 4 ? 80 : 40);
    if constexpr (P > 0)
        return &launch_query_fast<T, W, D, P>;
    else
        return nullptr;
}

// Shapes with a query kernel of their own, float32 and float64.
template <typename T>
QueryLauncher find_query_fast(int w, int d)
{
#define RPY_Q(W, D) \
    if (w == W && d == D) return query_entry<T, W, D>();
    RPY_Q(2, 2) RPY_Q(2, 3) RPY_Q(2, 4) RPY_Q(2, 5) RPY_Q(2, 6) RPY_Q(2, 7) RPY_Q(2, 8)
    RPY_Q(3, 2) RPY_Q(3, 3) RPY_Q(3, 4) RPY_Q(3, 5) RPY_Q(3, 6)
    RPY_Q(4, 2) RPY_Q(4, 3) RPY_Q(4, 4) RPY_Q(4, 5)
    RPY_Q(5, 2) RPY_Q(5, 3) RPY_Q(5, 4) RPY_Q(5, 5)
    RPY_Q(6, 2) RPY_Q(6, 3) RPY_Q(6, 4)
    RPY_Q(7, 2) RPY_Q(7, 3) RPY_Q(7, 4)
    RPY_Q(8, 2) RPY_Q(8, 3) RPY_Q(8, 4)
#undef RPY_Q
    return nullptr;
}

uint32_t blocks_for(int64_t n)
{
    return uint32_t(std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 65536)));
}

int check_args(int32_t width, int32_t depth, int32_t resolution, int32_t levels, rpy_dtype dtype,
               DevBasis& b)
{
    rpy_tensor_basis tb{width, depth, nullptr};
    int rc = make_basis(&tb, b);
    if (rc) return rc;
    if (depth < 1) return fail(RPY_ERR_INVALID_ARGUMENT, "the cache needs depth >= 1");
    if (width > 32) return fail(RPY_ERR_UNSUPPORTED, "the dyadic cache supports width <= 32");
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (resolution < 0 || resolution > kMaxRes)
        return fail(RPY_ERR_INVALID_ARGUMENT, "resolution must be in [0, %d]", kMaxRes);
    if (levels < 0 || levels > resolution || levels > 24)
        return fail(RPY_ERR_INVALID_ARGUMENT,
                    "stored levels must be in [0, min(resolution, 24)], got %d", levels);
    return RPY_OK;
}

}  // namespace

int64_t dyadic_cache_elements(int64_t n_streams, int32_t width, int32_t depth, int32_t levels)
{
    const int64_t tsize = tensor_size_checked(width, depth);
    if (tsize < 0 || n_streams < 0 || levels < 0 || levels > 24) return -1;
    return ((int64_t(1) << (levels + 1)) - 1) * n_streams * tsize;
}

int dyadic_cache_build(void* cache, const void* increments, const double* params,
                       const int64_t* offsets, int64_t n_streams, int32_t width, int32_t depth,
                       int32_t resolution, int32_t levels, rpy_dtype dtype, void* stream)
{
    DevBasis b;
    int rc = check_args(width, depth, resolution, levels, dtype, b);
    if (rc) return rc;
    if (n_streams < 0) return fail(RPY_ERR_SHAPE, "negative number of streams");
    if (n_streams == 0) return RPY_OK;
    if (!cache || !increments || !params || !offsets)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int J = levels;
    const int64_t tsize = b.size();
    const int64_t nseg = n_streams << J;
    const size_t es = dtype == RPY_F64 ? 8 : 4;

    int64_t* bounds = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&bounds, size_t(2 * nseg) * sizeof(int64_t), st));
    hipLaunchKernelGGL(segment_bounds_kernel, dim3(blocks_for(nseg)), dim3(kBlock), 0, st, params,
                       offsets, n_streams, J, bounds, bounds + nseg);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "segment launch failed: %s", hipGetErrorString(err));
    // level J: one signature per segment, written in place
    char* base = static_cast<char*>(cache);
    if (!rc)
        rc = stream_signature_ranges(base + size_t(node_offset(J, n_streams, 0, 0) * tsize) * es,
                                     increments, bounds, bounds + nseg, nseg, width, depth, dtype,
                                     stream);
    // levels J-1 .. 0: node(j, s, k) = node(j+1, s, 2k) (x) node(j+1, s, 2k+1)
    rpy_tensor_basis tb{width, depth, nullptr};
    for (int j = J - 1; j >= 0 && !rc; --j) {
        const int64_t batch = n_streams << j;
        char* outp = base + size_t(node_offset(j, n_streams, 0, 0) * tsize) * es;
        const char* lhs = base + size_t(node_offset(j + 1, n_streams, 0, 0) * tsize) * es;
        // fixed shapes write the parents directly; others clear them and add
        if (has_fixed_mul(width, depth, dtype)) {
            rc = dense_ft_mul_new(outp, lhs, lhs + size_t(tsize) * es, dtype, &tb, batch, tsize,
                                  2 * tsize, 2 * tsize, stream);
            continue;
        }
        if (dtype == RPY_F64)
            hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(blocks_for(batch * tsize)), dim3(kBlock),
                               0, st, (double*)outp, batch * tsize);
        else
            hipLaunchKernelGGL(fill_zero_kernel<float>, dim3(blocks_for(batch * tsize)), dim3(kBlock),
                               0, st, (float*)outp, batch * tsize);
        err = hipGetLastError();
        if (err != hipSuccess) {
            rc = fail(RPY_ERR_HIP, "zero launch failed: %s", hipGetErrorString(err));
            break;
        }
        rc = dense_ft_fma(outp, lhs, lhs + size_t(tsize) * es, dtype, &tb, {0, -1}, {0, -1}, {0, -1},
                          batch, tsize, 2 * tsize, 2 * tsize, stream);
    }
    hipError_t ferr = hipFreeAsync(bounds, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

int dyadic_cache_query(void* out, const void* cache, const void* increments, const double* params,
                       const int64_t* offsets, int64_t n_streams, int32_t width, int32_t depth,
                       int32_t resolution, int32_t levels, const int64_t* q_stream,
                       const double* q_inf, const double* q_sup, int64_t n_queries,
                       rpy_dtype dtype, void* stream)
{
    DevBasis b;
    int rc = check_args(width, depth, resolution, levels, dtype, b);
    if (rc) return rc;
    if (n_queries < 0 || n_streams < 0) return fail(RPY_ERR_SHAPE, "negative size");
    if (n_queries == 0) return RPY_OK;
    if (!out || !cache || !increments || !params || !offsets || !q_stream || !q_inf || !q_sup)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const int64_t tsize = b.size();
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    const Pow pw = make_pow(width, depth);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const uint32_t grid = uint32_t(std::min<int64_t>(n_queries, 16384));
    const QueryLauncher fast = dtype == RPY_F64 ? find_query_fast<double>(width, depth)
                                                : find_query_fast<float>(width, depth);
    if (fast && fast(out, cache, increments, params, offsets, n_streams, resolution, levels,
                     q_stream, q_inf, q_sup, n_queries, grid, st)) {
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    const size_t lds = 2 * size_t(tsize) * es;
    if (lds > 64 * 1024)
        return fail(RPY_ERR_UNSUPPORTED, "queries need %zu bytes of LDS (limit 65536)", lds);
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(dyadic_query_kernel<double>, dim3(grid), dim3(kBlock), lds, st,
                           (double*)out, (const double*)cache, (const double*)increments, params,
                           offsets, n_streams, b, pw, resolution, levels, q_stream, q_inf, q_sup,
                           n_queries);
    else
        hipLaunchKernelGGL(dyadic_query_kernel<float>, dim3(grid), dim3(kBlock), lds, st,
                           (float*)out, (const float*)cache, (const float*)increments, params,
                           offsets, n_streams, b, pw, resolution, levels, q_stream, q_inf, q_sup,
                           n_queries);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

// Host mirror of the device cover, for CPU tests of the dissection.
int dyadic_cover_host(int64_t A, int64_t B, int r, int64_t* ks, int* ns, int max_pieces)
{
    int64_t k[kMaxPieces];
    int n[kMaxPieces];
    if (r < 0 || r > kMaxRes || A < 0 || B > (int64_t(1) << r))
        return -1;
    const int cnt = dyadic_cover(A, B, r, k, n);
    for (int i = 0; i < cnt && i < max_pieces; ++i) {
        ks[i] = k[i];
        ns[i] = n[i];
    }
    return cnt;
}

}  // namespace rpy_amd
