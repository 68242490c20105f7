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
// One workgroup owns one stream; its four working tensors live in a
// per-workgroup global scratch slot (L2-resident).
//
// Also the adjoint of right multiplication, R*_B(g)(u) = sum_v g(uv) B(v),
// the right-hand counterpart of free_tensor_adjoint_left_mul.hpp:10-77.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "chen.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"

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

template <typename T>
__global__ __launch_bounds__(kBlock) void sig_vjp_kernel(
        T* __restrict__ grad, const T* __restrict__ incr, const int64_t* __restrict__ offsets,
        int64_t n_streams, const T* __restrict__ sig, const T* __restrict__ gsig,
        T* __restrict__ scratch, DevBasis b, Pow pw)
{
    const int W = b.width;
    const int D = b.depth;
    const int64_t tsize = b.size();
    T* A = scratch + int64_t(blockIdx.x) * 4 * tsize;
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

}  // namespace

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
    const Pow pw = make_pow(width, depth);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t grid = std::min<int64_t>(n_streams, 2048);
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    void* scratch = nullptr;
    RPY_HIP_CHECK(hipMallocAsync(&scratch, size_t(grid) * 4 * size_t(b.size()) * es, st));
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(sig_vjp_kernel<double>, dim3(uint32_t(grid)), dim3(kBlock), 0, st,
                           (double*)grad, (const double*)increments, offsets, n_streams,
                           (const double*)signatures, (const double*)grad_signatures,
                           (double*)scratch, b, pw);
    else
        hipLaunchKernelGGL(sig_vjp_kernel<float>, dim3(uint32_t(grid)), dim3(kBlock), 0, st,
                           (float*)grad, (const float*)increments, offsets, n_streams,
                           (const float*)signatures, (const float*)grad_signatures,
                           (float*)scratch, b, pw);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "vjp launch failed: %s", hipGetErrorString(err));
    hipError_t ferr = hipFreeAsync(scratch, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
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

}  // namespace rpy_amd
