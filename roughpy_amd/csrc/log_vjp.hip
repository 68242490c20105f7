// log_vjp.hip -- cotangent of the truncated tensor logarithm, the log step of
// the hand-written log-signature backward.  The reference has no derivatives;
// the forward map is the log of dense_ft_log
// (roughpy_compute/dense/intermediate/free_tensor_log.hpp:12-40, restated in
// dense_ops.hip's horner_kernel): with Y = arg on degrees >= 1 (arg's degree 0
// is not read) and c_m = (-1)^(m+1) / m,
//   A_D = c_D 1,   A_k = A_{k+1} (x) Y + c_k 1  (k = D-1 .. 1),   log = A_1 (x) Y.
// Reverse mode over that Horner chain, deg = 1 .. D, with G the cotangent of
// A_deg (x) Y (G = the given cotangent at deg = 1):
//   Ybar += L*_{A_deg} G    Ybar[w] += sum_u A_deg[u] G[u w]
//   G     = R*_Y G          G[u] = sum_{|v| >= 1} Y[v] G[u v]
// and the cotangent of arg is Ybar on degrees >= 1, 0 on degree 0.  A_deg
// has degrees 0 .. D - deg and is rebuilt in place from A_D each step
// (D(D-1)/2 small products) rather than stored.
//
// One workgroup per tensor, arithmetic in float64 whatever the data type.
// Y, A, G and Ybar sit in LDS when the four tensors fit (width 8 depth 4:
// 146 KiB), else in a global workspace.  Contractions whose outputs each
// sum many terms (low degrees of L* and R*) give each output a wave and
// reduce across its lanes in a fixed order, so results do not depend on
// timing.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "chen.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"
#include "strided.hpp"

namespace rpy_amd {

namespace {

constexpr int kLvThreads = 512;
constexpr int kLvWaves = kLvThreads / 64;

// Sum over the 64 lanes of a wave; every lane gets the total.
RPY_FI double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// A <- A (x) Y + c 1 on degrees 1 .. ml (A's degrees above ml - 1 are zero
// on entry), top-down in place: degree m reads only degrees < m.
RPY_FI void horner_step(double* A, const double* Y, const DevBasis& b, const Pow& pw, int ml,
                        double c)
{
    for (int m = ml; m >= 1; --m) {
        const uint32_t n = uint32_t(b.level_size(m));
        for (uint32_t idx = threadIdx.x; idx < n; idx += kLvThreads) {
            double acc = 0.0;
            for (int ld = m - 1; ld >= 0; --ld) {
                const int rd = m - ld;
                const uint32_t i = fdiv(idx, pw.pw[rd]);
                const uint32_t j = idx - i * pw.pw[rd].d;
                acc += A[b.db[ld] + i] * Y[b.db[rd] + j];
            }
            A[b.db[m] + idx] = acc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) A[0] = c;
    __syncthreads();
}

// Ybar[w] += sum_{|u| <= amax} A[u] G[u w] for every word w of degree >= 1.
RPY_FI void lstar_acc(double* Yb, const double* A, const double* G, const DevBasis& b,
                      const Pow& pw, int amax)
{
    const int D = b.depth;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int m = 1; m <= D; ++m) {
        const int jmax = min(amax, D - m);
        const uint32_t n = uint32_t(b.level_size(m));
        const uint32_t wm = pw.pw[m].d;
        // terms per output: words u of degree 0 .. jmax
        const int64_t K = b.db[jmax + 1];
        if (K > 32) {
            for (uint32_t idx = wave; idx < n; idx += kLvWaves) {
                double acc = 0.0;
                for (int64_t t = lane; t < K; t += 64) {
                    int j = 0;
                    while (t >= b.db[j + 1]) ++j;
                    const int64_t u = t - b.db[j];
                    acc += A[t] * G[b.db[j + m] + u * wm + idx];
                }
                acc = wave_sum(acc);
                if (lane == 0) Yb[b.db[m] + idx] += acc;
            }
        } else {
            for (uint32_t idx = threadIdx.x; idx < n; idx += kLvThreads) {
                double acc = 0.0;
                for (int j = 0; j <= jmax; ++j) {
                    const uint32_t nu = uint32_t(b.level_size(j));
                    for (uint32_t u = 0; u < nu; ++u)
                        acc += A[b.db[j] + u] * G[b.db[j + m] + int64_t(u) * wm + idx];
                }
                Yb[b.db[m] + idx] += acc;
            }
        }
    }
    __syncthreads();
}

// G[u] <- sum_{|v| >= 1} Y[v] G[u v] in place, degrees 1 .. D-1 bottom-up
// (degree m reads only degrees > m); degree D becomes 0 and degree 0, whose
// cotangent no later step reads, is left as it is.
RPY_FI void rstar_inplace(double* G, const double* Y, const DevBasis& b, const Pow& pw)
{
    const int D = b.depth;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int m = 1; m < D; ++m) {
        const uint32_t n = uint32_t(b.level_size(m));
        // terms per output: words v of degree 1 .. D - m
        const int64_t K = b.db[D - m + 1] - 1;
        if (K > 32) {
            for (uint32_t idx = wave; idx < n; idx += kLvWaves) {
                double acc = 0.0;
                for (int64_t t = 1 + lane; t <= K; t += 64) {
                    int k = 1;
                    while (t >= b.db[k + 1]) ++k;
                    const int64_t v = t - b.db[k];
                    acc += Y[t] * G[b.db[m + k] + int64_t(idx) * pw.pw[k].d + v];
                }
                acc = wave_sum(acc);
                if (lane == 0) G[b.db[m] + idx] = acc;
            }
        } else {
            for (uint32_t idx = threadIdx.x; idx < n; idx += kLvThreads) {
                double acc = 0.0;
                for (int k = 1; k <= D - m; ++k) {
                    const uint32_t nv = uint32_t(b.level_size(k));
                    const int64_t base = b.db[m + k] + int64_t(idx) * pw.pw[k].d;
                    for (uint32_t v = 0; v < nv; ++v) acc += Y[b.db[k] + v] * G[base + v];
                }
                G[b.db[m] + idx] = acc;
            }
        }
        __syncthreads();
    }
    const uint32_t nD = uint32_t(b.level_size(D));
    for (uint32_t idx = threadIdx.x; idx < nD; idx += kLvThreads) G[b.db[D] + idx] = 0.0;
    __syncthreads();
}

template <typename IO, bool InLds>
__global__ __launch_bounds__(kLvThreads) void ft_log_vjp_kernel(
        IO* __restrict__ out, const IO* __restrict__ arg, const IO* __restrict__ cot, DevBasis b,
        Pow pw, int64_t batch, int64_t os, int64_t as, int64_t cs, double* __restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = b.depth;
    const int64_t n = b.size();
    double* base = InLds ? reinterpret_cast<double*>(smem) : ws + int64_t(blockIdx.x) * 4 * n;
    double* Y = base;
    double* A = base + n;
    double* G = base + 2 * n;
    double* Yb = base + 3 * n;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const IO* x = arg + e * as;
        const IO* h = cot + e * cs;
        for (int64_t i = threadIdx.x; i < n; i += kLvThreads) {
            Y[i] = i == 0 ? 0.0 : double(x[i]);
            G[i] = double(h[i]);
            Yb[i] = 0.0;
        }
        __syncthreads();
        for (int deg = 1; deg <= D; ++deg) {
            // A_deg: degrees 0 .. D - deg
            for (int64_t i = threadIdx.x; i < n; i += kLvThreads)
                A[i] = i == 0 ? ((D % 2) ? 1.0 : -1.0) / double(D) : 0.0;
            __syncthreads();
            for (int k = D - 1; k >= deg; --k)
                horner_step(A, Y, b, pw, D - k, ((k % 2) ? 1.0 : -1.0) / double(k));
            lstar_acc(Yb, A, G, b, pw, D - deg);
            if (deg < D) rstar_inplace(G, Y, b, pw);
        }
        IO* o = out + e * os;
        for (int64_t i = threadIdx.x; i < n; i += kLvThreads) o[i] = i == 0 ? IO(0) : IO(Yb[i]);
        __syncthreads();
    }
}

template <typename IO>
int launch_log_vjp(void* out, const void* arg, const void* cot, const DevBasis& b, int64_t batch,
                   int64_t os, int64_t as, int64_t cs, hipStream_t st)
{
    const int64_t n = b.size();
    const Pow pw = make_pow(b.width, b.depth);
    const size_t lds = size_t(4) * size_t(n) * sizeof(double);
    int dev = 0, max_lds = 0, cus = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeSharedMemPerBlockOptin,
                                        dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (lds <= size_t(max_lds)) {
        // one or more workgroups per CU, as LDS allows
        const int per_cu = std::max(1, int(size_t(max_lds) / lds));
        const int64_t grid = std::min<int64_t>(batch, int64_t(cus) * per_cu * 4);
        RPY_HIP_CHECK(hipFuncSetAttribute((const void*)ft_log_vjp_kernel<IO, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        hipLaunchKernelGGL((ft_log_vjp_kernel<IO, true>), dim3(uint32_t(grid)),
                           dim3(kLvThreads), lds, st, (IO*)out, (const IO*)arg, (const IO*)cot, b,
                           pw, batch, os, as, cs, (double*)nullptr);
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    const int64_t grid = std::min<int64_t>(batch, int64_t(cus) * 2);
    double* ws = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&ws, size_t(grid) * lds, st));
    hipLaunchKernelGGL((ft_log_vjp_kernel<IO, false>), dim3(uint32_t(grid)), dim3(kLvThreads), 0,
                       st, (IO*)out, (const IO*)arg, (const IO*)cot, b, pw, batch, os, as, cs, ws);
    hipError_t err = hipGetLastError();
    hipError_t ferr = hipFreeAsync(ws, st);
    if (err != hipSuccess)
        return fail(RPY_ERR_HIP, "log cotangent launch failed: %s", hipGetErrorString(err));
    if (ferr != hipSuccess) return fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return RPY_OK;
}

}  // namespace

int dense_ft_log_vjp(void* out, const void* arg, const void* cotangent, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t as,
                     int64_t cs, void* stream)
{
    if (!basis) return fail(RPY_ERR_INVALID_ARGUMENT, "null basis");
    DevBasis b;
    int rc = make_basis(basis, b);
    if (rc) return rc;
    if (dtype != RPY_F32 && dtype != RPY_F64) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch");
    if (batch == 0) return RPY_OK;
    if (!out || !arg || !cotangent) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const int64_t n = b.size();
    if (n > int64_t(INT32_MAX)) return fail(RPY_ERR_SHAPE, "tensor too large");
    if ((batch > 1) && (os < n || as < n || cs < n))
        return fail(RPY_ERR_SHAPE, "batch stride smaller than the tensor size %lld",
                    (long long)n);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    return dtype == RPY_F64 ? launch_log_vjp<double>(out, arg, cotangent, b, batch, os, as, cs, st)
                            : launch_log_vjp<float>(out, arg, cotangent, b, batch, os, as, cs, st);
}

}  // namespace rpy_amd
