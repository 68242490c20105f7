// log_map.hip -- the log-signature's last two steps in one kernel for the
// fixed shapes: the float64 tensor log of E signatures (the Horner of
// dense_ops.hip's horner_fixed_kernel, restating the log of
// roughpy_compute/dense/intermediate/free_tensor_log.hpp:12-40) kept in LDS,
// then their projection to the Hall basis by the t2l matrix.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "chen_core.hpp"
#include "dense_ops.hpp"
#include "horner_fixed.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kLogMapE = 4;  // tensors per workgroup, at most

// Threads and LDS per workgroup by depth: up to depth 4 one tensor per
// workgroup of 256 threads, more workgroups per CU ((8,4) x 4096: 0.444 ->
// 0.413 ms per log-signature call, float64); deeper shapes two tensors per
// workgroup of 512 ((5,5) x 16384: 0.80 against 0.91 ms with one).
constexpr int log_map_block(int D) { return D <= 4 ? 256 : 512; }
constexpr size_t log_map_lds(int D) { return D <= 4 ? 64 * 1024 : 128 * 1024; }

// ---------------------------------------------------------------------------
// The log-signature's last two steps in one kernel: the float64 tensor log of
// E signatures (the fixed-shape Horner of horner_fixed.hpp, its result kept
// in LDS instead of written out), then their projection by the t2l matrix
// from LDS.  The log tensors never go to HBM, and the matrix is read from L2
// once per E tensors.
// Each projected coefficient sums its row's entries in order, as
// csr_map_lds_kernel does.
// ---------------------------------------------------------------------------
template <int W, int D, typename Tin, int E>
__global__ __launch_bounds__(log_map_block(D)) void log_t2l_fixed_kernel(
        double* __restrict__ out, const Tin* __restrict__ arg, int64_t batch, int64_t os,
        int64_t as, const int64_t* __restrict__ offs, const int64_t* __restrict__ idx,
        const double* __restrict__ val, int64_t rows)
{
    using T = double;
    constexpr int kLogMapBlock = log_map_block(D);
    constexpr uint32_t ns = uint32_t(fdb<W>(D));      // levels 0..D-1
    constexpr uint32_t np = (ns + 15) / 16 * 16;
    constexpr uint32_t nf = uint32_t(fdb<W>(D + 1));  // the whole tensor
    constexpr uint32_t nfp = (nf + 1) / 2 * 2;
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    T* xs = a + np;
    T* slot = xs + np;  // slot q: the log of tensor e0 + q
    constexpr int kIt = int((ns + kLogMapBlock - 1) / kLogMapBlock);
    for (int64_t e0 = int64_t(blockIdx.x) * E; e0 < batch; e0 += int64_t(gridDim.x) * E) {
        const int ne = int(batch - e0 < E ? batch - e0 : E);
        for (int q = 0; q < ne; ++q) {
            const Tin* x = arg + (e0 + q) * as;
            T* o = slot + q * nfp;
            Tin xv[kIt];
            chen::static_for<kIt>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const uint32_t i = threadIdx.x + uint32_t(k) * kLogMapBlock;
                if (i < ns) xv[k] = x[i];
            });
            chen::static_for<kIt>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const uint32_t i = threadIdx.x + uint32_t(k) * kLogMapBlock;
                if (i < ns) {
                    xs[i] = T(xv[k]);
                    a[i] = T(0);
                }
            });
            __syncthreads();
            chen::static_for<D>([&](auto di) {
                constexpr int deg = D - int(decltype(di)::value);
                constexpr int ml = D - deg + 1;
                if (threadIdx.x == 0) {
                    const T inv = T(1) / T(deg);
                    a[0] = (deg % 2 == 0) ? a[0] - inv : a[0] + inv;
                }
                __syncthreads();
                wg_mul_by_nounit_fixed<T, W, D, ml>(a, xs, T(1), x, o, nullptr);
            });
            for (uint32_t i = threadIdx.x; i < ns; i += kLogMapBlock) o[i] = a[i];
            __syncthreads();
        }
        for (int64_t r = threadIdx.x; r < rows; r += kLogMapBlock) {
            T acc[E];
            chen::static_for<E>([&](auto q) { acc[q] = T(0); });
            const int64_t pe = offs[r + 1];
            int64_t p = offs[r];
            for (; p + 4 <= pe; p += 4) {
                int64_t c[4];
                T v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    c[j] = idx[p + j];
                    v[j] = val[p + j];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    chen::static_for<E>([&](auto q) { acc[q] += v[j] * slot[q * nfp + c[j]]; });
            }
            if (p < pe) {
                int64_t c[3];
                T v[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    c[j] = p + j < pe ? idx[p + j] : 0;
                    v[j] = p + j < pe ? val[p + j] : T(0);
                }
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (p + j < pe)
                        chen::static_for<E>([&](auto q) { acc[q] += v[j] * slot[q * nfp + c[j]]; });
            }
            chen::static_for<E>([&](auto q) {
                if (int(q) < ne) out[(e0 + int(q)) * os + r] = acc[q];
            });
        }
        __syncthreads();  // the slots' readers are done
    }
}

template <int W, int D, typename Tin>
bool launch_log_t2l_fixed(double* out, const void* arg, int64_t batch, int64_t os, int64_t as,
                          const rpy_sparse_matrix* m, hipStream_t st)
{
    constexpr size_t np = (size_t(fdb<W>(D)) + 15) / 16 * 16;
    constexpr size_t nfp = (size_t(fdb<W>(D + 1)) + 1) / 2 * 2;
    // E = tensors per workgroup: the most (up to kLogMapE) whose slots fit
    // kLogMapLds with the Horner accumulator and the staged levels
    constexpr size_t kLogMapLds = log_map_lds(D);
    constexpr size_t per = nfp * sizeof(double), base = 2 * np * sizeof(double);
    constexpr int E = base + kLogMapE * per <= kLogMapLds ? kLogMapE
                      : base + 2 * per <= kLogMapLds       ? 2
                      : base + per <= kLogMapLds           ? 1
                                                           : 0;
    if constexpr (E == 0) {
        return false;
    } else {
        constexpr size_t lds = base + size_t(E) * per;
        auto* kern = log_t2l_fixed_kernel<W, D, Tin, E>;
        if constexpr (lds > 64 * 1024) {
            static const hipError_t err = hipFuncSetAttribute(
                    (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (err != hipSuccess) return false;
        }
        const int64_t groups = (batch + E - 1) / E;
        hipLaunchKernelGGL(kern, dim3(uint32_t(std::min<int64_t>(groups, 8192))),
                           dim3(log_map_block(D)), lds, st, out, (const Tin*)arg, batch, os, as,
                           m->offsets, m->indices, (const double*)m->data, m->rows);
        return true;
    }
}

using LogMapFixed = bool (*)(double*, const void*, int64_t, int64_t, int64_t,
                             const rpy_sparse_matrix*, hipStream_t);

template <typename Tin>
LogMapFixed find_log_t2l_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_log_t2l_fixed<W, D, Tin>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

}  // namespace

// log + t2l of float32 or float64 signatures for the fixed shapes;
// RPY_ERR_UNSUPPORTED where there is no fused kernel (the caller then runs
// the log and the map separately).
int dense_log_t2l_fixed(double* out, const void* sig, rpy_dtype sig_dtype, int width, int depth,
                        const rpy_sparse_matrix* t2l, int64_t batch, int64_t os, int64_t as,
                        void* stream)
{
    if (!t2l || !t2l->compressed_rows || !t2l->offsets || (t2l->nnz > 0 && (!t2l->indices ||
                                                                          !t2l->data)))
        return RPY_ERR_UNSUPPORTED;
    if (sig_dtype != RPY_F32 && sig_dtype != RPY_F64) return RPY_ERR_UNSUPPORTED;
    LogMapFixed f = sig_dtype == RPY_F32 ? find_log_t2l_fixed<float>(width, depth)
                                         : find_log_t2l_fixed<double>(width, depth);
    if (!f || !f(out, sig, batch, os, as, t2l, reinterpret_cast<hipStream_t>(stream))) return RPY_ERR_UNSUPPORTED;
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

}  // namespace rpy_amd
