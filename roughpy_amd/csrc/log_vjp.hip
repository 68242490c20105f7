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
// has degrees 0 .. D - deg; G_deg has degrees 1 .. D - deg + 1 (degree 0 is
// never read), so every sum stops where its operands do.
//
// One workgroup per tensor, arithmetic in float64 whatever the data type.
// The chain A_D .. A_1 is built once per tensor (one barrier per step) and
// kept; each reverse step deg then reads only A_deg, Y and G_deg and writes
// Ybar and G_{deg+1} into the other of two cotangent buffers, so it is one
// barrier phase.  Only what is read sits in LDS: Y below the top degree (its
// top degree is never read), G (full at deg 1, below the top degree after),
// Ybar below the top degree (its top degree is A_1[0] G = the given
// cotangent, written straight to the output) and the chain -- 56.8 KiB at
// width 8 depth 4, two workgroups per CU.  Shapes whose arrays exceed LDS
// run the same code on a global workspace.
//
// Each contraction is cut into wave tiles; the lanes of a tile read
// consecutive words of the large operand and the partial sums of one output
// are added by a fixed xor-shuffle tree, so results do not depend on timing.

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
constexpr int kLvPre = 14;

// Per-tensor arrays, offsets in doubles: G0 (n: the cotangent at odd
// steps), G1 (db[D]: at even steps), Y (db[D]), Ybar (db[D]) and the chain,
// A_k at a + ak[k] with db[D - k + 1] coefficients (degrees 0 .. D - k).
struct LvLayout {
    int64_t g0, g1, y, yb, a, size;
    int64_t ak[kMaxDepth + 1];
};

LvLayout make_layout(const DevBasis& b)
{
    const int D = b.depth;
    const int64_t n = b.size(), nl = b.db[D];
    LvLayout L{};
    L.g0 = 0;
    L.g1 = n;
    L.y = n + nl;
    L.yb = n + 2 * nl;
    L.a = n + 3 * nl;
    int64_t off = 0;
    for (int k = D; k >= 1; --k) {
        L.ak[k] = off;
        off += b.db[D - k + 1];
    }
    L.size = L.a + off;
    return L;
}

// c_k = (-1)^(k+1) / k
RPY_FI double log_coef(int k) { return ((k % 2) ? 1.0 : -1.0) / double(k); }

// smallest power of two >= x, at most 64
RPY_FI int pow2_cap64(uint32_t x)
{
    int p = 1;
    while (p < 64 && uint32_t(p) < x) p <<= 1;
    return p;
}

// first tile index >= t0 that wave `wave` takes (tiles go round robin)
RPY_FI int first_tile(int t0, int wave)
{
    return t0 + ((wave - t0) % kLvWaves + kLvWaves) % kLvWaves;
}

// Workgroup barrier.  In LDS only the LDS operations have to be complete;
// pending global stores (the previous tensor's output) are left in flight.
template <bool InLds>
RPY_FI void lv_barrier()
{
    if constexpr (InLds)
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else
        __syncthreads();
}

template <typename IO, bool InLds>
__global__ __launch_bounds__(kLvThreads) void ft_log_vjp_kernel(
        IO* __restrict__ out, const IO* __restrict__ arg, const IO* __restrict__ cot, DevBasis b,
        Pow pw, LvLayout lay, int64_t batch, int64_t os, int64_t as, int64_t cs,
        double* __restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = b.depth;
    const uint32_t w = uint32_t(b.width);
    const FastDiv fw = pw.pw[1];
    const int64_t n = b.size();
    const int64_t nl = b.db[D];  // coefficients below the top degree
    const int64_t nin = n + nl;  // words read per tensor: the cotangent, then arg below the top
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    double* base = InLds ? reinterpret_cast<double*>(smem) : ws + int64_t(blockIdx.x) * lay.size;
    double* G0 = base + lay.g0;
    double* G1 = base + lay.g1;
    double* Y = base + lay.y;
    double* Yb = base + lay.yb;
    double* As = base + lay.a;
    // the next tensor's first kLvPre * kLvThreads words are loaded into
    // registers while the current one is computed
    double pre[kLvPre];
    auto fetch = [&](int64_t e) {
        const IO* x = arg + e * as;
        const IO* h = cot + e * cs;
#pragma unroll
        for (int q = 0; q < kLvPre; ++q) {
            const int64_t i = int64_t(q) * kLvThreads + tid;
            pre[q] = i < n ? double(h[i]) : i < nin ? double(x[i - n]) : 0.0;
        }
    };
    auto stage = [&](int64_t i, double v) {
        if (i < n)
            G0[i] = v;
        else
            Y[i - n] = i == n ? 0.0 : v;
    };
    if (int64_t(blockIdx.x) < batch) fetch(blockIdx.x);
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        // G_1 = the given cotangent, Y = arg below the top degree (degree 0 unused)
#pragma unroll
        for (int q = 0; q < kLvPre; ++q) {
            const int64_t i = int64_t(q) * kLvThreads + tid;
            if (i < nin) stage(i, pre[q]);
        }
        {
            const IO* x = arg + e * as;
            const IO* h = cot + e * cs;
            for (int64_t i = int64_t(kLvPre) * kLvThreads + tid; i < nin; i += kLvThreads)
                stage(i, i < n ? double(h[i]) : double(x[i - n]));
        }
        for (int64_t i = tid; i < nl; i += kLvThreads) Yb[i] = 0.0;
        if (tid == 0 && D > 0) As[lay.ak[D]] = log_coef(D);
        lv_barrier<InLds>();
        if (e + gridDim.x < batch) fetch(e + gridDim.x);
        // A_k = A_{k+1} (x) Y + c_k 1 on degrees 0 .. D - k, k = D-1 .. 1;
        // degree m: sum over splits idx = i w^rd + j, rd = 1 .. m
        for (int k = D - 1; k >= 1; --k) {
            const double* An = As + lay.ak[k + 1];
            double* Ak = As + lay.ak[k];
            if (tid == 0) Ak[0] = log_coef(k);
            uint32_t wm = 1;
            int64_t dbm = 0;
            for (int m = 1; m <= D - k; ++m) {
                dbm += wm;  // db[m]
                wm *= w;    // w^m
                for (uint32_t idx = tid; idx < wm; idx += kLvThreads) {
                    uint32_t i = idx, j = 0, wr = 1, wl = wm;
                    int64_t dbr = 1, dbl = dbm;
                    double acc = 0.0;
                    for (int rd = 1; rd <= m; ++rd) {
                        const uint32_t q = fdiv(i, fw);
                        j += (i - q * w) * wr;  // idx mod w^rd
                        i = q;                  // idx / w^rd
                        wl = fdiv(wl, fw);      // w^(m - rd)
                        dbl -= wl;              // db[m - rd]
                        acc += An[dbl + i] * Y[dbr + j];
                        dbr += int64_t(wr) * w;  // db[rd + 1]
                        wr *= w;
                    }
                    Ak[dbm + idx] = acc;
                }
            }
            lv_barrier<InLds>();
        }
        // reverse steps; G_deg has degrees 1 .. D - deg + 1
        for (int deg = 1; deg <= D; ++deg) {
            const double* A = As + lay.ak[deg];
            const double* Gs = (deg & 1) ? G0 : G1;
            double* Gd = (deg & 1) ? G1 : G0;
            const int top = D - deg + 1;
            int tile0 = 0;
            // Ybar[w] += sum_{|u| <= top - |w|} A[u] G[u w]: a tile is P
            // consecutive outputs w (lane % P) times S = 64 / P lanes (lane / P)
            // taking the words u = s (mod S)
            uint32_t wm = 1;
            int64_t dbm = 0;
            for (int m = 1; m <= min(top, D - 1); ++m) {
                dbm += wm;
                wm *= w;
                const int P = pow2_cap64(wm), S = 64 / P, lgS = __builtin_ctz(uint32_t(S));
                const int ntiles = int((wm + uint32_t(P) - 1) / uint32_t(P));
                const int L = top - m;
                const int p = lane & (P - 1), s = lane / P;
                for (int t = first_tile(tile0, wave); t < tile0 + ntiles; t += kLvWaves) {
                    const uint32_t w_ = uint32_t(t - tile0) * uint32_t(P) + uint32_t(p);
                    const bool valid = w_ < wm;
                    double acc = 0.0;
                    if (valid) {
                        uint32_t nu = 1;
                        int64_t dbj = 0, dbjm = dbm;  // db[j], db[j + m]
                        for (int j = 0; j <= L; ++j) {
                            if (nu > uint32_t(s)) {
                                const uint32_t c = (nu - uint32_t(s) + uint32_t(S) - 1) >> lgS;
                                const double* Aj = A + dbj + s;
                                const double* Gj = Gs + dbjm + w_ + int64_t(s) * wm;
                                const int64_t gstep = int64_t(wm) << lgS;
#pragma unroll 4
                                for (uint32_t z = 0; z < c; ++z)
                                    acc += Aj[z << lgS] * Gj[int64_t(z) * gstep];
                            }
                            dbj += nu;
                            dbjm += int64_t(nu) * wm;
                            nu *= w;
                        }
                    }
                    for (int off = 32; off >= P; off >>= 1) acc += __shfl_xor(acc, off);
                    if (s == 0 && valid) Yb[dbm + w_] += acc;
                }
                tile0 += ntiles;
            }
            // G_{deg+1}[u] = sum_{1 <= |v| <= top - |u|} Y[v] G[u v]: a tile is
            // P outputs u (lane % P) times g = 64 / P lanes (lane / P) taking
            // the words v = r (mod g); each output's lanes start at a
            // different point of their sequence (rotated by the output's
            // lane), so the rows of one tile fall on different LDS banks
            wm = 1;
            dbm = 0;
            for (int m = 1; m <= D - deg; ++m) {
                dbm += wm;
                wm *= w;
                const int L = top - m;
                const int P = pow2_cap64(wm), g = 64 / P, lg = __builtin_ctz(uint32_t(g));
                const int ntiles = int((wm + uint32_t(P) - 1) / uint32_t(P));
                const int q = lane & (P - 1), r = lane / P;
                for (int t = first_tile(tile0, wave); t < tile0 + ntiles; t += kLvWaves) {
                    const uint32_t u = uint32_t(t - tile0) * uint32_t(P) + uint32_t(q);
                    const bool valid = u < wm;
                    double acc = 0.0;
                    if (valid) {
                        uint32_t nv = w;
                        int64_t dbk = 1, dbmk = dbm + wm;  // db[k], db[m + k]
                        for (int k = 1; k <= L; ++k) {
                            if (nv > uint32_t(r)) {
                                const uint32_t c = (nv - uint32_t(r) + uint32_t(g) - 1) >> lg;
                                const double* Yk = Y + dbk + r;
                                const double* Gk = Gs + dbmk + int64_t(u) * nv + r;
                                uint32_t i = uint32_t(q) % c;
#pragma unroll 4
                                for (uint32_t z = 0; z < c; ++z) {
                                    acc += Yk[i << lg] * Gk[i << lg];
                                    i = (i + 1 == c) ? 0 : i + 1;
                                }
                            }
                            dbk += nv;
                            dbmk += int64_t(nv) * wm;
                            nv *= w;
                        }
                    }
                    for (int off = 32; off >= P; off >>= 1) acc += __shfl_xor(acc, off);
                    if (r == 0 && valid) Gd[dbm + u] = acc;
                }
                tile0 += ntiles;
            }
            lv_barrier<InLds>();
        }
        // output: Ybar below the top degree, the given cotangent on it (G0's
        // top degree is never overwritten), 0 on degree 0
        IO* o = out + e * os;
        for (int64_t i = tid; i < n; i += kLvThreads)
            o[i] = i == 0 ? IO(0) : IO(i < nl ? Yb[i] : G0[i]);
        lv_barrier<InLds>();
    }
}

template <typename IO>
int launch_log_vjp(void* out, const void* arg, const void* cot, const DevBasis& b, int64_t batch,
                   int64_t os, int64_t as, int64_t cs, hipStream_t st)
{
    const Pow pw = make_pow(b.width, b.depth);
    const LvLayout lay = make_layout(b);
    const size_t lds = size_t(lay.size) * sizeof(double);
    int dev = 0, max_lds = 0, cus = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeSharedMemPerBlockOptin,
                                        dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (lds <= size_t(max_lds)) {
        // as many resident workgroups per CU as LDS allows, each looping over tensors
        const int per_cu = std::max(1, int(size_t(max_lds) / lds));
        const int64_t grid = std::min<int64_t>(batch, int64_t(cus) * per_cu);
        RPY_HIP_CHECK(hipFuncSetAttribute((const void*)ft_log_vjp_kernel<IO, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        hipLaunchKernelGGL((ft_log_vjp_kernel<IO, true>), dim3(uint32_t(grid)),
                           dim3(kLvThreads), lds, st, (IO*)out, (const IO*)arg, (const IO*)cot, b,
                           pw, lay, batch, os, as, cs, (double*)nullptr);
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    const int64_t grid = std::min<int64_t>(batch, int64_t(cus) * 2);
    double* ws = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&ws, size_t(grid) * lds, st));
    hipLaunchKernelGGL((ft_log_vjp_kernel<IO, false>), dim3(uint32_t(grid)), dim3(kLvThreads), 0,
                       st, (IO*)out, (const IO*)arg, (const IO*)cot, b, pw, lay, batch, os, as, cs,
                       ws);
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
