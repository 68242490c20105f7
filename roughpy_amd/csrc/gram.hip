// gram.hip -- signature-kernel Gram matrix G = S S^T on MFMA.
//
// The signature kernel by signature dot is <S_i, S_j> = np.inner of dense
// truncated signatures (examples/signature-kernel-by-signature-dot.py:15-23).
// For n signatures this is one GEMM.  Each float32 entry is split as
// s = hi + lo with hi = bf16(s), lo = bf16(s - hi), and
//   G ~= Hi Hi^T + Hi Lo^T + Lo Hi^T
// is accumulated in float32 with v_mfma_f32_32x32x16_bf16 (the Lo Lo^T term
// is below float32 resolution).  G is symmetric, so only upper-triangle
// 64x64 tiles are computed and mirrored.
//
// MFMA 32x32x16 bf16 operand layout (wave64): lane l holds 8 consecutive k
// of row (l % 32): A[l % 32][8 (l / 32) + 0..7], and likewise for B^T; the
// accumulator register r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "chen.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 64;   // output tile per workgroup (2 x 2 waves of 32 x 32)
constexpr int kK = 16;      // k per MFMA

// bf16 round-to-nearest-even of a float, as bits
RPY_FI uint16_t to_bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

RPY_FI float bf16_bits_to_float(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// hi/lo split into row-major (n x kp) bf16 arrays, zero padded to kp.
__global__ void split_kernel(const float* __restrict__ S, int64_t n, int64_t k, int64_t ld,
                             int64_t kp, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
    const int64_t total = n * kp;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t r = t / kp;
        const int64_t c = t - r * kp;
        const float v = c < k ? S[r * ld + c] : 0.0f;
        const uint16_t h = to_bf16_bits(v);
        hi[t] = h;
        lo[t] = to_bf16_bits(v - bf16_bits_to_float(h));
    }
}

RPY_FI bf16x8 load8(const uint16_t* p)
{
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

// One workgroup = one upper-triangle 64x64 tile; wave (wr, wc) owns the
// 32x32 sub-tile at (32 wr, 32 wc).
__global__ __launch_bounds__(256) void gram_kernel(const uint16_t* __restrict__ hi,
                                                   const uint16_t* __restrict__ lo, int64_t n,
                                                   int64_t kp, float* __restrict__ G,
                                                   int64_t ldg, int tiles)
{
    // map the linear block id to (bi <= bj)
    int64_t t = blockIdx.x;
    int bi = 0;
    while (t >= tiles - bi) {
        t -= tiles - bi;
        ++bi;
    }
    const int bj = bi + int(t);
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    // the lower sub-tile of a diagonal tile is the mirror of the upper one
    if (bi == bj && wave == 2) return;
    const int64_t row0 = int64_t(bi) * kTile + 32 * (wave / 2);
    const int64_t col0 = int64_t(bj) * kTile + 32 * (wave % 2);
    const int64_t ra = row0 + lane % 32;  // A operand row of this lane
    const int64_t rb = col0 + lane % 32;  // B operand row (of S) of this lane
    const int koff = 8 * (lane / 32);
    const bool va = ra < n, vb = rb < n;
    const uint16_t* ah = hi + (va ? ra : 0) * kp + koff;
    const uint16_t* al = lo + (va ? ra : 0) * kp + koff;
    const uint16_t* bh = hi + (vb ? rb : 0) * kp + koff;
    const uint16_t* bl = lo + (vb ? rb : 0) * kp + koff;
    f32x16 acc = {};
    const bf16x8 zero = {};
    for (int64_t k = 0; k < kp; k += kK) {
        const bf16x8 a_h = va ? load8(ah + k) : zero;
        const bf16x8 a_l = va ? load8(al + k) : zero;
        const bf16x8 b_h = vb ? load8(bh + k) : zero;
        const bf16x8 b_l = vb ? load8(bl + k) : zero;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, b_h, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, b_l, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_l, b_h, acc, 0, 0, 0);
    }
    const int64_t j = col0 + lane % 32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t i = row0 + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        // on diagonal tiles only i <= j is kept, so G is exactly symmetric
        if (i < n && j < n && (bi != bj || i <= j)) {
            G[i * ldg + j] = acc[r];
            G[j * ldg + i] = acc[r];
        }
    }
}

}  // namespace

int gram_bf16x3(float* G, const float* S, int64_t n, int64_t k, int64_t lds, int64_t ldg,
                void* stream)
{
    if (n < 0 || k < 0) return fail(RPY_ERR_SHAPE, "negative size");
    if (n == 0) return RPY_OK;
    if (!G || !S) return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    if (lds < k || ldg < n) return fail(RPY_ERR_SHAPE, "leading dimension too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t kp = std::max<int64_t>(kK, (k + kK - 1) / kK * kK);
    uint16_t* buf = nullptr;
    RPY_HIP_CHECK(hipMallocAsync((void**)&buf, size_t(2 * n * kp) * sizeof(uint16_t), st));
    const int64_t total = n * kp;
    const uint32_t blocks = uint32_t(std::min<int64_t>((total + 255) / 256, 65536));
    hipLaunchKernelGGL(split_kernel, dim3(blocks), dim3(256), 0, st, S, n, k, lds, kp, buf,
                       buf + n * kp);
    int rc = RPY_OK;
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "split launch failed: %s", hipGetErrorString(err));
    if (!rc) {
        const int tiles = int((n + kTile - 1) / kTile);
        const int64_t nblocks = int64_t(tiles) * (tiles + 1) / 2;
        hipLaunchKernelGGL(gram_kernel, dim3(uint32_t(nblocks)), dim3(256), 0, st, buf, buf + n * kp,
                           n, kp, G, ldg, tiles);
        err = hipGetLastError();
        if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "gram launch failed: %s", hipGetErrorString(err));
    }
    hipError_t ferr = hipFreeAsync(buf, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

}  // namespace rpy_amd
