// gram.hip -- signature-kernel Gram matrix G = S S^T on MFMA.
//
// The signature kernel by signature dot is <S_i, S_j> = np.inner of dense
// truncated signatures (examples/signature-kernel-by-signature-dot.py:15-23).
// For n signatures this is one GEMM.  Each float32 entry is split as
// s = hi + lo with hi = bf16(s), lo = bf16(s - hi), and
//   G ~= Hi Hi^T + Hi Lo^T + Lo Hi^T
// is accumulated in float32 with v_mfma_f32_32x32x16_bf16 (the Lo Lo^T term
// is below float32 resolution).
//
// Tiling.  G is symmetric: only the upper-triangle 256 x 256 output tiles
// are computed.  The float32 signatures are split once into bf16 hi/lo
// copies (k padded to 32).  A workgroup (8 waves, each a 64 x 128 block =
// 2 x 4 MFMA blocks) owns one tile and one slice of k; per 32-wide k step
// the four operand tiles are staged in LDS, the next step's rows loading
// into registers while the current step's 48 MFMAs per wave run.  When the tiles alone do
// not fill the chip, k is split; the slices' partial tiles go to a workspace
// and a second kernel sums them in a fixed order (deterministic) and writes
// the tile and its mirror (through an LDS transpose, so both writes
// coalesce).  With one slice the tile kernel writes G itself.
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

constexpr int kBT = 256;             // output tile edge
constexpr int kBK = 32;              // k per LDS stage
constexpr int kRow = kBK + 8;        // LDS row stride in bf16 (80 B: 16 B reads conflict-free)
constexpr int kMat = kBT * kRow;     // one operand tile in LDS (bf16 elements)
constexpr size_t kLdsBytes = size_t(4) * kMat * sizeof(uint16_t);  // A hi, A lo, B hi, B lo: 80 KiB
constexpr int kThreads = 512;        // 8 waves: 4 row quarters x 2 column halves
constexpr int kChunks = 4 * kBT * kBK / 8 / kThreads;   // 16-byte chunks per thread per stage (8)
constexpr int64_t kTileElems = int64_t(kBT) * kBT;

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
// grid.y walks rows, grid.x and the block walk columns.
__global__ void split_kernel(const float* __restrict__ S, int64_t n, int64_t k, int64_t ld,
                             int64_t kp, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
    for (int64_t r = blockIdx.y; r < n; r += gridDim.y) {
        for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < kp;
             c += int64_t(gridDim.x) * blockDim.x) {
            const float v = c < k ? S[r * ld + c] : 0.0f;
            const uint16_t h = to_bf16_bits(v);
            hi[r * kp + c] = h;
            lo[r * kp + c] = to_bf16_bits(v - bf16_bits_to_float(h));
        }
    }
}

RPY_FI bf16x8 lds8(const uint16_t* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// upper-triangle tile t -> (bi, bj), bi <= bj, row-major over the triangle
RPY_FI void tile_coords(int t, int tiles, int& bi, int& bj)
{
    int i = 0;
    while (t >= tiles - i) {
        t -= tiles - i;
        ++i;
    }
    bi = i;
    bj = i + t;
}

// One workgroup = one upper-triangle tile and one k slice.  Wave (wr, wc)
// owns rows 64 wr .. 64 wr + 63 and columns 128 wc .. 128 wc + 127 of the
// tile: 2 x 4 MFMA blocks.  The next stage's bf16 hi/lo rows are loaded
// into registers while the current stage's MFMAs run.
template <bool Direct>
__global__ __launch_bounds__(kThreads) void gram_tile_kernel(
        const uint16_t* __restrict__ hi, const uint16_t* __restrict__ lo, int64_t n, int64_t kp,
        int tiles, int ntiles, int splits, float* __restrict__ out, int64_t ldg)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint16_t* sm = reinterpret_cast<uint16_t*>(smem_raw);
    // the linear block id walks the k slices fastest, so the slices of one
    // tile start together
    const int bid = blockIdx.x;
    const int split = bid % splits;
    const int t = bid / splits;
    int bi, bj;
    tile_coords(t, tiles, bi, bj);
    const int tid = threadIdx.x;
    const int wave = tid / 64;
    const int lane = tid % 64;
    const int wr = wave / 2, wc = wave % 2;
    // on a diagonal tile the blocks wholly below the diagonal are mirrors
    const bool idle = (bi == bj) && (wr * 64 >= wc * 128 + 128);

    const int64_t steps = kp / kBK;
    const int64_t s0 = steps * split / splits;
    const int64_t s1 = steps * (split + 1) / splits;

    // chunk c of this thread: id = tid + kThreads c; matrix id / 1024 (A hi,
    // A lo, B hi, B lo), row (id % 1024) / 4, 8-element k chunk id % 4
    uint4 regs[kChunks];
    auto load_stage = [&](int64_t step) {
        const int64_t k0 = step * kBK;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int id = tid + kThreads * c;
            const int m = id / (kBT * 4);
            const int row = (id % (kBT * 4)) / 4;
            const int kc = id % 4;
            const uint16_t* src = (m % 2 == 0) ? hi : lo;
            const int64_t grow = int64_t(m < 2 ? bi : bj) * kBT + row;
            regs[c] = grow < n ? *reinterpret_cast<const uint4*>(src + grow * kp + k0 + kc * 8)
                               : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto store_stage = [&]() {
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int id = tid + kThreads * c;
            const int m = id / (kBT * 4);
            const int row = (id % (kBT * 4)) / 4;
            const int kc = id % 4;
            *reinterpret_cast<uint4*>(sm + m * kMat + row * kRow + kc * 8) = regs[c];
        }
    };

    f32x16 acc[2][4] = {};
    if (s0 < s1) load_stage(s0);
    for (int64_t step = s0; step < s1; ++step) {
        __syncthreads();  // every wave is done with the previous stage's LDS
        store_stage();
        __syncthreads();
        if (step + 1 < s1) load_stage(step + 1);
        if (!idle) {
#pragma unroll
            for (int kk = 0; kk < kBK; kk += 16) {
                const int koff = kk + 8 * (lane / 32);
                bf16x8 ah[2], al[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int arow = wr * 64 + r * 32 + lane % 32;
                    ah[r] = lds8(sm + 0 * kMat + arow * kRow + koff);
                    al[r] = lds8(sm + 1 * kMat + arow * kRow + koff);
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int brow = wc * 128 + c * 32 + lane % 32;
                    const bf16x8 bh = lds8(sm + 2 * kMat + brow * kRow + koff);
                    const bf16x8 bl = lds8(sm + 3 * kMat + brow * kRow + koff);
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bh, acc[r][c],
                                                                             0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bl, acc[r][c],
                                                                             0, 0, 0);
                        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[r], bh, acc[r][c],
                                                                             0, 0, 0);
                    }
                }
            }
        }
    }
    if (idle) return;

    // epilogue
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int li = wr * 64 + r * 32 + 8 * (e / 4) + 4 * (lane / 32) + (e % 4);
                const int lj = wc * 128 + c * 32 + lane % 32;
                const float v = acc[r][c][e];
                if (Direct) {
                    const int64_t i = int64_t(bi) * kBT + li;
                    const int64_t j = int64_t(bj) * kBT + lj;
                    // on diagonal tiles only i <= j is kept, so G is exactly symmetric
                    if (i < n && j < n && (bi != bj || i <= j)) {
                        out[i * ldg + j] = v;
                        out[j * ldg + i] = v;
                    }
                } else {
                    out[(int64_t(split) * ntiles + t) * kTileElems + li * kBT + lj] = v;
                }
            }
}

// Sums the k slices of one 32 x 32 block of an upper-triangle tile and
// writes it and its mirror.  grid: (tiles, 64 blocks per tile).
__global__ __launch_bounds__(256) void gram_reduce_kernel(const float* __restrict__ ws,
                                                          int tiles, int ntiles, int splits,
                                                          int64_t n, float* __restrict__ G,
                                                          int64_t ldg)
{
    __shared__ float blk[32][33];
    const int t = blockIdx.x;
    int bi, bj;
    tile_coords(t, tiles, bi, bj);
    const int br = blockIdx.y / 8, bc = blockIdx.y % 8;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    const bool skip = bi == bj && br > bc;  // a block wholly below the diagonal
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = ty + 8 * q;
        const int li = br * 32 + r, lj = bc * 32 + tx;
        float v = 0.0f;
        if (!skip)
            for (int s = 0; s < splits; ++s)
                v += ws[(int64_t(s) * ntiles + t) * kTileElems + li * kBT + lj];
        blk[r][tx] = v;
        const int64_t i = int64_t(bi) * kBT + li, j = int64_t(bj) * kBT + lj;
        if (!skip && i < n && j < n && (bi != bj || i <= j)) G[i * ldg + j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = ty + 8 * q;  // column of the block = row of the mirror
        const int li = br * 32 + tx, lj = bc * 32 + c;
        const int64_t i = int64_t(bi) * kBT + li, j = int64_t(bj) * kBT + lj;
        // mirror element G[j][i] = block[tx][c]
        if (!skip && i < n && j < n && (bi != bj || i < j)) G[j * ldg + i] = blk[tx][c];
    }
}

}  // namespace

// The tile kernel's 80 KiB of dynamic LDS is above the default limit; the
// attribute is set once per process and kernel.
template <bool Direct>
hipError_t reserve_lds()
{
    static const hipError_t err = hipFuncSetAttribute(
            (const void*)gram_tile_kernel<Direct>, hipFuncAttributeMaxDynamicSharedMemorySize,
            int(kLdsBytes));
    return err;
}

int gram_bf16x3(float* G, const float* S, int64_t n, int64_t k, int64_t lds, int64_t ldg,
                void* stream)
{
    if (n < 0 || k < 0) return fail(RPY_ERR_SHAPE, "negative size");
    if (n == 0) return RPY_OK;
    if (!G || !S) return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    if (lds < k || ldg < n) return fail(RPY_ERR_SHAPE, "leading dimension too small");
    const int64_t tiles64 = (n + kBT - 1) / kBT;
    const int64_t ntiles64 = tiles64 * (tiles64 + 1) / 2;
    if (ntiles64 > 0x7fffffffLL / 64) return fail(RPY_ERR_SHAPE, "Gram matrix too large");
    const int tiles = int(tiles64);
    const int ntiles = int(ntiles64);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t kp = std::max<int64_t>(kBK, (k + kBK - 1) / kBK * kBK);
    const int64_t steps = kp / kBK;
    // one workgroup per CU (80 KiB of LDS each): split k so the grid fills
    // the 256 CUs about once, keeping at least 8 k steps per slice
    int splits = 1;
    if (ntiles < 256) splits = int(std::min<int64_t>(256 / ntiles, std::max<int64_t>(1, steps / 8)));
    splits = std::max(1, std::min(splits, 64));

    // one allocation: hi | lo | slice workspace
    const size_t split_bytes = size_t(2 * n * kp) * sizeof(uint16_t);
    const size_t ws_bytes =
            splits > 1 ? size_t(splits) * size_t(ntiles) * kTileElems * sizeof(float) : 0;
    unsigned char* buf = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&buf, split_bytes + ws_bytes, st));
    uint16_t* hi = reinterpret_cast<uint16_t*>(buf);
    uint16_t* lo = hi + n * kp;
    float* ws = splits > 1 ? reinterpret_cast<float*>(buf + split_bytes) : nullptr;

    int rc = RPY_OK;
    const dim3 sgrid = dim3(uint32_t(std::min<int64_t>((kp + 255) / 256, 64)),
                            uint32_t(std::min<int64_t>(n, 65535)));
    hipLaunchKernelGGL(split_kernel, sgrid, dim3(256), 0, st, S, n, k, lds, kp, hi, lo);
    hipError_t err = hipGetLastError();
    const dim3 grid = dim3(uint32_t(ntiles) * uint32_t(splits));
    if (err == hipSuccess) {
        if (splits == 1) {
            err = reserve_lds<true>();
            if (err == hipSuccess)
                hipLaunchKernelGGL(gram_tile_kernel<true>, grid, dim3(kThreads), kLdsBytes, st,
                                   hi, lo, n, kp, tiles, ntiles, 1, G, ldg);
        } else {
            err = reserve_lds<false>();
            if (err == hipSuccess) {
                hipLaunchKernelGGL(gram_tile_kernel<false>, grid, dim3(kThreads), kLdsBytes, st,
                                   hi, lo, n, kp, tiles, ntiles, splits, ws, int64_t(0));
                err = hipGetLastError();
                if (err == hipSuccess)
                    hipLaunchKernelGGL(gram_reduce_kernel, dim3(uint32_t(ntiles), 64), dim3(256),
                                       0, st, ws, tiles, ntiles, splits, n, G, ldg);
            }
        }
    }
    if (err == hipSuccess) err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "gram launch failed: %s", hipGetErrorString(err));
    hipError_t ferr = hipFreeAsync(buf, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

}  // namespace rpy_amd
