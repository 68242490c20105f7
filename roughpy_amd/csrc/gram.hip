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
// are computed.  A workgroup (8 waves, each a 64 x 128 block = 2 x 4 MFMA
// blocks) owns one tile and one slice of k.  Per 32-wide k step it reads
// the float32 rows of its two row blocks and splits each value into bf16
// hi/lo on the way into LDS (the four operand tiles A hi, A lo, B hi, B lo:
// 80 KiB).  LDS holds two such stages: while the MFMAs of step s read one,
// step s + 1 is split into the other and step s + 2 loads into registers,
// with one barrier per step.  Workgroups are ordered so that the ones on
// one XCD (one L2) read the same k range of the same row blocks.  When the
// tiles alone do not fill the chip, k is split; the slices' partial tiles
// go to a workspace and a second kernel sums them in a fixed order
// (deterministic) and writes the tile and its mirror (through an LDS
// transpose, so both writes coalesce).  With one slice the tile kernel
// writes G itself and nothing is allocated.
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
// four floats at 4-byte alignment
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

constexpr int kBT = 256;             // output tile edge
constexpr int kBK = 32;              // k per LDS stage
constexpr int kRow = kBK + 8;        // LDS row stride in bf16 (80 B: 16 B reads conflict-free)
constexpr int kMat = kBT * kRow;     // one operand tile in LDS (bf16 elements)
constexpr int kStage = 4 * kMat;     // A hi, A lo, B hi, B lo of one k stage: 80 KiB
constexpr size_t kLdsBytes = size_t(2) * kStage * sizeof(uint16_t);  // two stages: 160 KiB
constexpr int kThreads = 512;        // 8 waves: 4 row quarters x 2 column halves
constexpr int kLoads = 2 * kBT * kBK / 8 / kThreads;   // 8-float row pieces per thread per stage (4)
constexpr int64_t kTileElems = int64_t(kBT) * kBT;

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Splits 8 floats into bf16 hi = rne(v) and lo = rne(v - hi), packed as
// 8 consecutive bf16 each (the hardware conversion rounds to nearest even,
// two values per instruction).
RPY_FI void split8(const float (&v)[8], uint4& hi, uint4& lo)
{
    uint32_t h[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const f32x2 x = {v[2 * i], v[2 * i + 1]};
        const bf16x2 hb = __builtin_convertvector(x, bf16x2);
        const bf16x2 lb = __builtin_convertvector(x - __builtin_convertvector(hb, f32x2), bf16x2);
        h[i] = __builtin_bit_cast(uint32_t, hb);
        l[i] = __builtin_bit_cast(uint32_t, lb);
    }
    hi = make_uint4(h[0], h[1], h[2], h[3]);
    lo = make_uint4(l[0], l[1], l[2], l[3]);
}

// Workgroup -> (tile, k slice).  Consecutive workgroup ids go to the XCDs in
// turn (id % 8), and each XCD has an L2 of its own; the logical index gives
// each XCD a contiguous run of (slice-major) work, so the workgroups sharing
// an L2 read the same k range of the same row blocks.
constexpr int kXcds = 8;

RPY_FI int logical_block(int bid, int nblocks)
{
    const int per = (nblocks + kXcds - 1) / kXcds;
    const int xcd = bid % kXcds;
    const int slot = bid / kXcds;
    // XCDs whose runs are full come first; the last runs may be one short
    const int full = nblocks - (per - 1) * kXcds;  // XCDs with `per` blocks
    if (xcd < full) return xcd * per + slot;
    return full * per + (xcd - full) * (per - 1) + slot;
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
// tile: 2 x 4 MFMA blocks.  The next stage's float32 rows are loaded into
// registers while the current stage's MFMAs run.
template <bool Direct>
__global__ __launch_bounds__(kThreads) void gram_tile_kernel(
        const float* __restrict__ S, int64_t n, int64_t k, int64_t lds, int tiles, int ntiles,
        int splits, float* __restrict__ out, int64_t ldg)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint16_t* sm = reinterpret_cast<uint16_t*>(smem_raw);
    // slice-major logical order, contiguous per XCD
    const int lb = logical_block(int(blockIdx.x), int(gridDim.x));
    const int split = lb / ntiles;
    const int t = lb % ntiles;
    int bi, bj;
    tile_coords(t, tiles, bi, bj);
    const int tid = threadIdx.x;
    const int wave = tid / 64;
    const int lane = tid % 64;
    const int wr = wave / 2, wc = wave % 2;
    // on a diagonal tile the blocks wholly below the diagonal are mirrors,
    // and the B rows are the A rows (read once, staged twice)
    const bool diag = bi == bj;
    const bool idle = diag && (wr * 64 >= wc * 128 + 128);

    const int64_t steps = (k + kBK - 1) / kBK;
    const int64_t s0 = steps * split / splits;
    const int64_t s1 = steps * (split + 1) / splits;

    // this thread's rows: q = 2 m + h is row rr + 128 h of block m (A, B);
    // its 8 columns of the stage are 4 kc .. 4 kc + 3 and 16 + 4 kc .. 16 +
    // 4 kc + 3, so the four threads of a row read two contiguous 64-byte
    // runs.  The columns sit in LDS slot kc in that order; A and B use the
    // same order, so the k sum is unchanged.
    const int kc = tid % 4;
    const int rr = tid / 4;
    const float* rowp[kLoads];
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
        const int64_t grow = int64_t(q < 2 ? bi : bj) * kBT + rr + 128 * (q % 2);
        rowp[q] = grow < n ? S + grow * lds + 4 * kc : nullptr;
    }
    float regs[kLoads][8];
    auto load_stage = [&](int64_t step) {
        const int64_t k0 = step * kBK;
        const bool full = k0 + kBK <= k;
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            if (diag && q >= 2) continue;
            if (!rowp[q]) {
#pragma unroll
                for (int i = 0; i < 8; ++i) regs[q][i] = 0.0f;
            } else if (full) {
                // rows of S are 4-byte aligned only: 16-byte loads at 4-byte alignment
                const f32x4u a = *reinterpret_cast<const f32x4u*>(rowp[q] + k0);
                const f32x4u b = *reinterpret_cast<const f32x4u*>(rowp[q] + k0 + 16);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    regs[q][i] = a[i];
                    regs[q][4 + i] = b[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int64_t c = k0 + 4 * kc + (i % 4) + 16 * (i / 4);
                    regs[q][i] = c < k ? rowp[q][k0 + (i % 4) + 16 * (i / 4)] : 0.0f;
                }
            }
        }
    };
    auto store_stage = [&](uint16_t* buf) {
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            const int m = q / 2;
            const int row = rr + 128 * (q % 2);
            uint4 h, l;
            split8(regs[diag ? q % 2 : q], h, l);
            *reinterpret_cast<uint4*>(buf + (2 * m) * kMat + row * kRow + kc * 8) = h;
            *reinterpret_cast<uint4*>(buf + (2 * m + 1) * kMat + row * kRow + kc * 8) = l;
        }
    };

    f32x16 acc[2][4] = {};
    // the MFMAs of one 16-wide k half of a stage
    auto mfma_half = [&](const uint16_t* buf, int kk) {
        const int koff = kk + 8 * (lane / 32);
        bf16x8 ah[2], al[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int arow = wr * 64 + r * 32 + lane % 32;
            ah[r] = lds8(buf + 0 * kMat + arow * kRow + koff);
            al[r] = lds8(buf + 1 * kMat + arow * kRow + koff);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int brow = wc * 128 + c * 32 + lane % 32;
            const bf16x8 bh = lds8(buf + 2 * kMat + brow * kRow + koff);
            const bf16x8 bl = lds8(buf + 3 * kMat + brow * kRow + koff);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bh, acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[r], bl, acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[r], bh, acc[r][c], 0, 0, 0);
            }
        }
    };

    // Two LDS stages: step s computes on stage (s - s0) % 2 while its other
    // stage is filled with step s + 1 (split from the registers loaded one
    // step earlier) and step s + 2 loads into the registers.  One barrier
    // per step: it orders this step's reads of one stage before the next
    // step's writes to it, and this step's writes before the next step's
    // reads.
    if (s0 < s1) {
        load_stage(s0);
        store_stage(sm);
        if (s0 + 1 < s1) load_stage(s0 + 1);
        __syncthreads();
    }
    for (int64_t step = s0; step < s1; ++step) {
        const uint16_t* cur = sm + ((step - s0) & 1) * kStage;
        uint16_t* nxt = sm + ((step - s0 + 1) & 1) * kStage;
        if (!idle) mfma_half(cur, 0);
        if (step + 1 < s1) {
            store_stage(nxt);
            if (step + 2 < s1) load_stage(step + 2);
        }
        if (!idle) mfma_half(cur, 16);
        __syncthreads();
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
// writes it and its mirror.  grid: (tiles, 64 blocks per tile).  A thread
// owns 4 consecutive columns of one row; its slice loads are issued four at
// a time and added in slice order 0, 1, ... (the sum is the same whatever
// the timing).
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
    const bool skip = bi == bj && br > bc;  // a block wholly below the diagonal
    const int r = threadIdx.x / 8;          // row of the block
    const int c4 = 4 * (threadIdx.x % 8);   // first of this thread's 4 columns
    const int li = br * 32 + r, lj = bc * 32 + c4;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!skip) {
        const float* p = ws + int64_t(t) * kTileElems + li * kBT + lj;
        const int64_t slice = int64_t(ntiles) * kTileElems;
        int s = 0;
        for (; s + 4 <= splits; s += 4) {
            float4 a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = *reinterpret_cast<const float4*>(p + (s + u) * slice);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[0] += a[u].x;
                v[1] += a[u].y;
                v[2] += a[u].z;
                v[3] += a[u].w;
            }
        }
        for (; s < splits; ++s) {
            const float4 a = *reinterpret_cast<const float4*>(p + s * slice);
            v[0] += a.x;
            v[1] += a.y;
            v[2] += a.z;
            v[3] += a.w;
        }
    }
    const int64_t i = int64_t(bi) * kBT + li;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        blk[r][c4 + e] = v[e];
        const int64_t j = int64_t(bj) * kBT + lj + e;
        if (!skip && i < n && j < n && (bi != bj || i <= j)) G[i * ldg + j] = v[e];
    }
    __syncthreads();
    // mirror: row c of the block's transpose, columns r4 .. r4 + 3
    const int c = threadIdx.x / 8;
    const int r4 = 4 * (threadIdx.x % 8);
    const int64_t mj = int64_t(bj) * kBT + bc * 32 + c;  // column index of G = mirror row
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t mi = int64_t(bi) * kBT + br * 32 + r4 + e;
        // mirror element G[mj][mi] = block[r4 + e][c]
        if (!skip && mi < n && mj < n && (bi != bj || mi < mj)) G[mj * ldg + mi] = blk[r4 + e][c];
    }
}

}  // namespace

// The tile kernel's 160 KiB of dynamic LDS is above the default limit; the
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
    const int64_t steps = std::max<int64_t>(1, (k + kBK - 1) / kBK);
    // one workgroup per CU (160 KiB of LDS each): split k so the grid fills
    // the 256 CUs about once, keeping at least 8 k steps per slice
    int splits = 1;
    if (ntiles < 256) splits = int(std::min<int64_t>(256 / ntiles, std::max<int64_t>(1, steps / 8)));
    splits = std::max(1, std::min(splits, 64));

    const dim3 grid = dim3(uint32_t(ntiles) * uint32_t(splits));
    int rc = RPY_OK;
    hipError_t err = hipSuccess;
    if (splits == 1) {
        err = reserve_lds<true>();
        if (err == hipSuccess)
            hipLaunchKernelGGL(gram_tile_kernel<true>, grid, dim3(kThreads), kLdsBytes, st, S, n,
                               k, lds, tiles, ntiles, 1, G, ldg);
        if (err == hipSuccess) err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "gram launch failed: %s", hipGetErrorString(err));
        return RPY_OK;
    }
    // the slices' partial tiles
    float* ws = nullptr;
    RPY_HIP_CHECK(workspace_alloc((void**)&ws,
                                  size_t(splits) * size_t(ntiles) * kTileElems * sizeof(float), st));
    err = reserve_lds<false>();
    if (err == hipSuccess) {
        hipLaunchKernelGGL(gram_tile_kernel<false>, grid, dim3(kThreads), kLdsBytes, st, S, n, k,
                           lds, tiles, ntiles, splits, ws, int64_t(0));
        err = hipGetLastError();
        if (err == hipSuccess)
            hipLaunchKernelGGL(gram_reduce_kernel, dim3(uint32_t(ntiles), 64), dim3(256), 0, st,
                               ws, tiles, ntiles, splits, n, G, ldg);
    }
    if (err == hipSuccess) err = hipGetLastError();
    if (err != hipSuccess) rc = fail(RPY_ERR_HIP, "gram launch failed: %s", hipGetErrorString(err));
    hipError_t ferr = hipFreeAsync(ws, st);
    if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    return rc;
}

}  // namespace rpy_amd
