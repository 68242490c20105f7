// chen_block.hip -- block-compensated float32 signatures.
//
// The plain float32 signature kernel adds every Chen step's increment into
// running coefficients that are much larger than the increment, so each
// step loses the increment's low bits and the loss grows with the stream
// length.  The compensated mode splits the running product into blocks:
//
//   1. every stream is cut into at most kSegSlots segments of T rows (T
//      grows for very long streams so the slot count holds), and the
//      float32 signature kernel computes each segment's signature B_j from
//      the identity -- short products of small increments, so their
//      rounding error is small relative to B_j itself;
//   2. a merge kernel multiplies the segments of each stream in order,
//      S <- S (x) B_j, keeping S as an unevaluated float32 pair (hi, c).
//      The level-k coefficients of S (x) B_j are S_k + t with
//      t = B_k + sum_{0<i<k} S_i B_{k-i} (S_0 = B_0 = 1); t is formed in
//      float32 from hi (its size is that of the block's contribution), and
//      S_k absorbs it by Kahan's compensated summation, the low part that
//      hi loses at each rounding carried in a second float32 (once per
//      block and coefficient rather than once per step).
//
// The result, rounded to float, is Context::compute_signature
// (algebra/src/lite_context.h:545-563) of the stream, like the other modes.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "chen.hpp"
#include "chen_core.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kSegSlots = 16;  // segments per stream at most

// Segments of a stream of `len` rows for the nominal segment length T:
// count = ceil(len / max(T, ceil(len / kSegSlots))) segments of balanced
// lengths (segment j covers rows [j len / count, (j + 1) len / count)), so
// a stream's segments differ in length by one row at most and the lanes of
// a chunk that share a stream finish together.
RPY_HD inline int64_t seg_count(int64_t len, int64_t T)
{
    if (len <= 0) return 0;
    int64_t r = (len + kSegSlots - 1) / kSegSlots;
    if (r < T) r = T;
    return (len + r - 1) / r;
}

// Thread i = s * kSegSlots + j: the row range of segment j of stream s
// (empty past the stream's segment count); thread j == 0 also writes the
// stream's segment count.  Thread 0 clears the queue counter and the empty
// slot at index n_streams * kSegSlots that pads the queue.
__global__ void block_prep_kernel(const int64_t* __restrict__ begins,
                                  const int64_t* __restrict__ ends, int64_t n_streams, int64_t T,
                                  int64_t* __restrict__ seg_begin, int64_t* __restrict__ seg_end,
                                  int32_t* __restrict__ counts, uint32_t* __restrict__ counter)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t slots = n_streams * kSegSlots;
    if (i == 0) {
        *counter = 0u;
        seg_begin[slots] = 0;
        seg_end[slots] = 0;
    }
    if (i >= slots) return;
    const int64_t s = i / kSegSlots;
    const int64_t j = i - s * kSegSlots;
    const int64_t b = begins[s];
    const int64_t len = ends[s] - b;
    const int64_t n = seg_count(len, T);
    int64_t lo = b, hi = b;
    if (j < n) {
        lo = b + (j * len) / n;
        hi = b + ((j + 1) * len) / n;
    }
    seg_begin[i] = lo;
    seg_end[i] = hi;
    if (j == 0) counts[s] = int32_t(n);
}

// The work queue: the non-empty segments stream by stream (bases = the
// exclusive scan of the counts), then the padding slot.  Deterministic, and
// every empty entry comes after every non-empty one.
__global__ void block_queue_kernel(const int32_t* __restrict__ counts,
                                   const int32_t* __restrict__ bases, int64_t n_streams,
                                   int32_t* __restrict__ order)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t slots = n_streams * kSegSlots;
    if (i >= slots) return;
    const int64_t s = i / kSegSlots;
    const int64_t j = i - s * kSegSlots;
    if (j < counts[s]) order[bases[s] + j] = int32_t(i);
    const int64_t total = int64_t(bases[n_streams - 1]) + counts[n_streams - 1];
    if (i >= total) order[i] = int32_t(slots);
}

// Kahan's compensated addition hi += t: c carries the (negated) low part
// that hi lost at its last rounding, and is taken off the next term first.
// No term here is a product, so contraction into fused multiply-adds cannot
// change it.  The sum is hi - c.
RPY_FI void kahan_add(float& hi, float& c, float t)
{
    const float y = t - c;
    const float s = hi + y;
    c = (s - hi) - y;
    hi = s;
}

template <int W, int D, int NT>
struct MergeShape {
    static constexpr int db(int k)
    {
        int s = 0;
        for (int i = 0; i < k; ++i) s += int(ipow(W, i));
        return s;
    }
    static constexpr int LOW = db(D);                    // coefficients of levels 0..D-1
    static constexpr int TILES = int(ipow(W, D - 1));    // top-level tiles of W words
    static constexpr int TPT = (TILES + NT - 1) / NT;    // tiles per thread
    static constexpr int PRE = (LOW + NT - 1) / NT;      // low values each thread stages
    static constexpr int LOWP = PRE * NT;                // LDS rows, padded: stores need no test
    // low-level words (levels 1..D-1) per thread, level by level
    static constexpr int per_level(int k) { return (int(ipow(W, k)) + NT - 1) / NT; }
    // first register slot of level k's words (k = 1..D)
    static constexpr int low_off(int k)
    {
        int n = 0;
        for (int i = 1; i < k; ++i) n += per_level(i);
        return n;
    }
    static constexpr int LPT = low_off(D) > 0 ? low_off(D) : 1;
};

// One workgroup per stream: S = B_0 (x) B_1 (x) ... (x) B_{n-1} over the
// stream's segment signatures seg[s * kSegSlots + j], levels 0..D-1 of S
// (hi and c) and of the current B in LDS, the top level of S in registers
// (thread r owns the tiles r, r + NT, ... of W consecutive top words).
// NT is a multiple of W^(D-2) that divides the W^(D-1) tiles, so a tile's
// prefixes and suffixes move by constants from one of the thread's tiles to
// the next, and the loop has no bounds tests: low-level loads past level
// D-1 read the segment's own top level, LDS rows are padded to whole passes
// of the workgroup, and words past a level are written to a per-thread
// dummy slot.
template <int W, int D, int NT>
__global__ __launch_bounds__(NT) void block_merge_kernel(float* __restrict__ out,
                                                         const float* __restrict__ seg,
                                                         const int64_t* __restrict__ begins,
                                                         const int64_t* __restrict__ ends,
                                                         int64_t n_streams, int64_t T)
{
    using M = MergeShape<W, D, NT>;
    constexpr int LOW = M::LOW, TILES = M::TILES, TPT = M::TPT, PRE = M::PRE, LOWP = M::LOWP;
    constexpr int TSIZE = LOW + TILES * W;
    static_assert(TILES % NT == 0 && NT % int(ipow(W, D - 2)) == 0,
                  "the workgroup must tile the top level in whole passes");
    // + NT: one dummy slot per thread for the writes of words past a level,
    // and room for the reads of those words
    __shared__ float shi[LOWP + NT];
    __shared__ float sc[LOWP + NT];
    __shared__ float bl[LOWP + NT];

    const int tid = threadIdx.x;
    const int64_t s = blockIdx.x;
    if (s >= n_streams) return;
    const int64_t count = seg_count(ends[s] - begins[s], T);
    float* o = out + s * TSIZE;
    if (count == 0) {
        for (int i = tid; i < TSIZE; i += NT) o[i] = i == 0 ? 1.0f : 0.0f;
        return;
    }
    const float* segs = seg + s * int64_t(kSegSlots) * TSIZE;

    float hi[TPT][W], cc[TPT][W];
    float pre_low[PRE], pre_top[TPT][W];
    auto fetch = [&](const float* B) {
#pragma unroll
        for (int i = 0; i < PRE; ++i) pre_low[i] = B[tid + NT * i];
#pragma unroll
        for (int i = 0; i < TPT; ++i) {
            const float* bt = B + LOW + (tid + NT * i) * W;
#pragma unroll
            for (int c = 0; c < W; ++c) pre_top[i][c] = bt[c];
        }
    };

    // S = B_0
    fetch(segs);
#pragma unroll
    for (int i = 0; i < PRE; ++i) {
        shi[tid + NT * i] = pre_low[i];
        sc[tid + NT * i] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < TPT; ++i)
#pragma unroll
        for (int c = 0; c < W; ++c) {
            hi[i][c] = pre_top[i][c];
            cc[i][c] = 0.0f;
        }
    if (count > 1) fetch(segs + TSIZE);

    for (int64_t j = 1; j < count; ++j) {
        // B_j: low levels to LDS, top level kept in registers
#pragma unroll
        for (int i = 0; i < PRE; ++i) bl[tid + NT * i] = pre_low[i];
        __syncthreads();

        // Top level: t(p c) = B(p c) + sum_{0<l<D} S(p_<l) B(p_>=l c).  For a
        // split l, the thread's tiles ti = tid + NT i share the suffix index
        // tid mod W^(D-1-l) (NT is a multiple of it), so each split is an
        // outer product of TPT prefix values of S and W suffix values of B.
        float t[TPT][W];
#pragma unroll
        for (int i = 0; i < TPT; ++i)
#pragma unroll
            for (int c = 0; c < W; ++c) t[i][c] = pre_top[i][c];
        chen::static_for<D - 1>([&](auto l0) {
            constexpr int l = int(decltype(l0)::value) + 1;  // prefix length
            constexpr int R = D - 1 - l;                     // tile letters after it
            constexpr int PR = int(ipow(W, R));
            const int col = tid % PR;
            float bv[W];
#pragma unroll
            for (int c = 0; c < W; ++c) bv[c] = bl[M::db(R + 1) + col * W + c];
            const float* srow = shi + M::db(l) + tid / PR;
#pragma unroll
            for (int i = 0; i < TPT; ++i) {
                const float sv = srow[(NT / PR) * i];
#pragma unroll
                for (int c = 0; c < W; ++c) t[i][c] = __builtin_fmaf(sv, bv[c], t[i][c]);
            }
        });
#pragma unroll
        for (int i = 0; i < TPT; ++i)
#pragma unroll
            for (int c = 0; c < W; ++c) kahan_add(hi[i][c], cc[i][c], t[i][c]);
        // the next segment's loads fly while the low levels are done
        if (j + 1 < count) fetch(segs + (j + 1) * TSIZE);

        // levels 1..D-1, read before any of them is overwritten; the same
        // outer products over the thread's words w = tid + NT r of a level
        float nhi[M::LPT], nc[M::LPT];
        int nidx[M::LPT];
        chen::static_for<D - 1>([&](auto k0) {
            constexpr int k = int(decltype(k0)::value) + 1;
            constexpr int NK = int(ipow(W, k));
            constexpr int PK = M::per_level(k);
            float tk[PK];
#pragma unroll
            for (int r = 0; r < PK; ++r) tk[r] = bl[M::db(k) + tid + NT * r];
            chen::static_for<k - 1>([&](auto l0) {
                constexpr int l = int(decltype(l0)::value) + 1;
                constexpr int PR = int(ipow(W, k - l));
                const float bv = bl[M::db(k - l) + tid % PR];
                const float* srow = shi + M::db(l) + tid / PR;
#pragma unroll
                for (int r = 0; r < PK; ++r) tk[r] = __builtin_fmaf(srow[(NT / PR) * r], bv, tk[r]);
            });
#pragma unroll
            for (int r = 0; r < PK; ++r) {
                const int wi = tid + NT * r;
                const int at = M::db(k) + wi;
                float h = shi[at], c = sc[at];
                kahan_add(h, c, tk[r]);
                nhi[M::low_off(k) + r] = h;
                nc[M::low_off(k) + r] = c;
                nidx[M::low_off(k) + r] = wi < NK ? at : LOWP + tid;
            }
        });
        __syncthreads();
#pragma unroll
        for (int r = 0; r < M::low_off(D); ++r) {
            shi[nidx[r]] = nhi[r];
            sc[nidx[r]] = nc[r];
        }
    }
    __syncthreads();
    for (int v = tid; v < LOW; v += NT) o[v] = shi[v] - sc[v];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int ti = tid + NT * i;
        if (ti < TILES) {
#pragma unroll
            for (int c = 0; c < W; ++c) o[LOW + ti * W + c] = hi[i][c] - cc[i][c];
        }
    }
}

using MergeLauncher = void (*)(float* out, const float* seg, const int64_t* begins,
                               const int64_t* ends, int64_t n_streams, int64_t T, hipStream_t st);

template <int W, int D, int NT>
void launch_merge(float* out, const float* seg, const int64_t* begins, const int64_t* ends,
                  int64_t n_streams, int64_t T, hipStream_t st)
{
    hipLaunchKernelGGL((block_merge_kernel<W, D, NT>), dim3(uint32_t(n_streams)), dim3(NT), 0, st,
                       out, seg, begins, ends, n_streams, T);
}

struct MergeEntry {
    int32_t width, depth;
    MergeLauncher launch;
};

// workgroup sizes: multiples of W^(D-2) dividing the W^(D-1) top-level
// tiles (625 tiles in 5 passes of 125 at (5,5), 1296 in 3 of 432 at (6,5))
const MergeEntry kMerges[] = {
        {3, 4, &launch_merge<3, 4, 27>},  {4, 6, &launch_merge<4, 6, 256>},
        {5, 5, &launch_merge<5, 5, 125>}, {8, 4, &launch_merge<8, 4, 128>},
        {6, 5, &launch_merge<6, 5, 432>},
};

MergeLauncher find_merge(int32_t w, int32_t d)
{
    for (const MergeEntry& e : kMerges)
        if (e.width == w && e.depth == d) return e.launch;
    return nullptr;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

bool block_merge_supported(int32_t width, int32_t depth)
{
    return find_merge(width, depth) != nullptr;
}

int64_t block_segment_slots() { return kSegSlots; }

int block_plan(int64_t n_sub, int64_t tsize, BlockPlan& p, hipStream_t st)
{
    const size_t slots = size_t(n_sub) * kSegSlots;
    p.n_sub = n_sub;
    p.slots = int64_t(slots);
    p.scan_bytes = 0;
    RPY_HIP_CHECK(rocprim::exclusive_scan(nullptr, p.scan_bytes, (const int32_t*)nullptr,
                                          (int32_t*)nullptr, int32_t(0), size_t(n_sub),
                                          rocprim::plus<int32_t>(), st));
    p.c_bytes = align256(size_t(n_sub) * sizeof(int32_t));
    p.o_bytes = align256(slots * sizeof(int32_t));
    p.r_bytes = align256((slots + 1) * sizeof(int64_t));
    p.seg_bytes = align256(slots * size_t(tsize) * sizeof(float));
    p.total = 2 * p.c_bytes + p.o_bytes + 2 * p.r_bytes + 256 + align256(p.scan_bytes) +
              p.seg_bytes;
    return RPY_OK;
}

int block_run(float* out, const float* increments, const int64_t* begins, const int64_t* ends,
              int64_t n_streams, const DevBasis& b, int64_t T, SigLauncher seg_launch,
              const BlockPlan& p, char* ws, hipStream_t st)
{
    const MergeLauncher merge = find_merge(b.width, b.depth);
    if (!merge || !seg_launch)
        return fail(RPY_ERR_UNSUPPORTED, "no block-compensated kernels for width %d depth %d",
                    b.width, b.depth);
    if (T < 1) return fail(RPY_ERR_INVALID_ARGUMENT, "segment length must be positive");
    char* w = ws;
    int32_t* counts = (int32_t*)w;
    w += p.c_bytes;
    int32_t* bases = (int32_t*)w;
    w += p.c_bytes;
    int32_t* order = (int32_t*)w;
    w += p.o_bytes;
    int64_t* seg_begin = (int64_t*)w;
    w += p.r_bytes;
    int64_t* seg_end = (int64_t*)w;
    w += p.r_bytes;
    uint32_t* counter = (uint32_t*)w;
    w += 256;
    void* scan_tmp = w;
    w += align256(p.scan_bytes);
    float* segbuf = (float*)w;
    const int64_t tsize = b.size();
    for (int64_t s0 = 0; s0 < n_streams; s0 += p.n_sub) {
        const int64_t n = std::min<int64_t>(p.n_sub, n_streams - s0);
        const int64_t slots = n * kSegSlots;
        const int pb = 256;
        const dim3 grid(uint32_t((slots + pb - 1) / pb));
        hipLaunchKernelGGL(block_prep_kernel, grid, dim3(pb), 0, st, begins + s0, ends + s0, n, T,
                           seg_begin, seg_end, counts, counter);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "segment prep launch failed: %s", hipGetErrorString(err));
        size_t scan_bytes = p.scan_bytes;
        err = rocprim::exclusive_scan(scan_tmp, scan_bytes, counts, bases, int32_t(0), size_t(n),
                                      rocprim::plus<int32_t>(), st);
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "segment count scan failed: %s", hipGetErrorString(err));
        hipLaunchKernelGGL(block_queue_kernel, grid, dim3(pb), 0, st, counts, bases, n, order);
        err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "segment queue launch failed: %s", hipGetErrorString(err));
        int rc = seg_launch(segbuf, increments, seg_begin, seg_end, order, slots, b, counter,
                            kSkipEmpty | kSortedDesc, st);
        if (rc) return rc;
        merge(out + s0 * tsize, segbuf, begins + s0, ends + s0, n, T, st);
        err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "segment merge launch failed: %s", hipGetErrorString(err));
    }
    return RPY_OK;
}

}  // namespace rpy_amd
