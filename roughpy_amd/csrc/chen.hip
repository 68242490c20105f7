// chen.hip -- batched Chen signatures of uneven piecewise-linear streams.
//
// One persistent launch: streams are sorted by length (longest first), every
// stream is split across W^P lanes (chen_core.hpp), and each wave64 pulls the
// next 64 (stream, prefix) items from a global work counter until the queue
// is empty.  Each lane keeps its share of the running signature in
// registers for the whole stream and writes it once at the end, so the only
// HBM traffic is the increments in and the signatures out.
//
// Reference: Context::compute_signature, algebra/src/lite_context.h:545-563
// (result = 1; result.fmexp_inplace(l2t(row)) per increment), with the fmexp
// Horner of roughpy_compute/dense/intermediate/free_tensor_fmexp.hpp:15-40.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "chen.hpp"
#include "chen_core.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kWave = 64;

// keys[i] = length of stream i, vals[i] = i; also clears the queue counter.
// Stream i owns rows [begins[i], ends[i]) (CSR offsets: ends = offsets + 1).
__global__ void prep_lengths_kernel(const int64_t* __restrict__ begins,
                                    const int64_t* __restrict__ ends, int64_t n,
                                    uint32_t* __restrict__ keys, int32_t* __restrict__ vals,
                                    uint32_t* __restrict__ counter)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0u;
    if (i < n) {
        const int64_t len = ends[i] - begins[i];
        keys[i] = uint32_t(len < 0 ? 0 : (len > 0xffffffffLL ? 0xffffffffLL : len));
        vals[i] = int32_t(i);
    }
}

// Increment staging.  A chunk of 64 consecutive (stream, prefix) items spans
// at most kSlots streams; their rows are copied block by block (kBlk steps)
// into LDS with coalesced loads, the next block being fetched into registers
// while the current one is consumed, so the step loop reads only LDS.
template <typename T, int W, int NP>
struct Staging {
    static constexpr int kSlots = (kWave + NP - 2) / NP + 1;
    // at most 3 prefetched values per lane: the lane state owns the VGPRs
    static constexpr int kMaxPerLane = 3;
    static constexpr int kBlkFit = kMaxPerLane * kWave / (kSlots * W);
    static constexpr int kBlk = kBlkFit < 32 ? (kBlkFit < 1 ? 1 : kBlkFit) : 32;
    static constexpr int kVals = kSlots * kBlk * W;        // raw values per block
    static constexpr int kPerLane = (kVals + kWave - 1) / kWave;
};

// T: arithmetic type; IO: type of the increments and signatures in memory
// (float data with double arithmetic is the compensated float32 mode).
template <typename T, int W, int D, int P, typename IO = T>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2)))
void chen_persistent_kernel(
        IO* __restrict__ out, const IO* __restrict__ incr, const int64_t* __restrict__ begins,
        const int64_t* __restrict__ ends, const int32_t* __restrict__ order, int64_t n_items,
        int64_t tsize, DevBasis b, uint32_t* __restrict__ counter)
{
    using Sh = chen::Shape<W, D, P>;
    constexpr int NP = Sh::NP;
    using St = Staging<T, W, NP>;
    constexpr int kSlots = St::kSlots;
    constexpr int kBlk = St::kBlk;
    constexpr int kBlkVals = kBlk * W;       // raw values of one slot's block
    // Scaled rows (x_a / d, d = 1..D) when a lane does at least 4 FMAs per
    // table value it reads; below that the LDS bandwidth (128 B/clk/CU) is
    // the limit and raw rows scaled in registers are cheaper.
    constexpr bool kScaled = chen::lane_fmas<W, D, P>() >= 4 * Sh::ROW;
    constexpr int kRow = kScaled ? Sh::ROW : W;

    // stage[(slot * kBlk + r) * kRow + (d-1) * W + a] = x_a(row r) / d
    __shared__ T stage[kSlots * kBlk * kRow];
    __shared__ int64_t slot_row0[kSlots];
    __shared__ int64_t slot_len[kSlots];
    __shared__ int32_t slot_stream[kSlots];

    const int lane = threadIdx.x;
    while (true) {
        uint32_t chunk = 0;
        if (lane == 0) chunk = atomicAdd(counter, 1u);
        chunk = __shfl(chunk, 0);
        const int64_t first_item = int64_t(chunk) * kWave;
        if (first_item >= n_items) break;
        const int64_t last_item = min(first_item + kWave, n_items) - 1;
        const int64_t si_first = first_item / NP;
        const int n_slots = int(last_item / NP - si_first) + 1;  // <= kSlots

        __syncthreads();  // previous chunk is done with the slot table
        if (lane < kSlots) {
            int64_t r0 = 0, ln = 0;
            int32_t s = 0;
            if (lane < n_slots) {
                s = order[si_first + lane];
                r0 = begins[s];
                ln = ends[s] - r0;
            }
            slot_row0[lane] = r0;
            slot_len[lane] = ln;
            slot_stream[lane] = s;
        }
        __syncthreads();

        const int64_t item = first_item + lane;
        const bool active = item < n_items;
        const int slot = active ? int(item / NP - si_first) : 0;
        const int q = active ? int(item % NP) : 0;
        const int len = active ? int(slot_len[slot]) : 0;
        // the chunk runs as long as its longest stream (with the length sort,
        // slot 0)
        int maxlen = 0;
        for (int i = 0; i < n_slots; ++i) maxlen = max(maxlen, int(slot_len[i]));

        int u[P];
        {
            int t = q;
            for (int j = P - 1; j >= 0; --j) {
                u[j] = t % W;
                t /= W;
            }
        }

        chen::Lane<T, W, D, P> st;
        st.init_identity();

        // staging assignment of this lane: values v = lane + 64 i
        T pre[St::kPerLane];
        auto fetch = [&](int t0) {
#pragma unroll
            for (int i = 0; i < St::kPerLane; ++i) {
                const int v = lane + kWave * i;
                const int sl = v / kBlkVals;
                const int rem = v - sl * kBlkVals;
                const int r = rem / W;
                T val = T(0);
                if (v < St::kVals && sl < n_slots && t0 + r < slot_len[sl])
                    val = T(incr[(slot_row0[sl] + t0) * W + rem]);
                pre[i] = val;
            }
        };

        fetch(0);
        for (int t0 = 0; t0 < maxlen; t0 += kBlk) {
#pragma unroll
            for (int i = 0; i < St::kPerLane; ++i) {
                const int v = lane + kWave * i;
                if (v < St::kVals) {
                    const int sl = v / kBlkVals;
                    const int rem = v - sl * kBlkVals;
                    const int r = rem / W;
                    const int a = rem - r * W;
                    T* dst = stage + (sl * kBlk + r) * kRow + a;
                    chen::static_for<kScaled ? D : 1>([&](auto d0) {
                        constexpr int d = int(decltype(d0)::value) + 1;
                        dst[(d - 1) * W] = (d == 1) ? pre[i] : pre[i] * (T(1) / T(d));
                    });
                }
            }
            __syncthreads();
            if (t0 + kBlk < maxlen) fetch(t0 + kBlk);
            const int nsteps = min(kBlk, len - t0);
            const T* rows = stage + slot * kBlk * kRow;
            for (int tt = 0; tt < nsteps; ++tt, rows += kRow) {
                if constexpr (kScaled)
                    st.step(chen::ScaledRow<T, W, D, P>{rows, u});
                else
                    st.step(chen::RawRow<T, W, D, P>(rows, u));
            }
            __syncthreads();
        }

        const int s = slot_stream[slot];
        if (active) {
            IO* o = out + int64_t(s) * tsize;
            // subtree levels: lane q owns words q*W^(k-P) .. (q+1)*W^(k-P)-1
            chen::static_for<D - P + 1>([&](auto i) {
                constexpr int k = P + int(decltype(i)::value);
                constexpr int nk = int(ipow(W, k - P));
                IO* ok = o + b.db[k] + int64_t(q) * nk;
                chen::static_for<nk>([&](auto v) { ok[v] = IO(st.sub[Sh::off(k) + v]); });
            });
            // path levels: written by the first lane of each prefix group
            chen::static_for<P>([&](auto i) {
                constexpr int k = int(decltype(i)::value);
                constexpr int grp = int(ipow(W, P - k));
                if (q % grp == 0) o[b.db[k] + q / grp] = IO(st.path[k]);
            });
        }
    }
}

// Stream-per-workgroup variant for prefix splits of 64 lanes or more: a
// workgroup of ceil(W^P / 64) waves takes one stream at a time from the
// length-sorted queue, so every lane of it steps the same increment.  Rows
// are staged block by block (kBlk steps, rows past the end of the stream
// zero, which leaves the signature unchanged) into LDS as
//   uni[t][a]            = x_a                      (read by every lane)
//   tab[t][a][d - 1]     = x_a / d,  d = 1..D       (the lane's path letters)
// and the block's steps are unrolled, so every LDS read takes its step's
// offset as an immediate.  Subtree partials fold their 1/d (UniformRow), so
// a step issues the lane's Horner FMAs and no scaling of the increment.
template <typename T, int W, int D, int P>
struct StreamCfg {
    static constexpr int NP = int(ipow(W, P));
    static constexpr int NW = (NP + 63) / 64;
    static constexpr int NT = NW * 64;
    static constexpr int LS = D % 2 ? D + 1 : D;    // slots per letter, even for b128 pairs
    static constexpr int US = W % 2 ? W + 1 : W;    // uniform slots per step
    static constexpr int STEP = W * LS + US;        // values per staged step
    static constexpr int kBlk = 32;
    static constexpr int kRaw = kBlk * W;           // raw values per block
    static constexpr int kPer = (kRaw + NT - 1) / NT;
};

template <typename T, int W, int D, int P, typename IO = T>
__global__ __launch_bounds__((StreamCfg<T, W, D, P>::NT))
void chen_stream_kernel(IO* __restrict__ out, const IO* __restrict__ incr,
                        const int64_t* __restrict__ begins, const int64_t* __restrict__ ends,
                        const int32_t* __restrict__ order, int64_t n_streams, int64_t tsize,
                        DevBasis b, uint32_t* __restrict__ counter)
{
    using Sh = chen::Shape<W, D, P>;
    using C = StreamCfg<T, W, D, P>;
    constexpr int NP = C::NP, LS = C::LS, STEP = C::STEP, kBlk = C::kBlk;
    __shared__ __align__(16) T stage[2][kBlk * STEP];
    __shared__ int64_t s_row0, s_len;
    __shared__ int32_t s_stream;

    const int tid = threadIdx.x;
    const bool active = tid < NP;
    const int q = active ? tid : 0;
    int u[P];
    {
        int t = q;
        for (int j = P - 1; j >= 0; --j) {
            u[j] = t % W;
            t /= W;
        }
    }
    T pre[C::kPer];
    // raw value v of the block at step t0: row t0 + v / W, letter v % W
    auto fetch = [&](int64_t row0, int64_t len, int64_t t0) {
#pragma unroll
        for (int i = 0; i < C::kPer; ++i) {
            const int v = tid + C::NT * i;
            const int r = v / W;
            T val = T(0);
            if (v < C::kRaw && t0 + r < len) val = T(incr[(row0 + t0) * W + v]);
            pre[i] = val;
        }
    };
    auto store = [&](T* st_) {
#pragma unroll
        for (int i = 0; i < C::kPer; ++i) {
            const int v = tid + C::NT * i;
            if (v < C::kRaw) {
                const int r = v / W;
                const int a = v - r * W;
                T* step = st_ + r * STEP;
                step[W * LS + a] = pre[i];
                chen::static_for<D>([&](auto d0) {
                    constexpr int d = int(decltype(d0)::value) + 1;
                    step[a * LS + d - 1] = (d == 1) ? pre[i] : pre[i] * (T(1) / T(d));
                });
            }
        }
    };

    while (true) {
        __syncthreads();  // the previous stream is done with the slot values
        if (tid == 0) {
            const uint32_t idx = atomicAdd(counter, 1u);
            int32_t s = -1;
            int64_t r0 = 0, ln = 0;
            if (int64_t(idx) < n_streams) {
                s = order[idx];
                r0 = begins[s];
                ln = ends[s] - r0;
                if (ln < 0) ln = 0;
            }
            s_stream = s;
            s_row0 = r0;
            s_len = ln;
        }
        __syncthreads();
        const int32_t s = s_stream;
        if (s < 0) break;
        const int64_t row0 = s_row0, len = s_len;

        chen::Lane<T, W, D, P> st;
        st.init_identity();
        fetch(row0, len, 0);
        store(stage[0]);
        __syncthreads();
        int buf = 0;
        for (int64_t t0 = 0; t0 < len; t0 += kBlk) {
            const bool more = t0 + kBlk < len;
            if (more) fetch(row0, len, t0 + kBlk);
            const T* blk = stage[buf];
#pragma unroll 1
            for (int tt = 0; tt < kBlk; ++tt) {
                const T* step = blk + tt * STEP;
                T raw[W];
                chen::static_for<W>([&](auto a) { raw[a] = step[W * LS + int(decltype(a)::value)]; });
                const T* lt[P];
                chen::static_for<P>([&](auto j) {
                    constexpr int jj = int(decltype(j)::value);
                    lt[jj] = step + u[jj] * LS;
                });
                st.step(chen::UniformRow<T, W, D, P, LS>{raw, lt});
                // keep each step's loads with its step: hoisting a whole
                // unrolled block's reads runs out of registers
                __builtin_amdgcn_sched_barrier(0);
            }
            if (more) store(stage[buf ^ 1]);
            __syncthreads();
            buf ^= 1;
        }

        if (active) {
            IO* o = out + int64_t(s) * tsize;
            chen::static_for<D - P + 1>([&](auto i) {
                constexpr int k = P + int(decltype(i)::value);
                constexpr int nk = int(ipow(W, k - P));
                IO* ok = o + b.db[k] + int64_t(q) * nk;
                chen::static_for<nk>([&](auto v) { ok[v] = IO(st.sub[Sh::off(k) + v]); });
            });
            chen::static_for<P>([&](auto i) {
                constexpr int k = int(decltype(i)::value);
                constexpr int grp = int(ipow(W, P - k));
                if (q % grp == 0) o[b.db[k] + q / grp] = IO(st.path[k]);
            });
        }
    }
}

// Fallback for shapes without a specialised kernel: one workgroup per
// stream, the running signature in the output buffer, per-coefficient Horner
// with a barrier between levels.
// With IO != T the running signature lives in a per-workgroup slot of
// `work` and is rounded into `out` at the end of each stream.
template <typename T, typename IO = T>
__global__ __launch_bounds__(256) void chen_generic_kernel(IO* __restrict__ out,
                                                          const IO* __restrict__ incr,
                                                          const int64_t* __restrict__ begins,
                                                          const int64_t* __restrict__ ends,
                                                          int64_t n_streams, int64_t tsize,
                                                          DevBasis b, T* __restrict__ work)
{
    const int W = b.width, D = b.depth;
    for (int64_t s = blockIdx.x; s < n_streams; s += gridDim.x) {
        T* S;
        if constexpr (std::is_same<T, IO>::value)
            S = out + s * tsize;
        else
            S = work + int64_t(blockIdx.x) * tsize;
        for (int64_t i = threadIdx.x; i < tsize; i += blockDim.x) S[i] = (i == 0) ? T(1) : T(0);
        __syncthreads();
        for (int64_t r = begins[s]; r < ends[s]; ++r) {
            const IO* x = incr + r * W;
            for (int m = D; m >= 1; --m) {
                const int64_t n = b.level_size(m);
                for (int64_t idx = threadIdx.x; idx < n; idx += blockDim.x) {
                    // letters of the word, most significant first
                    T p = S[0];
                    int64_t rest = idx, scale = n / W;
                    for (int K = 1; K <= m; ++K) {
                        const int64_t pre = idx / scale;  // prefix of length K
                        const int letter = int(rest / scale);
                        rest -= int64_t(letter) * scale;
                        p = p * (T(x[letter]) / T(m - K + 1)) + S[b.db[K] + pre];
                        scale = (scale >= W) ? scale / W : 1;
                    }
                    // only this thread reads S_m[idx]; lower levels stay old
                    // until the barrier below.
                    S[b.db[m] + idx] = p;
                }
                __syncthreads();
            }
        }
        if constexpr (!std::is_same<T, IO>::value) {
            for (int64_t i = threadIdx.x; i < tsize; i += blockDim.x) out[s * tsize + i] = IO(S[i]);
            __syncthreads();
        }
    }
}

using Launcher = int (*)(void* out, const void* incr, const int64_t* begins, const int64_t* ends,
                         const int32_t* order, int64_t n_streams, const DevBasis& b,
                         uint32_t* counter, hipStream_t st);

template <typename T, int W, int D, int P, typename IO = T>
int launch_persistent(void* out, const void* incr, const int64_t* begins, const int64_t* ends,
                      const int32_t* order,
                      int64_t n_streams, const DevBasis& b, uint32_t* counter, hipStream_t st)
{
    constexpr int NP = chen::Shape<W, D, P>::NP;
    const int64_t n_items = n_streams * NP;
    int dev = 0, cus = 0, per_cu = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RPY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, chen_persistent_kernel<T, W, D, P, IO>, kWave, 0));
    if (per_cu < 1) per_cu = 1;
    const int64_t chunks = (n_items + kWave - 1) / kWave;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(chunks, int64_t(cus) * per_cu));
    hipLaunchKernelGGL((chen_persistent_kernel<T, W, D, P, IO>), dim3(uint32_t(grid)), dim3(kWave),
                       0, st, (IO*)out, (const IO*)incr, begins, ends, order, n_items, b.size(), b,
                       counter);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

template <typename T, int W, int D, int P, typename IO = T>
int launch_stream(void* out, const void* incr, const int64_t* begins, const int64_t* ends,
                  const int32_t* order, int64_t n_streams, const DevBasis& b, uint32_t* counter,
                  hipStream_t st)
{
    using C = StreamCfg<T, W, D, P>;
    int dev = 0, cus = 0, per_cu = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RPY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, chen_stream_kernel<T, W, D, P, IO>, C::NT, 0));
    if (per_cu < 1) per_cu = 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_streams, int64_t(cus) * per_cu));
    hipLaunchKernelGGL((chen_stream_kernel<T, W, D, P, IO>), dim3(uint32_t(grid)), dim3(C::NT),
                       0, st, (IO*)out, (const IO*)incr, begins, ends, order, n_streams, b.size(),
                       b, counter);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

using chen::choose_p;
using chen::kBudgetF32;
using chen::kBudgetF64;

// Kernel variants of one shape: [0] the lane-chunk kernel (64 (stream,
// prefix) items per wave), [1] the stream-per-workgroup kernel (prefix
// splits of 64 lanes or more only, nullptr otherwise).
struct Entry {
    int w, d, p32, p64;
    Launcher f32[2], f64[2], f32c[2];
};

template <typename T, int W, int D, int P, typename IO = T>
constexpr Launcher stream_or_null()
{
    if constexpr (ipow(W, P) >= 64 && ipow(W, P) <= 1024)
        return &launch_stream<T, W, D, P, IO>;
    else
        return nullptr;
}

template <int W, int D>
Entry make_entry()
{
    constexpr int P32 = choose_p(W, D, kBudgetF32);
    constexpr int P64 = choose_p(W, D, kBudgetF64);
    return Entry{W, D, P32, P64,
                 {&launch_persistent<float, W, D, P32>, stream_or_null<float, W, D, P32>()},
                 {&launch_persistent<double, W, D, P64>, stream_or_null<double, W, D, P64>()},
                 {&launch_persistent<double, W, D, P64, float>,
                  stream_or_null<double, W, D, P64, float>()}};
}

// Which variant runs: the stream kernel where it exists, unless the
// environment variable ROUGHPY_AMD_SIGNATURE_KERNEL is "chunk" (or "stream"
// to insist), read once.
int variant_override()
{
    static const int v = [] {
        const char* e = std::getenv("ROUGHPY_AMD_SIGNATURE_KERNEL");
        if (!e) return -1;
        if (std::strcmp(e, "chunk") == 0) return 0;
        if (std::strcmp(e, "stream") == 0) return 1;
        return -1;
    }();
    return v;
}

Launcher pick(const Launcher (&f)[2])
{
    const int v = variant_override();
    if (v == 0 || !f[1]) return f[0];
    return f[1];
}

const std::vector<Entry>& table()
{
    static const std::vector<Entry> t = {
            make_entry<2, 1>(), make_entry<2, 2>(), make_entry<2, 3>(), make_entry<2, 4>(),
            make_entry<2, 5>(), make_entry<2, 6>(), make_entry<2, 7>(), make_entry<2, 8>(),
            make_entry<3, 1>(), make_entry<3, 2>(), make_entry<3, 3>(), make_entry<3, 4>(),
            make_entry<3, 5>(), make_entry<3, 6>(), make_entry<4, 1>(), make_entry<4, 2>(),
            make_entry<4, 3>(), make_entry<4, 4>(), make_entry<4, 5>(), make_entry<4, 6>(),
            make_entry<5, 1>(), make_entry<5, 2>(), make_entry<5, 3>(), make_entry<5, 4>(),
            make_entry<5, 5>(), make_entry<6, 1>(), make_entry<6, 2>(), make_entry<6, 3>(),
            make_entry<6, 4>(), make_entry<6, 5>(), make_entry<7, 1>(), make_entry<7, 2>(),
            make_entry<7, 3>(), make_entry<7, 4>(), make_entry<8, 1>(), make_entry<8, 2>(),
            make_entry<8, 3>(), make_entry<8, 4>(),
    };
    return t;
}

const Entry* find_entry(int w, int d)
{
    for (const auto& e : table())
        if (e.w == w && e.d == d) return &e;
    return nullptr;
}

}  // namespace

int chen_prefix_length(int32_t width, int32_t depth, int dtype_is_f64)
{
    const Entry* e = find_entry(width, depth);
    if (!e) return 0;
    return dtype_is_f64 ? e->p64 : e->p32;
}

namespace {

// Device workspace of one signature call: keys in/out, values in/out, the
// queue counter and the radix sort's scratch (specialised shapes), or the
// float64 running signatures of the generic kernel's compensated mode.
struct SigPlan {
    rpy_tensor_basis tb;
    DevBasis b;
    const Entry* e;
    size_t sort_bytes;  // radix sort scratch
    size_t k_bytes, v_bytes;
    size_t total;       // bytes of workspace the call needs
    int generic_grid;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Validates the arguments and sizes the workspace; *skip = true when there
// is nothing to launch.
int plan_signature(const void* out, const void* increments, const int64_t* begins,
                   const int64_t* ends, int64_t n_streams, int32_t width, int32_t depth,
                   rpy_dtype dtype, hipStream_t st, SigPlan& p, bool* skip,
                   int max_generic_grid = 4096)
{
    *skip = false;
    if (dtype != RPY_F32 && dtype != RPY_F64 && dtype != RPY_F32_COMPENSATED)
        return fail(RPY_ERR_DTYPE, "unsupported dtype");
    p.tb = rpy_tensor_basis{width, depth, nullptr};
    int rc = make_basis(&p.tb, p.b);
    if (rc) return rc;
    if (depth < 1) return fail(RPY_ERR_INVALID_ARGUMENT, "signatures need depth >= 1");
    if (n_streams < 0) return fail(RPY_ERR_SHAPE, "negative number of streams");
    if (n_streams == 0) {
        *skip = true;
        return RPY_OK;
    }
    if (n_streams > INT32_MAX) return fail(RPY_ERR_SHAPE, "too many streams");
    if (!out || !begins || !ends || !increments)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    p.e = find_entry(width, depth);
    p.sort_bytes = p.k_bytes = p.v_bytes = 0;
    p.generic_grid = int(std::min<int64_t>(n_streams, max_generic_grid));
    if (!p.e) {
        p.total = dtype == RPY_F32_COMPENSATED
                          ? align256(size_t(p.generic_grid) * size_t(p.b.size()) * 8)
                          : 0;
        return RPY_OK;
    }
    const size_t n = size_t(n_streams);
    RPY_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, p.sort_bytes, (uint32_t*)nullptr,
                                                 (uint32_t*)nullptr, (int32_t*)nullptr,
                                                 (int32_t*)nullptr, n, 0, 32, st));
    p.k_bytes = align256(n * sizeof(uint32_t));
    p.v_bytes = align256(n * sizeof(int32_t));
    p.total = 2 * p.k_bytes + 2 * p.v_bytes + 256 + align256(p.sort_bytes);
    return RPY_OK;
}

// Launches the call on workspace `ws` (p.total bytes, device memory).
int run_signature(void* out, const void* increments, const int64_t* begins, const int64_t* ends,
                  int64_t n_streams, rpy_dtype dtype, const SigPlan& p, char* ws,
                  hipStream_t st)
{
    const DevBasis& b = p.b;
    if (!p.e) {
        const int grid = p.generic_grid;
        if (dtype == RPY_F64)
            hipLaunchKernelGGL((chen_generic_kernel<double>), dim3(grid), dim3(256), 0, st,
                               (double*)out, (const double*)increments, begins, ends, n_streams,
                               b.size(), b, (double*)nullptr);
        else if (dtype == RPY_F32)
            hipLaunchKernelGGL((chen_generic_kernel<float>), dim3(grid), dim3(256), 0, st,
                               (float*)out, (const float*)increments, begins, ends, n_streams,
                               b.size(), b, (float*)nullptr);
        else
            hipLaunchKernelGGL((chen_generic_kernel<double, float>), dim3(grid), dim3(256), 0, st,
                               (float*)out, (const float*)increments, begins, ends, n_streams,
                               b.size(), b, (double*)ws);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "signature launch failed: %s", hipGetErrorString(err));
        return RPY_OK;
    }
    const size_t n = size_t(n_streams);
    uint32_t* keys_in = (uint32_t*)ws;
    uint32_t* keys_out = (uint32_t*)(ws + p.k_bytes);
    int32_t* vals_in = (int32_t*)(ws + 2 * p.k_bytes);
    int32_t* vals_out = (int32_t*)(ws + 2 * p.k_bytes + p.v_bytes);
    uint32_t* counter = (uint32_t*)(ws + 2 * p.k_bytes + 2 * p.v_bytes);
    void* sort_tmp = ws + 2 * p.k_bytes + 2 * p.v_bytes + 256;
    size_t sort_bytes = p.sort_bytes;

    // the length keys also clear the queue counter (a kernel, not a memset,
    // so a captured graph replays it)
    const int pb = 256;
    hipLaunchKernelGGL(prep_lengths_kernel, dim3(uint32_t((n + pb - 1) / pb)), dim3(pb), 0, st,
                       begins, ends, n_streams, keys_in, vals_in, counter);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return fail(RPY_ERR_HIP, "length prep launch failed: %s", hipGetErrorString(err));
    // Longest-first order only matters when the chunks run in more than one
    // round; a batch whose waves are all resident at once (a few per CU)
    // keeps the input order and skips the sort's launches.
    const int np = (dtype == RPY_F64 || dtype == RPY_F32_COMPENSATED) ? p.e->p64 : p.e->p32;
    const int64_t lanes = int64_t(n_streams) * int64_t(ipow(p.tb.width, np));
    const bool sort = (lanes + 63) / 64 > 1024;
    if (sort) {
        err = rocprim::radix_sort_pairs_desc(sort_tmp, sort_bytes, keys_in, keys_out, vals_in,
                                             vals_out, n, 0, 32, st);
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "radix sort failed: %s", hipGetErrorString(err));
    }
    Launcher f = (dtype == RPY_F64) ? pick(p.e->f64) : (dtype == RPY_F32) ? pick(p.e->f32)
                                                                         : pick(p.e->f32c);
    return f(out, increments, begins, ends, sort ? vals_out : vals_in, n_streams, b, counter, st);
}

// Scratch of the entry point that allocates nothing: one array per GPU,
// resolved to a device address once per device (its name is only a host
// stand-in).  Calls that overlap on one GPU share it and race.
constexpr size_t kScratchBytes = size_t(8) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_pointer(char** out)
{
    int dev = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(RPY_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail(RPY_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(err));
        g_scratch_ptr[dev] = static_cast<char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return RPY_OK;
}

}  // namespace

int stream_signature_ranges(void* out, const void* increments, const int64_t* begins,
                            const int64_t* ends, int64_t n_streams, int32_t width, int32_t depth,
                            rpy_dtype dtype, void* stream)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    SigPlan p;
    bool skip = false;
    int rc = plan_signature(out, increments, begins, ends, n_streams, width, depth, dtype, st, p,
                            &skip);
    if (rc || skip) return rc;
    char* ws = nullptr;
    if (p.total) RPY_HIP_CHECK(workspace_alloc((void**)&ws, p.total, st));
    rc = run_signature(out, increments, begins, ends, n_streams, dtype, p, ws, st);
    if (ws) {
        hipError_t ferr = hipFreeAsync(ws, st);
        if (!rc && ferr != hipSuccess)
            rc = fail(RPY_ERR_HIP, "hipFreeAsync failed: %s", hipGetErrorString(ferr));
    }
    return rc;
}

int stream_signature_batch(void* out, const void* increments, const int64_t* offsets,
                           int64_t n_streams, int32_t width, int32_t depth, rpy_dtype dtype,
                           void* stream)
{
    if (!offsets && n_streams > 0) return fail(RPY_ERR_INVALID_ARGUMENT, "null offsets");
    // CSR offsets: stream s owns rows [offsets[s], offsets[s+1])
    return stream_signature_ranges(out, increments, offsets, offsets ? offsets + 1 : nullptr,
                                   n_streams, width, depth, dtype, stream);
}

int64_t stream_signature_scratch_bytes() { return int64_t(kScratchBytes); }

int stream_signature_batch_noalloc(void* out, const void* increments, const int64_t* offsets,
                                   int64_t n_streams, int32_t width, int32_t depth,
                                   rpy_dtype dtype, void* stream)
{
    if (!offsets && n_streams > 0) return fail(RPY_ERR_INVALID_ARGUMENT, "null offsets");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t* begins = offsets;
    const int64_t* ends = offsets ? offsets + 1 : nullptr;
    SigPlan p;
    bool skip = false;
    // the generic kernel's compensated slots must fit the scratch as well
    int64_t tsize = tensor_size_checked(width, depth);
    int max_grid = 4096;
    if (dtype == RPY_F32_COMPENSATED && tsize > 0)
        max_grid = int(std::max<int64_t>(1, std::min<int64_t>(4096, int64_t(kScratchBytes) /
                                                                        (tsize * 8))));
    int rc = plan_signature(out, increments, begins, ends, n_streams, width, depth, dtype, st, p,
                            &skip, max_grid);
    if (rc || skip) return rc;
    if (p.total > kScratchBytes)
        return fail(RPY_ERR_SHAPE,
                    "%lld streams need %zu bytes of scratch, more than the %zu this entry point "
                    "holds; use rpy_stream_signature_batch", (long long)n_streams, p.total,
                    kScratchBytes);
    char* ws = nullptr;
    if ((rc = scratch_pointer(&ws))) return rc;
    return run_signature(out, increments, begins, ends, n_streams, dtype, p, ws, st);
}

// Runs the lane algorithm of the specialised kernel on the host for one
// stream (double precision) -- used by CPU tests to check the decomposition
// and the output index map without a GPU.
template <int W, int D, int P>
static void emulate(double* out, const double* incr, int64_t len, const DevBasis& b)
{
    using Sh = chen::Shape<W, D, P>;
    for (int q = 0; q < Sh::NP; ++q) {
        int u[P];
        int t = q;
        for (int j = P - 1; j >= 0; --j) {
            u[j] = t % W;
            t /= W;
        }
        chen::Lane<double, W, D, P> st;
        st.init_identity();
        double row[Sh::ROW];
        // letter-major table of the stream kernel: tab[a * D + d - 1] = x_a / d
        double tab[W * D];
        for (int64_t r = 0; r < len; ++r) {
            chen::build_row<double, W, D>(row, incr + r * W);
            if (r % 3 == 1) {
                st.step(chen::ScaledRow<double, W, D, P>{row, u});
            } else if (r % 3 == 2) {
                for (int a = 0; a < W; ++a)
                    for (int d = 1; d <= D; ++d) tab[a * D + d - 1] = row[(d - 1) * W + a];
                const double* lt[P];
                for (int j = 0; j < P; ++j) lt[j] = tab + u[j] * D;
                double raw[W];
                for (int a = 0; a < W; ++a) raw[a] = incr[r * W + a];
                st.step(chen::UniformRow<double, W, D, P, D>{raw, lt});
            } else {
                st.step(chen::RawRow<double, W, D, P>(row, u));
            }
        }
        for (int k = P; k <= D; ++k) {
            const int nk = int(ipow(W, k - P));
            for (int v = 0; v < nk; ++v) out[b.db[k] + int64_t(q) * nk + v] = st.sub[Sh::off(k) + v];
        }
        for (int k = 0; k < P; ++k) {
            const int grp = int(ipow(W, P - k));
            if (q % grp == 0) out[b.db[k] + q / grp] = st.path[k];
        }
    }
}

int signature_host_emulation(double* out, const double* incr, int64_t len, int32_t width,
                             int32_t depth, int32_t prefix)
{
    rpy_tensor_basis tb{width, depth, nullptr};
    DevBasis b;
    int rc = make_basis(&tb, b);
    if (rc) return rc;
#define RPY_EMU(W, D, P) \
    if (width == W && depth == D && prefix == P) { emulate<W, D, P>(out, incr, len, b); return RPY_OK; }
    RPY_EMU(2, 3, 1) RPY_EMU(2, 3, 2) RPY_EMU(2, 3, 3)
    RPY_EMU(3, 4, 1) RPY_EMU(3, 4, 2) RPY_EMU(3, 4, 3)
    RPY_EMU(5, 5, 2) RPY_EMU(5, 5, 3)
    RPY_EMU(4, 6, 3) RPY_EMU(4, 6, 4)
    RPY_EMU(8, 4, 2) RPY_EMU(8, 4, 3)
    RPY_EMU(6, 5, 3)
#undef RPY_EMU
    return fail(RPY_ERR_UNSUPPORTED, "no host emulation for this (width, depth, prefix)");
}

}  // namespace rpy_amd
