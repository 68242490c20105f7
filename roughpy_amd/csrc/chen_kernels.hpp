// chen_kernels.hpp -- the specialised signature kernels and their launchers.
//
// Two kernels step the lane decomposition of chen_core.hpp over uneven
// streams taken from a length-sorted queue:
//   * chen_persistent_kernel: each wave64 takes 64 consecutive (stream,
//     prefix) items ("lane chunks"); tunable: prefix length P and the
//     occupancy hint WPE (waves per EU, which caps the VGPRs per lane);
//   * chen_stream_kernel: a workgroup of ceil(W^P / 64) waves takes one
//     stream (W^P >= 64); tunable: P and the rows per LDS block BLK.
// The templates have internal linkage, so every translation unit that
// instantiates a variant owns its copy; launchers are handed around as
// SigLauncher pointers (chen.hpp).

#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "chen.hpp"
#include "chen_core.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kWave = 64;

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

// T: arithmetic type; IO: type of the increments and signatures in memory;
// Comp: Kahan-compensated lowest levels (chen_core.hpp, the compensated
// float32 mode).
template <typename T, int W, int D, int P, typename IO = T, int WPE = 2, bool Comp = false>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WPE)))
void chen_persistent_kernel(
        IO* __restrict__ out, const IO* __restrict__ incr, const int64_t* __restrict__ begins,
        const int64_t* __restrict__ ends, const int32_t* __restrict__ order, int64_t n_items,
        int64_t tsize, DevBasis b, uint32_t* __restrict__ counter, int flags)
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
        if ((flags & kSkipEmpty) && maxlen == 0) {
            // longest first: every later chunk is empty as well
            if (flags & kSortedDesc) break;
            continue;
        }

        int u[P];
        {
            int t = q;
            for (int j = P - 1; j >= 0; --j) {
                u[j] = t % W;
                t /= W;
            }
        }

        chen::Lane<T, W, D, P, Comp> st;
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
        if (active && !((flags & kSkipEmpty) && len == 0)) {
            IO* o = out + int64_t(s) * tsize;
            // subtree levels: lane q owns words q*W^(k-P) .. (q+1)*W^(k-P)-1
            chen::static_for<D - P + 1>([&](auto i) {
                constexpr int k = P + int(decltype(i)::value);
                constexpr int nk = int(ipow(W, k - P));
                IO* ok = o + b.db[k] + int64_t(q) * nk;
                chen::static_for<nk>([&](auto v) {
                    ok[v] = IO(st.template sub_out<Sh::off(k) + int(decltype(v)::value)>());
                });
            });
            // path levels: written by the first lane of each prefix group
            chen::static_for<P>([&](auto i) {
                constexpr int k = int(decltype(i)::value);
                constexpr int grp = int(ipow(W, P - k));
                if (q % grp == 0) o[b.db[k] + q / grp] = IO(st.template path_out<k>());
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
template <typename T, int W, int D, int P, int BLK = 32>
struct StreamCfg {
    static constexpr int NP = int(ipow(W, P));
    static constexpr int NW = (NP + 63) / 64;
    static constexpr int NT = NW * 64;
    static constexpr int LS = D % 2 ? D + 1 : D;    // slots per letter, even for b128 pairs
    static constexpr int US = W % 2 ? W + 1 : W;    // uniform slots per step
    static constexpr int STEP = W * LS + US;        // values per staged step
    static constexpr int kBlk = BLK;
    static constexpr int kRaw = kBlk * W;           // raw values per block
    static constexpr int kPer = (kRaw + NT - 1) / NT;
};

// PT: the path partials of each step are computed once per prefix group
// u_0..u_{P-2} (W^(P-1) builder threads) for the whole block ahead of the
// step loop and read from LDS by the lanes, which then do only their
// subtree work (at P = 3 the path work is 9 of a lane-step's 48 operations).
template <typename T, int W, int D, int P, typename IO = T, int BLK = 32, bool Comp = false,
          bool PT = false>
__global__ __launch_bounds__((StreamCfg<T, W, D, P, BLK>::NT))
void chen_stream_kernel(IO* __restrict__ out, const IO* __restrict__ incr,
                        const int64_t* __restrict__ begins, const int64_t* __restrict__ ends,
                        const int32_t* __restrict__ order, int64_t n_streams, int64_t tsize,
                        DevBasis b, uint32_t* __restrict__ counter, int flags)
{
    using Sh = chen::Shape<W, D, P>;
    using C = StreamCfg<T, W, D, P, BLK>;
    constexpr int NP = C::NP, LS = C::LS, STEP = C::STEP, kBlk = C::kBlk;
    __shared__ __align__(16) T stage[2][kBlk * STEP];
    __shared__ int64_t s_row0, s_len;
    __shared__ int32_t s_stream;
    constexpr int NG = int(ipow(W, P - 1));  // prefix groups u_0..u_{P-2}
    constexpr int NPT = D - P + 1;            // subtree targets
    __shared__ __align__(16) T ptab[PT ? kBlk * NPT * NG : 1];

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
        if ((flags & kSkipEmpty) && len == 0) {
            // longest first: every later stream is empty as well
            if (flags & kSortedDesc) break;
            continue;
        }

        chen::Lane<T, W, D, P, Comp> st;
        st.init_identity();
        // PT: the path levels live in the builder threads (tid < NG, the
        // prefix group tid with letter u_{P-1} = 0)
        chen::Lane<T, W, D, P, Comp> pst;
        int ug[P];
        if constexpr (PT) {
            pst.init_identity();
            int t = (tid < NG ? tid : 0) * W;
            for (int j = P - 1; j >= 0; --j) {
                ug[j] = t % W;
                t /= W;
            }
        }
        fetch(row0, len, 0);
        store(stage[0]);
        __syncthreads();
        int buf = 0;
        for (int64_t t0 = 0; t0 < len; t0 += kBlk) {
            const bool more = t0 + kBlk < len;
            if (more) fetch(row0, len, t0 + kBlk);
            const T* blk = stage[buf];
            if constexpr (PT) {
                // rows past the end of the stream are zero: they leave the
                // path levels unchanged
                if (tid < NG) {
#pragma unroll 1
                    for (int tt = 0; tt < kBlk; ++tt) {
                        const T* step = blk + tt * STEP;
                        T raw[W];
                        chen::static_for<W>(
                                [&](auto a) { raw[a] = step[W * LS + int(decltype(a)::value)]; });
                        const T* lt[P];
                        chen::static_for<P>([&](auto j) {
                            constexpr int jj = int(decltype(j)::value);
                            lt[jj] = step + ug[jj] * LS;
                        });
                        pst.path_step(chen::UniformRow<T, W, D, P, LS>{raw, lt},
                                      ptab + tt * NPT * NG + tid, NG);
                    }
                }
                __syncthreads();
            }
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
                if constexpr (PT)
                    st.step_pp(chen::UniformRow<T, W, D, P, LS>{raw, lt},
                               ptab + tt * NPT * NG + q / W, NG);
                else
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
                chen::static_for<nk>([&](auto v) {
                    ok[v] = IO(st.template sub_out<Sh::off(k) + int(decltype(v)::value)>());
                });
            });
            if constexpr (PT) {
                // path levels: written by the builder of each prefix group
                chen::static_for<P>([&](auto i) {
                    constexpr int k = int(decltype(i)::value);
                    constexpr int grp = int(ipow(W, P - 1 - k));
                    if (tid < NG && tid % grp == 0)
                        o[b.db[k] + tid / grp] = IO(pst.template path_out<k>());
                });
            } else {
                chen::static_for<P>([&](auto i) {
                    constexpr int k = int(decltype(i)::value);
                    constexpr int grp = int(ipow(W, P - k));
                    if (q % grp == 0) o[b.db[k] + q / grp] = IO(st.template path_out<k>());
                });
            }
        }
    }
}

template <typename T, int W, int D, int P, typename IO = T, int WPE = 2, bool Comp = false>
int launch_persistent(void* out, const void* incr, const int64_t* begins, const int64_t* ends,
                      const int32_t* order, int64_t n_streams, const DevBasis& b,
                      uint32_t* counter, int flags, hipStream_t st)
{
    constexpr int NP = chen::Shape<W, D, P>::NP;
    const int64_t n_items = n_streams * NP;
    int dev = 0, cus = 0, per_cu = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RPY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, chen_persistent_kernel<T, W, D, P, IO, WPE, Comp>, kWave, 0));
    if (per_cu < 1) per_cu = 1;
    const int64_t chunks = (n_items + kWave - 1) / kWave;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(chunks, int64_t(cus) * per_cu));
    hipLaunchKernelGGL((chen_persistent_kernel<T, W, D, P, IO, WPE, Comp>), dim3(uint32_t(grid)), dim3(kWave),
                       0, st, (IO*)out, (const IO*)incr, begins, ends, order, n_items, b.size(), b,
                       counter, flags);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

template <typename T, int W, int D, int P, typename IO = T, int BLK = 32, bool Comp = false,
          bool PT = false>
int launch_stream(void* out, const void* incr, const int64_t* begins, const int64_t* ends,
                  const int32_t* order, int64_t n_streams, const DevBasis& b, uint32_t* counter,
                  int flags, hipStream_t st)
{
    using C = StreamCfg<T, W, D, P, BLK>;
    int dev = 0, cus = 0, per_cu = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    RPY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RPY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, chen_stream_kernel<T, W, D, P, IO, BLK, Comp, PT>, C::NT, 0));
    if (per_cu < 1) per_cu = 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_streams, int64_t(cus) * per_cu));
    hipLaunchKernelGGL((chen_stream_kernel<T, W, D, P, IO, BLK, Comp, PT>), dim3(uint32_t(grid)), dim3(C::NT),
                       0, st, (IO*)out, (const IO*)incr, begins, ends, order, n_streams, b.size(),
                       b, counter, flags);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

}  // namespace

}  // namespace rpy_amd
