// chen_tune.hpp -- extra signature kernel variants for the tuned shapes.
//
// Each chen_tune_W_D.hip instantiates, for float32, float64 and the
// compensated float32 mode (float32 lanes with Kahan-compensated lowest
// levels), the candidate
// variants the autotuner times for one shape, skipping the ones chen.hip
// already compiles as defaults (prefix choose_p, 2 waves per EU, 32-row
// blocks).
#pragma once

#include <vector>

#include "chen_kernels.hpp"

namespace rpy_amd {

namespace {

// lane-chunk kernel at prefix P with the occupancy hint WPE
template <int W, int D, int P, int WPE>
void add_chunk(std::vector<SigVariant>& v)
{
    constexpr int P32 = chen::choose_p(W, D, chen::kBudgetF32);
    constexpr int P64 = chen::choose_p(W, D, chen::kBudgetF64);
    constexpr rpy_kernel_variant k{RPY_KIND_LANE_CHUNK, P, 0, WPE};
    if constexpr (!(P == P32 && WPE == 2)) {
        v.push_back({W, D, RPY_F32, k, &launch_persistent<float, W, D, P, float, WPE>});
        v.push_back({W, D, RPY_F32_COMPENSATED, k,
                     &launch_persistent<float, W, D, P, float, WPE, true>});
    }
    if constexpr (!(P == P64 && WPE == 2))
        v.push_back({W, D, RPY_F64, k, &launch_persistent<double, W, D, P, double, WPE>});
}

// stream kernel at prefix P (64 <= W^P <= 1024) with BLK rows per block
template <int W, int D, int P, int BLK>
void add_stream(std::vector<SigVariant>& v)
{
    static_assert(ipow(W, P) >= 64 && ipow(W, P) <= 1024, "stream kernel needs 64..1024 lanes");
    constexpr int P32 = chen::choose_p(W, D, chen::kBudgetF32);
    constexpr int P64 = chen::choose_p(W, D, chen::kBudgetF64);
    constexpr rpy_kernel_variant k{RPY_KIND_STREAM_GROUP, P, BLK, 0};
    if constexpr (!(P == P32 && BLK == 32)) {
        v.push_back({W, D, RPY_F32, k, &launch_stream<float, W, D, P, float, BLK>});
        v.push_back({W, D, RPY_F32_COMPENSATED, k,
                     &launch_stream<float, W, D, P, float, BLK, true>});
    }
    if constexpr (!(P == P64 && BLK == 32))
        v.push_back({W, D, RPY_F64, k, &launch_stream<double, W, D, P, double, BLK>});
}

// the stream kernel with the block's path partials tabulated per prefix group
template <int W, int D, int P, int BLK>
void add_stream_table(std::vector<SigVariant>& v)
{
    static_assert(ipow(W, P) >= 64 && ipow(W, P) <= 1024, "stream kernel needs 64..1024 lanes");
    constexpr rpy_kernel_variant k{RPY_KIND_STREAM_TABLE, P, BLK, 0};
    v.push_back({W, D, RPY_F32, k, &launch_stream<float, W, D, P, float, BLK, false, true>});
    v.push_back({W, D, RPY_F64, k, &launch_stream<double, W, D, P, double, BLK, false, true>});
}

}  // namespace

}  // namespace rpy_amd
