// chen_tune_5_5.hip -- signature kernel variants timed by the autotuner at
// width 5, depth 5.

#include "chen_tune.hpp"

namespace rpy_amd {

SigVariantList sig_tune_variants_5_5()
{
    static const std::vector<SigVariant> v = [] {
        std::vector<SigVariant> v;
        add_chunk<5, 5, 2, 2>(v);
        add_chunk<5, 5, 2, 1>(v);
        add_chunk<5, 5, 3, 2>(v);
        add_stream<5, 5, 3, 16>(v);
        add_stream<5, 5, 3, 32>(v);
        add_stream<5, 5, 3, 64>(v);
        add_stream<5, 5, 4, 32>(v);
        add_stream_table<5, 5, 3, 16>(v);
        add_stream_table<5, 5, 3, 32>(v);
        return v;
    }();
    return {v.data(), v.size()};
}

}  // namespace rpy_amd
