// chen_tune_6_5.hip -- signature kernel variants timed by the autotuner at
// width 6, depth 5.

#include "chen_tune.hpp"

namespace rpy_amd {

SigVariantList sig_tune_variants_6_5()
{
    static const std::vector<SigVariant> v = [] {
        std::vector<SigVariant> v;
        add_chunk<6, 5, 3, 2>(v);
        add_chunk<6, 5, 4, 2>(v);
        add_stream<6, 5, 3, 16>(v);
        add_stream<6, 5, 3, 32>(v);
        add_stream<6, 5, 3, 64>(v);
        add_stream_table<6, 5, 3, 16>(v);
        add_stream_table<6, 5, 3, 32>(v);
        return v;
    }();
    return {v.data(), v.size()};
}

}  // namespace rpy_amd
