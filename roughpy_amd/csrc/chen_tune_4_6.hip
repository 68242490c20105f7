// chen_tune_4_6.hip -- signature kernel variants timed by the autotuner at
// width 4, depth 6.

#include "chen_tune.hpp"

namespace rpy_amd {

SigVariantList sig_tune_variants_4_6()
{
    static const std::vector<SigVariant> v = [] {
        std::vector<SigVariant> v;
        add_chunk<4, 6, 3, 2>(v);
        add_chunk<4, 6, 3, 1>(v);
        add_chunk<4, 6, 4, 2>(v);
        add_stream<4, 6, 3, 16>(v);
        add_stream<4, 6, 3, 32>(v);
        add_stream<4, 6, 3, 64>(v);
        add_stream<4, 6, 4, 32>(v);
        add_stream<4, 6, 5, 32>(v);
        return v;
    }();
    return {v.data(), v.size()};
}

}  // namespace rpy_amd
