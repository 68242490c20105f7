// chen_tune_8_4.hip -- signature kernel variants timed by the autotuner at
// width 8, depth 4.

#include "chen_tune.hpp"

namespace rpy_amd {

SigVariantList sig_tune_variants_8_4()
{
    static const std::vector<SigVariant> v = [] {
        std::vector<SigVariant> v;
        add_chunk<8, 4, 2, 2>(v);
        add_chunk<8, 4, 2, 1>(v);
        add_chunk<8, 4, 3, 2>(v);
        add_stream<8, 4, 2, 16>(v);
        add_stream<8, 4, 2, 32>(v);
        add_stream<8, 4, 2, 64>(v);
        add_stream<8, 4, 3, 32>(v);
        return v;
    }();
    return {v.data(), v.size()};
}

}  // namespace rpy_amd
