// tuning.cpp -- the tuned table and the per-process kernel selections.

#include "tuning.hpp"

#include <map>
#include <mutex>
#include <tuple>

namespace rpy_amd {

namespace {

struct TunedRow {
    int32_t op, width, depth;
    rpy_dtype dtype;
    rpy_kernel_variant v;
};

// {op, width, depth, dtype, {kind, prefix, block, waves_per_eu}}; the file
// is generated (roughpy_amd/tune.py --write) and ends with a sentinel row.
const TunedRow kTuned[] = {
#include "signature_tuned.inc"
        {-1, 0, 0, RPY_F32, {0, 0, 0, 0}},
};

using Key = std::tuple<int32_t, int32_t, int32_t, int32_t>;

std::mutex g_mutex;
std::map<Key, rpy_kernel_variant>& selections()
{
    static std::map<Key, rpy_kernel_variant> m;
    return m;
}

}  // namespace

bool tuned_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                   rpy_kernel_variant* out)
{
    for (const TunedRow& r : kTuned) {
        if (r.op == op && r.width == width && r.depth == depth && r.dtype == dtype) {
            *out = r.v;
            return true;
        }
    }
    return false;
}

bool selected_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                      rpy_kernel_variant* out)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    const auto& m = selections();
    if (m.empty()) return false;
    auto it = m.find(Key{op, width, depth, int32_t(dtype)});
    if (it == m.end()) return false;
    *out = it->second;
    return true;
}

void set_selected_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                          const rpy_kernel_variant* v)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    const Key key{op, width, depth, int32_t(dtype)};
    if (v)
        selections()[key] = *v;
    else
        selections().erase(key);
}

}  // namespace rpy_amd
