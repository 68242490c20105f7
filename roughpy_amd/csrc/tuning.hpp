// tuning.hpp -- which compiled kernel variant a call runs.
//
// Precedence: a variant chosen with rpy_tune_select in this process, then
// the row of the tuned table (signature_tuned.inc, written by
// roughpy_amd/tune.py from timings on MI355X), then the caller's built-in
// default.  The kernels themselves register their variants (chen.hip,
// backward.hip); this file only keeps the choices.
#pragma once

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

inline bool same_variant(const rpy_kernel_variant& a, const rpy_kernel_variant& b)
{
    return a.kind == b.kind && a.prefix == b.prefix && a.block == b.block &&
           a.waves_per_eu == b.waves_per_eu;
}

// The tuned table's row for the shape, if any.
bool tuned_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                   rpy_kernel_variant* out);

// The rpy_tune_select choice for the shape, if any.
bool selected_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                      rpy_kernel_variant* out);

// Records (or with v == nullptr clears) the choice; the caller has checked
// that the variant exists.
void set_selected_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                          const rpy_kernel_variant* v);

// Resolves a call's variant: the selection or the tuned row when `exists`
// accepts it, else `fallback`.  *source: 0 default, 1 tuned, 2 selected.
template <typename Exists>
rpy_kernel_variant resolve_variant(int32_t op, int32_t width, int32_t depth, rpy_dtype dtype,
                                   const rpy_kernel_variant& fallback, Exists exists,
                                   int32_t* source)
{
    rpy_kernel_variant v;
    if (selected_variant(op, width, depth, dtype, &v) && exists(v)) {
        if (source) *source = 2;
        return v;
    }
    if (tuned_variant(op, width, depth, dtype, &v) && exists(v)) {
        if (source) *source = 1;
        return v;
    }
    if (source) *source = 0;
    return fallback;
}

}  // namespace rpy_amd
