// kernel_util.hpp -- small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rpy_common.hpp"

namespace rpy_amd {

// Exact unsigned division by a runtime divisor d >= 1 for n < 2^31
// (Granlund-Montgomery round-up method: q = (umulhi(m, n) + n) >> s).
struct FastDiv {
    uint32_t d, m, s;
};

RPY_HD inline FastDiv make_fastdiv(uint32_t d)
{
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d) ++s;
    const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
    return FastDiv{d, uint32_t(m), s};
}

RPY_FI uint32_t fdiv(uint32_t n, const FastDiv& f) { return (__umulhi(f.m, n) + n) >> f.s; }

// Powers of the width as fast divisors: pw[k] divides by w^k.
struct Pow {
    FastDiv pw[kMaxDepth + 1];
};

inline Pow make_pow(int w, int D)
{
    Pow p{};
    uint64_t v = 1;
    for (int k = 0; k <= D && k <= kMaxDepth; ++k) {
        p.pw[k] = make_fastdiv(uint32_t(v));
        v *= uint64_t(w);
    }
    return p;
}

}  // namespace rpy_amd
