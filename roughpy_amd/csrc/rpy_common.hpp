// rpy_common.hpp -- shared host/device definitions of the gfx950 library.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/roughpy_amd.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RPY_HD __host__ __device__
#define RPY_DEV __device__
#define RPY_FI __device__ __forceinline__
#else
#define RPY_HD
#define RPY_DEV
#define RPY_FI inline
#endif

namespace rpy_amd {

// Deepest truncation the device kernels accept (degree_begin is passed by
// value in kernel arguments).
constexpr int kMaxDepth = 16;

// Sets the thread-local error message and returns `code`.
int fail(int code, const char* fmt, ...);

// Host-side helpers implemented in lie_basis.cpp.
int64_t tensor_size_checked(int32_t width, int32_t depth);
int64_t lie_size(int32_t width, int32_t depth);
int build_hall(int32_t width, int32_t depth, int64_t* db, int64_t* data);
int check_hall(const rpy_lie_basis* b);
int build_l2t(const rpy_lie_basis* b, int64_t* nnz, int64_t* offsets, int64_t* indices,
              double* data);
int build_t2l(const rpy_lie_basis* b, int64_t* nnz, int64_t* offsets, int64_t* indices,
              double* data);

// Degree offsets of a tensor basis, passed by value to kernels.
struct DevBasis {
    int32_t width;
    int32_t depth;
    int64_t db[kMaxDepth + 2];

    RPY_HD int64_t size() const { return db[depth + 1]; }
    RPY_HD int64_t level_size(int k) const { return db[k + 1] - db[k]; }
};

// Validates width/depth (and a caller-supplied degree_begin) and fills a
// DevBasis; returns RPY_OK or an error code with the message set.
int make_basis(const rpy_tensor_basis* b, DevBasis& out);

// b^e by a loop (not recursion, so device code inlines it wherever it is not
// folded at compile time)
constexpr RPY_HD inline int64_t ipow(int64_t b, int e)
{
    int64_t r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

}  // namespace rpy_amd

#ifdef __HIPCC__
#define RPY_HIP_CHECK(expr)                                                          \
    do {                                                                             \
        hipError_t err_ = (expr);                                                    \
        if (err_ != hipSuccess)                                                      \
            return ::rpy_amd::fail(RPY_ERR_HIP, "%s failed: %s", #expr,              \
                                   hipGetErrorString(err_));                         \
    } while (0)
#endif
