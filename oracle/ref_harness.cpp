// ref_harness.cpp -- our own thin C-ABI around the reference's header-only
// roughpy_compute kernels, compiled where they lie (-I <reference root>) by
// oracle/Makefile into oracle/_ref/libref.so.
//
// TEST INFRASTRUCTURE ONLY: it serves the tests as a second, independent
// anchor for the oracle, and the benchmark as the reference CPU baseline.
// No reference source is copied here; this file only calls the templates:
//   basic::ft_fma             roughpy_compute/dense/basic/free_tensor_fma.hpp:13-53
//   basic::ft_inplace_mul     roughpy_compute/dense/basic/free_tensor_inplace_mul.hpp:13-87
//   basic::ft_antipode        roughpy_compute/dense/basic/free_tensor_antipode.hpp:45-100
//   basic::ft_adj_lmul        roughpy_compute/dense/basic/free_tensor_adjoint_left_mul.hpp:10-77
//   intermediate::ft_exp      roughpy_compute/dense/intermediate/free_tensor_exp.hpp:13-37
//   intermediate::ft_fmexp    roughpy_compute/dense/intermediate/free_tensor_fmexp.hpp:15-40
//   intermediate::ft_log      roughpy_compute/dense/intermediate/free_tensor_log.hpp:12-40
//   basic::st_fma             roughpy_compute/dense/basic/shuffle_tensor_product.hpp:24-85
//   basic::st_adj_mul         roughpy_compute/dense/basic/shuffle_tensor_adjoint_mul.hpp:13-70
// The batched signature loop restates algebra/src/lite_context.h:545-563
// (result = 1; result = result * exp(row) per increment) on top of ft_fmexp.

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "roughpy_compute/dense/views.hpp"
#include "roughpy_compute/dense/basic/free_tensor_fma.hpp"
#include "roughpy_compute/dense/basic/free_tensor_inplace_mul.hpp"
#include "roughpy_compute/dense/basic/free_tensor_antipode.hpp"
#include "roughpy_compute/dense/basic/free_tensor_adjoint_left_mul.hpp"
#include "roughpy_compute/dense/intermediate/free_tensor_exp.hpp"
#include "roughpy_compute/dense/intermediate/free_tensor_fmexp.hpp"
#include "roughpy_compute/dense/intermediate/free_tensor_log.hpp"
#include "roughpy_compute/dense/basic/shuffle_tensor_product.hpp"
#include "roughpy_compute/dense/basic/shuffle_tensor_adjoint_mul.hpp"

namespace rc = rpy::compute;

namespace {

struct Basis {
    std::vector<std::ptrdiff_t> db;
    rc::TensorBasis tb;
    Basis(int w, int d) : db(d + 2)
    {
        db[0] = 0;
        for (int k = 0; k <= d; ++k) db[k + 1] = 1 + w * db[k];
        tb = rc::TensorBasis{{db.data(), w, d}};
    }
    std::ptrdiff_t size() const { return db.back(); }
};

template <typename T>
using View = rc::DenseTensorView<T*>;

template <typename T>
void fma_batch(T* out, const T* lhs, const T* rhs, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> l(const_cast<T*>(lhs) + e * b.size(), b.tb, 0, d);
        View<T> r(const_cast<T*>(rhs) + e * b.size(), b.tb, 0, d);
        rc::basic::ft_fma(o, l, r);
    }
}

template <typename T>
void inplace_mul_batch(T* lhs, const T* rhs, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> l(lhs + e * b.size(), b.tb, 0, d);
        View<T> r(const_cast<T*>(rhs) + e * b.size(), b.tb, 0, d);
        rc::basic::ft_inplace_mul(l, r);
    }
}

template <typename T>
void exp_batch(T* out, const T* arg, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        std::memset(out + e * b.size(), 0, sizeof(T) * b.size());
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> a(const_cast<T*>(arg) + e * b.size(), b.tb, 0, d);
        rc::intermediate::ft_exp(o, a);
    }
}

template <typename T>
void fmexp_batch(T* out, const T* mul, const T* expo, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> m(const_cast<T*>(mul) + e * b.size(), b.tb, 0, d);
        View<T> x(const_cast<T*>(expo) + e * b.size(), b.tb, 0, d);
        rc::intermediate::ft_fmexp(o, m, x);
    }
}

template <typename T>
void log_batch(T* out, const T* arg, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        std::memset(out + e * b.size(), 0, sizeof(T) * b.size());
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> a(const_cast<T*>(arg) + e * b.size(), b.tb, 0, d);
        rc::intermediate::ft_log(o, a);
    }
}

template <typename T>
void antipode_batch(T* out, const T* arg, int w, int d, int64_t n)
{
    Basis b(w, d);
    rc::basic::BasicAntipodeConfig cfg;
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> a(const_cast<T*>(arg) + e * b.size(), b.tb, 0, d);
        rc::basic::ft_antipode(o, a, cfg, rc::basic::DefaultSigner{});
    }
}

template <typename T>
void adj_lmul_batch(T* out, const T* op, const T* arg, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> p(const_cast<T*>(op) + e * b.size(), b.tb, 0, d);
        View<T> a(const_cast<T*>(arg) + e * b.size(), b.tb, 0, d);
        rc::basic::ft_adj_lmul(o, p, a);
    }
}

template <typename T>
void st_fma_batch(T* out, const T* lhs, const T* rhs, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> l(const_cast<T*>(lhs) + e * b.size(), b.tb, 0, d);
        View<T> r(const_cast<T*>(rhs) + e * b.size(), b.tb, 0, d);
        rc::basic::st_fma(o, l, r);
    }
}

template <typename T>
void st_adj_mul_batch(T* out, const T* op, const T* arg, int w, int d, int64_t n)
{
    Basis b(w, d);
    for (int64_t e = 0; e < n; ++e) {
        View<T> o(out + e * b.size(), b.tb, 0, d);
        View<T> p(const_cast<T*>(op) + e * b.size(), b.tb, 0, d);
        View<T> a(const_cast<T*>(arg) + e * b.size(), b.tb, 0, d);
        rc::basic::st_adj_mul(o, p, a);
    }
}

template <typename T>
void signature_range(T* out, const T* incr, const int64_t* offsets,
                     int64_t first, int64_t last, int w, int d)
{
    Basis b(w, d);
    const auto sz = b.size();
    std::vector<T> cur(sz), nxt(sz), x(sz);
    for (int64_t s = first; s < last; ++s) {
        std::fill(cur.begin(), cur.end(), T(0));
        std::fill(x.begin(), x.end(), T(0));
        cur[0] = T(1);
        for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) {
            for (int a = 0; a < w; ++a) x[1 + a] = incr[i * w + a];
            View<T> o(nxt.data(), b.tb, 0, d);
            View<T> m(cur.data(), b.tb, 0, d);
            View<T> e(x.data(), b.tb, 0, d);
            rc::intermediate::ft_fmexp(o, m, e);
            std::swap(cur, nxt);
        }
        std::memcpy(out + s * sz, cur.data(), sizeof(T) * sz);
    }
}

template <typename T>
void signature_batch(T* out, const T* incr, const int64_t* offsets,
                     int64_t n_streams, int w, int d, int nthreads)
{
    if (nthreads <= 1) {
        signature_range(out, incr, offsets, 0, n_streams, w, d);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t per = (n_streams + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int64_t f = t * per;
        const int64_t l = std::min<int64_t>(n_streams, f + per);
        if (f >= l) break;
        pool.emplace_back([=] { signature_range(out, incr, offsets, f, l, w, d); });
    }
    for (auto& th : pool) th.join();
}

}// namespace

#define REF_EXPORT(T, SFX)                                                     \
    extern "C" void ref_ft_fma_##SFX(T* o, const T* l, const T* r, int w,     \
                                     int d, int64_t n)                         \
    { fma_batch(o, l, r, w, d, n); }                                           \
    extern "C" void ref_ft_inplace_mul_##SFX(T* l, const T* r, int w, int d,   \
                                             int64_t n)                        \
    { inplace_mul_batch(l, r, w, d, n); }                                      \
    extern "C" void ref_ft_exp_##SFX(T* o, const T* a, int w, int d,          \
                                     int64_t n)                                \
    { exp_batch(o, a, w, d, n); }                                              \
    extern "C" void ref_ft_fmexp_##SFX(T* o, const T* m, const T* x, int w,   \
                                       int d, int64_t n)                       \
    { fmexp_batch(o, m, x, w, d, n); }                                         \
    extern "C" void ref_ft_log_##SFX(T* o, const T* a, int w, int d,          \
                                     int64_t n)                                \
    { log_batch(o, a, w, d, n); }                                              \
    extern "C" void ref_antipode_##SFX(T* o, const T* a, int w, int d,        \
                                       int64_t n)                              \
    { antipode_batch(o, a, w, d, n); }                                         \
    extern "C" void ref_adj_lmul_##SFX(T* o, const T* p, const T* a, int w,   \
                                       int d, int64_t n)                       \
    { adj_lmul_batch(o, p, a, w, d, n); }                                      \
    extern "C" void ref_st_fma_##SFX(T* o, const T* l, const T* r, int w,     \
                                     int d, int64_t n)                         \
    { st_fma_batch(o, l, r, w, d, n); }                                        \
    extern "C" void ref_st_adj_mul_##SFX(T* o, const T* p, const T* a, int w, \
                                         int d, int64_t n)                     \
    { st_adj_mul_batch(o, p, a, w, d, n); }                                    \
    extern "C" void ref_signature_batch_##SFX(T* o, const T* incr,            \
                                              const int64_t* off, int64_t ns,  \
                                              int w, int d, int nt)            \
    { signature_batch(o, incr, off, ns, w, d, nt); }

REF_EXPORT(double, f64)
REF_EXPORT(float, f32)
