// dense_ops.hip -- batched dense free-tensor operations for gfx950.
//
// One workgroup (256 lanes = 4 wave64) owns one batch element at a time
// (grid-stride over the batch).  The lower levels of the operands, which
// every output coefficient re-reads, are staged in LDS; the top level --
// read once per coefficient -- streams straight from HBM, so an fma moves
// 4 * size * sizeof(T) bytes per element and no more.
//
// Reference kernels mirrored (same per-coefficient term order where it
// matters for rounding):
//   ft_fma          roughpy_compute/dense/basic/free_tensor_fma.hpp:13-53
//   ft_inplace_mul  roughpy_compute/dense/basic/free_tensor_inplace_mul.hpp:13-87
//   ft_antipode     roughpy_compute/dense/basic/free_tensor_antipode.hpp:45-100
//   ft_adj_lmul     roughpy_compute/dense/basic/free_tensor_adjoint_left_mul.hpp:10-77
//   ft_exp          roughpy_compute/dense/intermediate/free_tensor_exp.hpp:13-37
//   ft_fmexp        roughpy_compute/dense/intermediate/free_tensor_fmexp.hpp:15-40
//   ft_log          roughpy_compute/dense/intermediate/free_tensor_log.hpp:12-40
//   sparse map      roughpy_compute/dense/basic/apply_sparse_linear_map.hpp:16-55

#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <atomic>

#include "chen.hpp"
#include "chen_core.hpp"
#include "dense_ops.hpp"
#include "horner_fixed.hpp"
#include "kernel_util.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

constexpr int kBlock = 256;
constexpr size_t kLdsBudget = 64 * 1024;

struct Bounds {
    int32_t lo, hi;
};

int grid_for(int64_t batch)
{
    // Enough workgroups to fill 256 CUs several times over; the kernels
    // grid-stride over the rest of the batch.
    return int(std::min<int64_t>(batch, 8192));
}

// ---------------------------------------------------------------------------
// out += lhs (x) rhs
// ---------------------------------------------------------------------------
// Acc is the accumulator type: T itself, or double for float data in the
// compensated mode (products of two floats are exact in double, so each
// output coefficient is rounded to float once).
template <typename T, typename Acc, bool Stage>
__global__ __launch_bounds__(kBlock) void ft_fma_kernel(
        T* __restrict__ out, const T* __restrict__ lhs, const T* __restrict__ rhs,
        DevBasis b, Pow pw, Bounds ob, Bounds lb, Bounds rb, int64_t batch, int64_t os,
        int64_t ls, int64_t rs, int32_t top)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* sl = reinterpret_cast<T*>(smem);
    T* sr = sl + b.db[top];
    const int omin = ob.lo > 1 ? ob.lo : 1;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* l = lhs + e * ls;
        const T* r = rhs + e * rs;
        if (Stage) {
            // levels below `top` are re-read by many outputs: stage them.
            for (int64_t i = threadIdx.x; i < b.db[top]; i += kBlock) {
                sl[i] = l[i];
                sr[i] = r[i];
            }
            __syncthreads();
        }
        for (int od = ob.hi; od >= omin; --od) {
            int lmax = min(lb.hi, od - rb.lo);
            int lmin = max(lb.lo, od - rb.hi);
            if (lmax < lmin) continue;
            const uint32_t n = uint32_t(b.level_size(od));
            // With staging, only the two end terms of the top level read
            // global memory (lhs_top * rhs_0 and lhs_0 * rhs_top); every other
            // operand comes from LDS.  LDS and global reads use distinct
            // pointers so each compiles to its own (non-generic) load.
            const bool top_hi = Stage && od == top && lmax == top;
            const bool top_lo = Stage && od == top && lmin == 0;
            if (top_hi) --lmax;
            if (top_lo) ++lmin;
            // kU outputs per thread per pass: their global loads are issued
            // together before any arithmetic, so HBM latency overlaps.
            constexpr int kU = 4;
            for (uint32_t base = threadIdx.x; base < n; base += kU * kBlock) {
                Acc acc[kU];
                T lt[kU], rt[kU];
#pragma unroll
                for (int q = 0; q < kU; ++q) {
                    const uint32_t idx = base + q * kBlock;
                    const bool ok = idx < n;
                    acc[q] = ok ? Acc(o[b.db[od] + idx]) : Acc(0);
                    lt[q] = (ok && top_hi) ? l[b.db[od] + idx] : T(0);
                    rt[q] = (ok && top_lo) ? r[b.db[od] + idx] : T(0);
                }
#pragma unroll
                for (int q = 0; q < kU; ++q) {
                    const uint32_t idx = base + q * kBlock;
                    if (idx >= n) continue;
                    if (top_hi) acc[q] += Acc(lt[q]) * Acc(sr[0]);
                    for (int ld = lmax; ld >= lmin; --ld) {
                        const int rd = od - ld;
                        const uint32_t i = fdiv(idx, pw.pw[rd]);
                        const uint32_t j = idx - i * pw.pw[rd].d;
                        T lv, rv;
                        if (Stage) {
                            lv = sl[b.db[ld] + i];
                            rv = sr[b.db[rd] + j];
                        } else {
                            lv = l[b.db[ld] + i];
                            rv = r[b.db[rd] + j];
                        }
                        acc[q] += Acc(lv) * Acc(rv);
                    }
                    if (top_lo) acc[q] += Acc(sl[0]) * Acc(rt[q]);
                }
#pragma unroll
                for (int q = 0; q < kU; ++q) {
                    const uint32_t idx = base + q * kBlock;
                    if (idx < n) o[b.db[od] + idx] = T(acc[q]);
                }
            }
        }
        if (threadIdx.x == 0 && ob.lo == 0 && lb.lo == 0 && rb.lo == 0)
            o[0] = T(Acc(o[0]) + Acc(l[0]) * Acc(r[0]));
        if (Stage) __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// lhs <- lhs (x) rhs   (old lower levels of lhs are staged before writing)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void ft_inplace_mul_kernel(
        T* __restrict__ lhs, const T* __restrict__ rhs, DevBasis b, Pow pw, Bounds lb,
        Bounds rb, int64_t batch, int64_t ls, int64_t rs, int32_t top)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // Levels [0, max(top, 1)) are staged, so the unit is always in LDS.
    const int64_t nstage = b.db[top > 0 ? top : 1];
    T* sl = reinterpret_cast<T*>(smem);
    T* sr = sl + nstage;
    const int lo = lb.lo > 1 ? lb.lo : 1;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* l = lhs + e * ls;
        const T* r = rhs + e * rs;
        for (int64_t i = threadIdx.x; i < nstage; i += kBlock) {
            sl[i] = l[i];
            sr[i] = r[i];
        }
        __syncthreads();
        const T r0 = sr[0];
        for (int od = lb.hi; od >= lo; --od) {
            const int lmin = max(lb.lo, od - rb.hi);
            int lmax = od - max(1, rb.lo);
            const uint32_t n = uint32_t(b.level_size(od));
            // Levels below `top` come from LDS.  At the top level the old
            // coefficient and the rhs_top operand of the ld = 0 term are read
            // from global memory, through their own pointers.
            const bool at_top = od >= top;
            const bool top_lo = at_top && lmin == 0 && lmax >= 0;
            const int lmin_lds = top_lo ? 1 : lmin;
            // the two variants are separate code so each `old` load has a
            // known address space
            auto level = [&](auto at_top_c) {
                constexpr bool kTop = decltype(at_top_c)::value;
                for (uint32_t idx = threadIdx.x; idx < n; idx += kBlock) {
                    const T old = kTop ? l[b.db[od] + idx] : sl[b.db[od] + idx];
                    T acc = (rb.lo == 0) ? old * r0 : T(0);
                    for (int ld = lmax; ld >= lmin_lds; --ld) {
                        const int rd = od - ld;
                        const uint32_t i = fdiv(idx, pw.pw[rd]);
                        const uint32_t j = idx - i * pw.pw[rd].d;
                        acc += sl[b.db[ld] + i] * sr[b.db[rd] + j];
                    }
                    if (top_lo) acc += sl[0] * r[b.db[od] + idx];
                    l[b.db[od] + idx] = acc;
                }
            };
            if (at_top)
                level(std::true_type{});
            else
                level(std::false_type{});
        }
        if (threadIdx.x == 0) {
            if (lb.lo == 0 && rb.lo == 0)
                l[0] = sl[0] * r0;
            else if (lb.lo == 0)
                l[0] = T(0);
        }
        __syncthreads();
    }
}

// dst[e, i] = src[e, i] for keep_begin <= i < keep_end, else 0, over n
// coefficients per element (dst contiguous, src with batch stride ss): an
// operand whose degree bounds cut inside the levels a call reads.
template <typename T>
__global__ __launch_bounds__(kBlock) void truncate_copy_kernel(T* __restrict__ dst,
                                                              const T* __restrict__ src,
                                                              int64_t batch, int64_t n,
                                                              int64_t ss, int64_t keep_begin,
                                                              int64_t keep_end)
{
    const int64_t total = batch * n;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * kBlock) {
        const int64_t e = t / n;
        const int64_t i = t - e * n;
        dst[t] = (i >= keep_begin && i < keep_end) ? src[e * ss + i] : T(0);
    }
}

// ---------------------------------------------------------------------------
// Horner forms of exp / fmexp / log on one element held by one workgroup.
// `a` is the working tensor (LDS or the output buffer), `x` the argument.
// ---------------------------------------------------------------------------

// a.trunc(ml) <- (a.trunc(ml) (x) x.trunc(ml, min 1)) * f, top-down with a
// barrier per level; afterwards a[0] = 0 (free_tensor_inplace_mul.hpp:21-86
// with an implicitly zero unit on the right).
template <typename T>
RPY_FI void wg_mul_by_nounit(T* a, const T* x, const DevBasis& b, const Pow& pw, int ml, T f)
{
    for (int m = ml; m >= 1; --m) {
        const uint32_t n = uint32_t(b.level_size(m));
        for (uint32_t idx = threadIdx.x; idx < n; idx += kBlock) {
            T acc = T(0);
            for (int ld = m - 1; ld >= 0; --ld) {
                const int rd = m - ld;
                const uint32_t i = fdiv(idx, pw.pw[rd]);
                const uint32_t j = idx - i * pw.pw[rd].d;
                acc += (a[b.db[ld] + i] * x[b.db[rd] + j]) * f;
            }
            a[b.db[m] + idx] = acc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) a[0] = T(0);
    __syncthreads();
}

enum class Horner { Exp, FmExp, Log };

template <typename T, Horner H, bool InLds>
__global__ __launch_bounds__(kBlock) void horner_kernel(
        T* __restrict__ out, const T* __restrict__ mul, const T* __restrict__ arg, DevBasis b,
        Pow pw, int64_t batch, int64_t os, int64_t ms, int64_t as)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = b.depth;
    const int64_t n = b.size();
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* x = arg + e * as;
        const T* m = (H == Horner::FmExp) ? mul + e * ms : nullptr;
        T* a = InLds ? reinterpret_cast<T*>(smem) : o;
        const T* xs = x;
        if (InLds) {
            T* sx = reinterpret_cast<T*>(smem) + n;
            for (int64_t i = threadIdx.x; i < n; i += kBlock) sx[i] = x[i];
            xs = sx;
        }
        for (int64_t i = threadIdx.x; i < n; i += kBlock) {
            if (H == Horner::FmExp)
                a[i] = m[i];
            else
                a[i] = (H == Horner::Exp && i == 0) ? T(1) : T(0);
        }
        __syncthreads();
        for (int deg = D; deg > 0; --deg) {
            const int ml = D - deg + 1;
            if (H == Horner::Log) {
                if (threadIdx.x == 0) {
                    const T inv = T(1) / T(deg);
                    a[0] = (deg % 2 == 0) ? a[0] - inv : a[0] + inv;
                }
                __syncthreads();
                wg_mul_by_nounit(a, xs, b, pw, ml, T(1));
            } else {
                wg_mul_by_nounit(a, xs, b, pw, ml, T(1) / T(deg));
                if (H == Horner::Exp) {
                    if (threadIdx.x == 0) a[0] += T(1);
                } else {
                    for (int64_t i = threadIdx.x; i < n; i += kBlock) a[i] += m[i];
                }
                __syncthreads();
            }
        }
        if (InLds) {
            for (int64_t i = threadIdx.x; i < n; i += kBlock) o[i] = a[i];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// antipode: out[rev(w)] = (-1)^|w| arg[w]
// ---------------------------------------------------------------------------
// Writes the levels lo..hi of each element (the whole tensor when lo = 0
// and hi = depth); other levels of out are left as they are.
template <typename T>
__global__ __launch_bounds__(kBlock) void antipode_kernel(T* __restrict__ out,
                                                         const T* __restrict__ arg, DevBasis b,
                                                         Pow pw, int64_t batch, int64_t os,
                                                         int64_t as, int lo, int hi)
{
    const int64_t g0 = b.db[lo];
    const int64_t n = b.db[hi + 1] - g0;
    const int64_t total = batch * n;
    const uint32_t w = uint32_t(b.width);
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * kBlock) {
        const int64_t e = t / n;
        const int64_t g = g0 + (t - e * n);
        int d = 0;
        while (g >= b.db[d + 1]) ++d;
        uint32_t i = uint32_t(g - b.db[d]);
        uint32_t rev = 0;
        for (int k = 0; k < d; ++k) {
            const uint32_t q = fdiv(i, pw.pw[1]);
            rev = rev * w + (i - q * w);
            i = q;
        }
        const T v = arg[e * as + g];
        out[e * os + b.db[d] + rev] = (d & 1) ? -v : v;
    }
}

// ---------------------------------------------------------------------------
// out_k[i] += sum_{|u| = p} op_p[u] arg_{k+p}[u i]
// ---------------------------------------------------------------------------
// Degree ranges of the adjoint: output degrees olo..ohi, op degrees
// plo..phi, and argument degrees amin..amax, the latter computed as
// free_tensor_adjoint_left_mul.hpp:39-40 does from the three operands' bounds
// (amax = min(arg max - op min, out max), amin = max(arg min - op max,
// out min)).
struct AdjRange {
    int olo, ohi, plo, phi, amin, amax;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void adj_lmul_kernel(T* __restrict__ out,
                                                         const T* __restrict__ op,
                                                         const T* __restrict__ arg, DevBasis b,
                                                         int64_t batch, int64_t os, int64_t ps,
                                                         int64_t as, AdjRange r)
{
    const int64_t g0 = b.db[r.olo];
    const int64_t g1 = b.db[r.ohi + 1];
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const T* p = op + e * ps;
        const T* a = arg + e * as;
        T* o = out + e * os;
        for (int64_t g = g0 + threadIdx.x; g < g1; g += kBlock) {
            int od = 0;
            while (g >= b.db[od + 1]) ++od;
            const int64_t i = g - b.db[od];
            const int64_t osz = b.level_size(od);
            T acc = o[g];
            // argument degree descending == op degree descending
            const int pd_hi = min(r.phi, r.amax - od);
            const int pd_lo = max(r.plo, r.amin - od);
            for (int pd = pd_hi; pd >= pd_lo; --pd) {
                const int ad = od + pd;
                const int64_t psz = b.level_size(pd);
                for (int64_t u = 0; u < psz; ++u)
                    acc += p[b.db[pd] + u] * a[b.db[ad] + u * osz + i];
            }
            o[g] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// CSR sparse map: out[e, r] += scale * sum_p data[p] arg[e, idx[p]]
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void csr_map_kernel(T* __restrict__ out,
                                                        const T* __restrict__ arg,
                                                        const int64_t* __restrict__ offs,
                                                        const int64_t* __restrict__ idx,
                                                        const T* __restrict__ val, int64_t rows,
                                                        T scale, int64_t batch, int64_t os,
                                                        int64_t as, bool accumulate)
{
    const int64_t total = batch * rows;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * kBlock) {
        const int64_t e = t / rows;
        const int64_t r = t - e * rows;
        const T* a = arg + e * as;
        T acc = T(0);
        for (int64_t p = offs[r]; p < offs[r + 1]; ++p) acc += (val[p] * a[idx[p]]) * scale;
        out[e * os + r] = accumulate ? out[e * os + r] + acc : acc;
    }
}

// The same map with E tensors of the batch staged in LDS per workgroup: one
// thread per row walks the row's entries once and applies each to all E
// tensors, so the matrix is read from L2 once per E tensors and the gathers
// hit LDS.  Same per-row order and arithmetic as csr_map_kernel.
template <typename T, int E, int Block>
__global__ __launch_bounds__(Block) void csr_map_lds_kernel(T* __restrict__ out,
                                                            const T* __restrict__ arg,
                                                            const int64_t* __restrict__ offs,
                                                            const int64_t* __restrict__ idx,
                                                            const T* __restrict__ val,
                                                            int64_t rows, int64_t cols, T scale,
                                                            int64_t batch, int64_t os, int64_t as,
                                                            bool accumulate)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* s = reinterpret_cast<T*>(smem);
    // float32: column c of the E tensors sits at s[c E .. c E + E), so one
    // entry's E values are one vector read (t2l (5,5) 0.169 -> 0.149 ms, l2t
    // 0.141 -> 0.138 ms for 16384 tensors); float64 keeps tensor q at
    // s[q cols ..] (interleaved measured 0.265 -> 0.297 ms for t2l)
    constexpr bool kIL = sizeof(T) == 4;
    const FastDiv fcols = make_fastdiv(uint32_t(cols > 0 ? cols : 1));
    auto at = [&](int64_t c, int q) { return kIL ? c * E + q : int64_t(q) * cols + c; };
    for (int64_t e0 = int64_t(blockIdx.x) * E; e0 < batch; e0 += int64_t(gridDim.x) * E) {
        const int ne = int(batch - e0 < E ? batch - e0 : E);
        __syncthreads();  // the previous tensors' readers are done
        // eight loads per thread in flight before their LDS stores; a
        // contiguous batch is staged as one flat range
        if (as == cols) {
            const T* a = arg + e0 * as;
            const int64_t tot = int64_t(ne) * cols;
            for (int64_t j0 = 0; j0 < tot; j0 += 8 * Block) {
                T v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int64_t j = j0 + threadIdx.x + int64_t(k) * Block;
                    if (j < tot) v[k] = a[j];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int64_t j = j0 + threadIdx.x + int64_t(k) * Block;
                    if (j < tot) {
                        if constexpr (kIL) {
                            const uint32_t q = fdiv(uint32_t(j), fcols);
                            const uint32_t c = uint32_t(j) - q * uint32_t(cols);
                            s[c * E + q] = v[k];
                        } else {
                            s[j] = v[k];
                        }
                    }
                }
            }
        } else {
            for (int q = 0; q < ne; ++q) {
                const T* a = arg + (e0 + q) * as;
                for (int64_t c0 = 0; c0 < cols; c0 += 8 * Block) {
                    T v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int64_t c = c0 + threadIdx.x + int64_t(k) * Block;
                        if (c < cols) v[k] = a[c];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int64_t c = c0 + threadIdx.x + int64_t(k) * Block;
                        if (c < cols) s[at(c, q)] = v[k];
                    }
                }
            }
        }
        __syncthreads();
        // the output values a row adds to are loaded one row-pass ahead
        T ovn[E];
#pragma unroll
        for (int q = 0; q < E; ++q)
            ovn[q] = (accumulate && q < ne && int64_t(threadIdx.x) < rows)
                             ? out[(e0 + q) * os + threadIdx.x]
                             : T(0);
        for (int64_t r = threadIdx.x; r < rows; r += Block) {
            T ov[E], acc[E];
            const int64_t rn = r + Block;
#pragma unroll
            for (int q = 0; q < E; ++q) {
                ov[q] = ovn[q];
                ovn[q] = (accumulate && q < ne && rn < rows) ? out[(e0 + q) * os + rn] : T(0);
                acc[q] = T(0);
            }
            const int64_t pe = offs[r + 1];
            int64_t p = offs[r];
            // entries in groups of 4: the group's index and value loads are
            // issued together, then applied in order
            for (; p + 4 <= pe; p += 4) {
                int64_t c[4];
                T v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    c[j] = idx[p + j];
                    v[j] = val[p + j];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int q = 0; q < E; ++q) acc[q] += (v[j] * s[at(c[j], q)]) * scale;
            }
            // the last 1..3 entries: their loads also go out together
            if (p < pe) {
                int64_t c[3];
                T v[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    c[j] = p + j < pe ? idx[p + j] : 0;
                    v[j] = p + j < pe ? val[p + j] : T(0);
                }
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (p + j < pe) {
#pragma unroll
                        for (int q = 0; q < E; ++q) acc[q] += (v[j] * s[at(c[j], q)]) * scale;
                    }
            }
#pragma unroll
            for (int q = 0; q < E; ++q)
                if (q < ne) out[(e0 + q) * os + r] = accumulate ? ov[q] + acc[q] : acc[q];
        }
    }
}

// LDS index with bits 5..9 folded into bits 0..4: a 32-aligned block stays in
// place and power-of-two strides spread over the banks.
RPY_FI uint32_t swz(uint32_t j) { return j ^ ((j >> 5) & 31u); }

// antipode with the tensor staged in LDS: coalesced read, then each output
// word gathers its reversed word from LDS and is written coalesced (and the
// call may be in place).
template <typename T>
__global__ __launch_bounds__(kBlock) void antipode_lds_kernel(T* __restrict__ out,
                                                             const T* __restrict__ arg,
                                                             DevBasis b, Pow pw, int64_t batch,
                                                             int64_t os, int64_t as)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* sx = reinterpret_cast<T*>(smem);
    const uint32_t n = uint32_t(b.size());
    const uint32_t w = uint32_t(b.width);
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const T* a = arg + e * as;
        __syncthreads();  // the previous tensor's readers are done
        // eight loads per thread in flight before their LDS stores
        for (uint32_t j0 = 0; j0 < n; j0 += 8 * kBlock) {
            T v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t j = j0 + threadIdx.x + uint32_t(k) * kBlock;
                if (j < n) v[k] = a[j];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t j = j0 + threadIdx.x + uint32_t(k) * kBlock;
                if (j < n) sx[swz(j)] = v[k];
            }
        }
        __syncthreads();
        T* o = out + e * os;
        for (int d = 0; d <= b.depth; ++d) {
            const uint32_t base = uint32_t(b.db[d]);
            const uint32_t sz = uint32_t(b.level_size(d));
            for (uint32_t i = threadIdx.x; i < sz; i += kBlock) {
                uint32_t x = i, rev = 0;
                for (int k = 0; k < d; ++k) {
                    const uint32_t q = fdiv(x, pw.pw[1]);
                    rev = rev * w + (x - q * w);
                    x = q;
                }
                const T v = sx[swz(base + rev)];
                o[base + i] = (d & 1) ? -v : v;
            }
        }
    }
}

// Work split of the adjoint of left multiplication,
// out(v) += sum_{u} op(u) arg(u v): output level od sums db[D - od + 1]
// terms per word, so the words of low levels are split over groups of
// g[od] threads (a power of two <= 64, aligned inside a wave) whose partial
// sums meet in a fixed butterfly; vbase[od] is the level's first virtual
// thread.
// Levels whose words need no group (g = 1) give each virtual thread the W
// words q c, c < W, when W <= kAdjTile (tile[od] = W): the terms of a tile
// share their op value and read W contiguous arg values.
constexpr int kAdjTile = 8;

struct AdjLPlan {
    int g[kMaxDepth + 1];
    int tile[kMaxDepth + 1];
    int vbase[kMaxDepth + 2];
};

inline AdjLPlan make_adjl_plan(const DevBasis& b)
{
    AdjLPlan p{};
    int64_t v = 0;
    for (int od = 0; od <= b.depth; ++od) {
        const int64_t terms = b.db[b.depth - od + 1];
        int g = 1;
        while (g < 64 && terms / (2 * g) >= 16) g *= 2;
        p.g[od] = g;
        p.tile[od] = (g == 1 && od >= 1 && b.width <= kAdjTile) ? b.width : 1;
        p.vbase[od] = int(v);
        v += int64_t(g) * b.level_size(od) / p.tile[od];
    }
    p.vbase[b.depth + 1] = int(v);
    return p;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void adj_lmul_lds_kernel(T* __restrict__ out,
                                                             const T* __restrict__ op,
                                                             const T* __restrict__ arg,
                                                             DevBasis b, Pow pw, AdjLPlan plan,
                                                             int64_t batch, int64_t os,
                                                             int64_t ps, int64_t as)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t n = uint32_t(b.size());
    const uint32_t np = (n + 3) / 4 * 4;
    T* P = reinterpret_cast<T*>(smem);
    T* A = P + np;
    const int D = b.depth;
    const int V = plan.vbase[D + 1];
    const int Vr = (V + kBlock - 1) / kBlock * kBlock;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const T* p = op + e * ps;
        const T* a = arg + e * as;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < n; j += kBlock) {
            P[j] = p[j];
            A[j] = a[j];
        }
        __syncthreads();
        T* o = out + e * os;
        for (int vt = threadIdx.x; vt < Vr; vt += kBlock) {
            // every lane of a wave runs the same iterations; lanes past V
            // hold zero and write nothing
            int od = 0;
            while (od < D && vt >= plan.vbase[od + 1]) ++od;
            const bool live = vt < V;
            const int g = live ? plan.g[od] : 1;
            const int r = live ? vt - plan.vbase[od] : 0;
            if (live && plan.tile[od] > 1) {
                // W words q c of level od, every term of each
                const uint32_t wod = pw.pw[od].d;
                const uint32_t q = uint32_t(r);
                const uint32_t W = uint32_t(b.width);
                T accs[kAdjTile];
#pragma unroll
                for (int c = 0; c < kAdjTile; ++c) accs[c] = T(0);
                for (int pd = D - od; pd >= 0; --pd) {
                    const uint32_t psz = uint32_t(b.level_size(pd));
                    const T* Pp = P + b.db[pd];
                    const T* Ap = A + b.db[od + pd] + q * W;
                    for (uint32_t u = 0; u < psz; ++u) {
                        const T pv = Pp[u];
                        const T* ar = Ap + int64_t(u) * wod;
#pragma unroll
                        for (int c = 0; c < kAdjTile; ++c)
                            if (uint32_t(c) < W) accs[c] += pv * ar[c];
                    }
                }
                T* oo = o + b.db[od] + q * W;
#pragma unroll
                for (int c = 0; c < kAdjTile; ++c)
                    if (uint32_t(c) < W) oo[c] += accs[c];
            }
            const bool grouped = live && plan.tile[od] == 1;
            const uint32_t vw = uint32_t(r / g);        // output word of level od
            const int k = r - int(vw) * g;               // member of its group
            T acc = T(0);
            if (grouped) {
                const uint32_t wod = pw.pw[od].d;       // W^od
                const int64_t terms = b.db[D - od + 1];
                const int64_t chunk = (terms + g - 1) / g;
                int64_t f = int64_t(k) * chunk;
                int64_t fe = f + chunk < terms ? f + chunk : terms;
                // flat term f -> (pd, u), pd descending from D - od
                int pd = D - od;
                while (pd > 0 && f >= b.level_size(pd)) {
                    f -= b.level_size(pd);
                    fe -= b.level_size(pd);
                    --pd;
                }
                uint32_t u = uint32_t(f);
                for (int64_t t = fe - f; t > 0; --t) {
                    acc += P[b.db[pd] + u] * A[b.db[od + pd] + int64_t(u) * wod + vw];
                    if (++u == uint32_t(b.level_size(pd)) && pd > 0) {
                        --pd;
                        u = 0;
                    }
                }
            }
            // group butterfly (groups are aligned powers of two inside a wave),
            // skipped by waves with no group
            const int gw = grouped ? g : 1;
            if (__any(gw > 1)) {
                for (int off = 1; off < 64; off *= 2) {
                    const T other = __shfl_xor(acc, off, 64);
                    if (off < gw) acc += other;
                }
            }
            if (grouped && k == 0) o[b.db[od] + vw] += acc;
        }
    }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// max_degree -1 or above the depth means the depth, as update_algebra_params
// (roughpy/compute/_src/call_config.cpp:11-31) resolves it; a min_degree
// above the resolved max_degree (or a negative one) is an error.
int check_bounds(const DevBasis& b, rpy_degree_bounds in, Bounds& out, const char* name)
{
    out.lo = in.min_degree;
    out.hi = (in.max_degree == -1 || in.max_degree > b.depth) ? b.depth : in.max_degree;
    if (out.lo < 0 || out.hi < 0 || out.lo > out.hi)
        return fail(RPY_ERR_INVALID_ARGUMENT,
                    "invalid degree bounds for %s: min_degree %d, max_degree %d (depth %d)",
                    name, in.min_degree, in.max_degree, b.depth);
    return RPY_OK;
}

int check_batch(int64_t batch, std::initializer_list<int64_t> strides, int64_t need)
{
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch size");
    for (int64_t s : strides)
        if (batch > 1 && s < need)
            return fail(RPY_ERR_SHAPE, "batch stride %lld smaller than the tensor size %lld",
                        (long long)s, (long long)need);
    return RPY_OK;
}

// ---------------------------------------------------------------------------
// out += lhs (x) rhs at full depth for a compile-time shape: level offsets and
// divisors are constants, every term is unrolled and all loads of an output
// batch are issued before its arithmetic.  Levels below D are staged in LDS
// already converted to the accumulator type; the top level of lhs and rhs
// and every level of out are read once from global memory.  Same summation
// order as ft_fma_kernel (free_tensor_fma.hpp:13-53).
// ---------------------------------------------------------------------------
// One output of level m: lhs degree m (rhs unit), then m-1 .. 1, then 0 (lhs unit).
template <typename Acc, int W, int m>
RPY_FI Acc fma_output(Acc acc, Acc top_l, Acc top_r, int idx, const Acc* sl, const Acc* sr)
{
    Acc a = acc + top_l * sr[0];
    Acc lv[m > 1 ? m - 1 : 1], rv[m > 1 ? m - 1 : 1];
    chen::static_for<m - 1>([&](auto k0) {
        constexpr int ld = m - 1 - int(decltype(k0)::value);
        constexpr int rd = m - ld;
        constexpr int wr = int(ipow(W, rd));
        const int i = idx / wr;
        lv[k0] = sl[fdb<W>(ld) + i];
        rv[k0] = sr[fdb<W>(rd) + (idx - i * wr)];
    });
    chen::static_for<m - 1>([&](auto k0) { a += lv[k0] * rv[k0]; });
    return a + sl[0] * top_r;
}

// Accum = false computes out = lhs (x) rhs without reading out; with out equal
// to lhs that is the in-place product: each output word is written by the
// thread that read its own lhs top-level value, lower lhs levels come from the
// LDS copy, and the unit term reads lhs[0] before out[0] is written.
template <typename T, typename Acc, int W, int D, bool Accum = true>
__global__ __launch_bounds__(kBlock) void ft_fma_fixed_kernel(T* out, const T* lhs,
                                                             const T* __restrict__ rhs,
                                                             int64_t batch, int64_t os,
                                                             int64_t ls, int64_t rs)
{
    constexpr int kStaged = fdb<W>(D);
    constexpr int kN = int(ipow(W, D));          // top-level outputs
    constexpr int kTop = (kN + kBlock - 1) / kBlock;
    constexpr int kTopU = kTop < 16 ? kTop : 16;  // top outputs held per thread per pass
    constexpr int dbD = fdb<W>(D);
    __shared__ Acc sl[kStaged];
    __shared__ Acc sr[kStaged];
    constexpr int kU = 4;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const T* l = lhs + e * ls;
        const T* r = rhs + e * rs;
        // the first top-level pass's loads go out before the staging loads
        T t_o[kTopU], t_l[kTopU], t_r[kTopU];
        auto load_top = [&](int base) {
#pragma unroll
            for (int q = 0; q < kTopU; ++q) {
                const int idx = base + q * kBlock;
                const bool ok = idx < kN;
                t_o[q] = (ok && Accum) ? o[dbD + idx] : T(0);
                t_l[q] = ok ? l[dbD + idx] : T(0);
                t_r[q] = ok ? r[dbD + idx] : T(0);
            }
        };
        load_top(threadIdx.x);
        // staging loads in groups of up to 8 per operand, each group's loads
        // issued before its LDS stores
        constexpr int kSIt = (kStaged + kBlock - 1) / kBlock;
        constexpr int kSG = kSIt < 8 ? kSIt : 8;
        chen::static_for<(kSIt + kSG - 1) / kSG>([&](auto g0) {
            constexpr int g = int(decltype(g0)::value);
            T lv[kSG], rv[kSG];
            chen::static_for<kSG>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + (g * kSG + k) * kBlock;
                if (g * kSG + k < kSIt && i < kStaged) {
                    lv[k] = l[i];
                    rv[k] = r[i];
                }
            });
            chen::static_for<kSG>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + (g * kSG + k) * kBlock;
                if (g * kSG + k < kSIt && i < kStaged) {
                    sl[i] = Acc(lv[k]);
                    sr[i] = Acc(rv[k]);
                }
            });
        });
        __syncthreads();
        for (int base = threadIdx.x; base < kN; base += kTopU * kBlock) {
            if (base != int(threadIdx.x)) load_top(base);
            // the operands are in registers already; two outputs at a time keep
            // the LDS loads of a pair in flight without inflating the VGPRs
#pragma unroll 2
            for (int q = 0; q < kTopU; ++q) {
                const int idx = base + q * kBlock;
                if (idx < kN)
                    o[dbD + idx] = T(fma_output<Acc, W, D>(Acc(t_o[q]), Acc(t_l[q]), Acc(t_r[q]),
                                                            idx, sl, sr));
            }
        }
        chen::static_for<D - 1>([&](auto m0) {
            constexpr int m = D - 1 - int(decltype(m0)::value);
            constexpr int n = int(ipow(W, m));
            constexpr int dbm = fdb<W>(m);
            for (int base = threadIdx.x; base < n; base += kU * kBlock) {
                Acc acc[kU];
#pragma unroll
                for (int q = 0; q < kU; ++q) {
                    const int idx = base + q * kBlock;
                    acc[q] = (idx < n && Accum) ? Acc(o[dbm + idx]) : Acc(0);
                }
#pragma unroll
                for (int q = 0; q < kU; ++q) {
                    const int idx = base + q * kBlock;
                    if (idx < n)
                        o[dbm + idx] = T(fma_output<Acc, W, m>(acc[q], sl[dbm + idx],
                                                               sr[dbm + idx], idx, sl, sr));
                }
            }
        });
        if (threadIdx.x == 0) {
            const Acc prod = Acc(l[0]) * Acc(r[0]);
            o[0] = Accum ? T(Acc(o[0]) + prod) : T(prod);
        }
        __syncthreads();
    }
}

template <typename T, typename Acc, int W, int D>
void launch_fma_fixed(void* out, const void* lhs, const void* rhs, int64_t batch, int64_t os,
                      int64_t ls, int64_t rs, hipStream_t st)
{
    hipLaunchKernelGGL((ft_fma_fixed_kernel<T, Acc, W, D>), dim3(grid_for(batch)),
                       dim3(kBlock), 0, st, (T*)out, (const T*)lhs, (const T*)rhs, batch, os, ls,
                       rs);
}

using FmaFixed = void (*)(void*, const void*, const void*, int64_t, int64_t, int64_t, int64_t,
                          hipStream_t);

// in place: lhs <- lhs (x) rhs at full depth
template <typename T, int W, int D>
void launch_inplace_fixed(void* lhs, const void* rhs, int64_t batch, int64_t ls, int64_t rs,
                          hipStream_t st)
{
    hipLaunchKernelGGL((ft_fma_fixed_kernel<T, T, W, D, false>), dim3(grid_for(batch)),
                       dim3(kBlock), 0, st, (T*)lhs, (const T*)lhs, (const T*)rhs, batch, ls, ls,
                       rs);
}

using InplaceFixed = void (*)(void*, const void*, int64_t, int64_t, int64_t, hipStream_t);

// out = lhs (x) rhs at full depth, out not read (a buffer of its own)
template <typename T, int W, int D>
void launch_mul_fixed(void* out, const void* lhs, const void* rhs, int64_t batch, int64_t os,
                      int64_t ls, int64_t rs, hipStream_t st)
{
    hipLaunchKernelGGL((ft_fma_fixed_kernel<T, T, W, D, false>), dim3(grid_for(batch)),
                       dim3(kBlock), 0, st, (T*)out, (const T*)lhs, (const T*)rhs, batch, os, ls,
                       rs);
}

template <typename T>
FmaFixed find_mul_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_mul_fixed<T, W, D>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

template <typename T>
InplaceFixed find_inplace_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_inplace_fixed<T, W, D>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

template <typename T, typename Acc>
FmaFixed find_fma_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_fma_fixed<T, Acc, W, D>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

template <typename T, typename Acc = T>
int launch_fma(void* out, const void* lhs, const void* rhs, const DevBasis& b, Bounds ob,
               Bounds lb, Bounds rb, int64_t batch, int64_t os, int64_t ls, int64_t rs,
               hipStream_t st)
{
    const bool full = ob.lo == 0 && lb.lo == 0 && rb.lo == 0 && ob.hi == b.depth &&
                      lb.hi == b.depth && rb.hi == b.depth;
    if (full) {
        if (FmaFixed f = find_fma_fixed<T, Acc>(b.width, b.depth)) {
            f(out, lhs, rhs, batch, os, ls, rs, st);
            RPY_HIP_CHECK(hipGetLastError());
            return RPY_OK;
        }
    }
    const Pow pw = make_pow(b.width, b.depth);
    const int top = ob.hi;
    const size_t lds = size_t(2) * size_t(b.db[top]) * sizeof(T);
    dim3 grid(grid_for(batch)), block(kBlock);
    if (lds <= kLdsBudget) {
        hipLaunchKernelGGL((ft_fma_kernel<T, Acc, true>), grid, block, lds, st, (T*)out,
                           (const T*)lhs, (const T*)rhs, b, pw, ob, lb, rb, batch, os, ls, rs,
                           top);
    } else {
        hipLaunchKernelGGL((ft_fma_kernel<T, Acc, false>), grid, block, 0, st, (T*)out,
                           (const T*)lhs, (const T*)rhs, b, pw, ob, lb, rb, batch, os, ls, rs,
                           0);
    }
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

template <typename T>
int launch_inplace(void* lhs, const void* rhs, const DevBasis& b, Bounds lb, Bounds rb,
                   int64_t batch, int64_t ls, int64_t rs, hipStream_t st)
{
    const Pow pw = make_pow(b.width, b.depth);
    const int top = lb.hi;
    const size_t lds = size_t(2) * size_t(b.db[top > 0 ? top : 1]) * sizeof(T);
    if (lds > kLdsBudget)
        return fail(RPY_ERR_UNSUPPORTED,
                    "in-place multiplication needs %zu bytes of LDS (limit %zu)", lds,
                    kLdsBudget);
    dim3 grid(grid_for(batch)), block(kBlock);
    hipLaunchKernelGGL((ft_inplace_mul_kernel<T>), grid, block, lds, st, (T*)lhs,
                       (const T*)rhs, b, pw, lb, rb, batch, ls, rs, top);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

// ---------------------------------------------------------------------------
// exp / fmexp / log for a compile-time shape.  The same truncated in-place
// Horner as horner_kernel (products a <- a (x) x with x's unit dropped, levels
// top-down, each output summed over lhs degrees high to low as
// (a x) f), but every level offset and divisor is a constant and the thread
// of a prefix u computes the W outputs u c together: for lhs degree ld the
// lhs word is the first ld letters of u (one read) and the rhs words are
// the rest of u followed by c (W contiguous reads).
// ---------------------------------------------------------------------------

template <typename T, Horner H, int W, int D, typename Tin = T>
__global__ __launch_bounds__(kBlock) void horner_fixed_kernel(T* __restrict__ out,
                                                             const T* __restrict__ mul,
                                                             const Tin* __restrict__ arg,
                                                             int64_t batch, int64_t os, int64_t ms,
                                                             int64_t as)
{
    // levels 0..D-1 of the accumulator and of x live in LDS; only the last
    // product writes the top level, and it writes it to out directly
    constexpr uint32_t ns = uint32_t(fdb<W>(D));
    constexpr uint32_t np = (ns + 15) / 16 * 16;
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    T* xs = a + np;
    T* mls = xs + np;  // fmexp: levels 0..D-1 of the multiplier, added after every step
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        T* o = out + e * os;
        const Tin* x = arg + e * as;
        const T* m = (H == Horner::FmExp) ? mul + e * ms : nullptr;
        // every staging load of the tensor goes out before its LDS stores
        constexpr int kIt = int((ns + kBlock - 1) / kBlock);
        Tin xv[kIt];
        T mv[H == Horner::FmExp ? kIt : 1];
        chen::static_for<kIt>([&](auto k0) {
            constexpr int k = int(decltype(k0)::value);
            const uint32_t i = threadIdx.x + uint32_t(k) * kBlock;
            if (i < ns) {
                xv[k] = x[i];
                if constexpr (H == Horner::FmExp) mv[k] = m[i];
            }
        });
        chen::static_for<kIt>([&](auto k0) {
            constexpr int k = int(decltype(k0)::value);
            const uint32_t i = threadIdx.x + uint32_t(k) * kBlock;
            if (i < ns) {
                xs[i] = T(xv[k]);
                if constexpr (H == Horner::FmExp) {
                    a[i] = mv[k];
                    mls[i] = mv[k];
                } else {
                    a[i] = (H == Horner::Exp && i == 0) ? T(1) : T(0);
                }
            }
        });
        __syncthreads();
        chen::static_for<D>([&](auto di) {
            constexpr int deg = D - int(decltype(di)::value);
            constexpr int ml = D - deg + 1;
            if constexpr (H == Horner::Log) {
                if (threadIdx.x == 0) {
                    const T inv = T(1) / T(deg);
                    a[0] = (deg % 2 == 0) ? a[0] - inv : a[0] + inv;
                }
                __syncthreads();
                wg_mul_by_nounit_fixed<T, W, D, ml>(a, xs, T(1), x, o, nullptr);
            } else {
                wg_mul_by_nounit_fixed<T, W, D, ml>(a, xs, T(1) / T(deg), x, o,
                                                    H == Horner::FmExp ? m : nullptr);
                if constexpr (H == Horner::Exp) {
                    if (threadIdx.x == 0) a[0] += T(1);
                } else {
                    for (uint32_t i = threadIdx.x; i < ns; i += kBlock) a[i] += mls[i];
                }
                __syncthreads();
            }
        });
        for (uint32_t i = threadIdx.x; i < ns; i += kBlock) o[i] = a[i];
        __syncthreads();
    }
}

template <typename T, Horner H, int W, int D, typename Tin = T>
bool launch_horner_fixed(void* out, const void* mul, const void* arg, int64_t batch, int64_t os,
                         int64_t ms, int64_t as, hipStream_t st)
{
    constexpr size_t np = (size_t(fdb<W>(D)) + 15) / 16 * 16;
    constexpr size_t lds = (H == Horner::FmExp ? 3 : 2) * np * sizeof(T);
    if constexpr (lds > kLdsBudget) {
        return false;
    } else {
        hipLaunchKernelGGL((horner_fixed_kernel<T, H, W, D, Tin>), dim3(grid_for(batch)),
                           dim3(kBlock), lds, st, (T*)out, (const T*)mul, (const Tin*)arg, batch,
                           os, ms, as);
        return true;
    }
}

template <typename T, Horner H>
using HornerFixed = bool (*)(void*, const void*, const void*, int64_t, int64_t, int64_t, int64_t,
                             hipStream_t);

template <typename T, Horner H, typename Tin = T>
HornerFixed<T, H> find_horner_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_horner_fixed<T, H, W, D, Tin>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

template <typename T, Horner H>
int launch_horner(void* out, const void* mul, const void* arg, const DevBasis& b,
                  int64_t batch, int64_t os, int64_t ms, int64_t as, hipStream_t st)
{
    if (auto f = find_horner_fixed<T, H>(b.width, b.depth)) {
        if (f(out, mul, arg, batch, os, ms, as, st)) {
            RPY_HIP_CHECK(hipGetLastError());
            return RPY_OK;
        }
    }
    const Pow pw = make_pow(b.width, b.depth);
    const size_t lds = size_t(2) * size_t(b.size()) * sizeof(T);
    dim3 grid(grid_for(batch)), block(kBlock);
    if (lds <= kLdsBudget) {
        hipLaunchKernelGGL((horner_kernel<T, H, true>), grid, block, lds, st, (T*)out,
                           (const T*)mul, (const T*)arg, b, pw, batch, os, ms, as);
    } else {
        hipLaunchKernelGGL((horner_kernel<T, H, false>), grid, block, 0, st, (T*)out,
                           (const T*)mul, (const T*)arg, b, pw, batch, os, ms, as);
    }
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

int dtype_size(rpy_dtype dt)
{
    switch (dt) {
        case RPY_F32: return 4;
        case RPY_F64: return 8;
        default: return 0;
    }
}

int plain_dtype(rpy_dtype dt)
{
    if (dt != RPY_F32 && dt != RPY_F64)
        return fail(RPY_ERR_DTYPE, "only float32 and float64 are supported for this operation");
    return RPY_OK;
}

int64_t blocks_elementwise(int64_t total)
{
    return std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 65536));
}

}  // namespace

int dense_ft_fma(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                 const rpy_tensor_basis* basis, rpy_degree_bounds obi, rpy_degree_bounds lbi,
                 rpy_degree_bounds rbi, int64_t batch, int64_t os, int64_t ls, int64_t rs,
                 void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b))) return rc;
    if (dtype != RPY_F32_COMPENSATED && (rc = plain_dtype(dtype))) return rc;
    Bounds ob, lb, rb;
    if ((rc = check_bounds(b, obi, ob, "out")) || (rc = check_bounds(b, lbi, lb, "lhs")) ||
        (rc = check_bounds(b, rbi, rb, "rhs")))
        return rc;
    if ((rc = check_batch(batch, {os, ls, rs}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !lhs || !rhs) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    switch (dtype) {
        case RPY_F64:
            return launch_fma<double>(out, lhs, rhs, b, ob, lb, rb, batch, os, ls, rs,
                                      as_stream(stream));
        case RPY_F32_COMPENSATED:
            return launch_fma<float, double>(out, lhs, rhs, b, ob, lb, rb, batch, os, ls, rs,
                                             as_stream(stream));
        default:
            return launch_fma<float>(out, lhs, rhs, b, ob, lb, rb, batch, os, ls, rs,
                                     as_stream(stream));
    }
}

int dense_ft_inplace_mul(void* lhs, const void* rhs, rpy_dtype dtype,
                         const rpy_tensor_basis* basis, rpy_degree_bounds lbi,
                         rpy_degree_bounds rbi, int64_t batch, int64_t ls, int64_t rs,
                         void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b)) || (rc = plain_dtype(dtype))) return rc;
    Bounds lb, rb;
    if ((rc = check_bounds(b, lbi, lb, "lhs")) || (rc = check_bounds(b, rbi, rb, "rhs")))
        return rc;
    if ((rc = check_batch(batch, {ls, rs}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!lhs || !rhs) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    if (lb.lo == 0 && rb.lo == 0 && lb.hi == b.depth && rb.hi == b.depth && lhs != rhs) {
        const InplaceFixed f = dtype == RPY_F64 ? find_inplace_fixed<double>(b.width, b.depth)
                                                : find_inplace_fixed<float>(b.width, b.depth);
        if (f) {
            f(lhs, rhs, batch, ls, rs, as_stream(stream));
            RPY_HIP_CHECK(hipGetLastError());
            return RPY_OK;
        }
    }
    return dtype == RPY_F64 ? launch_inplace<double>(lhs, rhs, b, lb, rb, batch, ls, rs,
                                                     as_stream(stream))
                            : launch_inplace<float>(lhs, rhs, b, lb, rb, batch, ls, rs,
                                                    as_stream(stream));
}

int dense_ft_antipode(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                      rpy_degree_bounds obi, rpy_degree_bounds abi, int64_t batch, int64_t os,
                      int64_t as, void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b)) || (rc = plain_dtype(dtype))) return rc;
    Bounds ob, ab;
    if ((rc = check_bounds(b, obi, ob, "out")) || (rc = check_bounds(b, abi, ab, "arg")))
        return rc;
    if ((rc = check_batch(batch, {os, as}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !arg) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const Pow pw = make_pow(b.width, 1);
    // Levels written: max(mins) .. min(maxes); level 1 whenever that range
    // starts at or below 1 (free_tensor_antipode.hpp:60-80 writes it
    // unconditionally then).
    const int lo = std::max(ob.lo, ab.lo);
    int hi = std::min(ob.hi, ab.hi);
    if (lo <= 1 && b.depth >= 1) hi = std::max(hi, 1);
    if (lo > hi) return RPY_OK;
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    const size_t lds = size_t((b.size() + 31) / 32 * 32) * es;
    if (lo == 0 && hi == b.depth && lds <= kLdsBudget) {
        dim3 grid(grid_for(batch)), block(kBlock);
        if (dtype == RPY_F64)
            hipLaunchKernelGGL(antipode_lds_kernel<double>, grid, block, lds, as_stream(stream),
                               (double*)out, (const double*)arg, b, pw, batch, os, as);
        else
            hipLaunchKernelGGL(antipode_lds_kernel<float>, grid, block, lds, as_stream(stream),
                               (float*)out, (const float*)arg, b, pw, batch, os, as);
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    dim3 grid(blocks_elementwise(batch * (b.db[hi + 1] - b.db[lo]))), block(kBlock);
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(antipode_kernel<double>, grid, block, 0, as_stream(stream),
                           (double*)out, (const double*)arg, b, pw, batch, os, as, lo, hi);
    else
        hipLaunchKernelGGL(antipode_kernel<float>, grid, block, 0, as_stream(stream),
                           (float*)out, (const float*)arg, b, pw, batch, os, as, lo, hi);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of left multiplication for a compile-time shape,
// out(v) += sum_{u} op(u) arg(u v), both operands staged in LDS.
// Output level od >= 1 is cut into tiles of the W words q c (c < W): the terms
// of a tile are the op words u of degrees pd = D-od .. 0 (u ascending within
// a degree, the reference's order, free_tensor_adjoint_left_mul.hpp:47-74),
// each one op value times W contiguous arg values.  Level 0 is the dot
// product of op and arg.  A tile (or level 0) with many terms is split over
// an aligned group of G threads (a power of two <= 64) whose partial sums
// meet in a butterfly; groups of all levels are packed into one virtual
// thread range, heaviest levels first.
// ---------------------------------------------------------------------------
template <int W, int D>
constexpr int adjf_terms(int od)
{
    return fdb<W>(D - od + 1);  // op words of degree 0 .. D - od
}

template <int W, int D>
constexpr int adjf_group(int od)
{
    const int t = adjf_terms<W, D>(od);
    int g = 1;
    while (g < 64 && (t + 2 * g - 1) / (2 * g) >= 8) g *= 2;
    return g;
}

template <int W>
constexpr int adjf_tiles(int od)
{
    return od == 0 ? 1 : int(ipow(W, od - 1));
}

template <int W, int D>
constexpr int adjf_vbase(int od)
{
    int v = 0;
    for (int j = 0; j < od; ++j) v += adjf_group<W, D>(j) * adjf_tiles<W>(j);
    return v;
}

template <typename T, int W, int D>
__global__ __launch_bounds__(kBlock) void adj_lmul_fixed_kernel(T* __restrict__ out,
                                                               const T* __restrict__ op,
                                                               const T* __restrict__ arg,
                                                               int64_t batch, int64_t os,
                                                               int64_t ps, int64_t as)
{
    constexpr int kN = fdb<W>(D + 1);
    constexpr int kNP = (kN + 3) / 4 * 4;
    constexpr int kPer = (kN + kBlock - 1) / kBlock;   // staged values per thread
    constexpr int kGrp = kPer < 16 ? kPer : 16;        // values per staging group
    constexpr int kGroups = (kPer + kGrp - 1) / kGrp;
    constexpr int kV = adjf_vbase<W, D>(D + 1);
    constexpr int kPasses = (kV + kBlock - 1) / kBlock;
    extern __shared__ __align__(16) unsigned char smem[];
    T* P = reinterpret_cast<T*>(smem);
    T* A = P + kNP;
    for (int64_t e = blockIdx.x; e < batch; e += gridDim.x) {
        const T* p = op + e * ps;
        const T* a = arg + e * as;
        T* o = out + e * os;
        // the global loads of a staging group (and of the output words this
        // thread adds to) go out together, ahead of their LDS stores
        T sp[kGrp], sa[kGrp];
        auto load_grp = [&](auto g0) {
            constexpr int g = int(decltype(g0)::value);
            chen::static_for<kGrp>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + (g * kGrp + k) * kBlock;
                if (g * kGrp + k < kPer && i < kN) {
                    sp[k] = p[i];
                    sa[k] = a[i];
                }
            });
        };
        auto store_grp = [&](auto g0) {
            constexpr int g = int(decltype(g0)::value);
            chen::static_for<kGrp>([&](auto k0) {
                constexpr int k = int(decltype(k0)::value);
                const int i = threadIdx.x + (g * kGrp + k) * kBlock;
                if (g * kGrp + k < kPer && i < kN) {
                    P[i] = sp[k];
                    A[i] = sa[k];
                }
            });
        };
        load_grp(std::integral_constant<int, 0>{});
        // pass ps_ covers virtual threads ps_ * kBlock + threadIdx.x; a level
        // takes part in a pass only if their ranges meet (known at compile time)
        T ov[kPasses][W];
        int dst[kPasses];
        chen::static_for<kPasses>([&](auto p0) {
            constexpr int ps_ = int(decltype(p0)::value);
            const int vt = ps_ * kBlock + int(threadIdx.x);
            dst[ps_] = -1;
            chen::static_for<D + 1>([&](auto od0) {
                constexpr int od = int(decltype(od0)::value);
                constexpr int G = adjf_group<W, D>(od);
                constexpr int vb = adjf_vbase<W, D>(od);
                constexpr int nv = G * adjf_tiles<W>(od);
                if constexpr (vb < (ps_ + 1) * kBlock && vb + nv > ps_ * kBlock) {
                    if (vt >= vb && vt < vb + nv) {
                        const int r = vt - vb;
                        const int q = r / G;
                        if (r - q * G == 0) dst[ps_] = od == 0 ? 0 : fdb<W>(od) + q * W;
                    }
                }
            });
            const int lvl_w = dst[ps_] == 0 ? 1 : W;  // level 0 has one word
            chen::static_for<W>([&](auto c) {
                constexpr int ci = int(decltype(c)::value);
                ov[ps_][ci] = (dst[ps_] >= 0 && ci < lvl_w) ? o[dst[ps_] + ci] : T(0);
            });
        });
        __syncthreads();  // the previous tensor's readers are done
        store_grp(std::integral_constant<int, 0>{});
        chen::static_for<kGroups - 1>([&](auto g0) {
            constexpr int g = int(decltype(g0)::value) + 1;
            load_grp(std::integral_constant<int, g>{});
            store_grp(std::integral_constant<int, g>{});
        });
        __syncthreads();
        chen::static_for<kPasses>([&](auto p0) {
            constexpr int ps_ = int(decltype(p0)::value);
            const int vt = ps_ * kBlock + int(threadIdx.x);
            T acc[W];
            chen::static_for<W>([&](auto c) { acc[c] = T(0); });
            int gw = 1;  // the lane's group size
            chen::static_for<D + 1>([&](auto od0) {
                constexpr int od = int(decltype(od0)::value);
                constexpr int G = adjf_group<W, D>(od);
                constexpr int vb = adjf_vbase<W, D>(od);
                constexpr int nv = G * adjf_tiles<W>(od);
                constexpr int T_ = adjf_terms<W, D>(od);
                constexpr int chunk = (T_ + G - 1) / G;
                constexpr int wod = int(ipow(W, od));
                if constexpr (vb < (ps_ + 1) * kBlock && vb + nv > ps_ * kBlock) {
                    if (vt >= vb && vt < vb + nv) {
                        const int r = vt - vb;
                        const int q = r / G;          // tile
                        const int k = r - q * G;      // member of its group
                        const int f0 = k * chunk;
                        const int f1 = f0 + chunk < T_ ? f0 + chunk : T_;
                        // flat term f -> degree pd (descending from D - od), word u
                        chen::static_for<D - od + 1>([&](auto i0) {
                            constexpr int pd = D - od - int(decltype(i0)::value);
                            constexpr int F = fdb<W>(D - od + 1) - fdb<W>(pd + 1);
                            constexpr int S = int(ipow(W, pd));
                            const int lo = f0 - F > 0 ? f0 - F : 0;
                            const int hi = f1 - F < S ? f1 - F : S;
                            const T* Pp = P + fdb<W>(pd);
                            if constexpr (od == 0) {
                                const T* Ap = A + fdb<W>(pd);
                                for (int u = lo; u < hi; ++u) acc[0] += Pp[u] * Ap[u];
                            } else {
                                const T* Ap = A + fdb<W>(od + pd) + q * W;
                                for (int u = lo; u < hi; ++u) {
                                    const T pv = Pp[u];
                                    const T* ar = Ap + u * wod;
                                    chen::static_for<W>([&](auto c) {
                                        acc[c] += pv * ar[int(decltype(c)::value)];
                                    });
                                }
                            }
                        });
                        gw = G;
                    }
                }
            });
            // group butterfly (groups are aligned powers of two inside a
            // wave), skipped by waves with no group
            if (__any(gw > 1)) {
                for (int off = 1; off < 64; off *= 2) {
                    chen::static_for<W>([&](auto c) {
                        const T other = __shfl_xor(acc[c], off, 64);
                        if (off < gw) acc[c] += other;
                    });
                }
            }
            if (dst[ps_] >= 0) {
                const int lvl_w = dst[ps_] == 0 ? 1 : W;
                chen::static_for<W>([&](auto c) {
                    constexpr int ci = int(decltype(c)::value);
                    if (ci < lvl_w) o[dst[ps_] + ci] = ov[ps_][ci] + acc[ci];
                });
            }
        });
    }
}

template <typename T, int W, int D>
bool launch_adjl_fixed(void* out, const void* op, const void* arg, int64_t batch, int64_t os,
                       int64_t ps, int64_t as, hipStream_t st)
{
    constexpr size_t lds = 2 * size_t((fdb<W>(D + 1) + 3) / 4 * 4) * sizeof(T);
    if constexpr (lds > 128 * 1024) {
        return false;
    } else {
        if constexpr (lds > 64 * 1024) {
            // extended LDS (one workgroup per CU), reserved once per process
            static const hipError_t err = hipFuncSetAttribute(
                    (const void*)adj_lmul_fixed_kernel<T, W, D>,
                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (err != hipSuccess) return false;
        }
        hipLaunchKernelGGL((adj_lmul_fixed_kernel<T, W, D>), dim3(grid_for(batch)), dim3(kBlock),
                           lds, st, (T*)out, (const T*)op, (const T*)arg, batch, os, ps, as);
        return true;
    }
}

template <typename T>
using AdjlFixed = bool (*)(void*, const void*, const void*, int64_t, int64_t, int64_t, int64_t,
                           hipStream_t);

template <typename T>
AdjlFixed<T> find_adjl_fixed(int w, int d)
{
#define RPY_F(W, D) \
    if (w == W && d == D) return &launch_adjl_fixed<T, W, D>;
    RPY_F(2, 2) RPY_F(2, 3) RPY_F(2, 4) RPY_F(2, 5) RPY_F(2, 6) RPY_F(2, 7) RPY_F(2, 8)
    RPY_F(3, 2) RPY_F(3, 3) RPY_F(3, 4) RPY_F(3, 5) RPY_F(3, 6)
    RPY_F(4, 2) RPY_F(4, 3) RPY_F(4, 4) RPY_F(4, 5) RPY_F(4, 6)
    RPY_F(5, 2) RPY_F(5, 3) RPY_F(5, 4) RPY_F(5, 5)
    RPY_F(6, 2) RPY_F(6, 3) RPY_F(6, 4) RPY_F(6, 5)
    RPY_F(7, 2) RPY_F(7, 3) RPY_F(7, 4)
    RPY_F(8, 2) RPY_F(8, 3) RPY_F(8, 4)
#undef RPY_F
    return nullptr;
}

int dense_ft_adjoint_left_mul(void* out, const void* op, const void* arg, rpy_dtype dtype,
                              const rpy_tensor_basis* basis, rpy_degree_bounds obi,
                              rpy_degree_bounds pbi, rpy_degree_bounds abi, int64_t batch,
                              int64_t os, int64_t ps, int64_t as, void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b)) || (rc = plain_dtype(dtype))) return rc;
    Bounds ob, pb, ab;
    if ((rc = check_bounds(b, obi, ob, "out")) || (rc = check_bounds(b, pbi, pb, "op")) ||
        (rc = check_bounds(b, abi, ab, "arg")))
        return rc;
    if ((rc = check_batch(batch, {os, ps, as}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !op || !arg) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const bool full = ob.lo == 0 && pb.lo == 0 && ab.lo == 0 && ob.hi == b.depth &&
                      pb.hi == b.depth && ab.hi == b.depth;
    if (full) {
        const auto f64 = find_adjl_fixed<double>(b.width, b.depth);
        const auto f32 = find_adjl_fixed<float>(b.width, b.depth);
        const bool done = dtype == RPY_F64
                ? (f64 && f64(out, op, arg, batch, os, ps, as, as_stream(stream)))
                : (f32 && f32(out, op, arg, batch, os, ps, as, as_stream(stream)));
        if (done) {
            RPY_HIP_CHECK(hipGetLastError());
            return RPY_OK;
        }
    }
    dim3 grid(grid_for(batch)), block(kBlock);
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    const size_t lds = 2 * size_t((b.size() + 3) / 4 * 4) * es;
    if (full && lds <= kLdsBudget && b.depth <= kMaxDepth) {
        const Pow pw = make_pow(b.width, b.depth);
        const AdjLPlan plan = make_adjl_plan(b);
        if (dtype == RPY_F64)
            hipLaunchKernelGGL(adj_lmul_lds_kernel<double>, grid, block, lds, as_stream(stream),
                               (double*)out, (const double*)op, (const double*)arg, b, pw, plan,
                               batch, os, ps, as);
        else
            hipLaunchKernelGGL(adj_lmul_lds_kernel<float>, grid, block, lds, as_stream(stream),
                               (float*)out, (const float*)op, (const float*)arg, b, pw, plan,
                               batch, os, ps, as);
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    const AdjRange r{ob.lo, ob.hi, pb.lo, pb.hi, std::max(ab.lo - pb.hi, ob.lo),
                     std::min(ab.hi - pb.lo, ob.hi)};
    if (r.amin > r.amax) return RPY_OK;
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(adj_lmul_kernel<double>, grid, block, 0, as_stream(stream),
                           (double*)out, (const double*)op, (const double*)arg, b, batch, os, ps,
                           as, r);
    else
        hipLaunchKernelGGL(adj_lmul_kernel<float>, grid, block, 0, as_stream(stream),
                           (float*)out, (const float*)op, (const float*)arg, b, batch, os, ps,
                           as, r);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

// exp / fmexp / log of the first db[m + 1] coefficients, m = out's max
// degree: the computation of basis (width, m), whose dense layout is that
// prefix; out's coefficients above degree m are left as they are.  An
// argument (or multiplier) whose degree bounds cut inside the levels the
// series reads (1..m for the exponent and the log's argument, 0..m for the
// multiplier) is read through a copy with the coefficients outside its
// bounds set to zero.
static int horner_entry(Horner h, void* out, const void* mul, const void* arg, rpy_dtype dtype,
                        const rpy_tensor_basis* basis, rpy_degree_bounds obi,
                        rpy_degree_bounds mbi, rpy_degree_bounds abi, int64_t batch, int64_t os,
                        int64_t ms, int64_t as, void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b)) || (rc = plain_dtype(dtype))) return rc;
    Bounds ob, mb{0, b.depth}, ab;
    if ((rc = check_bounds(b, obi, ob, "out")) || (rc = check_bounds(b, abi, ab, "arg")) ||
        (h == Horner::FmExp && (rc = check_bounds(b, mbi, mb, "multiplier"))))
        return rc;
    if (ob.lo != 0)
        return fail(RPY_ERR_INVALID_ARGUMENT,
                    "the output of exp, fmexp and log starts at degree 0 (min_degree %d)", ob.lo);
    if ((rc = check_batch(batch, {os, as}, b.size()))) return rc;
    if (h == Horner::FmExp && (rc = check_batch(batch, {ms}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !arg || (h == Horner::FmExp && !mul))
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    hipStream_t st = as_stream(stream);
    const int m = ob.hi;
    DevBasis bm;
    {
        rpy_tensor_basis tb{b.width, m, nullptr};
        if ((rc = make_basis(&tb, bm))) return rc;
    }
    const int64_t nm = bm.size();
    const bool cut_arg = ab.lo > 1 || ab.hi < m;
    const bool cut_mul = h == Horner::FmExp && (mb.lo > 0 || mb.hi < m);
    const size_t es = dtype == RPY_F64 ? 8 : 4;
    char* ws = nullptr;
    const size_t seg = (size_t(batch) * size_t(nm) * es + 255) & ~size_t(255);
    if (cut_arg || cut_mul)
        RPY_HIP_CHECK(workspace_alloc((void**)&ws, seg * ((cut_arg ? 1 : 0) + (cut_mul ? 1 : 0)),
                                      st));
    const void* argp = arg;
    const void* mulp = mul;
    int64_t as2 = as, ms2 = ms;
    auto cut = [&](const void* src, int64_t ss, Bounds bd, void* dst) {
        const int64_t kb = bm.db[std::min(bd.lo, m + 1)];
        const int64_t ke = bm.db[std::min(bd.hi, m) + 1];
        dim3 grid(blocks_elementwise(batch * nm)), block(kBlock);
        if (dtype == RPY_F64)
            hipLaunchKernelGGL(truncate_copy_kernel<double>, grid, block, 0, st, (double*)dst,
                               (const double*)src, batch, nm, ss, kb, ke);
        else
            hipLaunchKernelGGL(truncate_copy_kernel<float>, grid, block, 0, st, (float*)dst,
                               (const float*)src, batch, nm, ss, kb, ke);
    };
    if (cut_arg) {
        cut(arg, as, ab, ws);
        argp = ws;
        as2 = nm;
    }
    if (cut_mul) {
        void* dst = ws + (cut_arg ? seg : 0);
        cut(mul, ms, mb, dst);
        mulp = dst;
        ms2 = nm;
    }
    rc = RPY_OK;
    if ((cut_arg || cut_mul) && hipGetLastError() != hipSuccess)
        rc = fail(RPY_ERR_HIP, "truncating copy launch failed");
    if (!rc) {
        if (dtype == RPY_F64) {
            switch (h) {
                case Horner::Exp: rc = launch_horner<double, Horner::Exp>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
                case Horner::FmExp: rc = launch_horner<double, Horner::FmExp>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
                case Horner::Log: rc = launch_horner<double, Horner::Log>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
            }
        } else {
            switch (h) {
                case Horner::Exp: rc = launch_horner<float, Horner::Exp>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
                case Horner::FmExp: rc = launch_horner<float, Horner::FmExp>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
                case Horner::Log: rc = launch_horner<float, Horner::Log>(out, mulp, argp, bm, batch, os, ms2, as2, st); break;
            }
        }
    }
    if (ws) {
        hipError_t ferr = hipFreeAsync(ws, st);
        if (!rc && ferr != hipSuccess) rc = fail(RPY_ERR_HIP, "hipFreeAsync failed");
    }
    return rc;
}

int dense_ft_exp(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                 rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                 int64_t as, void* stream)
{
    return horner_entry(Horner::Exp, out, nullptr, arg, dtype, basis, ob, {0, -1}, ab, batch, os,
                        0, as, stream);
}

int dense_ft_fmexp(void* out, const void* mul, const void* expo, rpy_dtype dtype,
                   const rpy_tensor_basis* basis, rpy_degree_bounds ob, rpy_degree_bounds mb,
                   rpy_degree_bounds eb, int64_t batch, int64_t os, int64_t ms, int64_t es,
                   void* stream)
{
    return horner_entry(Horner::FmExp, out, mul, expo, dtype, basis, ob, mb, eb, batch, os, ms,
                        es, stream);
}

int dense_ft_log(void* out, const void* arg, rpy_dtype dtype, const rpy_tensor_basis* basis,
                 rpy_degree_bounds ob, rpy_degree_bounds ab, int64_t batch, int64_t os,
                 int64_t as, void* stream)
{
    return horner_entry(Horner::Log, out, nullptr, arg, dtype, basis, ob, {0, -1}, ab, batch, os,
                        0, as, stream);
}

bool has_fixed_mul(int32_t width, int32_t depth, rpy_dtype dtype)
{
    if (dtype == RPY_F64) return find_mul_fixed<double>(width, depth) != nullptr;
    if (dtype == RPY_F32) return find_mul_fixed<float>(width, depth) != nullptr;
    return false;
}

int dense_ft_mul_new(void* out, const void* lhs, const void* rhs, rpy_dtype dtype,
                     const rpy_tensor_basis* basis, int64_t batch, int64_t os, int64_t ls,
                     int64_t rs, void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b)) || (rc = plain_dtype(dtype))) return rc;
    if ((rc = check_batch(batch, {os, ls, rs}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !lhs || !rhs) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    const FmaFixed f = dtype == RPY_F64 ? find_mul_fixed<double>(b.width, b.depth)
                                        : find_mul_fixed<float>(b.width, b.depth);
    if (!f)
        return fail(RPY_ERR_UNSUPPORTED, "no fixed-shape product for width %d depth %d",
                    b.width, b.depth);
    f(out, lhs, rhs, batch, os, ls, rs, as_stream(stream));
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

int dense_ft_log_from_f32(double* out, const float* arg, const rpy_tensor_basis* basis,
                          int64_t batch, int64_t os, int64_t as, void* stream)
{
    DevBasis b;
    int rc;
    if ((rc = make_basis(basis, b))) return rc;
    if ((rc = check_batch(batch, {os, as}, b.size()))) return rc;
    if (batch == 0) return RPY_OK;
    if (!out || !arg) return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    // no error text for the missing shape: the caller falls back to widening
    // first, and a successful fallback must leave no message behind
    auto f = find_horner_fixed<double, Horner::Log, float>(b.width, b.depth);
    if (!f || !f(out, nullptr, arg, batch, os, 0, as, as_stream(stream)))
        return RPY_ERR_UNSUPPORTED;
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

int dense_sparse_map(void* out, const void* arg, const rpy_sparse_matrix* m, rpy_dtype dtype,
                     double scale, int64_t batch, int64_t os, int64_t as, void* stream,
                     bool accumulate)
{
    int rc;
    if ((rc = plain_dtype(dtype))) return rc;
    if (!m) return fail(RPY_ERR_INVALID_ARGUMENT, "null matrix");
    if (!m->compressed_rows)
        return fail(RPY_ERR_UNSUPPORTED,
                    "the device sparse map takes CSR matrices; convert CSC on the host");
    if (m->rows < 0 || m->cols < 0 || m->nnz < 0)
        return fail(RPY_ERR_SHAPE, "negative matrix dimensions");
    if (batch < 0) return fail(RPY_ERR_SHAPE, "negative batch size");
    if (batch > 1 && (os < m->rows || as < m->cols))
        return fail(RPY_ERR_SHAPE, "batch strides smaller than the matrix dimensions");
    if (batch == 0 || m->rows == 0) return RPY_OK;
    if (!out || !arg || !m->offsets || (m->nnz > 0 && (!m->indices || !m->data)))
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    // tensors staged per workgroup: 8 when they fit 64 KiB of LDS (two or
    // more workgroups per CU), else up to 4 in at most 128 KiB with 512
    // threads (the matrix is re-read once per group, so reuse beats
    // occupancy: 2 tensors per group measured slower than 4 for t2l at (5,5))
    const int64_t es = dtype == RPY_F64 ? 8 : 4;
    const int64_t per = m->cols * es;
    int E = 0;
    if (per > 0 && 8 * per <= 64 * 1024)
        E = 8;
    else if (per > 0 && 4 * per <= 128 * 1024)
        E = 4;
    else if (per > 0 && 2 * per <= 128 * 1024)
        E = 2;
    else if (per > 0 && per <= 128 * 1024)
        E = 1;
    if (per == 0) E = 8;
    if (E > 0 && batch > 1) {
        const size_t lds = size_t(E) * size_t(per);
        const int64_t groups = (batch + E - 1) / E;
        dim3 grid(uint32_t(std::min<int64_t>(groups, 65536)));
        auto go = [&](auto tag, auto ec) {
            using T = decltype(tag);
            constexpr int EE = decltype(ec)::value;
            // 512 threads; 1024 when the staged tensors take more than 64 KiB
            // (one workgroup per CU): t2l (5,5) float64 365 -> 257 us, l2t
            // float32 171 -> 150 us (E = 8 at 256 -> 512 threads), 16384 tensors
            const bool big = lds > 64 * 1024;
            auto* kern = big ? csr_map_lds_kernel<T, EE, 1024> : csr_map_lds_kernel<T, EE, 512>;
            const int BL = big ? 1024 : 512;
            if (lds > 64 * 1024) {
                static std::atomic<bool> reserved{false};
                if (!reserved.load(std::memory_order_acquire)) {
                    (void)hipFuncSetAttribute((const void*)kern,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              128 * 1024);
                    reserved.store(true, std::memory_order_release);
                }
            }
            hipLaunchKernelGGL(kern, grid, dim3(BL), lds, as_stream(stream), (T*)out,
                               (const T*)arg, m->offsets, m->indices, (const T*)m->data, m->rows,
                               m->cols, T(scale), batch, os, as, accumulate);
        };
        auto by_e = [&](auto tag) {
            switch (E) {
                case 8: go(tag, std::integral_constant<int, 8>{}); break;
                case 4: go(tag, std::integral_constant<int, 4>{}); break;
                case 2: go(tag, std::integral_constant<int, 2>{}); break;
                default: go(tag, std::integral_constant<int, 1>{}); break;
            }
        };
        if (dtype == RPY_F64)
            by_e(double{});
        else
            by_e(float{});
        RPY_HIP_CHECK(hipGetLastError());
        return RPY_OK;
    }
    dim3 grid(blocks_elementwise(batch * m->rows)), block(kBlock);
    if (dtype == RPY_F64)
        hipLaunchKernelGGL(csr_map_kernel<double>, grid, block, 0, as_stream(stream),
                           (double*)out, (const double*)arg, m->offsets, m->indices,
                           (const double*)m->data, m->rows, scale, batch, os, as, accumulate);
    else
        hipLaunchKernelGGL(csr_map_kernel<float>, grid, block, 0, as_stream(stream),
                           (float*)out, (const float*)arg, m->offsets, m->indices,
                           (const float*)m->data, m->rows, float(scale), batch, os, as,
                           accumulate);
    RPY_HIP_CHECK(hipGetLastError());
    return RPY_OK;
}

}  // namespace rpy_amd
