// chen.hip -- batched Chen signatures of uneven piecewise-linear streams.
//
// One persistent launch: streams are sorted by length (longest first), every
// stream is split across W^P lanes (chen_core.hpp), and each wave64 pulls the
// next 64 (stream, prefix) items from a global work counter until the queue
// is empty.  Each lane keeps its share of the running signature in
// registers for the whole stream and writes it once at the end, so the only
// HBM traffic is the increments in and the signatures out.
//
// Reference: Context::compute_signature, algebra/src/lite_context.h:545-563
// (result = 1; result.fmexp_inplace(l2t(row)) per increment), with the fmexp
// Horner of roughpy_compute/dense/intermediate/free_tensor_fmexp.hpp:15-40.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "chen.hpp"
#include "chen_core.hpp"
#include "chen_kernels.hpp"
#include "tuning.hpp"
#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

// keys[i] = length of stream i, vals[i] = i; also clears the queue counter.
// Stream i owns rows [begins[i], ends[i]) (CSR offsets: ends = offsets + 1).
__global__ void prep_lengths_kernel(const int64_t* __restrict__ begins,
                                    const int64_t* __restrict__ ends, int64_t n,
                                    uint32_t* __restrict__ keys, int32_t* __restrict__ vals,
                                    uint32_t* __restrict__ counter)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0u;
    if (i < n) {
        const int64_t len = ends[i] - begins[i];
        keys[i] = uint32_t(len < 0 ? 0 : (len > 0xffffffffLL ? 0xffffffffLL : len));
        vals[i] = int32_t(i);
    }
}

// Fallback for shapes without a specialised kernel: one workgroup per
// stream, the running signature in the output buffer, per-coefficient Horner
// with a barrier between levels.
// With IO != T the running signature lives in a per-workgroup slot of
// `work` and is rounded into `out` at the end of each stream.
template <typename T, typename IO = T>
__global__ __launch_bounds__(256) void chen_generic_kernel(IO* __restrict__ out,
                                                          const IO* __restrict__ incr,
                                                          const int64_t* __restrict__ begins,
                                                          const int64_t* __restrict__ ends,
                                                          int64_t n_streams, int64_t tsize,
                                                          DevBasis b, T* __restrict__ work)
{
    const int W = b.width, D = b.depth;
    for (int64_t s = blockIdx.x; s < n_streams; s += gridDim.x) {
        T* S;
        if constexpr (std::is_same<T, IO>::value)
            S = out + s * tsize;
        else
            S = work + int64_t(blockIdx.x) * tsize;
        for (int64_t i = threadIdx.x; i < tsize; i += blockDim.x) S[i] = (i == 0) ? T(1) : T(0);
        __syncthreads();
        for (int64_t r = begins[s]; r < ends[s]; ++r) {
            const IO* x = incr + r * W;
            for (int m = D; m >= 1; --m) {
                const int64_t n = b.level_size(m);
                for (int64_t idx = threadIdx.x; idx < n; idx += blockDim.x) {
                    // letters of the word, most significant first
                    T p = S[0];
                    int64_t rest = idx, scale = n / W;
                    for (int K = 1; K <= m; ++K) {
                        const int64_t pre = idx / scale;  // prefix of length K
                        const int letter = int(rest / scale);
                        rest -= int64_t(letter) * scale;
                        p = p * (T(x[letter]) / T(m - K + 1)) + S[b.db[K] + pre];
                        scale = (scale >= W) ? scale / W : 1;
                    }
                    // only this thread reads S_m[idx]; lower levels stay old
                    // until the barrier below.
                    S[b.db[m] + idx] = p;
                }
                __syncthreads();
            }
        }
        if constexpr (!std::is_same<T, IO>::value) {
            for (int64_t i = threadIdx.x; i < tsize; i += blockDim.x) out[s * tsize + i] = IO(S[i]);
            __syncthreads();
        }
    }
}

using chen::choose_p;

// Segment lengths of the block-compensated variants, and the default.
constexpr int32_t kBlockRows[] = {128, 256, 512};
// The block-compensated segment buffer holds at most this many bytes; larger
// batches run in rounds of streams.
constexpr size_t kBlockSegBytes = size_t(16) << 30;
using chen::kBudgetF32;
using chen::kBudgetF64;

constexpr bool stream_ok(int W, int P) { return ipow(W, P) >= 64 && ipow(W, P) <= 1024; }

// Default variants of a specialised shape: the lane-chunk kernel at the
// register-budget prefix (choose_p) with 2 waves per EU, and where W^P is
// 64..1024 the stream kernel at the same prefix with 32-row blocks, which
// is the default choice when it exists.
template <int W, int D>
void add_defaults(std::vector<SigVariant>& v)
{
    constexpr int P32 = choose_p(W, D, kBudgetF32);
    constexpr int P64 = choose_p(W, D, kBudgetF64);
    v.push_back({W, D, RPY_F32, {RPY_KIND_LANE_CHUNK, P32, 0, 2},
                 &launch_persistent<float, W, D, P32>});
    v.push_back({W, D, RPY_F64, {RPY_KIND_LANE_CHUNK, P64, 0, 2},
                 &launch_persistent<double, W, D, P64>});
    v.push_back({W, D, RPY_F32_COMPENSATED, {RPY_KIND_LANE_CHUNK, P32, 0, 2},
                 &launch_persistent<float, W, D, P32, float, 2, true>});
    if constexpr (stream_ok(W, P32)) {
        v.push_back({W, D, RPY_F32, {RPY_KIND_STREAM_GROUP, P32, 32, 0},
                     &launch_stream<float, W, D, P32>});
        v.push_back({W, D, RPY_F32_COMPENSATED, {RPY_KIND_STREAM_GROUP, P32, 32, 0},
                     &launch_stream<float, W, D, P32, float, 32, true>});
    }
    if constexpr (stream_ok(W, P64)) {
        v.push_back({W, D, RPY_F64, {RPY_KIND_STREAM_GROUP, P64, 32, 0},
                     &launch_stream<double, W, D, P64>});

    }
}

const SigVariant* find_in(const std::vector<SigVariant>& all, int32_t w, int32_t d, rpy_dtype dt,
                          const rpy_kernel_variant& k)
{
    for (const SigVariant& x : all)
        if (x.width == w && x.depth == d && x.dtype == dt && same_variant(x.v, k)) return &x;
    return nullptr;
}

// Every compiled specialised variant: the defaults of all shapes, then the
// tuning units' extra variants of the tuned shapes.
const std::vector<SigVariant>& registry()
{
    static const std::vector<SigVariant> r = [] {
        std::vector<SigVariant> v;
        add_defaults<2, 1>(v); add_defaults<2, 2>(v); add_defaults<2, 3>(v); add_defaults<2, 4>(v);
        add_defaults<2, 5>(v); add_defaults<2, 6>(v); add_defaults<2, 7>(v); add_defaults<2, 8>(v);
        add_defaults<3, 1>(v); add_defaults<3, 2>(v); add_defaults<3, 3>(v); add_defaults<3, 4>(v);
        add_defaults<3, 5>(v); add_defaults<3, 6>(v); add_defaults<4, 1>(v); add_defaults<4, 2>(v);
        add_defaults<4, 3>(v); add_defaults<4, 4>(v); add_defaults<4, 5>(v); add_defaults<4, 6>(v);
        add_defaults<5, 1>(v); add_defaults<5, 2>(v); add_defaults<5, 3>(v); add_defaults<5, 4>(v);
        add_defaults<5, 5>(v); add_defaults<6, 1>(v); add_defaults<6, 2>(v); add_defaults<6, 3>(v);
        add_defaults<6, 4>(v); add_defaults<6, 5>(v); add_defaults<7, 1>(v); add_defaults<7, 2>(v);
        add_defaults<7, 3>(v); add_defaults<7, 4>(v); add_defaults<8, 1>(v); add_defaults<8, 2>(v);
        add_defaults<8, 3>(v); add_defaults<8, 4>(v);
        for (const SigVariantList& list : {sig_tune_variants_4_6(), sig_tune_variants_5_5(),
                                           sig_tune_variants_6_5(), sig_tune_variants_8_4()})
            for (size_t i = 0; i < list.n; ++i) {
                const SigVariant& x = list.items[i];
                if (!find_in(v, x.width, x.depth, x.dtype, x.v)) v.push_back(x);
            }
        // block-compensated float32 (chen_block.hip): no launcher of its
        // own, run_signature drives the float32 variant and the merge
        for (int32_t w = 2; w <= 8; ++w)
            for (int32_t d = 1; d <= 8; ++d)
                if (block_merge_supported(w, d))
                    for (int32_t T : kBlockRows)
                        v.push_back({w, d, RPY_F32_COMPENSATED,
                                     {RPY_KIND_BLOCK_MERGE, 0, T, 0}, nullptr});
        return v;
    }();
    return r;
}

bool specialised(int32_t w, int32_t d)
{
    for (const SigVariant& x : registry())
        if (x.width == w && x.depth == d) return true;
    return false;
}

rpy_kernel_variant generic_variant() { return rpy_kernel_variant{RPY_KIND_GENERIC, 0, 0, 0}; }

// The register-budget default of a dtype: the stream kernel at prefix
// choose_p where it exists, else the lane-chunk kernel, else generic.  The
// compensated mode keeps float32 lanes (with compensated lowest levels), so
// it takes the float32 budget.
rpy_kernel_variant default_variant(int32_t w, int32_t d, rpy_dtype dt)
{
    const int p = choose_p(w, d, dt == RPY_F64 ? kBudgetF64 : kBudgetF32);
    const rpy_kernel_variant stream{RPY_KIND_STREAM_GROUP, p, 32, 0};
    if (find_in(registry(), w, d, dt, stream)) return stream;
    const rpy_kernel_variant chunk{RPY_KIND_LANE_CHUNK, p, 0, 2};
    if (find_in(registry(), w, d, dt, chunk)) return chunk;
    return generic_variant();
}

bool variant_exists(int32_t w, int32_t d, rpy_dtype dt, const rpy_kernel_variant& k)
{
    return same_variant(k, generic_variant()) || find_in(registry(), w, d, dt, k) != nullptr;
}

// The variant a call runs: nullptr for the generic kernel.
const SigVariant* choose_variant(int32_t w, int32_t d, rpy_dtype dt, int32_t* source)
{
    if (!specialised(w, d)) {
        if (source) *source = 0;
        return nullptr;
    }
    const rpy_kernel_variant k = resolve_variant(
            RPY_OP_SIGNATURE, w, d, dt, default_variant(w, d, dt),
            [&](const rpy_kernel_variant& v) { return variant_exists(w, d, dt, v); }, source);
    return k.kind == RPY_KIND_GENERIC ? nullptr : find_in(registry(), w, d, dt, k);
}

bool valid_dtype(rpy_dtype dt)
{
    return dt == RPY_F32 || dt == RPY_F64 || dt == RPY_F32_COMPENSATED;
}

}  // namespace

int chen_prefix_length(int32_t width, int32_t depth, int dtype_is_f64)
{
    const SigVariant* v = choose_variant(width, depth, dtype_is_f64 ? RPY_F64 : RPY_F32, nullptr);
    return v ? v->v.prefix : 0;
}

int32_t sig_variants(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
                     int32_t max)
{
    if (!valid_dtype(dtype)) return 0;
    int32_t n = 0;
    for (const SigVariant& x : registry()) {
        if (x.width != width || x.depth != depth || x.dtype != dtype) continue;
        if (out && n < max) out[n] = x.v;
        ++n;
    }
    // the generic kernel runs every shape
    if (out && n < max) out[n] = generic_variant();
    return n + 1;
}

int sig_select(int32_t width, int32_t depth, rpy_dtype dtype, const rpy_kernel_variant* v)
{
    if (!valid_dtype(dtype)) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    if (v && !variant_exists(width, depth, dtype, *v))
        return fail(RPY_ERR_UNSUPPORTED,
                    "no signature kernel variant (kind %d, prefix %d, block %d, waves/EU %d) for "
                    "width %d depth %d",
                    v->kind, v->prefix, v->block, v->waves_per_eu, width, depth);
    set_selected_variant(RPY_OP_SIGNATURE, width, depth, dtype, v);
    return RPY_OK;
}

int sig_config(int32_t width, int32_t depth, rpy_dtype dtype, rpy_kernel_variant* out,
               int32_t* source)
{
    if (!out) return fail(RPY_ERR_INVALID_ARGUMENT, "null pointer");
    if (!valid_dtype(dtype)) return fail(RPY_ERR_DTYPE, "unsupported dtype");
    const SigVariant* v = choose_variant(width, depth, dtype, source);
    *out = v ? v->v : generic_variant();
    return RPY_OK;
}

namespace {

// Device workspace of one signature call: keys in/out, values in/out, the
// queue counter and the radix sort's scratch (specialised shapes), or the
// float64 running signatures of the generic kernel's compensated mode.
struct SigPlan {
    rpy_tensor_basis tb;
    DevBasis b;
    const SigVariant* var;  // nullptr: the generic kernel
    size_t sort_bytes;  // radix sort scratch
    size_t k_bytes, v_bytes;
    size_t total;       // bytes of workspace the call needs
    int generic_grid;
    // block-compensated float32: the float32 variant run on the segments
    const SigVariant* seg_var;
    BlockPlan bp;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Validates the arguments and sizes the workspace; *skip = true when there
// is nothing to launch.
int plan_signature(const void* out, const void* increments, const int64_t* begins,
                   const int64_t* ends, int64_t n_streams, int32_t width, int32_t depth,
                   rpy_dtype dtype, hipStream_t st, SigPlan& p, bool* skip,
                   int max_generic_grid = 4096, size_t ws_limit = 0)
{
    *skip = false;
    if (dtype != RPY_F32 && dtype != RPY_F64 && dtype != RPY_F32_COMPENSATED)
        return fail(RPY_ERR_DTYPE, "unsupported dtype");
    p.tb = rpy_tensor_basis{width, depth, nullptr};
    int rc = make_basis(&p.tb, p.b);
    if (rc) return rc;
    if (depth < 1) return fail(RPY_ERR_INVALID_ARGUMENT, "signatures need depth >= 1");
    if (n_streams < 0) return fail(RPY_ERR_SHAPE, "negative number of streams");
    if (n_streams == 0) {
        *skip = true;
        return RPY_OK;
    }
    if (n_streams > INT32_MAX) return fail(RPY_ERR_SHAPE, "too many streams");
    if (!out || !begins || !ends || !increments)
        return fail(RPY_ERR_INVALID_ARGUMENT, "null data pointer");
    p.var = choose_variant(width, depth, dtype, nullptr);
    p.sort_bytes = p.k_bytes = p.v_bytes = 0;
    p.generic_grid = int(std::min<int64_t>(n_streams, max_generic_grid));
    p.seg_var = nullptr;
    if (p.var && p.var->v.kind == RPY_KIND_BLOCK_MERGE) {
        // segments run on the compensated lanes a compensated call runs
        // without a selection: the tuned row when it is a lane variant, else
        // the default
        rpy_kernel_variant lanes = default_variant(width, depth, RPY_F32_COMPENSATED);
        rpy_kernel_variant t;
        if (tuned_variant(RPY_OP_SIGNATURE, width, depth, RPY_F32_COMPENSATED, &t) &&
            (t.kind == RPY_KIND_LANE_CHUNK || t.kind == RPY_KIND_STREAM_GROUP) &&
            find_in(registry(), width, depth, RPY_F32_COMPENSATED, t))
            lanes = t;
        p.seg_var = find_in(registry(), width, depth, RPY_F32_COMPENSATED, lanes);
        if (!p.seg_var || !p.seg_var->launch)
            return fail(RPY_ERR_UNSUPPORTED, "no float32 signature kernel for width %d depth %d",
                        width, depth);
        const int64_t per_stream = block_segment_slots() * p.b.size() * int64_t(sizeof(float));
        int64_t n_sub = std::min<int64_t>(n_streams, std::max<int64_t>(
                                                  1, int64_t(kBlockSegBytes) / per_stream));
        n_sub = std::min<int64_t>(n_sub, INT32_MAX / block_segment_slots());
        if ((rc = block_plan(n_sub, p.b.size(), p.bp, st))) return rc;
        // a caller with a fixed workspace gets rounds of fewer streams
        while (ws_limit && p.bp.total > ws_limit && n_sub > 1) {
            n_sub = (n_sub + 1) / 2;
            if ((rc = block_plan(n_sub, p.b.size(), p.bp, st))) return rc;
        }
        p.total = p.bp.total;
        return RPY_OK;
    }
    if (!p.var) {
        p.total = dtype == RPY_F32_COMPENSATED
                          ? align256(size_t(p.generic_grid) * size_t(p.b.size()) * 8)
                          : 0;
        return RPY_OK;
    }
    const size_t n = size_t(n_streams);
    RPY_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, p.sort_bytes, (uint32_t*)nullptr,
                                                 (uint32_t*)nullptr, (int32_t*)nullptr,
                                                 (int32_t*)nullptr, n, 0, 32, st));
    p.k_bytes = align256(n * sizeof(uint32_t));
    p.v_bytes = align256(n * sizeof(int32_t));
    p.total = 2 * p.k_bytes + 2 * p.v_bytes + 256 + align256(p.sort_bytes);
    return RPY_OK;
}

// Launches the call on workspace `ws` (p.total bytes, device memory).
int run_signature(void* out, const void* increments, const int64_t* begins, const int64_t* ends,
                  int64_t n_streams, rpy_dtype dtype, const SigPlan& p, char* ws,
                  hipStream_t st)
{
    const DevBasis& b = p.b;
    if (p.var && p.var->v.kind == RPY_KIND_BLOCK_MERGE)
        return block_run((float*)out, (const float*)increments, begins, ends, n_streams, b,
                         p.var->v.block, p.seg_var->launch, p.bp, ws, st);
    if (!p.var) {
        const int grid = p.generic_grid;
        if (dtype == RPY_F64)
            hipLaunchKernelGGL((chen_generic_kernel<double>), dim3(grid), dim3(256), 0, st,
                               (double*)out, (const double*)increments, begins, ends, n_streams,
                               b.size(), b, (double*)nullptr);
        else if (dtype == RPY_F32)
            hipLaunchKernelGGL((chen_generic_kernel<float>), dim3(grid), dim3(256), 0, st,
                               (float*)out, (const float*)increments, begins, ends, n_streams,
                               b.size(), b, (float*)nullptr);
        else
            hipLaunchKernelGGL((chen_generic_kernel<double, float>), dim3(grid), dim3(256), 0, st,
                               (float*)out, (const float*)increments, begins, ends, n_streams,
                               b.size(), b, (double*)ws);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "signature launch failed: %s", hipGetErrorString(err));
        return RPY_OK;
    }
    const size_t n = size_t(n_streams);
    uint32_t* keys_in = (uint32_t*)ws;
    uint32_t* keys_out = (uint32_t*)(ws + p.k_bytes);
    int32_t* vals_in = (int32_t*)(ws + 2 * p.k_bytes);
    int32_t* vals_out = (int32_t*)(ws + 2 * p.k_bytes + p.v_bytes);
    uint32_t* counter = (uint32_t*)(ws + 2 * p.k_bytes + 2 * p.v_bytes);
    void* sort_tmp = ws + 2 * p.k_bytes + 2 * p.v_bytes + 256;
    size_t sort_bytes = p.sort_bytes;

    // the length keys also clear the queue counter (a kernel, not a memset,
    // so a captured graph replays it)
    const int pb = 256;
    hipLaunchKernelGGL(prep_lengths_kernel, dim3(uint32_t((n + pb - 1) / pb)), dim3(pb), 0, st,
                       begins, ends, n_streams, keys_in, vals_in, counter);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return fail(RPY_ERR_HIP, "length prep launch failed: %s", hipGetErrorString(err));
    // Longest-first order only matters when the chunks run in more than one
    // round; a batch whose waves are all resident at once (a few per CU)
    // keeps the input order and skips the sort's launches.
    const int np = p.var->v.prefix;
    const int64_t lanes = int64_t(n_streams) * int64_t(ipow(p.tb.width, np));
    const bool sort = (lanes + 63) / 64 > 1024;
    if (sort) {
        err = rocprim::radix_sort_pairs_desc(sort_tmp, sort_bytes, keys_in, keys_out, vals_in,
                                             vals_out, n, 0, 32, st);
        if (err != hipSuccess)
            return fail(RPY_ERR_HIP, "radix sort failed: %s", hipGetErrorString(err));
    }
    return p.var->launch(out, increments, begins, ends, sort ? vals_out : vals_in, n_streams, b,
                         counter, 0, st);
}

// Scratch of the entry point that allocates nothing: one array per GPU,
// resolved to a device address once per device (its name is only a host
// stand-in).  Calls that overlap on one GPU share it and race.
constexpr size_t kScratchBytes = size_t(8) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_pointer(char** out)
{
    int dev = 0;
    RPY_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(RPY_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail(RPY_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(err));
        g_scratch_ptr[dev] = static_cast<char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return RPY_OK;
}

}  // namespace

int stream_signature_ranges(void* out, const void* increments, const int64_t* begins,
                            const int64_t* ends, int64_t n_streams, int32_t width, int32_t depth,
                            rpy_dtype dtype, void* stream)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    SigPlan p;
    bool skip = false;
    int rc = plan_signature(out, increments, begins, ends, n_streams, width, depth, dtype, st, p,
                            &skip);
    if (rc || skip) return rc;
    char* ws = nullptr;
    if (p.total) RPY_HIP_CHECK(workspace_alloc((void**)&ws, p.total, st));
    rc = run_signature(out, increments, begins, ends, n_streams, dtype, p, ws, st);
    if (ws) {
        hipError_t ferr = hipFreeAsync(ws, st);
        if (!rc && ferr != hipSuccess)
            rc = fail(RPY_ERR_HIP, "hipFreeAsync failed: %s", hipGetErrorString(ferr));
    }
    return rc;
}

int stream_signature_batch(void* out, const void* increments, const int64_t* offsets,
                           int64_t n_streams, int32_t width, int32_t depth, rpy_dtype dtype,
                           void* stream)
{
    if (!offsets && n_streams > 0) return fail(RPY_ERR_INVALID_ARGUMENT, "null offsets");
    // CSR offsets: stream s owns rows [offsets[s], offsets[s+1])
    return stream_signature_ranges(out, increments, offsets, offsets ? offsets + 1 : nullptr,
                                   n_streams, width, depth, dtype, stream);
}

int64_t stream_signature_scratch_bytes() { return int64_t(kScratchBytes); }

int stream_signature_batch_noalloc(void* out, const void* increments, const int64_t* offsets,
                                   int64_t n_streams, int32_t width, int32_t depth,
                                   rpy_dtype dtype, void* stream)
{
    if (!offsets && n_streams > 0) return fail(RPY_ERR_INVALID_ARGUMENT, "null offsets");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int64_t* begins = offsets;
    const int64_t* ends = offsets ? offsets + 1 : nullptr;
    SigPlan p;
    bool skip = false;
    // the generic kernel's compensated slots must fit the scratch as well
    int64_t tsize = tensor_size_checked(width, depth);
    int max_grid = 4096;
    if (dtype == RPY_F32_COMPENSATED && tsize > 0)
        max_grid = int(std::max<int64_t>(1, std::min<int64_t>(4096, int64_t(kScratchBytes) /
                                                                        (tsize * 8))));
    int rc = plan_signature(out, increments, begins, ends, n_streams, width, depth, dtype, st, p,
                            &skip, max_grid, kScratchBytes);
    if (rc || skip) return rc;
    if (p.total > kScratchBytes)
        return fail(RPY_ERR_SHAPE,
                    "%lld streams need %zu bytes of scratch, more than the %zu this entry point "
                    "holds; use rpy_stream_signature_batch", (long long)n_streams, p.total,
                    kScratchBytes);
    char* ws = nullptr;
    if ((rc = scratch_pointer(&ws))) return rc;
    return run_signature(out, increments, begins, ends, n_streams, dtype, p, ws, st);
}

// Runs the lane algorithm of the specialised kernel on the host for one
// stream (double precision) -- used by CPU tests to check the decomposition
// and the output index map without a GPU.
template <int W, int D, int P>
static void emulate(double* out, const double* incr, int64_t len, const DevBasis& b)
{
    using Sh = chen::Shape<W, D, P>;
    for (int q = 0; q < Sh::NP; ++q) {
        int u[P];
        int t = q;
        for (int j = P - 1; j >= 0; --j) {
            u[j] = t % W;
            t /= W;
        }
        chen::Lane<double, W, D, P> st;
        st.init_identity();
        double row[Sh::ROW];
        // letter-major table of the stream kernel: tab[a * D + d - 1] = x_a / d
        double tab[W * D];
        for (int64_t r = 0; r < len; ++r) {
            chen::build_row<double, W, D>(row, incr + r * W);
            if (r % 3 == 1) {
                st.step(chen::ScaledRow<double, W, D, P>{row, u});
            } else if (r % 3 == 2) {
                for (int a = 0; a < W; ++a)
                    for (int d = 1; d <= D; ++d) tab[a * D + d - 1] = row[(d - 1) * W + a];
                const double* lt[P];
                for (int j = 0; j < P; ++j) lt[j] = tab + u[j] * D;
                double raw[W];
                for (int a = 0; a < W; ++a) raw[a] = incr[r * W + a];
                st.step(chen::UniformRow<double, W, D, P, D>{raw, lt});
            } else {
                st.step(chen::RawRow<double, W, D, P>(row, u));
            }
        }
        for (int k = P; k <= D; ++k) {
            const int nk = int(ipow(W, k - P));
            for (int v = 0; v < nk; ++v) out[b.db[k] + int64_t(q) * nk + v] = st.sub[Sh::off(k) + v];
        }
        for (int k = 0; k < P; ++k) {
            const int grp = int(ipow(W, P - k));
            if (q % grp == 0) out[b.db[k] + q / grp] = st.path[k];
        }
    }
}

int signature_host_emulation(double* out, const double* incr, int64_t len, int32_t width,
                             int32_t depth, int32_t prefix)
{
    rpy_tensor_basis tb{width, depth, nullptr};
    DevBasis b;
    int rc = make_basis(&tb, b);
    if (rc) return rc;
#define RPY_EMU(W, D, P) \
    if (width == W && depth == D && prefix == P) { emulate<W, D, P>(out, incr, len, b); return RPY_OK; }
    RPY_EMU(2, 3, 1) RPY_EMU(2, 3, 2) RPY_EMU(2, 3, 3)
    RPY_EMU(3, 4, 1) RPY_EMU(3, 4, 2) RPY_EMU(3, 4, 3)
    RPY_EMU(5, 5, 2) RPY_EMU(5, 5, 3) RPY_EMU(5, 5, 4)
    RPY_EMU(4, 6, 3) RPY_EMU(4, 6, 4) RPY_EMU(4, 6, 5)
    RPY_EMU(8, 4, 2) RPY_EMU(8, 4, 3)
    RPY_EMU(6, 5, 3) RPY_EMU(6, 5, 4)
#undef RPY_EMU
    return fail(RPY_ERR_UNSUPPORTED, "no host emulation for this (width, depth, prefix)");
}

}  // namespace rpy_amd
