// streams_host.cpp -- host logic of stream queries: dyadic dissection of a
// clopen interval and the selection of the increment rows a query covers.
//
// Mirrors (integer dyadic arithmetic, same results):
//   DyadicInterval ops        intervals/src/dyadic_interval.cpp:47-183
//   to_dyadic_intervals       intervals/src/dyadic_interval.cpp:191-228
//   param_to_resolution       streams/include/roughpy/streams/stream_base.h:73-83
//   Stream::refine_interval   streams/src/stream.cpp:37-71
//   DyadicCachingLayer query  streams/src/dyadic_caching_layer.cpp:86-130
//   leaf rows [inf, sup)      streams/src/lie_increment_stream.cpp:174-199

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

// Clopen dyadic interval [k 2^-n, (k+1) 2^-n).
struct DI {
    int64_t k;
    int32_t n;
};

int64_t floor_div_pow2(int64_t k, int32_t s)
{
    // floor(k / 2^s) for s >= 0 (arithmetic shift floors for negatives)
    return s >= 63 ? (k < 0 ? -1 : 0) : (k >> s);
}

DI rebase(DI d, int32_t resolution)
{
    if (resolution >= d.n) return DI{d.k * (int64_t(1) << (resolution - d.n)), resolution};
    return DI{floor_div_pow2(d.k, d.n - resolution), resolution};
}

DI from_param(double v, int32_t resolution)
{
    return DI{int64_t(std::floor(std::ldexp(v, resolution))), resolution};
}

DI expand(DI d, int32_t arg = 1) { return rebase(d, d.n - arg); }

bool aligned(DI d)
{
    const DI parent = rebase(d, d.n - 1);
    const DI back = rebase(parent, d.n);
    return back.k == d.k;
}

bool contains(DI a, DI other)
{
    if (other.n < a.n) return false;
    const int32_t s = other.n - a.n;
    const int64_t p = int64_t(1) << s;
    return a.k * p == floor_div_pow2(other.k, s) * p;
}

DI shrink_contained(DI d) { return DI{d.k * 2, d.n + 1}; }
DI flip(DI d) { return DI{(d.k % 2 == 0) ? d.k + 1 : d.k - 1, d.n}; }
DI shrink_omitted(DI d) { return flip(shrink_contained(d)); }

}  // namespace

int32_t param_to_resolution(double arg)
{
    int e = 0;
    std::frexp(arg, &e);
    return -std::min(0, e - 1);
}

// Pieces of [inf, sup) at tolerance `tol`, left to right.
int dyadic_dissection(double inf, double sup, int32_t tol, std::vector<DI>& out)
{
    out.clear();
    if (!(std::isfinite(inf) && std::isfinite(sup)))
        return fail(RPY_ERR_INVALID_ARGUMENT, "interval end points must be finite");
    // negative resolutions are pieces longer than 1 (a stream whose
    // parameters are 2 or more apart infers one)
    if (tol < -60 || tol > 60) return fail(RPY_ERR_INVALID_ARGUMENT, "resolution out of range");
    if (std::fabs(std::ldexp(inf, tol)) > 9.0e15 || std::fabs(std::ldexp(sup, tol)) > 9.0e15)
        return fail(RPY_ERR_INVALID_ARGUMENT, "interval too large for the resolution");
    DI begin = from_param(inf, tol);
    DI end = from_param(sup, tol);
    // per-thread scratch, reused across calls
    static thread_local std::vector<DI> left, right;
    left.clear();
    right.clear();
    int guard = 0;
    while (!contains(begin, end)) {
        DI next = expand(begin);
        if (!aligned(begin)) {
            left.push_back(shrink_omitted(next));
            next.k += 1;
        }
        begin = next;
        if (++guard > 4096) return fail(RPY_ERR_INVALID_ARGUMENT, "dyadic dissection diverged");
    }
    DI next = end;
    while (true) {
        next = expand(next);
        if (!contains(begin, next)) break;
        if (!aligned(end)) right.push_back(shrink_contained(next));
        end = next;
        if (++guard > 8192) return fail(RPY_ERR_INVALID_ARGUMENT, "dyadic dissection diverged");
    }
    out.assign(left.begin(), left.end());
    out.insert(out.end(), right.rbegin(), right.rend());
    return RPY_OK;
}

namespace {

int64_t lower_bound_param(const double* p, int64_t n, double v)
{
    return int64_t(std::lower_bound(p, p + n, v) - p);
}

// First index i >= b with p[i] >= v (p sorted, p[b - 1] < v): steps of 1, 2,
// 4, .. from b, then a binary search inside the last step, so a short run
// costs a few cache lines instead of a search over the whole array.
int64_t gallop_param(const double* p, int64_t n, int64_t b, double v)
{
    int64_t lo = b, hi = b, step = 1;
    while (hi < n && p[hi] < v) {
        lo = hi + 1;
        hi = b + step;
        step *= 2;
    }
    if (hi > n) hi = n;
    return int64_t(std::lower_bound(p + lo, p + hi, v) - p);
}

}  // namespace

// Row ranges [b, e) of the sorted params that the query [inf, sup) at
// `query_res` covers for a stream of resolution `stream_res`.  A caller that
// walks consecutive queries passes the end point and end row of the previous
// one in (hint_hi, hint_e) (hint_e < 0: none) and gets this query's back.
int select_rows(const double* params, int64_t n, int32_t stream_res, double inf, double sup,
                int32_t query_res, std::vector<std::pair<int64_t, int64_t>>& ranges,
                double* hint_hi, int64_t* hint_e)
{
    ranges.clear();
    static thread_local std::vector<DI> pieces;
    int rc = dyadic_dissection(inf, sup, query_res, pieces);
    if (rc) return rc;
    // the pieces tile [inf, sup) left to right, so a piece starts where the
    // previous one ended and its first row is the previous end row
    double prev_hi = hint_hi ? *hint_hi : 0.0;
    int64_t prev_e = hint_e ? *hint_e : -1;
    for (const DI& d : pieces) {
        const double lo = std::ldexp(double(d.k), -d.n);
        const double hi = std::ldexp(double(d.k + 1), -d.n);
        const int64_t b = (prev_e >= 0 && lo == prev_hi) ? prev_e : lower_bound_param(params, n, lo);
        const int64_t e = gallop_param(params, n, b, hi);
        prev_hi = hi;
        prev_e = e;
        if (b == e) continue;  // empty piece contributes nothing
        if (d.n <= stream_res) {
            ranges.push_back({b, e});
        } else {
            // finer than the stream: the leaf counts iff the piece starts it
            const DI leaf = rebase(d, stream_res);
            const double llo = std::ldexp(double(leaf.k), -leaf.n);
            if (llo == lo) {
                const double lhi = std::ldexp(double(leaf.k + 1), -leaf.n);
                ranges.push_back({lower_bound_param(params, n, llo),
                                  lower_bound_param(params, n, lhi)});
            }
        }
    }
    if (hint_hi && hint_e) {
        *hint_hi = prev_hi;
        *hint_e = prev_e;
    }
    return RPY_OK;
}

namespace {

constexpr int64_t kMinQueriesPerThread = 2048;
constexpr int kMaxSelectThreads = 16;

struct BlockArgs {
    const double* params;
    int64_t n_params;
    int32_t stream_res;
    double support_inf, support_sup;
    const double* infs;
    const double* sups;
    int32_t query_res;
};

// params[i - 1] < params[i] for every i, checked in blocks on host threads
// when the array is long.
bool params_increasing(const double* p, int64_t n)
{
    constexpr int64_t kMinPerThread = int64_t(1) << 18;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = int(std::max<int64_t>(
            1, std::min<int64_t>({int64_t(hw ? hw : 1), int64_t(kMaxSelectThreads),
                                  n / kMinPerThread})));
    const int64_t chunk = (n + nt - 1) / nt;
    auto check = [&](int t) {
        const int64_t b = std::max<int64_t>(1, int64_t(t) * chunk);
        const int64_t e = std::min(n, int64_t(t + 1) * chunk);
        for (int64_t i = b; i < e; ++i)
            if (!(p[i - 1] < p[i])) return false;
        return true;
    };
    std::vector<char> ok(size_t(nt), 1);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) {
        try {
            pool.emplace_back([&, t] { ok[size_t(t)] = check(t); });
        } catch (...) {
            ok[size_t(t)] = check(t);
        }
    }
    ok[0] = check(0);
    for (auto& th : pool) th.join();
    for (char c : ok)
        if (!c) return false;
    return true;
}

// Queries [q0, q1) in order: with `ends`, the running count of selected rows
// after each query (from 0 at q0); with `out`, the row indices in order.
int select_block(const BlockArgs& a, int64_t q0, int64_t q1, int64_t* ends, int64_t* out)
{
    std::vector<std::pair<int64_t, int64_t>> r;
    double hint_hi = 0.0;  // consecutive queries (a partition) share end points
    int64_t hint_e = -1;
    int64_t total = 0;
    for (int64_t q = q0; q < q1; ++q) {
        const double qi = a.infs[q], qs = a.sups[q];
        if (!(qs < a.support_inf || qi > a.support_sup)) {
            const double lo = std::max(qi, a.support_inf), hi = std::min(qs, a.support_sup);
            if (hi - lo != 0.0) {
                int rc = select_rows(a.params, a.n_params, a.stream_res, lo, hi, a.query_res, r,
                                     &hint_hi, &hint_e);
                if (rc) return rc;
                for (const auto& pr : r) {
                    if (out)
                        for (int64_t i = pr.first; i < pr.second; ++i) out[total + i - pr.first] = i;
                    total += pr.second - pr.first;
                }
            }
        }
        if (ends) ends[q - q0] = total;
    }
    return RPY_OK;
}

}  // namespace

}  // namespace rpy_amd

using namespace rpy_amd;

extern "C" {

int32_t rpy_param_to_resolution(double arg) { return param_to_resolution(arg); }

int rpy_to_dyadic_intervals(double inf, double sup, int32_t tol, int64_t* k_out,
                            int32_t* n_out, int32_t max_pieces, int32_t* count)
{
    if (!count) return fail(RPY_ERR_INVALID_ARGUMENT, "null count");
    std::vector<DI> pieces;
    int rc = dyadic_dissection(inf, sup, tol, pieces);
    if (rc) return rc;
    *count = int32_t(pieces.size());
    if (int32_t(pieces.size()) > max_pieces)
        return fail(RPY_ERR_SHAPE, "%zu pieces do not fit in %d", pieces.size(), max_pieces);
    for (size_t i = 0; i < pieces.size(); ++i) {
        if (k_out) k_out[i] = pieces[i].k;
        if (n_out) n_out[i] = pieces[i].n;
    }
    return RPY_OK;
}

int rpy_stream_select_rows(const double* params, int64_t n_params, int32_t stream_resolution,
                           double inf, double sup, int32_t query_resolution, int64_t* ranges,
                           int64_t max_ranges, int64_t* n_ranges)
{
    if (!n_ranges || (n_params > 0 && !params))
        return fail(RPY_ERR_INVALID_ARGUMENT, "null argument");
    for (int64_t i = 1; i < n_params; ++i)
        if (!(params[i - 1] < params[i]))
            return fail(RPY_ERR_INVALID_ARGUMENT, "params must be strictly increasing");
    std::vector<std::pair<int64_t, int64_t>> r;
    int rc = select_rows(params, n_params, stream_resolution, inf, sup, query_resolution, r,
                         nullptr, nullptr);
    if (rc) return rc;
    *n_ranges = int64_t(r.size());
    if (int64_t(r.size()) > max_ranges)
        return fail(RPY_ERR_SHAPE, "%zu ranges do not fit in %lld", r.size(),
                    (long long)max_ranges);
    if (ranges)
        for (size_t i = 0; i < r.size(); ++i) {
            ranges[2 * i] = r[i].first;
            ranges[2 * i + 1] = r[i].second;
        }
    return RPY_OK;
}

int rpy_stream_select_rows_batch(const double* params, int64_t n_params,
                                 int32_t stream_resolution, double support_inf,
                                 double support_sup, const double* infs, const double* sups,
                                 int64_t n_queries, int32_t query_resolution, int64_t* offsets,
                                 int64_t* rows, int64_t max_rows)
{
    if (n_queries < 0) return fail(RPY_ERR_SHAPE, "negative query count");
    if (!offsets || (n_params > 0 && !params) || (n_queries > 0 && (!infs || !sups)))
        return fail(RPY_ERR_INVALID_ARGUMENT, "null argument");
    if (!params_increasing(params, n_params))
        return fail(RPY_ERR_INVALID_ARGUMENT, "params must be strictly increasing");
    offsets[0] = 0;
    // Queries are independent: blocks of them run on host threads, first
    // counting (each block's running counts land in its part of offsets),
    // then, after a prefix over the blocks, writing their rows in place.
    // The result is the same as one pass in query order.
    const BlockArgs args{params, n_params, stream_resolution, support_inf, support_sup,
                         infs, sups, query_resolution};
    unsigned hw = std::thread::hardware_concurrency();
    const int64_t by_work = (n_queries + kMinQueriesPerThread - 1) / kMinQueriesPerThread;
    const int nt = int(std::max<int64_t>(1, std::min<int64_t>({int64_t(hw ? hw : 1),
                                                                int64_t(kMaxSelectThreads),
                                                                by_work})));
    const int64_t chunk = (n_queries + nt - 1) / std::max(nt, 1);
    auto run = [&](auto&& body) {
        std::vector<std::thread> pool;
        std::vector<int> rcs(size_t(nt), RPY_OK);
        for (int t = 1; t < nt; ++t) {
            try {
                pool.emplace_back([&, t] { rcs[size_t(t)] = body(t); });
            } catch (...) {
                rcs[size_t(t)] = body(t);  // no thread to spare: run the block here
            }
        }
        rcs[0] = body(0);
        for (auto& th : pool) th.join();
        for (int t = 0; t < nt; ++t)
            if (rcs[size_t(t)] != RPY_OK) return t;
        return -1;
    };
    auto bounds = [&](int t) {
        const int64_t q0 = std::min(n_queries, int64_t(t) * chunk);
        return std::make_pair(q0, std::min(n_queries, q0 + chunk));
    };
    int bad = run([&](int t) {
        const auto [q0, q1] = bounds(t);
        return select_block(args, q0, q1, offsets + 1 + q0, nullptr);
    });
    if (bad >= 0) {
        // the message of a worker's failure lives in that thread: the same
        // block again on this thread fails the same way and sets it here
        const auto [q0, q1] = bounds(bad);
        return select_block(args, q0, q1, offsets + 1 + q0, nullptr);
    }
    int64_t base = 0;
    for (int t = 0; t < nt; ++t) {
        const auto [q0, q1] = bounds(t);
        for (int64_t q = q0; q < q1; ++q) offsets[q + 1] += base;
        if (q1 > q0) base = offsets[q1];
    }
    if (!rows) return RPY_OK;
    if (base > max_rows)
        return fail(RPY_ERR_SHAPE, "more than %lld selected rows", (long long)max_rows);
    bad = run([&](int t) {
        const auto [q0, q1] = bounds(t);
        return select_block(args, q0, q1, nullptr, rows + offsets[q0]);
    });
    // the filling pass repeats the counting pass's work, so it fails only if
    // that did
    if (bad >= 0) return fail(RPY_ERR_INVALID_ARGUMENT, "row selection failed while filling");
    return RPY_OK;
}

}  // extern "C"
