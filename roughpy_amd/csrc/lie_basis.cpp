// lie_basis.cpp -- host construction of the Hall basis and of the sparse
// lie_to_tensor / tensor_to_lie maps.
//
// Reference behaviour being mirrored (same basis order, same map values):
//   greedy Hall set          roughpy/compute/_src/lie_basis.c:201-286
//   Witt dimensions          roughpy/compute/_src/lie_basis.c:1788-1813
//   l2t (commutator expand)  roughpy/compute/_src/lie_basis.c:806-950
//   t2l (Dynkin r(u)/|u|)    roughpy/compute/_src/lie_basis.c:1066-1250
//   bracket of Hall keys     roughpy/compute/_src/lie_multiplication_cache.cpp:242-367
//   basis data validation    roughpy/compute/_src/lie_basis.c:1333-1419
//
// The maps are built once per basis on the host and handed to the device
// as CSR arrays; they are integer-valued (t2l: integer / degree).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "rpy_common.hpp"

namespace rpy_amd {

namespace {

using Key = int64_t;
using SparseInt = std::vector<std::pair<int64_t, int64_t>>;  // (index, coeff) sorted

int64_t ipow_i64(int64_t b, int e)
{
    int64_t r = 1;
    while (e-- > 0) r *= b;
    return r;
}

int mobius(int n)
{
    int result = 1;
    for (int p = 2; p * p <= n; ++p) {
        if (n % p == 0) {
            n /= p;
            if (n % p == 0) return 0;
            result = -result;
        }
    }
    return n > 1 ? -result : result;
}

int64_t witt(int64_t width, int degree)
{
    if (degree <= 0) return 0;
    if (degree == 1) return width;
    int64_t s = 0;
    for (int d = 1; d <= degree; ++d)
        if (degree % d == 0) s += mobius(d) * ipow_i64(width, degree / d);
    return s / degree;
}

struct Hall {
    int32_t width = 0, depth = 0;
    std::vector<int64_t> db;    // depth + 2
    std::vector<int64_t> data;  // 2 * (n + 1)
    std::map<std::pair<Key, Key>, Key> find;

    Key left(Key k) const { return data[2 * k]; }
    Key right(Key k) const { return data[2 * k + 1]; }
    int degree(Key k) const
    {
        for (int d = 0; d <= depth; ++d)
            if (db[d] <= k && k < db[d + 1]) return d;
        return -1;
    }
    Key size() const { return db[depth + 1] - 1; }
};

bool load_hall(const rpy_lie_basis* b, Hall& h)
{
    h.width = b->width;
    h.depth = b->depth;
    h.db.assign(b->degree_begin, b->degree_begin + b->depth + 2);
    const Key n = h.db[h.depth + 1];
    h.data.assign(b->data, b->data + 2 * n);
    for (Key k = 1; k < n; ++k)
        if (h.left(k) != 0) h.find[{h.left(k), h.right(k)}] = k;
    return true;
}

// [x, y] for Hall keys x, y, expanded in the basis (sparse, integer).  Uses
// the Jacobi rewriting [x,[y1,y2]] = [[x,y1],y2] - [[x,y2],y1] whenever
// (x, y) with x < y is not itself a Hall pair -- the identity the reference
// multiplication cache applies.
struct BracketCache {
    const Hall& h;
    std::map<std::pair<Key, Key>, SparseInt> memo;
    explicit BracketCache(const Hall& hh) : h(hh) {}

    const SparseInt& get(Key x, Key y)
    {
        auto it = memo.find({x, y});
        if (it != memo.end()) return it->second;
        SparseInt res = compute(x, y);
        return memo.emplace(std::make_pair(x, y), std::move(res)).first->second;
    }

    static void add(std::map<Key, int64_t>& acc, Key k, int64_t c)
    {
        auto& v = acc[k];
        v += c;
        if (v == 0) acc.erase(k);
    }

    SparseInt compute(Key x, Key y)
    {
        if (x == y) return {};
        int64_t sign = 1;
        if (x > y) { std::swap(x, y); sign = -1; }
        if (h.degree(x) + h.degree(y) > h.depth) return {};
        auto f = h.find.find({x, y});
        if (f != h.find.end()) return {{f->second, sign}};
        const Key y1 = h.left(y), y2 = h.right(y);
        std::map<Key, int64_t> acc;
        if (y1 != 0) {
            // [[x, y1], y2]
            const SparseInt a = get(x, y1);
            for (auto [k, c] : a) {
                const SparseInt b = get(k, y2);
                for (auto [k2, c2] : b) add(acc, k2, sign * c * c2);
            }
            // - [[x, y2], y1]
            const SparseInt a2 = get(x, y2);
            for (auto [k, c] : a2) {
                const SparseInt b = get(k, y1);
                for (auto [k2, c2] : b) add(acc, k2, -sign * c * c2);
            }
        }
        return SparseInt(acc.begin(), acc.end());
    }
};

void coo_to_csr(int64_t rows, std::vector<std::pair<std::pair<int64_t, int64_t>, double>>& coo,
                int64_t* offsets, int64_t* indices, double* data)
{
    std::sort(coo.begin(), coo.end(),
              [](const auto& a, const auto& b) { return a.first < b.first; });
    std::fill(offsets, offsets + rows + 1, 0);
    for (auto& e : coo) offsets[e.first.first + 1] += 1;
    for (int64_t r = 0; r < rows; ++r) offsets[r + 1] += offsets[r];
    for (size_t i = 0; i < coo.size(); ++i) {
        indices[i] = coo[i].first.second;
        data[i] = coo[i].second;
    }
}

}  // namespace

int64_t tensor_size_checked(int32_t width, int32_t depth)
{
    if (width < 1 || depth < 0) return -1;
    int64_t s = 1;
    for (int k = 1; k <= depth; ++k) {
        if (s > (INT64_MAX - 1) / width) return -1;
        s = 1 + width * s;
    }
    return s;
}

int64_t lie_size(int32_t width, int32_t depth)
{
    int64_t s = 0;
    for (int d = 1; d <= depth; ++d) s += witt(width, d);
    return s;
}

int build_hall(int32_t width, int32_t depth, int64_t* db, int64_t* data)
{
    int64_t size = 1;
    db[0] = 0;
    data[0] = 0;
    data[1] = 0;
    if (depth > 0) {
        db[1] = 1;
        for (int64_t l = 1; l <= width; ++l) {
            data[2 * l] = 0;
            data[2 * l + 1] = l;
        }
        size += width;
        db[2] = size;
    }
    for (int degree = 2; degree <= depth; ++degree) {
        for (int ld = 1; 2 * ld <= degree; ++ld) {
            const int rd = degree - ld;
            for (int64_t i = db[ld]; i < db[ld + 1]; ++i) {
                for (int64_t j = std::max(i + 1, db[rd]); j < db[rd + 1]; ++j) {
                    if (data[2 * j] <= i) {
                        data[2 * size] = i;
                        data[2 * size + 1] = j;
                        ++size;
                    }
                }
            }
        }
        db[degree + 1] = size;
    }
    return RPY_OK;
}

int check_hall(const rpy_lie_basis* b)
{
    const int64_t* db = b->degree_begin;
    const int64_t* data = b->data;
    if (b->depth < 0 || b->width < 1) return fail(RPY_ERR_BASIS, "invalid width/depth");
    if (db[0] != 0) return fail(RPY_ERR_BASIS, "degree_begin[0] must be 0");
    for (int d = 0; d <= b->depth; ++d)
        if (db[d + 1] < db[d]) return fail(RPY_ERR_BASIS, "degree_begin must be non-decreasing");
    if (db[1] > 0 && (data[0] != 0 || data[1] != 0))
        return fail(RPY_ERR_BASIS, "first element of data must be the neutral element");
    if (b->depth < 1) return RPY_OK;
    if (db[2] - db[1] != b->width)
        return fail(RPY_ERR_BASIS, "degree one must hold exactly `width` letters");
    for (int64_t l = db[1]; l < db[2]; ++l)
        if (data[2 * l] != 0 || data[2 * l + 1] != l)
            return fail(RPY_ERR_BASIS, "data must contain letters represented as pairs (0, l)");
    for (int degree = 2; degree <= b->depth; ++degree) {
        if (db[degree + 1] - db[degree] != witt(b->width, degree))
            return fail(RPY_ERR_BASIS, "number of elements of degree %d does not match the Witt formula",
                        degree);
        for (int64_t e = db[degree]; e < db[degree + 1]; ++e) {
            const int64_t a = data[2 * e], bc = data[2 * e + 1];
            if (a <= 0 || bc <= 0 || a >= e || bc >= e)
                return fail(RPY_ERR_BASIS, "parents of element %lld must be earlier keys",
                            (long long)e);
            if (a >= bc) return fail(RPY_ERR_BASIS, "elements h = [k,l] in the basis must satisfy k < l");
            const int64_t bb = data[2 * bc];
            if (bb != 0 && bb > a)
                return fail(RPY_ERR_BASIS, "when h=[a,[b,c]] we must have b <= a < [b,c]");
        }
    }
    return RPY_OK;
}

int build_l2t(const rpy_lie_basis* b, int64_t* nnz, int64_t* offsets, int64_t* indices,
              double* data)
{
    Hall h;
    load_hall(b, h);
    const int64_t tdim = tensor_size_checked(h.width, h.depth);
    std::vector<int64_t> tdb(h.depth + 2, 0);
    for (int k = 0; k <= h.depth; ++k) tdb[k + 1] = 1 + h.width * tdb[k];
    const Key n = h.db[h.depth + 1];
    // per key: sparse (within-level word index, coeff)
    std::vector<SparseInt> exp(n);
    for (Key k = 1; k < n; ++k) {
        if (h.left(k) == 0) {
            exp[k] = {{h.right(k) - 1, 1}};
            continue;
        }
        const Key a = h.left(k), c = h.right(k);
        const int64_t da = h.degree(a), dc = h.degree(c);
        const int64_t wa = ipow_i64(h.width, (int)da), wc = ipow_i64(h.width, (int)dc);
        std::map<int64_t, int64_t> acc;
        for (auto [i, ci] : exp[a])
            for (auto [j, cj] : exp[c]) {
                acc[i * wc + j] += ci * cj;
                acc[j * wa + i] -= ci * cj;
            }
        SparseInt v;
        for (auto [i, c2] : acc)
            if (c2 != 0) v.push_back({i, c2});
        exp[k] = std::move(v);
    }
    std::vector<std::pair<std::pair<int64_t, int64_t>, double>> coo;
    for (Key k = 1; k < n; ++k) {
        const int d = h.degree(k);
        for (auto [i, c] : exp[k]) coo.push_back({{tdb[d] + i, k - 1}, (double)c});
    }
    *nnz = (int64_t)coo.size();
    if (offsets && indices && data) coo_to_csr(tdim, coo, offsets, indices, data);
    return RPY_OK;
}

int build_t2l(const rpy_lie_basis* b, int64_t* nnz, int64_t* offsets, int64_t* indices,
              double* data)
{
    Hall h;
    load_hall(b, h);
    BracketCache cache(h);
    std::vector<int64_t> tdb(h.depth + 2, 0);
    for (int k = 0; k <= h.depth; ++k) tdb[k + 1] = 1 + h.width * tdb[k];
    std::vector<std::pair<std::pair<int64_t, int64_t>, double>> coo;
    // r(u) for the words of the previous degree, indexed by word index.
    std::vector<SparseInt> prev;
    for (int d = 1; d <= h.depth; ++d) {
        const int64_t nw = ipow_i64(h.width, d);
        const int64_t nprev = d > 1 ? ipow_i64(h.width, d - 1) : 1;
        std::vector<SparseInt> cur(nw);
        for (int64_t idx = 0; idx < nw; ++idx) {
            if (d == 1) {
                cur[idx] = {{idx + 1, 1}};  // letter key idx+1
            } else {
                const Key a = idx / nprev + 1;  // first letter as a key
                const SparseInt& r = prev[idx % nprev];
                std::map<Key, int64_t> acc;
                for (auto [k, c] : r) {
                    const SparseInt& br = cache.get(a, k);
                    for (auto [k2, c2] : br) {
                        auto& v = acc[k2];
                        v += c * c2;
                    }
                }
                SparseInt v;
                for (auto [k, c] : acc)
                    if (c != 0) v.push_back({k, c});
                cur[idx] = std::move(v);
            }
            for (auto [k, c] : cur[idx])
                coo.push_back({{k - 1, tdb[d] + idx}, (double)c * (1.0 / (double)d)});
        }
        prev = std::move(cur);
    }
    *nnz = (int64_t)coo.size();
    if (offsets && indices && data) coo_to_csr(h.size(), coo, offsets, indices, data);
    return RPY_OK;
}

}  // namespace rpy_amd
