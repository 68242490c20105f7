"""GPU parity of the dyadic signature cache and its interval queries.

Nodes are checked against the oracle signature of the rows they cover;
queries against the oracle's restatement of Stream.signature for a
LieIncrementStream of the cache resolution (oracle.IncrementStream).
float64 within 1e-12 normwise per level; float32 within 2e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu


def _streams(rng, n, w, lmin, lmax, dtype):
    lens = rng.integers(lmin, lmax + 1, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    params = np.concatenate([np.arange(L) / L for L in lens]).astype(np.float64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.repeat(np.maximum(lens, 1), lens))[:, None])
    return inc.astype(dtype), params, offs, lens


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
@pytest.mark.parametrize("w,d", [(3, 4), (2, 9)])
def test_cache_nodes_and_queries(gpu, rng, dtype, tol, w, d):
    """(3, 4) builds parent nodes with the fixed-shape product that writes
    them; (2, 9) has no fixed-shape kernel and clears them, then adds."""
    r, J = 8, 3
    inc, params, offs, lens = _streams(rng, 7, w, 0, 300, dtype)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J, dtype=dtype)
    inc64 = inc.astype(np.float64)
    # every stored node equals the signature of the rows it covers
    for s in range(7):
        p = params[offs[s]:offs[s + 1]]
        x = inc64[offs[s]:offs[s + 1]]
        for j in range(J + 1):
            for k in range(2 ** j):
                sel = (p >= k / 2 ** j) & (p < (k + 1) / 2 ** j)
                ref = O.signature(x[sel], w, d)
                got = cache.node(j, s, k).cpu().numpy()
                assert O.level_errors(got, ref, w, d).max() <= tol
    # random clopen queries, including empty, reversed, out-of-range and full ones
    nq = 300
    sid = rng.integers(0, 7, size=nq)
    a = rng.uniform(-0.2, 1.2, size=nq)
    b = rng.uniform(-0.2, 1.2, size=nq)
    a[:5] = [0.0, 0.0, 0.5, 0.3, 1.0]
    b[:5] = [1.0, 0.0, 0.5, 0.2, 1.5]
    out = cache.query(sid, a, b).cpu().numpy()
    for q in range(nq):
        s = sid[q]
        st = O.IncrementStream(inc64[offs[s]:offs[s + 1]], w, d,
                               indices=params[offs[s]:offs[s + 1]], resolution=r)
        if lens[s] == 0 or b[q] <= a[q]:
            ref = np.eye(1, O.tensor_size(w, d)).ravel()
        else:
            # the cache domain is [0, 1); the stream's support ends at max + 2^-r <= 1
            ref = st.signature(max(a[q], 0.0), min(b[q], 1.0), r)
        assert O.level_errors(out[q], ref, w, d).max() <= tol, (q, s, a[q], b[q])


@pytest.mark.parametrize("w,d,dtype,tol", [(5, 5, np.float64, 1e-12), (8, 4, np.float32, 2e-5),
                                           (2, 7, np.float32, 2e-5)])
def test_queries_of_long_runs_and_many_nodes(gpu, rng, w, d, dtype, tol):
    """Streams longer than one staging chunk per fine run (resolution 10,
    stored levels 2, so a fine run spans up to 256 leaves) and queries over
    many stored nodes, against the oracle signature of the selected rows."""
    r, J = 10, 2
    inc, params, offs, lens = _streams(rng, 4, w, 600, 1100, dtype)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J, dtype=dtype)
    nq = 16
    sid = rng.integers(0, 4, size=nq)
    a = np.floor(rng.uniform(0, 1, size=nq) * 2 ** r) / 2 ** r
    b = np.floor(rng.uniform(0, 1, size=nq) * 2 ** r) / 2 ** r
    a, b = np.minimum(a, b), np.maximum(a, b)
    a[0], b[0] = 1 / 2 ** r, 1 - 1 / 2 ** r
    out = cache.query(sid, a, b).cpu().numpy()
    inc64 = inc.astype(np.float64)
    for q in range(nq):
        s = sid[q]
        sel = (params[offs[s]:offs[s + 1]] >= a[q]) & (params[offs[s]:offs[s + 1]] < b[q])
        ref = O.signature(inc64[offs[s]:offs[s + 1]][sel], w, d)
        assert O.level_errors(out[q], ref, w, d).max() <= tol, q


def test_cache_benchmark_shape_sample(gpu):
    """Width 5, depth 5, resolution 12, stored levels 6 -- the benchmark's
    cache shape -- on 64 streams, with queries checked against the oracle."""
    rng = np.random.default_rng(12345 + 4)
    w, d, r, J = 5, 5, 12, 6
    inc, params, offs, lens = _streams(rng, 64, w, 64, 4096, np.float32)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J)
    nq = 40
    sid = rng.integers(0, 64, size=nq)
    a = np.floor(rng.uniform(0, 1, size=nq) * 4096) / 4096
    b = np.floor(rng.uniform(0, 1, size=nq) * 4096) / 4096
    a, b = np.minimum(a, b), np.maximum(a, b)
    out = cache.query(sid, a, b).cpu().numpy()
    inc64 = inc.astype(np.float64)
    for q in range(nq):
        s = sid[q]
        sel = (params[offs[s]:offs[s + 1]] >= a[q]) & (params[offs[s]:offs[s + 1]] < b[q])
        ref = O.signature(inc64[offs[s]:offs[s + 1]][sel], w, d)
        assert O.level_errors(out[q], ref, w, d).max() <= 2e-5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [40, 3000])
def test_cache_leaves_at_the_benchmark_shape(gpu, rng, dtype, n):
    """At width 5, depth 5 (stored levels 6) each level-J leaf is the
    signature of its segment's rows as the batched signature call computes
    it (float32 bit for bit; float64 within 1e-13), including empty
    streams, empty leading, inner and trailing segments, and rows outside
    [0, 1); 3000 streams take the length-sorted queue.  A sample of nodes of
    every level is checked against the oracle."""
    w, d, r, J = 5, 5, 12, 6
    lens = rng.integers(0, 200, size=n)
    lens[:4] = [0, 1, 2, 150]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    params = []
    for i, L in enumerate(lens):
        p = np.sort(rng.uniform(-0.05, 1.05, size=L)) if i % 3 == 0 else np.arange(L) / max(L, 1)
        if i % 5 == 1:
            p = 0.3 + 0.1 * p    # every row in a few segments: empty ones around them
        params.append(p)
    params = np.concatenate(params).astype(np.float64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / 8).astype(dtype)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J, dtype=dtype)
    # segment row ranges: rows with params in [k 2^-J, (k+1) 2^-J)
    nseg = 2 ** J
    bounds = np.empty((n, nseg + 1), dtype=np.int64)
    for s in range(n):
        p = params[offs[s]:offs[s + 1]]
        bounds[s] = offs[s] + np.searchsorted(p, np.arange(nseg + 1) / nseg, side="left")
    seg_offs_b = bounds[:, :-1].ravel()
    seg_offs_e = bounds[:, 1:].ravel()
    # the same segments through the batched signature call (rows gathered per segment)
    rows = np.concatenate([np.arange(b0, e0) for b0, e0 in zip(seg_offs_b, seg_offs_e)])
    seg_offs = np.concatenate([[0], np.cumsum(seg_offs_e - seg_offs_b)]).astype(np.int64)
    ref = rs.signature_batch(inc[rows], seg_offs, w, d)
    ts = O.tensor_size(w, d)
    first = (nseg - 1) * n * ts
    leaves_all = cache.cache[first:first + n * nseg * ts].cpu().numpy().reshape(n * nseg, ts)
    leaves = np.stack([cache.node(J, s, k).cpu().numpy() for s in range(3) for k in range(nseg)])
    if dtype == np.float32:
        np.testing.assert_array_equal(leaves_all, ref)
    else:
        err = O.level_errors(leaves_all, ref, w, d).max()
        assert err <= 1e-13, err
    assert np.array_equal(leaves, leaves_all[:len(leaves)])
    # a sample of every level against the oracle
    inc64 = inc.astype(np.float64)
    tol = 2e-5 if dtype == np.float32 else 1e-12
    for s in range(6):
        p = params[offs[s]:offs[s + 1]]
        x = inc64[offs[s]:offs[s + 1]]
        for j in range(J + 1):
            for k in sorted({0, 2 ** j - 1, (2 ** j) // 3}):
                sel = (p >= k / 2 ** j) & (p < (k + 1) / 2 ** j)
                want = O.signature(x[sel], w, d)
                got = cache.node(j, s, k).cpu().numpy()
                assert O.level_errors(got, want, w, d).max() <= tol, (s, j, k)


def test_queries_on_the_workgroup_path(gpu, rng):
    """Width 2, depth 9 has no query kernel of its own, so its queries take
    the all-workgroup kernel; same oracle check."""
    w, d, r, J = 2, 9, 8, 3
    inc, params, offs, lens = _streams(rng, 3, w, 20, 200, np.float64)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J,
                                    dtype=np.float64)
    nq = 24
    sid = rng.integers(0, 3, size=nq)
    a = np.floor(rng.uniform(0, 1, size=nq) * 2 ** r) / 2 ** r
    b = np.floor(rng.uniform(0, 1, size=nq) * 2 ** r) / 2 ** r
    a, b = np.minimum(a, b), np.maximum(a, b)
    out = cache.query(sid, a, b).cpu().numpy()
    for q in range(nq):
        s = sid[q]
        sel = (params[offs[s]:offs[s + 1]] >= a[q]) & (params[offs[s]:offs[s + 1]] < b[q])
        ref = O.signature(inc[offs[s]:offs[s + 1]][sel], w, d)
        assert O.level_errors(out[q], ref, w, d).max() <= 1e-12, q


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
@pytest.mark.parametrize("support,r", [((-3.7, 1000.2), 2), ((1e6 + 0.25, 1e6 + 9.5), 4),
                                       ((0.0, 1000.0), 0)])
def test_cache_over_any_support(gpu, rng, dtype, tol, support, r):
    """A cache whose streams live on an arbitrary parameter range (negative,
    far from zero, default integer indices): nodes below the root are the
    rows' signatures, and queries equal the oracle's Stream.signature at the
    stream resolution over the same rows."""
    w, d, n = 3, 4, 5
    lo, hi = support
    grid = np.arange(np.ceil(lo * 2 ** r), np.floor(hi * 2 ** r)) / 2 ** r
    streams = []
    for s in range(n):
        m = int(rng.integers(1, min(400, grid.size) + 1))
        p = np.sort(rng.choice(grid, size=m, replace=False)) + 2.0 ** -(r + 2)
        p = p[p < hi]
        streams.append((p, rng.normal(size=(p.size, w)) / np.sqrt(p.size)))
    offs = np.concatenate([[0], np.cumsum([p.size for p, _ in streams])]).astype(np.int64)
    params = np.concatenate([p for p, _ in streams])
    inc = np.concatenate([x for _, x in streams]).astype(dtype)
    J = 3
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J, dtype=dtype,
                                    support=support)
    o, k = cache.origin, cache.root_power
    assert o <= lo and o + 2.0 ** k >= hi and o * 2 ** r == np.floor(o * 2 ** r)
    assert np.array_equal(cache.params().cpu().numpy(), params)
    inc64 = inc.astype(np.float64)
    for s in range(n):
        p, x = params[offs[s]:offs[s + 1]], inc64[offs[s]:offs[s + 1]]
        for j in range(J + 1):
            wdt = 2.0 ** (k - j)
            for kk in range(2 ** j):
                sel = (p >= o + kk * wdt) & (p < o + (kk + 1) * wdt)
                got = cache.node(j, s, kk).cpu().numpy()
                assert O.level_errors(got, O.signature(x[sel], w, d), w, d).max() <= tol
    nq = 200
    sid = rng.integers(0, n, size=nq)
    span = hi - lo
    a = rng.uniform(lo - 0.1 * span, hi, size=nq)
    b = a + rng.uniform(0, 0.6 * span, size=nq)
    a[:3], b[:3] = [lo, lo, hi], [hi, lo, hi + 1]
    out = cache.query(sid, a, b).cpu().numpy()
    for q in range(nq):
        s = sid[q]
        st = O.IncrementStream(inc64[offs[s]:offs[s + 1]], w, d,
                               indices=params[offs[s]:offs[s + 1]], resolution=r)
        lo_q, hi_q = max(a[q], o), min(b[q], o + 2.0 ** k)
        if hi_q <= lo_q:
            ref = np.eye(1, O.tensor_size(w, d)).ravel()
        else:
            ref = st.signature(lo_q, hi_q, r)
        assert O.level_errors(out[q], ref, w, d).max() <= tol, (q, s, a[q], b[q])


def test_stream_queries_read_the_cache(gpu, rng):
    """LieIncrementStream.signature at the stream resolution goes through the
    stream's dyadic cache (built once, kept); other resolutions select rows
    on the host; both agree with the oracle."""
    w, d, n = 3, 4, 1000
    inc = rng.normal(0.0, np.sqrt(1.0 / n), size=(n, w))
    s = rs.LieIncrementStream.from_increments(inc, width=w, depth=d)
    so = O.IncrementStream(inc, w, d)
    for a, b in [(0, 1000), (10.5, 500.25), (999, 1000), (123, 124), (-5, 3), (-5, 2000)]:
        got = s.signature(a, b).data
        assert O.level_errors(got, so.signature(a, b), w, d).max() <= 1e-12, (a, b)
    cache = s._query_cache()
    assert cache is not None and cache.origin == 0.0 and cache.root_power == 10
    assert cache.levels == 4
    assert s._query_cache() is cache
    # a finer query resolution takes the host path, with the same answer
    got = s.signature(rs.RealInterval(10.0, 500.0), 3).data
    assert O.level_errors(got, so.signature(10.0, 500.0, 3), w, d).max() <= 1e-12
