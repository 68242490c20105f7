"""Serialised streams and caches answer queries exactly as the originals:
the reference's pickle test (tests/streams/test_stream.py:11-22), piecewise
abelian streams, and the dyadic signature cache dumped to a file and loaded
back without a rebuild (bit-identical nodes and query results)."""
import pickle

import numpy as np
import pytest

from roughpy_amd import serialization as ser
from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu


def test_stream_pickle_roundtrip(gpu):
    data = np.array([[0.4, 0.5, 0.6], [0.7, 0.8, 0.9]])
    s = rs.LieIncrementStream.from_increments(data, width=3, depth=4)
    s2 = pickle.loads(pickle.dumps(s))
    assert np.array_equal(s2.signature().data, s.signature().data)


def test_piecewise_abelian_roundtrip(gpu, rng):
    inc = rng.normal(size=(32, 2)) * 0.3
    s = rs.LieIncrementStream.from_increments(inc, depth=4)
    part = rs.Partition(rs.RealInterval(0.0, 32.0), [5.0, 11.0, 20.0])
    pas = s.simplify(part, 1)
    pas2 = pickle.loads(pickle.dumps(pas))
    for (i1, l1), (i2, l2) in zip(pas.pieces, pas2.pieces):
        assert (i1.inf, i1.sup) == (i2.inf, i2.sup)
        assert np.array_equal(l1, l2)
    for a, b, r in [(0.0, 32.0, 2), (3.0, 17.5, 4)]:
        assert np.array_equal(pas.log_signature(rs.RealInterval(a, b), r).data,
                              pas2.log_signature(rs.RealInterval(a, b), r).data)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dyadic_cache_dump_load(gpu, rng, tmp_path, dtype):
    w, d, n, r, J = 3, 3, 5, 8, 3
    lens = rng.integers(10, 200, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    params = np.concatenate([np.sort(rng.uniform(0, 1, size=L)) for L in lens])
    inc = (rng.normal(size=(int(offs[-1]), w)) / 10).astype(dtype)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J)
    path = tmp_path / "cache.rpyc"
    cache.dump(path)
    loaded = rs.DyadicSignatureCache.load(path)
    assert gpu.equal(loaded.cache, cache.cache)
    nq = 100
    sid = rng.integers(0, n, size=nq)
    a = np.floor(rng.uniform(0, 1, nq) * 2 ** r) / 2 ** r
    b = np.floor(rng.uniform(0, 1, nq) * 2 ** r) / 2 ** r
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    assert gpu.equal(cache.query(sid, lo, hi), loaded.query(sid, lo, hi))
    again = pickle.loads(pickle.dumps(cache))
    assert gpu.equal(again.cache, cache.cache)
