"""GPU parity of Stream.simplify and PiecewiseAbelianStream queries against
the oracle (oracle.simplify / oracle.PiecewiseAbelianStream), whose branch
semantics and Chen identity are pinned in test_oracle_pins.py.

Tolerances: pieces and queries within 1e-11 (float64 data) per level of the
oracle, 2e-5 relative for float32 data (the log-signature is computed in the
data type, the pieces are float64)."""
import numpy as np
import pytest

from oracle import oracle as O
from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu


def _lie_err(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref).max() / max(1.0, np.abs(ref).max())


def _streams(rng, w, d, n, dtype, uneven):
    inc = (rng.normal(size=(n, w)) * 0.25).astype(dtype)
    if uneven:
        idx = np.sort(rng.uniform(0, 4, size=n))
        prod = rs.LieIncrementStream.from_increments(inc, indices=idx, depth=d, resolution=6)
        ref = O.IncrementStream(inc.astype(np.float64), w, d, indices=idx, resolution=6)
    else:
        prod = rs.LieIncrementStream.from_increments(inc, depth=d)
        ref = O.IncrementStream(inc.astype(np.float64), w, d)
    return prod, ref


@pytest.mark.parametrize("w,d,dtype,uneven", [(3, 4, np.float64, False), (2, 5, np.float64, True),
                                              (4, 3, np.float32, True)])
def test_simplify_pieces_and_queries(gpu, rng, w, d, dtype, uneven):
    n = 40
    prod, ref = _streams(rng, w, d, n, dtype, uneven)
    lo, hi = prod.support
    cuts = np.sort(rng.uniform(lo, hi, size=5))
    part = rs.Partition(rs.RealInterval(lo, hi), cuts)
    res = prod.resolution
    pas = prod.simplify(part, res)
    ivs = [(part[i].inf, part[i].sup) for i in range(len(part))]
    opas = O.simplify(ref, ivs, res)
    tol = 1e-11 if dtype == np.float64 else 2e-5
    assert len(pas.pieces) == len(opas.pieces)
    for (iv, got), ((a, b), exp) in zip(pas.pieces, opas.pieces):
        assert (iv.inf, iv.sup) == (a, b)
        assert _lie_err(got, exp) <= tol
    # queries: whole support, inside one piece, across pieces, partial
    # overlaps at finer resolutions, outside the support
    qs = [(lo, hi, None), (lo, hi, 2), (cuts[0], cuts[3], 3), (cuts[1] + 0.01, cuts[2], 5),
          (lo - 1, lo + 0.3, 4), (hi + 1, hi + 2, None), ((lo + hi) / 2, hi, 1)]
    for a, b, r in qs:
        got = pas.log_signature(rs.RealInterval(a, b), r).data
        exp = opas.log_signature(a, b, r)
        assert _lie_err(got, exp) <= tol, (a, b, r)
    got = pas.signature(rs.RealInterval(lo, hi), 3).data
    exp = opas.signature(lo, hi, 3)
    assert _lie_err(got, exp) <= tol


def test_simplified_signature_matches_the_stream(gpu, rng):
    """Chen's identity through the GPU path: unit pieces at resolution 0
    reproduce the stream's signature over the support."""
    w, d, n = 3, 4, 16
    inc = rng.normal(size=(n, w)) * 0.3
    prod = rs.LieIncrementStream.from_increments(inc, depth=d)
    part = rs.Partition(rs.RealInterval(0.0, float(n)), np.arange(1, n))
    pas = prod.simplify(part, 0)
    got = pas.signature(rs.RealInterval(0.0, float(n)), 0).data
    assert _lie_err(got, O.signature(inc, w, d)) <= 1e-12
    assert _lie_err(got, prod.signature(rs.RealInterval(0.0, float(n)), 0).data) <= 1e-12


def test_resimplify(gpu, rng):
    """A piecewise abelian stream simplified again (its own queries become
    the pieces) against the oracle."""
    w, d, n = 2, 4, 24
    prod, ref = _streams(rng, w, d, n, np.float64, False)
    part = rs.Partition(rs.RealInterval(0.0, float(n)), [3.0, 7.5, 12.0, 20.0])
    pas = prod.simplify(part, 1)
    opas = O.simplify(ref, [(part[i].inf, part[i].sup) for i in range(len(part))], 1)
    coarse = rs.Partition(rs.RealInterval(0.0, float(n)), [8.0, 16.0])
    pas2 = pas.simplify(coarse, 2)
    pieces2 = [((coarse[i].inf, coarse[i].sup),
                opas.log_signature(coarse[i].inf, coarse[i].sup, 2)) for i in range(len(coarse))]
    for (iv, got), (_, exp) in zip(pas2.pieces, pieces2):
        assert _lie_err(got, exp) <= 1e-11
