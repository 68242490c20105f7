"""GPU parity of the shuffle product and its adjoint against the oracle,
the reference kernels' vectors (tests/golden/ref_vectors.npz) and the
reference's shuffle goldens (tests/compute/test_shuffle.py:58-112).

Tolerances: float64 within 1e-13 normwise per level, float32 within 1e-5 of
the float64 oracle on the same inputs (the GPU fuses multiply-adds)."""
import os

import numpy as np
import pytest
from numpy.testing import assert_array_equal

import roughpy_amd.compute as rc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REF = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
_KEYS = sorted({k.split("/")[0] for k in REF.files})


def _level_err(x, ref, w, d):
    return O.level_errors(np.asarray(x, np.float64), np.asarray(ref, np.float64), w, d).max()


def _flat(w, *letters):
    db = O.tensor_degree_begin(w, len(letters))
    i = 0
    for a in letters:
        i = i * w + a
    return int(db[len(letters)]) + i


def test_shuffle_goldens_on_gpu(gpu):
    w, d = 4, 3
    tb = rc.TensorBasis(w, d)
    size = tb.size()
    lhs = np.zeros(size)
    rhs = np.zeros(size)
    lhs[_flat(w, 0, 1)] = 1.0
    rhs[_flat(w, 2)] = 1.0
    got = rc.st_mul(rc.ShuffleTensor(lhs, tb), rc.ShuffleTensor(rhs, tb)).data
    expected = np.zeros(size)
    for word in [(0, 1, 2), (0, 2, 1), (2, 0, 1)]:
        expected[_flat(w, *word)] = 1.0
    assert_array_equal(got, expected)


@pytest.mark.parametrize("key", _KEYS)
def test_shuffle_vs_reference_vectors(gpu, key):
    w = int(key.split("_")[0][1:])
    d = int(key.split("_")[1][1:])
    g = {name.split("/")[1]: REF[name] for name in REF.files if name.startswith(key + "/")}
    tol = 1e-13 if g["a"].dtype == np.float64 else 1e-5
    tb = rc.TensorBasis(w, d)
    out = rc.ShuffleTensor(g["c"].copy(), tb)
    rc.st_fma(out, rc.ShuffleTensor(g["a"], tb), rc.ShuffleTensor(g["b"], tb))
    assert _level_err(out.data, g["st_fma"], w, d) <= tol
    adj = rc.st_adj_mul(rc.ShuffleTensor(g["a"], tb), rc.FreeTensor(g["c"], tb)).data
    assert _level_err(adj, g["st_adj_mul"], w, d) <= tol


@pytest.mark.parametrize("w,d", [(2, 9), (3, 5), (4, 6), (5, 4), (8, 3), (6, 4), (7, 3), (3, 7)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_shuffle_batched_vs_oracle(gpu, rng, w, d, dtype):
    tb = rc.TensorBasis(w, d)
    size = tb.size()
    a = rng.uniform(-1, 1, size=(3, size)).astype(dtype)
    b = rng.uniform(-1, 1, size=(3, size)).astype(dtype)
    c = rng.uniform(-1, 1, size=(3, size)).astype(dtype)
    tol = 1e-13 if dtype == np.float64 else 2e-5
    ref = c.astype(np.float64)
    O.st_fma(ref, a.astype(np.float64), b.astype(np.float64), w, d)
    out = rc.ShuffleTensor(c.copy(), tb)
    rc.st_fma(out, rc.ShuffleTensor(a, tb), rc.ShuffleTensor(b, tb))
    assert _level_err(out.data, ref, w, d) <= tol
    if size <= 6000:
        ref_adj = O.st_adj_mul(a.astype(np.float64), c.astype(np.float64), w, d)
        got = rc.st_adj_mul(rc.ShuffleTensor(a, tb), rc.FreeTensor(c, tb)).data
        assert _level_err(got, ref_adj, w, d) <= tol


def test_shuffle_adjoint_identity_on_gpu(gpu, rng):
    """<st_adj_mul(op, arg), x> = <arg, op (shuffle) x> for random x."""
    w, d = 3, 4
    tb = rc.TensorBasis(w, d)
    size = tb.size()
    op, arg, x = (rng.normal(size=size) for _ in range(3))
    adj = rc.st_adj_mul(rc.ShuffleTensor(op, tb), rc.FreeTensor(arg, tb)).data
    sh = rc.st_mul(rc.ShuffleTensor(op, tb), rc.ShuffleTensor(x, tb)).data
    assert abs(np.dot(adj, x) - np.dot(arg, sh)) <= 1e-12 * max(1.0, abs(np.dot(arg, sh)))


@pytest.mark.parametrize("w,d,dtype", [(10, 2, np.float64), (9, 3, np.float32),
                                       (8, 6, np.float64), (4, 8, np.float32),
                                       (2, 12, np.float64)])
def test_shuffle_generic_and_global_paths(gpu, rng, w, d, dtype):
    """Widths 9 and 10 (outside the fixed-shape kernels), fixed shapes whose
    operands do not fit LDS (the global-memory variants: width 8 depth 6
    float64, width 4 depth 8 float32) and the deepest fixed shape."""
    tb = rc.TensorBasis(w, d)
    size = tb.size()
    a = rng.uniform(-1, 1, size=(2, size)).astype(dtype)
    b = rng.uniform(-1, 1, size=(2, size)).astype(dtype)
    tol = 1e-12 if dtype == np.float64 else 5e-5
    ref = O.st_mul(a.astype(np.float64), b.astype(np.float64), w, d)
    got = rc.st_mul(rc.ShuffleTensor(a, tb), rc.ShuffleTensor(b, tb)).data
    assert _level_err(got, ref, w, d) <= tol
    ref_adj = O.st_adj_mul(a.astype(np.float64), b.astype(np.float64), w, d)
    got_adj = rc.st_adj_mul(rc.ShuffleTensor(a, tb), rc.FreeTensor(b, tb)).data
    assert _level_err(got_adj, ref_adj, w, d) <= tol
