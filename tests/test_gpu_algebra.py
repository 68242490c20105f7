"""The algebra objects' GPU operations (roughpy_amd.algebra) against the
reference's own tests and the oracle: free tensor products and the golden
of tests/algebra/test_free_tensor.py:346-364, exp of a letter (:367-377)
and by Horner's construction (:380-395), fmexp (:398-405), log after exp
(:408-412); the Lie bracket (tests/algebra/test_lie.py Lie_mul); the
context's maps (tests/algebra/test_algebra_context.py), cbh against the
oracle's Context::cbh, compute_signature and to_logsignature against the
oracle signature and log-signature, and the shuffle product against the
oracle.  float64 within 1e-12 normwise per level unless stated."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal, assert_array_equal

from oracle import oracle as O
from roughpy_amd.algebra import DPReal, FreeTensor, Lie, ShuffleTensor, get_context

pytestmark = pytest.mark.gpu


def test_free_tensor_mul_golden(gpu):
    t1 = FreeTensor(np.array([1.0, 2.0, 3.0]), width=2, depth=2)
    t2 = FreeTensor(np.array([1.0, -1.0, -1.0]), width=2, depth=2)
    expected = FreeTensor(np.array([1.0, 1.0, 2.0, -2.0, -2.0, -3.0, -3.0]), width=2, depth=2)
    assert t1 * t2 == expected
    t1 *= t2
    assert t1 == expected


@pytest.mark.parametrize("width", [2, 3, 4, 5])
def test_exp_of_a_letter_and_by_horner(gpu, width):
    depth = 3
    t = FreeTensor(np.array([0.0, 1.0] + [0.0] * (width - 1)), width=width, depth=depth)
    expected = FreeTensor(np.array([1.0, 1.0] + [0.0] * (width - 1) + [0.5] +
                                   [0.0] * (width ** 2 - 1) + [1.0 / 6.0] +
                                   [0.0] * (width ** 3 - 1)), width=width, depth=depth)
    assert t.exp() == expected
    t = FreeTensor(np.array([1.0] * (1 + width)), width=width, depth=depth)
    tunit = FreeTensor(np.array([1.0]), width=width, depth=depth)
    i = float(t.max_degree)
    acc = FreeTensor(np.array([1.0]), width=width, depth=depth)
    assert acc.degree() == 0
    while i:
        acc *= (t / i)
        acc += tunit
        i -= 1.0
    assert acc.degree() == depth
    assert_array_almost_equal(t.exp(), acc)


@pytest.mark.parametrize("width,depth", [(2, 4), (3, 3), (5, 2)])
def test_fmexp_log_antipode(gpu, rng, width, depth):
    n = O.tensor_size(width, depth)
    d1 = rng.uniform(0.0, 1.0, size=n)
    d2 = rng.uniform(1.0, 2.0, size=n)
    d2[0] = 0.0
    t1 = FreeTensor(d1, width=width, depth=depth)
    t2 = FreeTensor(d2, width=width, depth=depth)
    expected = t1 * t2.exp()
    t1.fmexp(t2)
    assert_array_almost_equal(t1, expected)
    assert_array_almost_equal(t2.exp().log(), t2)
    # a group-like element: its antipode is its inverse
    g = get_context(width, depth, DPReal).compute_signature(rng.normal(size=(20, width)) / 5)
    unit = np.zeros(n)
    unit[0] = 1
    assert O.level_errors(np.array(g.antipode() * g), unit, width, depth).max() <= 1e-12
    assert O.level_errors(np.array(g.log()), O.ft_log(np.array(g), width, depth),
                          width, depth).max() <= 1e-12


@pytest.mark.parametrize("width", [2, 3, 4, 5])
def test_lie_bracket(gpu, width):
    ctx = get_context(width, 2, DPReal)
    l1 = Lie(np.array([1.0] + [0.0] * (width - 1)), width=width)
    l2 = Lie(np.array([0.0, 1.0] + [0.0] * (width - 2)), width=width)
    exp_data = np.zeros(ctx.lie_size(2))
    exp_data[width] = 1.0
    assert l1 * l2 == Lie(exp_data, width=width)
    l1 *= l2
    assert l1 == Lie(exp_data, width=width)


@pytest.mark.parametrize("width,depth", [(2, 3), (3, 4), (5, 2)])
def test_context_maps(gpu, rng, width, depth):
    l = Lie(np.array(range(1, width + 1), dtype=np.float64), width=width, depth=depth)
    ctx = get_context(width, depth, DPReal)
    t = ctx.lie_to_tensor(l)
    assert isinstance(t, FreeTensor)
    assert t == FreeTensor(np.array(range(width + 1), dtype=np.float64), width=width,
                           depth=depth)
    back = ctx.tensor_to_lie(t)
    assert isinstance(back, Lie) and back == l
    x = rng.normal(size=ctx.lie_size()) / 4
    lx = Lie(x, width=width, depth=depth)
    assert_array_almost_equal(ctx.tensor_to_lie(ctx.lie_to_tensor(lx)), lx, decimal=14)
    ref = O.lie_to_tensor(x, width, depth)
    assert_array_almost_equal(ctx.lie_to_tensor(lx), ref, decimal=14)


@pytest.mark.parametrize("width,depth", [(2, 4), (3, 4), (4, 3)])
def test_cbh_signature_logsignature(gpu, rng, width, depth):
    ctx = get_context(width, depth, DPReal)
    lies = [rng.normal(size=ctx.lie_size()) / 5 for _ in range(3)]
    got = ctx.cbh([Lie(x, width=width, depth=depth) for x in lies])
    ref = O.cbh(lies, width, depth)
    assert np.abs(np.array(got) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    one = Lie(lies[0], width=width, depth=depth)
    assert ctx.cbh([one]) == one
    assert ctx.cbh([]).size() == 0
    assert np.abs(np.array(ctx.cbh([one, -one]))).max() <= 1e-15
    inc = rng.normal(size=(200, width)) / 14
    sig = ctx.compute_signature(inc)
    assert isinstance(sig, FreeTensor)
    assert O.level_errors(np.array(sig), O.signature(inc, width, depth), width, depth).max() \
        <= 1e-12
    ls = ctx.to_logsignature(sig)
    lref = O.log_signature(inc, width, depth)
    assert np.abs(np.array(ls) - lref).max() <= 1e-12 * max(1.0, np.abs(lref).max())
    assert ctx.compute_signature(np.zeros((0, width))) == FreeTensor(1.0, ctx=ctx)


def test_shuffle_product(gpu, rng):
    w, d = 3, 4
    n = O.tensor_size(w, d)
    a, b = rng.normal(size=n), rng.normal(size=n)
    got = ShuffleTensor(a, width=w, depth=d) * ShuffleTensor(b, width=w, depth=d)
    ref = O.st_mul(a, b, w, d)
    assert O.level_errors(np.array(got), ref, w, d).max() <= 1e-13
