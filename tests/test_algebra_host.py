"""Host side of the algebra objects (roughpy_amd.algebra): construction,
sizes, degrees, formats and linear arithmetic as the reference's tests
state them (tests/algebra/test_free_tensor.py, test_lie.py,
test_algebra_context.py); no GPU call."""
import pickle
from array import array

import numpy as np
import pytest
from numpy.testing import assert_array_equal

from roughpy_amd.algebra import (DPReal, FreeTensor, Lie, SPReal, ShuffleTensor, TensorKey,
                                 VectorType, get_context)


def test_create_scalars_and_lists():
    for v in (1.0, 1):
        t = FreeTensor(v, width=2, depth=2)
        a = np.array(t)
        assert a.flags.owndata and a.dtype == np.float64
        assert_array_equal(a, [1., 0, 0, 0, 0, 0, 0])
        assert (t.width, t.max_degree) == (2, 2)
    t = FreeTensor([0, 1, 2, 3], width=3, depth=2)
    assert_array_equal(np.array(t), [0., 1., 2., 3.] + [0.] * 9)
    assert FreeTensor([0, 1, 2, 3], depth=2).width == 3


def test_create_buffers_and_dlpack():
    t = FreeTensor(array('f', [0., 1., 2., 3.]), width=3, depth=2)
    assert np.array(t).dtype == np.float32 and t.dtype is SPReal
    t = FreeTensor(array('d', [0., 1., 2., 3.]), width=3, depth=2)
    assert np.array(t).dtype == np.float64 and t.dtype is DPReal

    class Dummy:
        def __init__(self, a):
            self.array = a

        def __dlpack__(self, stream=None):
            return self.array.__dlpack__(stream=stream)

    t = FreeTensor(Dummy(np.array([0., 1., 2., 3.])), width=3, depth=2)
    assert_array_equal(np.array(t), [0., 1., 2., 3.] + [0.] * 9)


def test_create_key_value_data():
    assert FreeTensor((1, 1.0), width=2, depth=2).storage_type == VectorType.SparseVector
    t = FreeTensor([(1, 1.0), (2, 2.0)], width=2, depth=2,
                   vector_type=VectorType.DenseVector)
    assert t.storage_type == VectorType.DenseVector
    assert_array_equal(np.array(t), [0., 1., 2., 0, 0, 0, 0])
    k1 = TensorKey(1, width=2, depth=2)
    t = FreeTensor([(k1, 1.0), (TensorKey(2, width=2, depth=2), 2.0)], width=2, depth=2)
    assert t[k1] == 1.0 and t[2] == 2.0 and t[(2,)] == 2.0
    assert FreeTensor({1: 1., 2: 2.}, width=2, depth=2)[1] == 1.0
    with pytest.raises(ValueError):
        FreeTensor([{1: 1., 2: 2.}, {0: 1., 2: 1.}], width=2, depth=2)
    ctx = get_context(2, 3, DPReal)
    with pytest.raises(ValueError, match="unexpected string as tensor data"):
        FreeTensor("string_free_tensor_data", ctx=ctx)
    with pytest.raises(ValueError, match="unexpected string in key-scalar data"):
        FreeTensor(["string_free_tensor_data_item"], ctx=ctx)


@pytest.mark.parametrize("width", [2, 3, 4, 5])
def test_sizes_and_degrees(rng, width):
    data = rng.uniform(0.0, 1.0, size=1 + width)
    t = FreeTensor(data)
    assert (t.width, t.degree(), t.size()) == (width, 1, data.size)
    data = rng.uniform(0.0, 1.0, size=1 + width * (1 + width))
    t = FreeTensor(data, width=width)
    assert (t.degree(), t.size()) == (2, data.size)
    data = rng.uniform(1.0, 2.0, size=1 + width + 1)
    t = FreeTensor(data, width=width, depth=2)
    assert (t.degree(), t.size(), np.array(t).size) == (2, data.size, 1 + width * (1 + width))
    assert_array_equal(np.array(t), np.concatenate((data, np.zeros(width ** 2 - 1))))
    ctx = get_context(width, 2, DPReal)
    l = Lie(rng.uniform(0.0, 1.0, size=width))
    assert (l.width, l.degree(), l.size()) == (width, 1, width)
    l = Lie(rng.uniform(0.0, 1.0, size=ctx.lie_size(2)), width=width)
    assert (l.width, l.size(), l.degree()) == (width, ctx.lie_size(2), 2)


def test_formats():
    t = FreeTensor(np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]), width=2, depth=2)
    assert str(t) == "{ 1() 2(1) 3(2) 4(1,1) 5(1,2) 6(2,1) 7(2,2) }"
    assert repr(t) == "FreeTensor(width=2, depth=2, ctype=DPReal)"
    l = Lie(np.array([1.0, 2.0, 3.0]), width=2, depth=2)
    assert str(l) == "{ 1(1) 2(2) 3([1,2]) }"
    assert repr(l) == "Lie(width=2, depth=2, ctype=DPReal)"
    assert repr(ShuffleTensor([1.0, 2.0], width=1, depth=2)) == \
        "ShuffleTensor(width=1, depth=2, ctype=DPReal)"


def test_linear_arithmetic(rng):
    d1, d2 = rng.uniform(0, 1, size=13), rng.uniform(1, 2, size=13)
    for cls in (FreeTensor, ShuffleTensor):
        t1, t2 = cls(d1, width=3, depth=2), cls(d2, width=3, depth=2)
        assert t1 + t2 == cls(d1 + d2, width=3, depth=2)
        assert t1 - t2 == cls(d1 - d2, width=3, depth=2)
        assert t1 * 2.0 == cls(2.0 * d1, width=3, depth=2) == 2.0 * t1
        assert t1 / 2.0 == cls(d1 / 2.0, width=3, depth=2)
        assert -t1 == cls(-d1, width=3, depth=2)
        t = cls(d1, width=3, depth=2)
        t += t2
        t -= t2
        t *= 4.0
        t /= 2.0
        assert t == cls(((d1 + d2) - d2) * 4.0 / 2.0, width=3, depth=2)
        t = cls(d1, width=3, depth=2)
        t.add_scal_mul(t2, 3.0)
        t.sub_scal_div(t2, 2.0)
        assert t == cls(d1 + d2 * 3.0 - d2 / 2.0, width=3, depth=2)
    l1, l2 = Lie(d1[:6], width=3, depth=2), Lie(d2[:6], width=3, depth=2)
    assert l1 + l2 == Lie(d1[:6] + d2[:6], width=3, depth=2)
    assert l1 != l2
    with pytest.raises(TypeError):
        Lie(1.0, width=2, depth=2)
    with pytest.raises(ValueError):
        FreeTensor(d1, width=3, depth=2) + FreeTensor(d1[:4], width=3, depth=3)


def test_context():
    ctx = get_context(3, 4, DPReal)
    assert get_context(3, 4, "DPReal") is ctx and get_context(3, 4, SPReal) is not ctx
    assert (ctx.width, ctx.depth) == (3, 4)
    assert ctx.lie_size(1) == 3 and ctx.tensor_size(1) == 4
    assert ctx.lie_size() == 32 and ctx.tensor_size() == 121
    with pytest.raises(ValueError):
        ctx.tensor_size(5)
    z = ctx.zero_lie()
    assert isinstance(z, Lie) and (z.width, z.max_degree, z.size(), z.degree()) == (3, 4, 0, 0)
    with ctx as c:
        assert c is ctx
    with pytest.raises(TypeError):
        ctx.lie_to_tensor(FreeTensor([1.0, 2.0], width=1, depth=2))


def test_pickle_and_keys():
    t = FreeTensor([1.0, 2.0, 0.0, 4.0], width=3, depth=2)
    t2 = pickle.loads(pickle.dumps(t))
    assert t2 == t and t2.size() == 4 and t2.dtype is DPReal
    items = {str(i.key()): i.value() for i in t}
    assert items == {"()": 1.0, "(1)": 2.0, "(3)": 4.0}
    assert TensorKey(2, 1, width=3, depth=2).index() == 1 + 3 + 3
    assert TensorKey(2, 1, width=3, depth=2).degree() == 2
