"""GPU parity of the dense free-tensor operations against the oracle and the
reference kernels' vectors (tests/golden/ref_vectors.npz).

Tolerances: float64 within 1e-12 normwise per level (max|x-y| / max|y| per
level); float32 within 2e-5 of the float64 oracle on the same inputs.
"""
import json
import os

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

import roughpy_amd.compute as rc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
REF = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))

TOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 2e-5}


def _close(x, ref, w, d, dtype):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    err = O.level_errors(x, ref, w, d).max()
    assert err <= TOL[np.dtype(dtype)], f"normwise level error {err:.3e}"


def _rand(rng, w, d, batch, dtype, unit=None):
    db = O.tensor_degree_begin(w, d)
    x = rng.uniform(-1, 1, size=batch + (int(db[-1]),))
    f = 1.0
    for k in range(1, d + 1):
        f *= k
        x[..., db[k]:db[k + 1]] /= f
    if unit is not None:
        x[..., 0] = unit
    return x.astype(dtype)


def test_ft_fma_explicit_golden(gpu):
    g = GOLD["ft_fma_explicit"]
    tb = rc.TensorBasis(2, 2)
    for dt in (np.float32, np.float64):
        a = rc.FreeTensor(np.zeros(7, dtype=dt), tb)
        rc.ft_fma(a, rc.FreeTensor(np.array(g["b"], dt), tb), rc.FreeTensor(np.array(g["c"], dt), tb))
        assert_array_equal(a.data, np.array(g["expected"], dt))


def test_mul_and_exp_goldens(gpu):
    g = GOLD["free_tensor_mul"]
    tb = rc.TensorBasis(2, 2)
    r = rc.ft_mul(rc.FreeTensor(np.array(g["t1"]), tb), rc.FreeTensor(np.array(g["t2"]), tb))
    assert_array_equal(r.data, g["expected"])
    for name in ("exp_zero", "exp_letter"):
        e = rc.ft_exp(rc.FreeTensor(np.array(GOLD[name]["x"], float), tb))
        assert_array_equal(e.data, GOLD[name]["expected"])
    lg = rc.ft_log(rc.FreeTensor(np.array(GOLD["log_identity"]["x"], float), tb))
    assert_array_equal(lg.data, GOLD["log_identity"]["expected"])


_REF_KEYS = sorted({k.split("/")[0] for k in REF.files})


@pytest.mark.parametrize("key", _REF_KEYS)
def test_against_reference_kernel_vectors(gpu, key):
    w = int(key.split("_")[0][1:])
    d = int(key.split("_")[1][1:])
    g = {n.split("/")[1]: REF[n] for n in REF.files if n.startswith(key + "/")}
    dt = g["a"].dtype
    tb = rc.TensorBasis(w, d)
    F = lambda x: rc.FreeTensor(x.copy(), tb)  # noqa: E731
    out = F(g["c"])
    rc.ft_fma(out, F(g["a"]), F(g["b"]))
    _close(out.data, g["fma"], w, d, dt)
    im = F(g["a"])
    rc.ft_inplace_mul(im, F(g["b"]))
    _close(im.data, g["inplace_mul"], w, d, dt)
    _close(rc.ft_exp(F(g["x"])).data, g["exp"], w, d, dt)
    _close(rc.ft_fmexp(F(g["a"]), F(g["x"])).data, g["fmexp"], w, d, dt)
    _close(rc.ft_log(F(g["a"])).data, g["log"], w, d, dt)
    assert_array_equal(rc.antipode(F(g["c"])).data, g["antipode"])
    _close(rc.ft_adjoint_left_mul(F(g["a"]), rc.ShuffleTensor(g["c"].copy(), tb)).data,
           g["adj_lmul"], w, d, dt)


@pytest.mark.parametrize("w,d", [(2, 5), (3, 4), (4, 6), (5, 5), (8, 4), (6, 5), (2, 9)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dense_ops_batched_vs_oracle(gpu, rng, w, d, dtype):
    batch = (3, 2)
    tb = rc.TensorBasis(w, d)
    a = _rand(rng, w, d, batch, np.float64, unit=1.0)
    b = _rand(rng, w, d, batch, np.float64, unit=1.0)
    c = _rand(rng, w, d, batch, np.float64)
    x = _rand(rng, w, d, batch, np.float64, unit=0.0)
    A, B, C, X = (v.astype(dtype) for v in (a, b, c, x))
    a64, b64, c64, x64 = (V.astype(np.float64) for V in (A, B, C, X))
    out = rc.FreeTensor(C.copy(), tb)
    rc.ft_fma(out, rc.FreeTensor(A, tb), rc.FreeTensor(B, tb))
    ref = c64.copy()
    O.ft_fma(ref, a64, b64, w, d)
    _close(out.data, ref, w, d, dtype)
    im = rc.FreeTensor(A.copy(), tb)
    rc.ft_inplace_mul(im, rc.FreeTensor(B, tb))
    ref = a64.copy()
    O.ft_inplace_mul(ref, b64, w, d)
    _close(im.data, ref, w, d, dtype)
    if O.tensor_size(w, d) <= 5000:
        _close(rc.ft_exp(rc.FreeTensor(X, tb)).data, O.ft_exp(x64, w, d), w, d, dtype)
        _close(rc.ft_log(rc.FreeTensor(A, tb)).data, O.ft_log(a64, w, d), w, d, dtype)
        _close(rc.ft_fmexp(rc.FreeTensor(A, tb), rc.FreeTensor(X, tb)).data,
               O.ft_fmexp(a64, x64, w, d), w, d, dtype)
    assert_array_equal(rc.antipode(rc.FreeTensor(C, tb)).data, O.antipode(C, w, d))
    # adjoint of left multiplication (op A, argument C) against the oracle
    adj = rc.ft_adjoint_left_mul(rc.FreeTensor(A, tb), rc.ShuffleTensor(C.copy(), tb)).data
    _close(adj, O.ft_adjoint_left_mul(a64, c64, w, d), w, d, dtype)


def test_degree_bounds_and_torch_inputs(gpu, rng):
    import ctypes
    from roughpy_amd import _capi, _device
    w, d = 3, 4
    tb = rc.TensorBasis(w, d)
    n = tb.size()
    a, b, c = (rng.normal(size=(5, n)) for _ in range(3))
    ta, tbb, tc = (gpu.from_numpy(v).cuda() for v in (a, b, c))
    for ob, lb, rb in [((0, -1), (1, 3), (0, 2)), ((2, 4), (0, -1), (1, -1)),
                       ((0, 0), (0, 0), (0, 0)), ((1, 3), (1, 1), (1, 2))]:
        out = tc.clone()
        _capi.check(_capi.lib().rpy_dense_ft_fma(
            out.data_ptr(), ta.data_ptr(), tbb.data_ptr(), _capi.RPY_F64, tb.c_struct(),
            _capi.DegreeBoundsC(*ob), _capi.DegreeBoundsC(*lb), _capi.DegreeBoundsC(*rb), 5, n,
            n, n, 1, 1, 1, _device.stream_handle()))
        ref = c.copy()
        O.ft_fma(ref, a, b, w, d, out_bounds=ob, lhs_bounds=lb, rhs_bounds=rb)
        _close(out.cpu().numpy(), ref, w, d, np.float64)
    # in-place on device tensors, no copies back
    t = gpu.from_numpy(a.copy()).cuda()
    rc.ft_inplace_mul(rc.FreeTensor(t, tb), rc.FreeTensor(tbb, tb))
    ref = a.copy()
    O.ft_inplace_mul(ref, b, w, d)
    _close(t.cpu().numpy(), ref, w, d, np.float64)


def test_lie_maps_on_gpu(gpu, rng):
    lb = rc.LieBasis(4, 4)
    tb = rc.TensorBasis(4, 4)
    x = rng.uniform(-1, 1, size=(6, lb.size()))
    t = rc.lie_to_tensor(rc.Lie(x, lb), tb)
    assert_allclose(t.data, x @ O.l2t_dense(4, 4).T, atol=1e-15)
    y = rc.tensor_to_lie(t, lb)
    assert_allclose(y.data, x, atol=1e-15)  # test_l2t_and_t2l.py:8-56
    y2 = rc.tensor_to_lie(t, lb, scale_factor=0.5)
    assert_allclose(y2.data, 0.5 * x, atol=1e-15)
    t2 = rc.lie_to_tensor(rc.Lie(x, lb), tb, scale_factor=0.5)
    assert_allclose(rc.tensor_to_lie(t2, lb).data, 0.5 * x, atol=1e-15)


@pytest.mark.parametrize("w,d,dtype", [(5, 5, np.float64), (5, 5, np.float32),
                                       (2, 9, np.float64), (8, 4, np.float32)])
def test_lie_maps_batched(gpu, rng, w, d, dtype):
    """The batched maps (tensors staged in LDS, several per workgroup, with a
    ragged last group: 37 tensors) against the dense oracle matrices."""
    lb = rc.LieBasis(w, d)
    tb = rc.TensorBasis(w, d)
    x = rng.uniform(-1, 1, size=(37, lb.size())).astype(dtype)
    t = rc.lie_to_tensor(rc.Lie(x, lb), tb)
    ref_t = x.astype(np.float64) @ O.l2t_dense(w, d).T
    tol = 1e-13 if dtype == np.float64 else 2e-6
    assert np.abs(np.asarray(t.data, np.float64) - ref_t).max() <= tol * max(1, np.abs(ref_t).max())
    g = rng.uniform(-1, 1, size=(37, tb.size())).astype(dtype)
    y = rc.tensor_to_lie(rc.FreeTensor(g, tb), lb, scale_factor=0.5)
    ref_y = 0.5 * (g.astype(np.float64) @ O.t2l_dense(w, d).T)
    assert np.abs(np.asarray(y.data, np.float64) - ref_y).max() <= tol * max(1, np.abs(ref_y).max()) * 10


def test_errors(gpu):
    tb = rc.TensorBasis(2, 2)
    a = rc.FreeTensor(np.zeros(7), tb)
    with pytest.raises(TypeError):
        rc.ft_fma(a, rc.FreeTensor(np.zeros(7, np.float32), tb), a)
    with pytest.raises(ValueError):
        rc.ft_fma(a, rc.FreeTensor(np.zeros((2, 7)), tb), a)
    with pytest.raises(ValueError):
        rc.ft_fma(a, rc.FreeTensor(np.zeros(7), rc.TensorBasis(3, 2)), a)
    with pytest.raises(ValueError):
        rc.ft_exp(rc.FreeTensor(np.zeros(7, np.int64), tb))
    empty = rc.FreeTensor(np.zeros((0, 7)), tb)
    assert rc.ft_mul(empty, empty).data.shape == (0, 7)


def test_antipode_and_group_inverse(gpu, rng):
    w, d = 4, 3
    tb = rc.TensorBasis(w, d)
    x = np.zeros((4, tb.size()))
    x[:, 1:1 + w] = rng.normal(size=(4, w))
    sig = rc.ft_exp(rc.FreeTensor(x, tb))
    prod = rc.ft_mul(rc.antipode(sig), sig)
    ident = np.zeros_like(x)
    ident[:, 0] = 1
    assert_allclose(prod.data, ident, atol=1e-14)  # tests/compute/test_antipode.py:10-37


@pytest.mark.parametrize("w,d", [(2, 5), (4, 6), (5, 4)])
def test_ft_fma_compensated_is_rounded_once(gpu, rng, w, d):
    """Compensated float32 FMA: each coefficient is the float64 result on the
    same float32 inputs rounded once, so it is within one float32 ulp of the
    float64 oracle (and never worse than plain float32)."""
    batch = (17,)
    tb = rc.TensorBasis(w, d)
    A = _rand(rng, w, d, batch, np.float32, unit=1.0)
    B = _rand(rng, w, d, batch, np.float32, unit=1.0)
    C = _rand(rng, w, d, batch, np.float32, unit=1.0)
    ref = C.astype(np.float64)
    O.ft_fma(ref, A.astype(np.float64), B.astype(np.float64), w, d)
    out = rc.FreeTensor(C.copy(), tb)
    rc.ft_fma(out, rc.FreeTensor(A, tb), rc.FreeTensor(B, tb), compensated=True)
    got = out.data.astype(np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= ulp + 1e-30), np.abs(got - ref).max()
    plain = rc.FreeTensor(C.copy(), tb)
    rc.ft_fma(plain, rc.FreeTensor(A, tb), rc.FreeTensor(B, tb))
    assert np.abs(got - ref).max() <= np.abs(plain.data.astype(np.float64) - ref).max()
