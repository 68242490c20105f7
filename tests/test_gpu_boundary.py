"""GPU parity of the drop-in boundary's full argument surface, through the
C-ABI: per-operand degree bounds of exp / fmexp / log / antipode / adjoint
/ fma / in-place product, element strides other than 1 on every dense entry
point, the reference's bounds as roughpy.compute passes them, and the
LieIncrementStream.from_increments keyword forms.

Bounds are checked against the oracle: the layout of depth m is the first
db[m + 1] coefficients of any deeper layout, so a result truncated at
out max_degree m is the oracle at depth m on arguments whose coefficients
outside their bounds are zero, and coefficients outside an output's bounds
keep their values.  Strided calls must equal contiguous calls bit for bit
(they run the same kernels on gathered copies).  Tolerances: float64 1e-12
normwise per level, float32 2e-5.
"""
import ctypes

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

import roughpy_amd.compute as rc
from oracle import oracle as O
from roughpy_amd import _capi, _device
from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu

B = _capi.DegreeBoundsC


def _rand(rng, w, d, n, dtype, unit=None):
    db = O.tensor_degree_begin(w, d)
    x = rng.uniform(-1, 1, size=(n, int(db[-1])))
    f = 1.0
    for k in range(1, d + 1):
        f *= k
        x[:, db[k]:db[k + 1]] /= f
    if unit is not None:
        x[:, 0] = unit
    return x.astype(dtype)


def _code(dt):
    return _capi.RPY_F64 if np.dtype(dt) == np.float64 else _capi.RPY_F32


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _zero_outside(x, w, D, lo, hi):
    """x with the coefficients of degrees outside [lo, hi] set to zero."""
    db = O.tensor_degree_begin(w, D)
    y = x.copy()
    y[:, :db[lo]] = 0
    y[:, db[hi + 1]:] = 0
    return y


TOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 2e-5}


def _close(x, ref, w, d, dt):
    err = O.level_errors(np.asarray(x, np.float64), np.asarray(ref, np.float64), w, d).max()
    assert err <= TOL[np.dtype(dt)], err


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("w,D", [(3, 4), (5, 5), (4, 6)])
def test_exp_fmexp_log_degree_bounds(gpu, rng, dt, w, D):
    torch = gpu
    L = _capi.lib()
    n = 6
    tb = rc.TensorBasis(w, D)
    size = tb.size()
    db = O.tensor_degree_begin(w, D)
    x = _rand(rng, w, D, n, dt, unit=0.0)
    a = _rand(rng, w, D, n, dt, unit=1.0)
    # device copies stay bound for the whole test: a temporary's block could
    # be handed to the next allocation before the launch reads it
    xd, ad = _dev(torch, x), _dev(torch, a)
    for m, (alo, ahi) in [(D, (0, -1)), (D - 1, (0, -1)), (D - 2, (0, D)), (D, (1, D - 2)),
                          (D - 1, (2, D)), (0, (0, -1)), (1, (0, 0))]:
        mdb = O.tensor_degree_begin(w, m)
        ahi_r = D if ahi < 0 else ahi
        xt = _zero_outside(x, w, D, max(alo, 0), ahi_r)[:, :int(mdb[-1])]
        at = _zero_outside(a, w, D, max(alo, 0), ahi_r)[:, :int(mdb[-1])]
        sentinel = np.full((n, size), 7.0, dtype=dt)
        for name in ("exp", "log", "fmexp"):
            out = _dev(torch, sentinel)
            if name == "exp":
                _capi.check(L.rpy_dense_ft_exp(out.data_ptr(), xd.data_ptr(),
                                               _code(dt), tb.c_struct(), B(0, m), B(alo, ahi), n,
                                               size, size, 1, 1, _device.stream_handle()))
                ref = O.ft_exp(xt, w, m)
            elif name == "log":
                _capi.check(L.rpy_dense_ft_log(out.data_ptr(), ad.data_ptr(),
                                               _code(dt), tb.c_struct(), B(0, m), B(alo, ahi), n,
                                               size, size, 1, 1, _device.stream_handle()))
                at1 = at.copy()
                at1[:, 0] = 1.0
                ref = O.ft_log(at1, w, m)
            else:
                _capi.check(L.rpy_dense_ft_fmexp(out.data_ptr(), ad.data_ptr(),
                                                 xd.data_ptr(), _code(dt),
                                                 tb.c_struct(), B(0, m), B(alo, ahi), B(alo, ahi),
                                                 n, size, size, size, 1, 1, 1,
                                                 _device.stream_handle()))
                ref = O.ft_fmexp(at, xt, w, m)
            got = out.cpu().numpy()
            _close(got[:, :int(mdb[-1])], ref, w, m, dt)
            assert_array_equal(got[:, int(mdb[-1]):], sentinel[:, int(mdb[-1]):])
    # the output of exp / fmexp / log starts at degree 0
    out = torch.zeros((n, size), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _capi.check(L.rpy_dense_ft_exp(out.data_ptr(), out.data_ptr(), _capi.RPY_F64,
                                       tb.c_struct(), B(1, -1), B(0, -1), n, size, size, 1, 1,
                                       _device.stream_handle()))
    assert db[-1] == size


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_antipode_degree_bounds(gpu, rng, dt):
    """Levels max(mins)..min(maxes) are written, and level 1 whenever that
    range starts at or below 1 (free_tensor_antipode.hpp:60-80)."""
    torch = gpu
    L = _capi.lib()
    w, D, n = 4, 5, 5
    tb = rc.TensorBasis(w, D)
    size = tb.size()
    db = O.tensor_degree_begin(w, D)
    x = _rand(rng, w, D, n, dt)
    xd = _dev(torch, x)
    full = O.antipode(x, w, D)
    for ob, ab in [((0, -1), (0, -1)), ((2, 4), (0, -1)), ((0, -1), (0, 0)), ((3, 3), (1, 5)),
                   ((1, 2), (0, 5)), ((4, 5), (0, 3))]:
        out = _dev(torch, np.full((n, size), -3.0, dtype=dt))
        _capi.check(L.rpy_dense_ft_antipode(out.data_ptr(), xd.data_ptr(),
                                            _code(dt), tb.c_struct(), B(*ob), B(*ab), n, size,
                                            size, 1, 1, _device.stream_handle()))
        got = out.cpu().numpy()
        lo = max(ob[0], ab[0])
        hi = min(D if ob[1] < 0 else ob[1], D if ab[1] < 0 else ab[1])
        if lo <= 1:
            hi = max(hi, 1)
        for k in range(D + 1):
            seg = slice(int(db[k]), int(db[k + 1]))
            if lo <= k <= hi:
                assert_array_equal(got[:, seg], full[:, seg])
            else:
                assert_array_equal(got[:, seg], -3.0)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adjoint_left_mul_degree_bounds(gpu, rng, dt):
    """The adjoint's degree ranges follow free_tensor_adjoint_left_mul.hpp:
    39-51 (oracle orc_adj_lmul with the same bounds)."""
    torch = gpu
    L = _capi.lib()
    OL = O.lib()
    sfx = "f64" if np.dtype(dt) == np.float64 else "f32"
    w, D, n = 3, 5, 4
    tb = rc.TensorBasis(w, D)
    size = tb.size()
    op = _rand(rng, w, D, n, dt)
    arg = _rand(rng, w, D, n, dt)
    opd, argd = _dev(torch, op), _dev(torch, arg)
    for ob, pb, ab in [((0, 5), (0, 5), (0, 5)), ((0, 2), (0, 5), (0, 5)),
                       ((1, 4), (1, 3), (2, 5)), ((0, 5), (2, 2), (0, 5)),
                       ((3, 5), (0, 1), (0, 4))]:
        out0 = _rand(rng, w, D, n, dt)
        out = _dev(torch, out0)
        _capi.check(L.rpy_dense_ft_adjoint_left_mul(
            out.data_ptr(), opd.data_ptr(), argd.data_ptr(), _code(dt),
            tb.c_struct(), B(*ob), B(*pb), B(*ab), n, size, size, size, 1, 1, 1,
            _device.stream_handle()))
        ref = out0.copy()
        for e in range(n):
            getattr(OL, f"orc_adj_lmul_{sfx}")(O._p(ref[e]), O._p(op[e]), O._p(arg[e]), w,
                                               *ob, *pb, *ab)
        _close(out.cpu().numpy(), ref, w, D, dt)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_fma_and_inplace_degree_bounds(gpu, rng, dt):
    torch = gpu
    L = _capi.lib()
    w, D, n = 4, 5, 4
    tb = rc.TensorBasis(w, D)
    size = tb.size()
    a = _rand(rng, w, D, n, dt, unit=1.0)
    b = _rand(rng, w, D, n, dt, unit=1.0)
    adv, bdv = _dev(torch, a), _dev(torch, b)
    for ob, lb, rb in [((0, 3), (0, -1), (0, -1)), ((1, 5), (1, 4), (0, 2)),
                       ((0, 9), (0, 9), (0, 9))]:
        c0 = _rand(rng, w, D, n, dt)
        c = _dev(torch, c0)
        _capi.check(L.rpy_dense_ft_fma(c.data_ptr(), adv.data_ptr(),
                                       bdv.data_ptr(), _code(dt), tb.c_struct(),
                                       B(*ob), B(*lb), B(*rb), n, size, size, size, 1, 1, 1,
                                       _device.stream_handle()))
        clamp = lambda bd: (bd[0], min(D, bd[1]) if bd[1] >= 0 else -1)  # noqa: E731
        ref = c0.copy()
        O.ft_fma(ref, a, b, w, D, clamp(ob), clamp(lb), clamp(rb))
        _close(c.cpu().numpy(), ref, w, D, dt)
    for lb, rb in [((0, 3), (0, -1)), ((0, -1), (0, 2))]:
        x = _dev(torch, a)
        _capi.check(L.rpy_dense_ft_inplace_mul(x.data_ptr(), bdv.data_ptr(),
                                               _code(dt), tb.c_struct(), B(*lb), B(*rb), n,
                                               size, size, 1, 1, _device.stream_handle()))
        ref = a.copy()
        O.ft_inplace_mul(ref, b, w, D, lb, rb)
        _close(x.cpu().numpy(), ref, w, D, dt)


def _strided_view(torch, host, inc, fortran=False):
    """A device buffer holding `host` (n, size) at element stride `inc`
    (negative: reversed), or in column-major order; returns (buffer,
    data pointer of element 0 coefficient 0, batch stride, element stride)."""
    n, size = host.shape
    es = host.itemsize
    if fortran:
        buf = _dev(torch, np.asfortranarray(host).ravel(order="K"))
        return buf, buf.data_ptr(), 1, n
    span = size * abs(inc)
    raw = np.full(n * span, 11.0, dtype=host.dtype)
    for e in range(n):
        if inc > 0:
            raw[e * span:e * span + size * inc:inc] = host[e]
        else:
            start = e * span + (size - 1) * abs(inc)
            raw[start - np.arange(size) * abs(inc)] = host[e]
    buf = _dev(torch, raw)
    if inc > 0:
        return buf, buf.data_ptr(), span, inc
    return buf, buf.data_ptr() + (size - 1) * abs(inc) * es, span, inc


def _unstride(buf, n, size, stride, inc, fortran=False):
    raw = buf.cpu().numpy()
    if fortran:
        return raw.reshape(size, n).T.copy()
    out = np.empty((n, size), dtype=raw.dtype)
    for e in range(n):
        if inc > 0:
            out[e] = raw[e * stride:e * stride + size * inc:inc]
        else:
            start = e * stride + (size - 1) * abs(inc)
            out[e] = raw[start - np.arange(size) * abs(inc)]
    return out


@pytest.mark.parametrize("layout", ["inc2", "inc-3", "fortran"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_element_strides_equal_contiguous_calls(gpu, rng, layout, dt):
    """Every dense entry point on operands with element stride 2, -3 or in
    column-major order gives bit for bit the contiguous call's result; the
    slots between a strided output's coefficients are untouched."""
    torch = gpu
    L = _capi.lib()
    w, D, n = 3, 4, 5
    tb = rc.TensorBasis(w, D)
    size = tb.size()
    cs = tb.c_struct()
    st = _device.stream_handle()
    code = _code(dt)
    fortran = layout == "fortran"
    inc = {"inc2": 2, "inc-3": -3, "fortran": None}[layout]
    a = _rand(rng, w, D, n, dt, unit=1.0)
    b = _rand(rng, w, D, n, dt, unit=1.0)
    c = _rand(rng, w, D, n, dt)
    x = _rand(rng, w, D, n, dt, unit=0.0)
    full = B(0, -1)

    def run(op, outs, ins, strided):
        """op(ptrs, strides, incs) on device copies; returns the outputs."""
        bufs, ptrs, strides, incs = [], [], [], []
        for arr in outs + ins:
            if strided:
                buf, p, s_, i_ = _strided_view(torch, arr, inc if not fortran else 1,
                                               fortran=fortran)
            else:
                buf = _dev(torch, arr)
                p, s_, i_ = buf.data_ptr(), size, 1
            bufs.append(buf)
            ptrs.append(p)
            strides.append(s_)
            incs.append(i_)
        _capi.check(op(ptrs, strides, incs))
        torch.cuda.synchronize()
        res = []
        for k in range(len(outs)):
            if strided:
                res.append(_unstride(bufs[k], n, size, strides[k], incs[k], fortran))
                if not fortran:
                    raw = bufs[k].cpu().numpy()
                    mask = np.ones(raw.shape, bool)
                    for e in range(n):
                        for i in range(size):
                            pos = (e * strides[k] + i * incs[k] if incs[k] > 0 else
                                   e * strides[k] + (size - 1 - i) * abs(incs[k]))
                            mask[pos] = False
                    assert np.all(raw[mask] == 11.0)
            else:
                res.append(bufs[k].cpu().numpy())
        return res

    ops = {
        "fma": (lambda p, s, i: L.rpy_dense_ft_fma(p[0], p[1], p[2], code, cs, full, full, full,
                                                   n, s[0], s[1], s[2], i[0], i[1], i[2], st),
                [c], [a, b]),
        "inplace_mul": (lambda p, s, i: L.rpy_dense_ft_inplace_mul(
            p[0], p[1], code, cs, full, full, n, s[0], s[1], i[0], i[1], st), [a], [b]),
        "antipode": (lambda p, s, i: L.rpy_dense_ft_antipode(
            p[0], p[1], code, cs, full, full, n, s[0], s[1], i[0], i[1], st), [c], [a]),
        "adjoint_left_mul": (lambda p, s, i: L.rpy_dense_ft_adjoint_left_mul(
            p[0], p[1], p[2], code, cs, full, full, full, n, s[0], s[1], s[2], i[0], i[1], i[2],
            st), [c], [a, b]),
        "adjoint_right_mul": (lambda p, s, i: L.rpy_dense_ft_adjoint_right_mul(
            p[0], p[1], p[2], code, cs, n, s[0], s[1], s[2], i[0], i[1], i[2], st), [c], [a, b]),
        "st_fma": (lambda p, s, i: L.rpy_dense_st_fma(p[0], p[1], p[2], code, cs, n, s[0], s[1],
                                                      s[2], i[0], i[1], i[2], st), [c], [a, b]),
        "st_adj_mul": (lambda p, s, i: L.rpy_dense_st_adj_mul(
            p[0], p[1], p[2], code, cs, n, s[0], s[1], s[2], i[0], i[1], i[2], st), [c], [a, b]),
        "exp": (lambda p, s, i: L.rpy_dense_ft_exp(p[0], p[1], code, cs, full, full, n, s[0],
                                                   s[1], i[0], i[1], st), [c], [x]),
        "fmexp": (lambda p, s, i: L.rpy_dense_ft_fmexp(p[0], p[1], p[2], code, cs, full, full,
                                                       full, n, s[0], s[1], s[2], i[0], i[1],
                                                       i[2], st), [c], [a, x]),
        "log": (lambda p, s, i: L.rpy_dense_ft_log(p[0], p[1], code, cs, full, full, n, s[0],
                                                   s[1], i[0], i[1], st), [c], [a]),
    }
    for name, (op, outs, ins) in ops.items():
        want = run(op, outs, ins, False)
        got = run(op, outs, ins, True)
        for g, wv in zip(got, want):
            assert_array_equal(g, wv, err_msg=name)
    # the sparse map (tensor_to_lie) with strided operands
    lb = rc.LieBasis(w, D)
    t2l = lb.get_t2l_matrix(np.dtype(dt))
    lsize = lb.size()
    ys = {}
    for strided in (False, True):
        if strided:
            xb, xp, xs, xi = _strided_view(torch, a, inc if not fortran else 1, fortran=fortran)
        else:
            xb = _dev(torch, a)
            xp, xs, xi = xb.data_ptr(), size, 1
        y = torch.zeros((n, lsize), dtype=xb.dtype, device="cuda")
        _capi.check(L.rpy_dense_sparse_map(y.data_ptr(), xp, t2l.device_struct(), code, 1.0, n,
                                           lsize, xs, 1, xi, st))
        ys[strided] = y.cpu().numpy()
    assert_array_equal(ys[True], ys[False])


def test_compute_mirrors_reference_depth_arguments(gpu, rng):
    """roughpy.compute's own calls: ft_fma in b's basis truncated at c's
    depth (__init__.py:145), ft_mul with rhs_depth = b's depth (:161),
    ft_inplace_mul with out_depth = b's depth (:181), ft_adjoint_left_mul in
    op's basis with (arg, op, arg) depths (:333-334)."""
    w = 3
    t4, t2 = rc.TensorBasis(w, 4), rc.TensorBasis(w, 2)
    a4 = _rand(rng, w, 4, 3, np.float64, unit=1.0)
    b4 = _rand(rng, w, 4, 3, np.float64, unit=1.0)
    c2 = _rand(rng, w, 2, 3, np.float64, unit=1.0)
    out = _rand(rng, w, 4, 3, np.float64)
    acc = out.copy()
    rc.ft_fma(rc.FreeTensor(acc, t4), rc.FreeTensor(a4, t4), rc.FreeTensor(c2, t2))
    c2p = np.zeros_like(a4)
    c2p[:, :c2.shape[1]] = c2
    ref = out.copy()
    O.ft_fma(ref, a4, c2p, w, 4, out_bounds=(0, 2))
    assert_allclose(acc, ref, rtol=0, atol=1e-13)
    prod = rc.ft_mul(rc.FreeTensor(a4, t4), rc.FreeTensor(c2, t2)).data
    ref = np.zeros_like(a4)
    O.ft_fma(ref, a4, c2p, w, 4, rhs_bounds=(0, 2))
    assert_allclose(prod, ref, rtol=0, atol=1e-13)
    ip = a4.copy()
    rc.ft_inplace_mul(rc.FreeTensor(ip, t4), rc.FreeTensor(c2, t2))
    ref = a4.copy()
    O.ft_inplace_mul(ref, c2p, w, 4, lhs_bounds=(0, 2))
    assert_allclose(ip, ref, rtol=0, atol=1e-13)
    # op deeper than arg: the result has arg's shape and sums over op's basis
    arg2 = _rand(rng, w, 2, 3, np.float64)
    adj = rc.ft_adjoint_left_mul(rc.FreeTensor(a4, t4), rc.ShuffleTensor(arg2, t2)).data
    assert adj.shape == arg2.shape
    argp = np.zeros_like(a4)
    argp[:, :arg2.shape[1]] = arg2
    ref = np.zeros_like(a4)
    for e in range(3):
        O.lib().orc_adj_lmul_f64(O._p(ref[e]), O._p(a4[e]), O._p(argp[e]), w, 0, 2, 0, 4, 0, 2)
    assert_allclose(adj, ref[:, :arg2.shape[1]], rtol=0, atol=1e-13)


def test_from_increments_forms_on_gpu(gpu, rng):
    """Signatures of streams built with indices=<int>, a negative inferred
    resolution, include_time and support=, against the oracle's stream
    semantics on the same parameters and rows."""
    w, d = 3, 3
    data = np.abs(rng.normal(size=(20, w))) + 0.05
    # cumulative index column
    s = rs.LieIncrementStream.from_increments(data, indices=0, depth=d)
    idx = np.cumsum(data[:, 0])
    o = O.IncrementStream(data, w, d, indices=idx)
    for a_, b_ in [(idx[0], idx[-1] + 1.0), (idx[3], idx[12]), (0.0, 1e3)]:
        assert O.level_errors(s.signature(a_, b_).data, o.signature(a_, b_), w, d).max() <= 1e-12
    # indices 2 or more apart: negative resolution, pieces longer than 1
    idx2 = 4.0 + 2.5 * np.arange(20)
    s = rs.LieIncrementStream.from_increments(data, indices=idx2, depth=d)
    o = O.IncrementStream(data, w, d, indices=idx2)
    assert s.resolution == o.resolution < 0
    for a_, b_ in [(4.0, 60.0), (9.0, 31.5), (0.0, 100.0), (14.0, 15.0)]:
        assert O.level_errors(s.signature(a_, b_).data, o.signature(a_, b_), w, d).max() <= 1e-12
    # include_time: a last channel of parameter increments
    idx3 = np.sort(rng.uniform(0, 1, size=20))
    s = rs.LieIncrementStream.from_increments(data, indices=idx3, depth=d, include_time=True)
    aug = np.hstack([data, (idx3 - np.concatenate([[0.0], idx3[:-1]]))[:, None]])
    o = O.IncrementStream(aug, w + 1, d, indices=idx3)
    got = s.signature().data
    assert O.level_errors(got, o.signature(*o.support), w + 1, d).max() <= 1e-12
    # support= replaces the support
    s = rs.LieIncrementStream.from_increments(data, depth=d, support=rs.RealInterval(2.0, 7.0))
    ref = O.signature(data[2:7], w, d)
    assert O.level_errors(s.signature().data, ref, w, d).max() <= 1e-12


def test_float32_log_signature_single_stream_odd_size(gpu, rng):
    """One float32 stream at width 4, depth 6: n_streams x tensor_size is odd,
    so the float64 workspace segments must not follow the float32 one
    unaligned."""
    w, d = 4, 6
    inc = (rng.normal(size=(30, w)) / 6).astype(np.float32)
    got = rs.log_signature_batch(inc, np.array([0, 30]), w, d)
    ref = O.log_signature(inc.astype(np.float64), w, d)
    assert_allclose(got[0], ref, rtol=0, atol=2e-5 * max(1.0, np.abs(ref).max()))


def test_batch_and_cache_arguments_are_checked(gpu, rng):
    torch = gpu
    w, d = 3, 3
    inc = rng.normal(size=(10, w)) / 4
    with pytest.raises(ValueError, match="offsets"):
        rs.signature_batch(inc, np.array([0, 4, 3, 10]), w, d)
    with pytest.raises(ValueError, match="offsets"):
        rs.signature_batch(inc, np.array([0, 5, 11]), w, d)
    with pytest.raises(ValueError, match="columns"):
        rs.signature_batch(rng.normal(size=(10, w + 1)), np.array([0, 10]), w, d)
    with pytest.raises(ValueError, match="offsets"):
        rs.log_signature_batch(inc, np.array([-1, 10]), w, d)
    params = np.concatenate([np.arange(5) / 5.0, np.arange(5) / 5.0])
    cache = rs.DyadicSignatureCache(inc, params, np.array([0, 5, 10]), w, d, resolution=4,
                                    levels=2)
    with pytest.raises(IndexError):
        cache.query(np.array([2]), np.array([0.0]), np.array([0.5]))
    with pytest.raises(IndexError):
        cache.node(1, 2, 0)
    ok = cache.query(np.array([1]), np.array([0.0]), np.array([0.5]))
    for bad in (torch.empty((1, ok.shape[1] - 1), dtype=ok.dtype, device="cuda"),
                torch.empty((1, ok.shape[1]), dtype=torch.float32, device="cuda"),
                torch.empty((1, 2 * ok.shape[1]), dtype=ok.dtype, device="cuda")[:, ::2]):
        with pytest.raises((ValueError, TypeError)):
            cache.query(np.array([1]), np.array([0.0]), np.array([0.5]), out=bad)
    out = torch.empty_like(ok)
    cache.query(np.array([1]), np.array([0.0]), np.array([0.5]), out=out)
    assert torch.equal(out, ok)
