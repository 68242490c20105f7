"""The C-ABI library on the CPU: it loads without a GPU, exports every symbol
include/roughpy_amd.h declares, and its host-side logic (Hall basis, l2t /
t2l maps, dyadic dissection, stream row selection, and a host run of the
signature kernel's lane decomposition) agrees with the oracle.  No compute
entry point is called here."""
import ctypes
import json
import os
import re

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

import roughpy_amd as ra
from oracle import oracle as O
from roughpy_amd import _capi
from roughpy_amd.streams import (LieIncrementStream, RealInterval, param_to_resolution,
                                 to_dyadic_intervals)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "roughpy_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rpy_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
        assert name in _capi.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_sizes():
    L = _capi.lib()
    assert L.rpy_version() >= 1
    for w, d in [(3, 4), (4, 6), (5, 5), (8, 4), (6, 5)]:
        assert L.rpy_tensor_size(w, d) == O.tensor_size(w, d)
        assert L.rpy_lie_basis_size(w, d) == sum(O.witt_dim(w, k) for k in range(1, d + 1))
    assert L.rpy_tensor_size(0, 3) == -1


@pytest.mark.parametrize("w,d", [(2, 3), (3, 4), (5, 5), (8, 4), (6, 5), (4, 6), (2, 8)])
def test_hall_basis_matches_oracle(w, d):
    lb = ra.LieBasis(w, d)
    db, data = O.hall_basis(w, d)
    assert_array_equal(lb.degree_begin, db)
    assert_array_equal(lb.data, data)


def test_hall_golden_order():
    lb = ra.LieBasis(2, 3)
    assert [lb.key_to_str(k) for k in lb] == GOLD["hall_order_w2_d3"]["keys"]


@pytest.mark.parametrize("w,d", [(2, 3), (3, 4), (4, 4), (2, 6)])
def test_l2t_t2l_match_oracle(w, d):
    lb = ra.LieBasis(w, d)
    assert_array_equal(lb.get_l2t_matrix().to_dense(), O.l2t_dense(w, d))
    assert_allclose(lb.get_t2l_matrix().to_dense(), O.t2l_dense(w, d), rtol=0, atol=1e-16)


def test_l2t_golden_and_caching():
    lb = ra.LieBasis(2, 3)
    assert_array_equal(lb.get_l2t_matrix().to_dense(), GOLD["l2t_w2_d3"]["matrix"])
    # tests/compute/test_l2t_and_t2l.py:59-75: caches are per map and per dtype
    assert lb.get_l2t_matrix(np.float64) is lb.get_l2t_matrix(np.float64)
    assert lb.get_t2l_matrix(np.float64) is lb.get_t2l_matrix(np.float64)
    assert lb.get_t2l_matrix(np.float64) is not lb.get_l2t_matrix(np.float64)
    assert lb.get_l2t_matrix(np.float32).data.dtype == np.float32


def test_t2l_l2t_roundtrip_large():
    lb = ra.LieBasis(5, 5)
    l2t = lb.get_l2t_matrix().to_dense()
    t2l = lb.get_t2l_matrix().to_dense()
    assert_allclose(t2l @ l2t, np.eye(lb.size()), atol=1e-14)


def test_check_lie_basis_data():
    lb = ra.LieBasis(3, 4)
    assert ra.check_lie_basis_data(lb.data, lb.degree_begin, 3, 4)
    bad = np.array(lb.data)
    bad[5] = (bad[5, 1], bad[5, 0])          # [k, l] with k > l
    assert not ra.check_lie_basis_data(bad, lb.degree_begin, 3, 4)
    with pytest.raises(ValueError):
        ra.check_lie_basis_data(bad, lb.degree_begin, 3, 4, raise_on_fail=True)
    with pytest.raises(ValueError):
        ra.LieBasis(3, 4, degree_begin=lb.degree_begin, data=bad)
    same = ra.LieBasis(3, 4, degree_begin=lb.degree_begin, data=lb.data)
    assert same == lb and hash(same) == hash(lb)


def test_tensor_basis_mirror():
    tb = ra.TensorBasis(3, 4)
    assert list(tb.degree_begin) == [0, 1, 4, 13, 40, 121]
    assert tb.size() == 121
    assert tb.truncate(2) == ra.TensorBasis(3, 2)
    assert tb.truncate(0) == ra.TensorBasis(3, 0)
    # tensor_basis.c:177-191: new_depth must be in [0, depth)
    for bad in (4, 5, -1):
        with pytest.raises(ValueError):
            tb.truncate(bad)
    assert repr(tb) == "TensorBasis(3, 4)"
    assert tb != ra.TensorBasis(4, 4)
    assert len({ra.TensorBasis(2, 2), ra.TensorBasis(2, 2), ra.TensorBasis(2, 3)}) == 2
    with pytest.raises(ValueError):
        ra.TensorBasis(3, 4, degree_begin=[0, 1, 4, 13, 40, 120])


def test_tensor_basis_hash_is_fnv1a():
    """hash(TensorBasis) is the 64-bit FNV-1a of tensor_basis.c:144-171: the
    FNV function against published test vectors, then the hash of the
    reference's byte string (b"TensorBasis", int32 width, int32 depth,
    depth + 2 pointer-sized degree_begin entries)."""
    from roughpy_amd.bases import fnv1a64
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8
    for w, d in [(2, 2), (5, 5), (4, 6)]:
        tb = ra.TensorBasis(w, d)
        raw = (b"TensorBasis" + np.int32(w).tobytes() + np.int32(d).tobytes()
               + np.array(tb.degree_begin, dtype=np.intp).tobytes())
        v = fnv1a64(raw)
        assert hash(tb) == (v - (1 << 64) if v >= (1 << 63) else v)
    assert hash(ra.TensorBasis(5, 5)) != hash(ra.TensorBasis(5, 4))


@pytest.mark.parametrize("case", GOLD["to_dyadic_intervals"]["cases"])
def test_dyadic_goldens(case):
    assert to_dyadic_intervals(case["inf"], case["sup"], case["tol"]) == [
        tuple(p) for p in case["pieces"]]


def test_dyadic_matches_oracle_random(rng):
    for _ in range(300):
        a, b = sorted(rng.uniform(-3, 3, size=2))
        tol = int(rng.integers(0, 14))
        assert to_dyadic_intervals(a, b, tol) == O.to_dyadic_intervals(a, b, tol)


def test_param_to_resolution():
    for v in [0.1, 0.25, 0.5, 0.7, 1.0, 3.0, 1e-6, 0.0]:
        assert param_to_resolution(v) == O.param_to_resolution(v)


def _rows_oracle(s: O.IncrementStream, inf, sup, res):
    """Rows the oracle's recursive dyadic-caching semantics include."""
    picked = []
    q = s.query_resolution(inf, sup)
    if q is None:
        return picked
    a, b, r = q
    if res is not None:
        r = res
    params = s.params

    def piece_rows(k, n):
        lo, hi = k / 2.0 ** n, (k + 1) / 2.0 ** n
        inside = [i for i, p in enumerate(params) if lo <= p < hi]
        if not inside:
            return []
        if n == s.resolution:
            return inside
        if n > s.resolution:
            lk = k >> (n - s.resolution)
            if lk / 2.0 ** s.resolution == lo:
                return piece_rows(lk, s.resolution)
            return []
        return piece_rows(2 * k, n + 1) + piece_rows(2 * k + 1, n + 1)

    for k, n in O.to_dyadic_intervals(a, b, r):
        picked += piece_rows(k, n)
    return picked


def test_stream_row_selection_matches_oracle(rng):
    for trial in range(40):
        n = int(rng.integers(1, 30))
        idx = np.sort(rng.uniform(0, 2, size=n))
        res = int(rng.integers(0, 6)) if trial % 2 else None
        data = rng.normal(size=(n, 2))
        s_o = O.IncrementStream(data, 2, 2, indices=idx, resolution=res)
        s_p = LieIncrementStream.from_increments(data, indices=idx, width=2, depth=2,
                                                 resolution=res)
        assert s_p.resolution == s_o.resolution
        assert s_p.support == pytest.approx(s_o.support)
        for _ in range(10):
            a, b = sorted(rng.uniform(-0.5, 2.5, size=2))
            qres = int(rng.integers(0, 9)) if rng.uniform() < 0.5 else None
            got = list(s_p._select((a, b), qres))
            assert got == _rows_oracle(s_o, a, b, qres)


@pytest.mark.parametrize("w,d,p", [(2, 3, 1), (2, 3, 3), (3, 4, 1), (3, 4, 2), (5, 5, 2),
                                   (5, 5, 3), (5, 5, 4), (8, 4, 2), (8, 4, 3), (6, 5, 3),
                                   (6, 5, 4), (4, 6, 3), (4, 6, 4), (4, 6, 5)])
def test_kernel_lane_decomposition_on_host(rng, w, d, p):
    """The signature kernel's per-lane algorithm and output index map, run on
    the host, equals the oracle's Chen signature."""
    L = _capi.lib()
    inc = np.ascontiguousarray(rng.normal(size=(25, w)) / 5)
    out = np.full(O.tensor_size(w, d), np.nan)
    _capi.check(L.rpy_debug_signature_host_emulation(_capi.dblp(out), _capi.dblp(inc), 25, w,
                                                     d, p))
    assert not np.isnan(out).any(), "some coefficient was never written"
    assert O.level_errors(out, O.signature(inc, w, d), w, d).max() < 1e-13


def _tuned_rows():
    """Rows of the compiled tuned table, parsed from its source."""
    path = os.path.join(ROOT, "roughpy_amd", "csrc", "signature_tuned.inc")
    ops = {"RPY_OP_SIGNATURE": 0, "RPY_OP_SIGNATURE_VJP": 1}
    dts = {"RPY_F32": 0, "RPY_F64": 1, "RPY_F32_COMPENSATED": 2}
    pat = re.compile(r"^\s*\{(RPY_OP_\w+), (\d+), (\d+), (RPY_F\w+), "
                     r"\{(\d+), (\d+), (\d+), (\d+)\}\},")
    rows = []
    for line in open(path):
        m = pat.match(line)
        if m:
            rows.append((ops[m.group(1)], int(m.group(2)), int(m.group(3)), dts[m.group(4)],
                         tuple(int(m.group(i)) for i in range(5, 9))))
    return rows


def _config(L, op, w, d, dt):
    c = _capi.KernelVariantC()
    src = ctypes.c_int32()
    _capi.check(L.rpy_tune_config(op, w, d, dt, ctypes.byref(c), ctypes.byref(src)))
    return c.astuple(), src.value


def _variants(L, op, w, d, dt):
    n = L.rpy_tune_variants(op, w, d, dt, None, 0)
    arr = (_capi.KernelVariantC * n)()
    assert L.rpy_tune_variants(op, w, d, dt, arr, n) == n
    return [a.astuple() for a in arr]


def test_kernel_prefix_choice():
    """rpy_debug_signature_prefix reports the prefix of the variant that runs;
    without a tuned row that is the register-budget prefix."""
    L = _capi.lib()
    tuned = {(w, d, dt) for op, w, d, dt, _ in _tuned_rows() if op == 0}
    for w, d, f64, p in [(5, 5, 0, 2), (5, 5, 1, 3), (8, 4, 1, 2), (4, 6, 1, 4), (3, 4, 0, 1)]:
        got = L.rpy_debug_signature_prefix(w, d, f64)
        assert got == _config(L, 0, w, d, f64)[0][1]
        if (w, d, f64) not in tuned:
            assert got == p
    assert L.rpy_debug_signature_prefix(9, 3, 1) == 0      # generic kernel


def test_tuned_table_is_what_dispatch_runs():
    """Every row of the tuned table names a compiled variant and is the
    configuration a call of that shape runs (source 1 = tuned table)."""
    L = _capi.lib()
    for op, w, d, dt, v in _tuned_rows():
        assert v in _variants(L, op, w, d, dt), (op, w, d, dt, v)
        assert _config(L, op, w, d, dt) == (v, 1), (op, w, d, dt)


def test_kernel_variant_selection():
    """rpy_tune_select forces a compiled variant and NULL restores the tuned
    or default choice; variants that are not compiled are refused."""
    L = _capi.lib()
    base = _config(L, 0, 5, 5, 1)
    vs = _variants(L, 0, 5, 5, 1)
    assert vs[-1] == (0, 0, 0, 0)                    # the generic kernel runs every shape
    assert len(vs) >= 5 and len(set(vs)) == len(vs)
    for v in vs:
        _capi.check(L.rpy_tune_select(0, 5, 5, 1, ctypes.byref(_capi.KernelVariantC(*v))))
        assert _config(L, 0, 5, 5, 1) == (v, 2)
        assert L.rpy_debug_signature_prefix(5, 5, 1) == v[1]
    _capi.check(L.rpy_tune_select(0, 5, 5, 1, None))
    assert _config(L, 0, 5, 5, 1) == base
    bad = _capi.KernelVariantC(2, 3, 48, 0)        # no 48-row stream kernel is compiled
    assert L.rpy_tune_select(0, 5, 5, 1, ctypes.byref(bad)) == _capi.RPY_ERR_UNSUPPORTED
    assert L.rpy_tune_select(1, 5, 5, 1, ctypes.byref(_capi.KernelVariantC(3, 3, 0, 0))) == \
        _capi.RPY_ERR_UNSUPPORTED                    # no register backward at width 5
    assert L.rpy_tune_select(7, 5, 5, 1, None) == _capi.RPY_ERR_INVALID_ARGUMENT
    assert _variants(L, 0, 9, 3, 0) == [(0, 0, 0, 0)]
    vjp = _variants(L, 1, 8, 4, 0)
    assert (3, 2, 0, 0) in vjp and (3, 3, 0, 0) in vjp and vjp[-1] == (0, 0, 0, 0)
    assert _variants(L, 1, 5, 5, 2) == []            # the backward has no compensated mode


@pytest.mark.parametrize("w,d,p", [(8, 2, 2), (8, 3, 2), (8, 4, 2), (8, 4, 3), (4, 3, 3),
                                   (4, 5, 3), (4, 6, 3), (4, 6, 4), (2, 6, 6), (2, 7, 6),
                                   (2, 7, 7)])
def test_backward_lane_algebra_on_host(rng, w, d, p):
    """Host run of the register backward's lane algebra (W^P prefix lanes,
    path shares summed across lanes) against float64 central differences
    (step 1e-6) of the oracle signature: d<G, S>/dx for every increment."""
    L = _capi.lib()
    n = 7
    inc = np.ascontiguousarray(rng.normal(size=(n, w)) / 3)
    G = rng.normal(size=O.tensor_size(w, d))
    S = O.signature(inc, w, d)
    grad = np.full((n, w), np.nan)
    _capi.check(L.rpy_debug_vjp_host_emulation(_capi.dblp(grad), _capi.dblp(inc), n, w, d, p,
                                                _capi.dblp(S), _capi.dblp(G)))
    h = 1e-6
    for t in range(n):
        for a in range(w):
            ip, im = inc.copy(), inc.copy()
            ip[t, a] += h
            im[t, a] -= h
            fd = (np.dot(G, O.signature(ip, w, d)) - np.dot(G, O.signature(im, w, d))) / (2 * h)
            assert abs(grad[t, a] - fd) <= 1e-6 * max(1.0, abs(fd)), (t, a, grad[t, a], fd)


def test_compute_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from roughpy_amd import compute as rc
    tb = ra.TensorBasis(2, 2)
    a = rc.FreeTensor(np.zeros(7), tb)
    with pytest.raises(RuntimeError, match="GPU"):
        rc.ft_mul(a, a)
    with pytest.raises(RuntimeError, match="GPU"):
        LieIncrementStream.from_increments(np.ones((2, 2)), depth=2).signature()


def test_error_codes_map_to_exceptions():
    L = _capi.lib()
    tb = _capi.TensorBasisC(0, 2, None)
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 1, 1, 1, 1, 1, 1, 1, None)
    assert rc == _capi.RPY_ERR_BASIS
    with pytest.raises(ValueError, match="width"):
        _capi.check(rc)
    tb = _capi.TensorBasisC(2, 2, None)
    rc = L.rpy_dense_ft_fma(None, None, None, 7, ctypes.byref(tb), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1), 1, 7, 7, 7,
                            1, 1, 1, None)
    assert rc == _capi.RPY_ERR_DTYPE
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(2, 1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 1, 7, 7, 7, 1, 1, 1, None)
    assert rc == _capi.RPY_ERR_INVALID_ARGUMENT
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 4, 6, 7, 7, 1, 1, 1, None)
    assert rc == _capi.RPY_ERR_SHAPE
    # zero-size batch is a no-op that touches no pointer
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 0, 7, 7, 7, 1, 1, 1, None)
    assert rc == _capi.RPY_OK
    # a zero element stride is a shape error; a negative one is accepted
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 2, 7, 7, 7, 1, 0, 1, None)
    assert rc == _capi.RPY_ERR_SHAPE
    # an output whose elements would overlap under its strides
    rc = L.rpy_dense_ft_fma(None, None, None, _capi.RPY_F64, ctypes.byref(tb),
                            _capi.DegreeBoundsC(0, -1), _capi.DegreeBoundsC(0, -1),
                            _capi.DegreeBoundsC(0, -1), 3, 2, 7, 7, 2, 1, 1, None)
    assert rc == _capi.RPY_ERR_SHAPE


def test_log_to_lie_checks_arguments_on_the_host():
    """rpy_dense_log_to_lie rejects bad arguments before any device work: a
    missing or mis-shaped t2l matrix (width 2, depth 3: 5 x 15), an unknown
    dtype, an invalid basis, a negative batch, null data and short row
    strides; an empty batch is a no-op."""
    f = _capi.lib().rpy_dense_log_to_lie
    t2l = _capi.SparseMatrixC(5, 15, 0, None, None, None, 1)
    wrong = _capi.SparseMatrixC(5, 14, 0, None, None, None, 1)
    ref = ctypes.byref

    def call(m=t2l, dt=_capi.RPY_F64, w=2, d=3, n=2, os=5, ss=15, out=None, sig=None):
        return f(out, sig, dt, w, d, None if m is None else ref(m), n, os, ss, None)

    assert call(m=None) == _capi.RPY_ERR_INVALID_ARGUMENT
    assert call(dt=7) == _capi.RPY_ERR_DTYPE
    assert call(w=0) == _capi.RPY_ERR_BASIS
    assert call(m=wrong) == _capi.RPY_ERR_SHAPE
    assert call(n=-1) == _capi.RPY_ERR_SHAPE
    assert call(n=0) == _capi.RPY_OK
    assert call() == _capi.RPY_ERR_INVALID_ARGUMENT
    # strides are checked before anything is read through the (dummy) pointers
    assert call(os=4, out=8, sig=8) == _capi.RPY_ERR_SHAPE
    assert call(ss=14, out=8, sig=8, dt=_capi.RPY_F32) == _capi.RPY_ERR_SHAPE


def test_query_kernel_dyadic_cover_matches_reference_dissection(rng):
    """The query kernel's greedy cover of [A, B) (leaf units) equals
    to_dyadic_intervals of the clopen [A 2^-r, B 2^-r) (oracle restatement
    of intervals/src/dyadic_interval.cpp:191-228)."""
    L = _capi.lib()
    ks = np.zeros(64, dtype=np.int64)
    ns = np.zeros(64, dtype=np.int32)
    for r in (0, 1, 3, 7, 12):
        for _ in range(200):
            a, b = sorted(rng.integers(0, 2 ** r + 1, size=2))
            n = L.rpy_debug_dyadic_cover(int(a), int(b), r, _capi.i64p(ks),
                                         ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 64)
            got = [(int(ks[i]), int(ns[i])) for i in range(n)]
            want = O.to_dyadic_intervals(a / 2.0 ** r, b / 2.0 ** r, r) if a < b else []
            assert got == want, (a, b, r)


def test_select_rows_batch_matches_single_queries():
    """rpy_stream_select_rows_batch (the batched selection behind simplify)
    returns, per query, exactly the rows of the single-query selection after
    the support trim: random uneven streams, random and edge queries."""
    from roughpy_amd import streams as rs
    rng = np.random.default_rng(7)
    for res, n in ((0, 40), (4, 200), (7, 500)):
        params = np.unique(np.round(rng.uniform(0, 5, size=n), 3))
        inc = rng.normal(size=(params.shape[0], 2))
        st = rs.LieIncrementStream.from_increments(inc, indices=params, depth=2, resolution=res)
        lo, hi = st.support
        a = rng.uniform(lo - 1, hi + 1, size=60)
        b = a + rng.uniform(0, 2, size=60)
        a = np.concatenate([a, [lo, hi, hi + 1, lo - 2, 2.0]])
        b = np.concatenate([b, [hi, hi + 1, hi + 2, lo - 1, 2.0]])
        for qres in (res, res + 3):
            idx, offs = st._select_batch(a, b, qres)
            for q in range(a.shape[0]):
                single = st._select((a[q], b[q]), qres)
                assert np.array_equal(idx[offs[q]:offs[q + 1]], single), (res, qres, q)


def test_select_rows_batch_threaded_blocks():
    """Many queries run in blocks on host threads: the result equals one pass
    in query order, for a partition (consecutive queries sharing end points),
    for overlapping random queries (more rows than the stream holds, so the
    one-pass buffer overflows and the counting pass sizes the output), and
    for a failing block (its error message reaches the caller)."""
    from roughpy_amd import streams as rs
    rng = np.random.default_rng(11)
    n = 20000
    params = np.cumsum(rng.uniform(0.5, 1.5, size=n))
    inc = rng.normal(size=(n, 2))
    st = rs.LieIncrementStream.from_increments(inc, indices=params, depth=2, resolution=2)
    lo, hi = st.support
    cuts = np.sort(rng.uniform(lo, hi, size=9999))
    a = np.concatenate([[lo], cuts])
    b = np.concatenate([cuts, [hi]])
    qa = rng.uniform(lo, hi, size=6000)
    qb = qa + rng.uniform(0, 400, size=6000)
    for xa, xb in ((a, b), (qa, qb)):
        idx, offs = st._select_batch(xa, xb, 2)
        assert offs.shape[0] == xa.shape[0] + 1
        for q in range(0, xa.shape[0], 97):
            single = st._select((xa[q], xb[q]), 2)
            assert np.array_equal(idx[offs[q]:offs[q + 1]], single), q
        counts = np.diff(offs)
        ref = np.array([st._select((xa[q], xb[q]), 2).shape[0] for q in range(xa.shape[0])])
        assert np.array_equal(counts, ref)
    assert offs[-1] > n  # the overlapping queries select more rows than the stream has
    # a NaN query end fails inside a worker block; the message comes back
    bad_a = qa.copy()
    bad_a[5000] = np.nan
    with pytest.raises(ValueError, match="finite"):
        st._select_batch(bad_a, qb, 2)


# --- LieIncrementStream.from_increments keyword surface (host side) ----------

def _meta(s):
    return (s.width, s.depth, s.resolution, tuple(s.support), s._params.tolist(),
            s._rows.tolist())


def test_from_increments_requires_depth():
    """lie_increment_stream.cpp:103-109: without ctx, width and depth must be
    known (width defaults to the column count)."""
    with pytest.raises(ValueError, match="width and depth"):
        LieIncrementStream.from_increments(np.ones((3, 2)))
    with pytest.raises(ValueError, match="width and depth"):
        LieIncrementStream.from_increments(np.ones((3, 2)), width=2, depth=0)
    s = LieIncrementStream.from_increments(np.ones((3, 2)), depth=3)
    assert (s.width, s.depth, s.resolution) == (2, 3, 0)


def test_from_increments_index_column_is_cumulative(rng):
    """indices=<int> reads that column as parameter steps whose running sum
    gives the indices (lie_increment_stream.cpp:133-168); the column stays a
    data channel."""
    data = np.abs(rng.normal(size=(7, 3))) + 0.1
    by_col = LieIncrementStream.from_increments(data, indices=1, depth=2)
    explicit = LieIncrementStream.from_increments(data, indices=np.cumsum(data[:, 1]), depth=2)
    assert _meta(by_col) == _meta(explicit)
    assert by_col.width == 3
    with pytest.raises(RuntimeError):
        LieIncrementStream.from_increments(data, indices=3, depth=2)
    with pytest.raises(TypeError):
        LieIncrementStream.from_increments(data, indices="0", depth=2)
    with pytest.raises(ValueError, match="mismatch"):
        LieIncrementStream.from_increments(data, indices=[0.0, 1.0], depth=2)


def test_from_increments_inferred_resolution_matches_oracle():
    """Given indices without a resolution: py_param_to_resolution(min index) + 1
    (kwargs_to_path_metadata.cpp:224-229, lie_increment_stream.cpp:193-206),
    negative when every index is 2 or more; support [min, max + 2^-res)."""
    from roughpy_amd.streams import py_param_to_resolution
    for v, want in [(0.7, 1), (1.0, 0), (2.0, -1), (3.0, -1), (16.0, -4), (40.0, -5),
                    (1e6, -5), (0.1, 4)]:
        assert py_param_to_resolution(v) == want, v
    data = np.arange(12.0).reshape(4, 3)
    for idx in ([4.0, 5.0, 6.5, 9.0], [0.3, 0.5, 0.55, 2.0], [40.0, 41.0, 50.0, 64.0]):
        s = LieIncrementStream.from_increments(data, indices=np.array(idx), depth=2)
        o = O.IncrementStream(data, 3, 2, indices=np.array(idx))
        assert s.resolution == o.resolution
        assert tuple(s.support) == tuple(o.support)


def test_to_dyadic_intervals_negative_tolerance():
    """Pieces longer than 1 (tolerance below 0) dissect as the oracle's
    restatement of dyadic_interval.cpp:191-228 does."""
    for a, b, tol in [(4.0, 8.0, -1), (3.0, 17.0, -2), (0.0, 64.0, -5), (-6.0, 10.0, -1)]:
        assert to_dyadic_intervals(a, b, tol) == O.to_dyadic_intervals(a, b, tol)


def test_from_increments_include_time_vtype_schema_support(rng):
    data = rng.normal(size=(5, 2))
    idx = np.array([0.5, 1.0, 1.0, 2.5, 4.0])
    s = LieIncrementStream.from_increments(data, indices=idx, depth=2, include_time=True)
    # the time channel holds index - previous index (0 before the first row);
    # the repeated index keeps its first row
    steps = idx - np.concatenate([[0.0], idx[:-1]])
    keep = [0, 1, 3, 4]
    assert s.width == 3
    assert_allclose(s._rows, np.hstack([data, steps[:, None]])[keep])
    for vt in ("dense", "sparse", "Dense"):
        LieIncrementStream.from_increments(data, depth=2, vtype=vt)
    with pytest.raises(TypeError):
        LieIncrementStream.from_increments(data, depth=2, vtype="banded")
    t = LieIncrementStream.from_increments(data, depth=2, schema=["increment", "increment"])
    assert t.width == 2
    t = LieIncrementStream.from_increments(data, depth=2, channel_types={"a": "increment",
                                                                          "b": "increment"})
    assert t.width == 2
    with pytest.raises(NotImplementedError):
        LieIncrementStream.from_increments(data, depth=2, schema=["increment", "categorical"])
    with pytest.raises(ValueError, match="schema width"):
        LieIncrementStream.from_increments(data, width=2, depth=2, schema=["increment"])
    u = LieIncrementStream.from_increments(data, depth=2, support=RealInterval(-1.0, 3.0))
    assert tuple(u.support) == (-1.0, 3.0)
    with pytest.raises(TypeError, match="Interval"):
        LieIncrementStream.from_increments(data, depth=2, support=(0.0, 1.0))
