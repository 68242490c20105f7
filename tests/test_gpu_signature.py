"""GPU parity of batched signatures, log-signatures and stream queries.

float64: within 1e-12 normwise per level of the oracle on the same inputs.
float32: error against the float64 oracle on the same float32 inputs is
stated and bounded (2e-5 here, lengths up to 1000).
At the full benchmark size (16384 streams, lengths up to 4096) float32
signatures are held to 1e-5 (max normwise relative error per level) against
the float64 oracle on the 32 longest and 32 random streams, and on every
stream through the float64 kernel, which equals the oracle there within
1e-12; the other checks are size-independent properties: level one equals
the sum of increments and Chen's identity between two halves of a stream.
"""
import json
import os

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

import roughpy_amd.compute as rc
from oracle import oracle as O
from roughpy_amd import _capi
from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
REF = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


def _uneven(rng, n, lo, hi, w, dtype):
    lens = rng.integers(lo, hi + 1, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.repeat(np.maximum(lens, 1), lens))[:, None])
    return inc.astype(dtype), offs


@pytest.mark.parametrize("w,d", [(2, 3), (3, 4), (5, 5), (8, 4), (6, 5), (4, 6), (2, 8), (9, 3)])
def test_signature_batch_f64_vs_oracle(gpu, rng, w, d):
    inc, offs = _uneven(rng, 37, 0, 60, w, np.float64)
    got = rs.signature_batch(inc, offs, w, d)
    ref = O.signature_batch(inc, offs, w, d, nthreads=8)
    err = O.level_errors(got, ref, w, d).max()
    assert err <= 1e-12, err


@pytest.mark.parametrize("w,d", [(3, 4), (5, 5), (8, 4), (6, 5)])
def test_signature_batch_f32_error_vs_f64_oracle(gpu, rng, w, d):
    inc, offs = _uneven(rng, 40, 1, 1000, w, np.float32)
    got = rs.signature_batch(inc, offs, w, d)
    ref = O.signature_batch(inc.astype(np.float64), offs, w, d, nthreads=8)
    err = O.level_errors(got, ref, w, d).max()
    print(f"f32 signature (w={w}, d={d}) max normwise level error vs f64 oracle: {err:.3e}")
    assert err <= 2e-5, err


def test_reference_signature_vectors(gpu):
    for key in sorted({k.split("/")[0] for k in REF.files}):
        w = int(key.split("_")[0][1:])
        d = int(key.split("_")[1][1:])
        inc = REF[key + "/sig_increments"]
        offs = REF[key + "/sig_offsets"]
        got = rs.signature_batch(inc, offs, w, d)
        tol = 1e-12 if inc.dtype == np.float64 else 2e-6
        assert O.level_errors(got, REF[key + "/signature"], w, d).max() <= tol


def test_empty_and_single_streams(gpu, rng):
    w, d = 5, 5
    lens = np.array([0, 1, 0, 3, 0])
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w))
    got = rs.signature_batch(inc, offs, w, d)
    unit = np.zeros(O.tensor_size(w, d))
    unit[0] = 1
    for s in (0, 2, 4):
        assert_array_equal(got[s], unit)
    assert O.level_errors(got[1], O.ft_exp(np.concatenate([[0], inc[0], np.zeros(unit.size - 1 - w)]), w, d), w, d).max() < 1e-14
    assert rs.signature_batch(np.zeros((0, w)), np.zeros(1, np.int64), w, d).shape == (0, unit.size)


def test_log_signature_batch_vs_oracle(gpu, rng):
    for (w, d) in [(3, 4), (5, 5), (8, 4)]:
        inc, offs = _uneven(rng, 9, 1, 200, w, np.float64)
        got = rs.log_signature_batch(inc, offs, w, d)
        for s in range(9):
            ref = O.log_signature(inc[offs[s]:offs[s + 1]], w, d)
            assert_allclose(got[s], ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("w,d", [(5, 5), (4, 6), (2, 9)])
def test_log_signature_batch_float32_vs_oracle(gpu, rng, w, d):
    """float32 increments: float32 signatures, then the float64 log (which
    widens them as it loads them at (5,5) and (4,6); (2,9) has no fixed-shape
    log and widens first) and the t2l projection, which writes its output."""
    inc, offs = _uneven(rng, 4, 1, 60, w, np.float64)
    inc32 = (inc / 2).astype(np.float32)
    got = rs.log_signature_batch(inc32, offs, w, d)
    for s in range(4):
        ref = O.log_signature(inc32[offs[s]:offs[s + 1]].astype(np.float64), w, d)
        assert_allclose(got[s], ref, rtol=0, atol=2e-5 * max(1.0, np.abs(ref).max()))


def test_brownian_w3_d4_1000_increments(gpu):
    """One stream, width 3, 1000 increments ~ N(0, 1/1000), depth 4: signature
    (121) and log-signature (32) through LieIncrementStream."""
    rng = np.random.default_rng(12345 + 1)
    w, d, n = 3, 4, 1000
    inc = rng.normal(0.0, np.sqrt(1.0 / n), size=(n, w))
    s = rs.LieIncrementStream.from_increments(inc, width=w, depth=d)
    sig = s.signature()
    ref = O.signature(inc, w, d)
    assert sig.data.shape == (121,)
    assert O.level_errors(sig.data, ref, w, d).max() <= 1e-12
    ls = s.log_signature()
    lref = O.log_signature(inc, w, d)
    assert ls.data.shape == (32,)
    assert_allclose(ls.data, lref, rtol=0, atol=1e-12 * np.abs(lref).max())
    # oracle stream semantics agree on sub-intervals
    so = O.IncrementStream(inc, w, d)
    for a, b in [(0, 1000), (10.5, 500.25), (999, 1000), (123, 124), (0.3, 0.7)]:
        g = s.signature(a, b).data
        assert O.level_errors(g, so.signature(a, b), w, d).max() <= 1e-12


def test_stream_goldens_on_gpu(gpu):
    g = GOLD["stream_tick_accuracy"]
    s = rs.LieIncrementStream.from_increments(np.array(g["data"]), indices=np.array(g["indices"]),
                                              width=2, depth=2)
    unit = np.eye(1, 7).ravel()
    q0, q1 = g["queries"]
    assert_array_equal(s.signature(q0["inf"], q0["sup"], q0["resolution_param"]).data, unit)
    arg = np.zeros(7)
    arg[:3] = q1["exp_arg"]
    assert_allclose(s.signature(q1["inf"], q1["sup"], q1["resolution_param"]).data,
                    O.ft_exp(arg, 2, 2), rtol=1e-15)
    g = GOLD["stream_interval_alignment"]
    s = rs.LieIncrementStream.from_increments(np.array(g["data"]), depth=2)
    assert_allclose(s.log_signature(rs.RealInterval(0.0, 0.1)).data[:3], g["first"]["expected_lie"])
    for a, b in g["zero_windows"]:
        assert_array_equal(s.log_signature(rs.RealInterval(a, b)).data, 0.0)
    for c in GOLD["stream_from_array_level1"]["cases"]:
        s = rs.LieIncrementStream.from_increments(np.array(c["data"]), width=c["width"],
                                                  depth=c["depth"])
        assert_allclose(s.signature(depth=c["sig_depth"]).data, c["expected"])
    g = GOLD["linear_signature"]
    x = np.array(g["increment"])
    s = rs.LieIncrementStream.from_increments(x[None, :], indices=np.array([0.0]), width=3, depth=3)
    l2 = np.einsum("i,j->ij", x, x) / 2
    exp = np.hstack([[1.0], x, l2.ravel(), np.einsum("ij,k->ijk", l2, x).ravel() / 3])
    assert_allclose(s.signature(rs.RealInterval(0.0, 1.0)).data, exp, rtol=1e-14)


def test_randint_level_one(gpu, rng):
    # tests/streams/test_lie_increment_path.py:290-332
    arr = rng.integers(0, 5, size=(4, 3))
    s = rs.LieIncrementStream.from_increments(arr, width=3, depth=2)
    sig = s.signature(rs.RealInterval(0.0, 5.0), 2)
    assert_array_equal(sig.data[:4], np.hstack([[1.0], arr.sum(axis=0)]))


def test_benchmark_size_properties(gpu):
    """16384 uneven streams, width 5, depth 5, lengths 64..4096 (the benchmark
    workload) -- properties that need no oracle at this size, plus a sample."""
    rng = np.random.default_rng(12345 + 3)
    w, d, n = 5, 5, 16384
    lens = rng.integers(64, 4097, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.repeat(lens, lens))[:, None]).astype(np.float32)
    t = gpu
    dinc = t.from_numpy(inc).cuda()
    doffs = t.from_numpy(offs).cuda()
    s32 = rs.signature_batch(dinc, doffs, w, d)
    s64 = rs.signature_batch(dinc.double(), doffs, w, d)
    a32, a64 = s32.cpu().numpy(), s64.cpu().numpy()
    # level one = sum of increments (float64 kernel)
    sums = np.add.reduceat(inc.astype(np.float64), offs[:-1], axis=0)
    assert_allclose(a64[:, 1:1 + w], sums, rtol=1e-12, atol=1e-12)
    # float32 signatures within 1e-5 (max normwise relative error per level)
    # of the float64 oracle: the 32 longest streams and 32 random ones against
    # the oracle itself, on the float32 increments promoted to float64 ...
    longest = np.argsort(-lens, kind="stable")[:32]
    pick = np.concatenate([longest, rng.choice(n, size=32, replace=False)])
    sub_lens = lens[pick]
    sub_offs = np.concatenate([[0], np.cumsum(sub_lens)]).astype(np.int64)
    sub_inc = np.concatenate([inc[offs[s]:offs[s + 1]] for s in pick]).astype(np.float64)
    ref = O.signature_batch(sub_inc, sub_offs, w, d, nthreads=8)
    err_oracle = O.level_errors(a32[pick], ref, w, d).max()
    print(f"benchmark workload f32 vs f64 oracle (32 longest + 32 random streams): "
          f"{err_oracle:.3e}")
    assert err_oracle <= 1e-5
    # ... the float64 kernel equals the oracle there within 1e-12 ...
    assert O.level_errors(a64[pick], ref, w, d).max() <= 1e-12
    # ... so all 16384 streams are held to the same bar through it
    err32 = O.level_errors(a32, a64, w, d).max()
    print(f"benchmark workload f32 vs f64 kernel, all streams: {err32:.3e}")
    assert err32 <= 1e-5
    # block-compensated float32: the same bars, and below the plain kernel's
    c32 = rs.signature_batch(dinc, doffs, w, d, compensated=True).cpu().numpy()
    errc_oracle = O.level_errors(c32[pick], ref, w, d).max()
    errc = O.level_errors(c32, a64, w, d).max()
    print(f"benchmark workload compensated f32: vs oracle {errc_oracle:.3e}, "
          f"vs f64 kernel (all streams) {errc:.3e}")
    assert errc_oracle <= 1e-5 and errc <= 1e-5
    assert errc < err32
    # Chen identity: split every stream in two and multiply the halves
    mid = offs[:-1] + lens // 2
    offs2 = np.empty(2 * n + 1, dtype=np.int64)
    offs2[0:-1:2] = offs[:-1]
    offs2[1::2] = mid
    offs2[-1] = offs[-1]
    halves = rs.signature_batch(dinc.double(), t.from_numpy(offs2).cuda(), w, d)
    tb = rc.TensorBasis(w, d)
    prod = rc.ft_mul(rc.FreeTensor(halves[0::2].contiguous(), tb),
                     rc.FreeTensor(halves[1::2].contiguous(), tb))
    assert O.level_errors(prod.data.cpu().numpy(), a64, w, d).max() <= 1e-12


@pytest.mark.parametrize("w,d", [(9, 3), (3, 7)])
def test_compensated_generic_shapes_round_once(gpu, rng, w, d):
    """Shapes without a specialised kernel run the compensated mode in
    float64 arithmetic on the float32 data: within 2^-23 normwise per level
    of the float64 oracle on the same float32 increments."""
    lens = rng.integers(1, 300, size=9)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / 8).astype(np.float32)
    got = rs.signature_batch(inc, offs, w, d, compensated=True)
    assert got.dtype == np.float32
    for s in range(len(lens)):
        ref = O.signature(inc[offs[s]:offs[s + 1]].astype(np.float64), w, d)
        assert O.level_errors(got[s].astype(np.float64), ref, w, d).max() <= 2.0 ** -23


SHAPES = [(3, 4), (5, 5), (8, 4), (6, 5), (4, 6)]


def _variant(w, d, dt=None):
    from roughpy_amd import tune
    v, _ = tune.config(_capi.RPY_OP_SIGNATURE, w, d,
                       _capi.RPY_F32_COMPENSATED if dt is None else dt)
    return v


def _long_streams(rng, w, T, n=48):
    lens = np.concatenate([[40 * T + 3, 16 * T + 7], rng.integers(4 * T, 16 * T + 1, size=n - 2)])
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) /
           np.sqrt(np.repeat(lens, lens))[:, None]).astype(np.float32)
    return inc, offs


@pytest.mark.parametrize("w,d", SHAPES)
def test_compensated_lanes_vs_f64_oracle(gpu, rng, w, d):
    """The compensated mode's float32 lanes (levels 1 and 2 accumulated
    with Kahan's compensation): the same prefix as the float32 kernel, within
    1e-5 normwise per level of the float64 oracle on long streams, and on
    average well below the plain float32 kernel's error.  Per stream the two
    errors are random quantities of the same kind, so only the worst case
    and the averages are compared: the compensation removes the error that
    the lowest levels feed into the higher ones, not the higher levels' own
    rounding, so the gain is in the average more than in the worst case."""
    v = _variant(w, d)
    assert v[0] in (_capi.RPY_KIND_LANE_CHUNK, _capi.RPY_KIND_STREAM_GROUP), v
    assert v[1] == _variant(w, d, _capi.RPY_F32)[1]
    inc, offs = _long_streams(rng, w, 256)
    got = rs.signature_batch(inc, offs, w, d, compensated=True)
    plain = rs.signature_batch(inc, offs, w, d)
    ref = O.signature_batch(inc.astype(np.float64), offs, w, d, nthreads=8)
    e_comp = O.level_errors(got.astype(np.float64), ref, w, d).max(axis=1)
    e_plain = O.level_errors(plain.astype(np.float64), ref, w, d).max(axis=1)
    print(f"({w},{d}) compensated lanes: max {e_comp.max():.2e} mean {e_comp.mean():.2e}; "
          f"plain max {e_plain.max():.2e} mean {e_plain.mean():.2e}")
    assert e_comp.max() <= 1e-5
    assert e_comp.mean() <= 0.85 * e_plain.mean()


def _block(w, d, T=256):
    return (_capi.RPY_KIND_BLOCK_MERGE, 0, T, 0)


class _Selected:
    """Runs the block-merge variant of a shape's compensated mode."""

    def __init__(self, w, d, v):
        self.args = (_capi.RPY_OP_SIGNATURE, w, d, _capi.RPY_F32_COMPENSATED)
        self.v = v

    def __enter__(self):
        from roughpy_amd import tune
        tune.select(*self.args, self.v)

    def __exit__(self, *exc):
        from roughpy_amd import tune
        tune.select(*self.args, None)


@pytest.mark.parametrize("w,d", SHAPES)
def test_block_merge_one_segment_is_the_lane_kernel(gpu, rng, w, d):
    """A stream of at most one segment's rows is the compensated lanes'
    signature of it, bit for bit (the merge only rounds hi - 0); empty
    streams give the unit tensor."""
    T = 256
    lens = np.concatenate([[0, 1, T], rng.integers(0, T + 1, size=20)])
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / np.sqrt(T)).astype(np.float32)
    lanes = rs.signature_batch(inc, offs, w, d, compensated=True)
    with _Selected(w, d, _block(w, d, T)):
        got = rs.signature_batch(inc, offs, w, d, compensated=True)
    assert_array_equal(got, lanes)
    assert got[0, 0] == 1 and not got[0, 1:].any()


@pytest.mark.parametrize("w,d", SHAPES)
def test_block_merge_vs_f64_oracle(gpu, rng, w, d):
    """Block merge (segments of 256 rows on the compensated lanes, merged
    per stream with Kahan sums): within 1e-5 of the float64 oracle and on
    average below the plain float32 kernel's error."""
    inc, offs = _long_streams(rng, w, 256)
    with _Selected(w, d, _block(w, d)):
        got = rs.signature_batch(inc, offs, w, d, compensated=True)
    plain = rs.signature_batch(inc, offs, w, d)
    ref = O.signature_batch(inc.astype(np.float64), offs, w, d, nthreads=8)
    e_comp = O.level_errors(got.astype(np.float64), ref, w, d).max(axis=1)
    e_plain = O.level_errors(plain.astype(np.float64), ref, w, d).max(axis=1)
    print(f"({w},{d}) block merge: max {e_comp.max():.2e} mean {e_comp.mean():.2e}; "
          f"plain max {e_plain.max():.2e} mean {e_plain.mean():.2e}")
    assert e_comp.max() <= 1e-5, e_comp.max()
    assert e_comp.mean() <= 0.7 * e_plain.mean()


def test_block_merge_rounds_of_streams_agree(gpu, rng):
    """The allocation-free entry point runs the block merge in rounds of
    streams that fit its scratch; the result equals the allocating call's."""
    from roughpy_amd import _device
    t = gpu
    w, d = 5, 5
    lens = rng.integers(0, 3000, size=300)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / 40).astype(np.float32)
    dinc = t.from_numpy(inc).cuda()
    doffs = t.from_numpy(offs).cuda()
    tsize = O.tensor_size(w, d)
    a = t.empty((300, tsize), dtype=t.float32, device="cuda")
    b = t.empty_like(a)
    L = _capi.lib()
    with _Selected(w, d, _block(w, d)):
        _capi.check(L.rpy_stream_signature_batch(a.data_ptr(), dinc.data_ptr(), doffs.data_ptr(),
                                                 300, w, d, _capi.RPY_F32_COMPENSATED,
                                                 _device.stream_handle()))
        _capi.check(L.rpy_stream_signature_batch_noalloc(b.data_ptr(), dinc.data_ptr(),
                                                         doffs.data_ptr(), 300, w, d,
                                                         _capi.RPY_F32_COMPENSATED,
                                                         _device.stream_handle()))
    assert t.equal(a, b)


def test_sorted_and_unsorted_schedules_agree(gpu, rng):
    """A batch large enough to be length-sorted (more than one round of waves)
    and the same streams in two halves small enough to keep input order give
    bit-identical signatures: the schedule moves work between waves, never
    changes a lane's arithmetic."""
    w, d, n = 5, 5, 3000
    lens = rng.integers(0, 120, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / 10).astype(np.float32)
    whole = rs.signature_batch(inc, offs, w, d)
    h = n // 2
    first = rs.signature_batch(inc[:offs[h]], offs[:h + 1], w, d)
    second = rs.signature_batch(inc[offs[h]:], offs[h:] - offs[h], w, d)
    assert_array_equal(whole, np.concatenate([first, second]))


@pytest.mark.parametrize("w,d,n", [(3, 4, 9), (8, 4, 5), (5, 5, 7), (4, 6, 3)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_log_signature_equals_the_separate_launches(gpu, rng, w, d, n, dtype):
    """For the fixed shapes the batched log-signature runs the float64 log and
    the Hall projection as one kernel, the log kept in LDS (tensors per
    workgroup: 4 at (3,4), 1 at (8,4), 2 at (5,5) and (4,6); the batches
    leave a partial last group).  It equals the float64 log (rpy_dense_ft_log)
    and the projection (rpy_dense_sparse_map) launched separately on the same
    signatures, to float64 rounding (the separate projection multiplies each
    term by its scale, 1)."""
    torch = gpu
    from roughpy_amd import _device
    from roughpy_amd.bases import LieBasis
    inc, offs = _uneven(rng, n, 1, 80, w, dtype)
    got = np.asarray(rs.log_signature_batch(inc, offs, w, d))
    tb, lb = rc.TensorBasis(w, d), LieBasis(w, d)
    size = tb.size()
    sig = np.asarray(rs.signature_batch(inc, offs, w, d), dtype=np.float64)  # exact widening
    S = torch.from_numpy(sig).cuda()
    lg = torch.empty_like(S)
    L, st = _capi.lib(), _device.stream_handle()
    full = _capi.DegreeBoundsC(0, -1)
    _capi.check(L.rpy_dense_ft_log(lg.data_ptr(), S.data_ptr(), _capi.RPY_F64, tb.c_struct(),
                                   full, full, n, size, size, 1, 1, st))
    out = torch.zeros((n, lb.size()), dtype=torch.float64, device="cuda")
    t2l = lb.get_t2l_matrix(np.float64)
    _capi.check(L.rpy_dense_sparse_map(out.data_ptr(), lg.data_ptr(), t2l.device_struct(),
                                       _capi.RPY_F64, 1.0, n, lb.size(), size, 1, 1, st))
    ref = out.cpu().numpy()
    assert_allclose(got, ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("w,d", [(5, 5), (2, 9)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_log_to_lie_of_held_signatures(gpu, rng, w, d, dtype):
    """rpy_dense_log_to_lie on signatures the caller holds, with rows padded
    (row strides above the tensor and Lie sizes), equals the batched
    log-signature of the same streams: the fused kernel at (5,5), the
    separate log (widening float32 first) and projection at (2,9)."""
    torch = gpu
    from roughpy_amd import _device
    from roughpy_amd.bases import LieBasis
    n = 6
    inc, offs = _uneven(rng, n, 1, 50, w, dtype)
    ref = np.asarray(rs.log_signature_batch(inc, offs, w, d))
    sig = np.asarray(rs.signature_batch(inc, offs, w, d))
    tb, lb = rc.TensorBasis(w, d), LieBasis(w, d)
    size, lsize = tb.size(), lb.size()
    S = torch.zeros((n, size + 3), dtype=torch.float64 if dtype == np.float64 else torch.float32,
                    device="cuda")
    S[:, :size] = torch.from_numpy(sig).cuda()
    out = torch.full((n, lsize + 5), float("nan"), dtype=torch.float64, device="cuda")
    code = _capi.RPY_F64 if dtype == np.float64 else _capi.RPY_F32
    L = _capi.lib()
    t2l = lb.get_t2l_matrix(np.float64).device_struct()
    # a known error text first: a successful call (the fallback included)
    # leaves it in place
    assert L.rpy_dense_log_to_lie(0, 0, code, w, d, t2l, -1, 0, 0, None) != _capi.RPY_OK
    assert L.rpy_last_error() == b"negative batch"
    _capi.check(L.rpy_dense_log_to_lie(out.data_ptr(), S.data_ptr(), code, w, d, t2l, n,
                                       lsize + 5, size + 3, _device.stream_handle()))
    assert L.rpy_last_error() == b"negative batch"
    got = out.cpu().numpy()
    assert np.isnan(got[:, lsize:]).all()
    assert_allclose(got[:, :lsize], ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()))
