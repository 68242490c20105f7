"""Every compiled kernel variant of the tuned shapes against the oracle, and
the tuned table as the configuration that runs.

The signature forward of (4,6), (5,5), (6,5) and (8,4) has several compiled
variants per dtype (lane-chunk kernel at several prefix lengths and waves
per EU, stream-per-workgroup kernel at several prefix lengths and LDS block
lengths, the generic kernel; for the compensated mode also the block merge
at several segment lengths).  Each is forced through rpy_tune_select and
run on a seeded batch of uneven streams that includes empty and one-row
streams; the result must match the oracle's float64 Chen signature
(float64 1e-12, float32 and compensated float32 2e-5, max over degrees of
max|x - y| / max|y|).  The backward variants of (8,4) and (4,6) are checked
against the generic backward's float64 gradient.  A call without a selection
runs the tuned table's variant: its output is bit-equal to the one with that
variant selected explicitly.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from roughpy_amd import _capi, _device

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5), (8, 4), (6, 5), (4, 6)]
TOL = {_capi.RPY_F64: 1e-12, _capi.RPY_F32: 2e-5, _capi.RPY_F32_COMPENSATED: 2e-5}


def _variants(op, w, d, dt):
    L = _capi.lib()
    n = L.rpy_tune_variants(op, w, d, dt, None, 0)
    arr = (_capi.KernelVariantC * n)()
    L.rpy_tune_variants(op, w, d, dt, arr, n)
    return [a.astuple() for a in arr]


def _select(op, w, d, dt, v):
    L = _capi.lib()
    arg = None if v is None else ctypes.byref(_capi.KernelVariantC(*v))
    _capi.check(L.rpy_tune_select(op, w, d, dt, arg))


def _config(op, w, d, dt):
    c = _capi.KernelVariantC()
    src = ctypes.c_int32()
    _capi.check(_capi.lib().rpy_tune_config(op, w, d, dt, ctypes.byref(c), ctypes.byref(src)))
    return c.astuple(), src.value


def _batch(w, seed, n=48, lmax=90):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, lmax, size=n)
    lens[:3] = [0, 1, lmax + 37]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.maximum(lens, 1).repeat(lens))[:, None]
    return inc, offs


def _run_sig(torch, inc_d, offs_d, n, w, d, dt):
    tsize = int(_capi.lib().rpy_tensor_size(w, d))
    out = torch.full((n, tsize), float("nan"), dtype=inc_d.dtype, device="cuda")
    _capi.check(_capi.lib().rpy_stream_signature_batch(out.data_ptr(), inc_d.data_ptr(),
                                                       offs_d.data_ptr(), n, w, d, dt,
                                                       _device.stream_handle()))
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("w,d", SHAPES)
def test_every_signature_variant_matches_the_oracle(gpu, w, d):
    torch = gpu
    inc, offs = _batch(w, 900 + 10 * w + d)
    n = len(offs) - 1
    ref = O.signature_batch(inc, offs, w, d)
    offs_d = torch.from_numpy(offs).cuda()
    data = {_capi.RPY_F64: torch.from_numpy(inc).cuda()}
    data[_capi.RPY_F32] = data[_capi.RPY_F32_COMPENSATED] = data[_capi.RPY_F64].float()
    checked = 0
    try:
        for dt in (_capi.RPY_F32, _capi.RPY_F64, _capi.RPY_F32_COMPENSATED):
            for v in _variants(_capi.RPY_OP_SIGNATURE, w, d, dt):
                print(f"signature {w}x{d} dtype {dt} variant {v}", flush=True)
                _select(_capi.RPY_OP_SIGNATURE, w, d, dt, v)
                out = _run_sig(torch, data[dt], offs_d, n, w, d, dt).cpu().numpy()
                assert not np.isnan(out).any(), (dt, v)
                err = O.level_errors(out, ref, w, d).max()
                assert err <= TOL[dt], (dt, v, err)
                checked += 1
            _select(_capi.RPY_OP_SIGNATURE, w, d, dt, None)
    finally:
        for dt in TOL:
            _select(_capi.RPY_OP_SIGNATURE, w, d, dt, None)
    assert checked >= 12


@pytest.mark.parametrize("w,d", SHAPES)
def test_call_runs_the_configured_variant(gpu, w, d):
    """Without a selection a call runs rpy_tune_config's variant (the tuned
    row, else the default): bit-equal to forcing that variant."""
    torch = gpu
    inc, offs = _batch(w, 700 + w + d, n=200, lmax=300)
    n = len(offs) - 1
    offs_d = torch.from_numpy(offs).cuda()
    for dt, tdt in ((_capi.RPY_F32, torch.float32), (_capi.RPY_F64, torch.float64)):
        x = torch.from_numpy(inc).to(device="cuda", dtype=tdt)
        v, src = _config(_capi.RPY_OP_SIGNATURE, w, d, dt)
        assert src in (0, 1)
        plain = _run_sig(torch, x, offs_d, n, w, d, dt)
        _select(_capi.RPY_OP_SIGNATURE, w, d, dt, v)
        try:
            forced = _run_sig(torch, x, offs_d, n, w, d, dt)
        finally:
            _select(_capi.RPY_OP_SIGNATURE, w, d, dt, None)
        assert torch.equal(plain, forced), (dt, v)


@pytest.mark.parametrize("w,d", [(8, 4), (4, 6), (5, 5), (6, 5), (3, 4)])
def test_every_backward_variant_matches_the_generic_backward(gpu, w, d):
    torch = gpu
    rng = np.random.default_rng(31 + w + d)
    n, length = 40, 33
    inc = rng.normal(size=(n * length, w)) / np.sqrt(length)
    offs = np.arange(0, n * length + 1, length, dtype=np.int64)
    tsize = int(_capi.lib().rpy_tensor_size(w, d))
    G = rng.normal(size=(n, tsize))
    offs_d = torch.from_numpy(offs).cuda()
    L = _capi.lib()
    op = _capi.RPY_OP_SIGNATURE_VJP

    def vjp(dt, tdt):
        x = torch.from_numpy(inc).to(device="cuda", dtype=tdt)
        g = torch.from_numpy(G).to(device="cuda", dtype=tdt)
        S = _run_sig(torch, x, offs_d, n, w, d, dt)
        grad = torch.full_like(x, float("nan"))
        _capi.check(L.rpy_stream_signature_vjp(grad.data_ptr(), x.data_ptr(), offs_d.data_ptr(),
                                               n, w, d, S.data_ptr(), g.data_ptr(), dt,
                                               _device.stream_handle()))
        torch.cuda.synchronize()
        return grad.double().cpu().numpy()

    try:
        _select(op, w, d, _capi.RPY_F64, (0, 0, 0, 0))
        ref = vjp(_capi.RPY_F64, torch.float64)
        scale = np.abs(ref).max()
        for dt, tdt, tol in ((_capi.RPY_F64, torch.float64, 1e-10),
                             (_capi.RPY_F32, torch.float32, 1e-4)):
            vs = _variants(op, w, d, dt)
            assert len(vs) >= 2
            for v in vs:
                print(f"backward {w}x{d} dtype {dt} variant {v}", flush=True)
                _select(op, w, d, dt, v)
                got = vjp(dt, tdt)
                assert not np.isnan(got).any(), (dt, v)
                err = np.abs(got - ref).max() / scale
                assert err <= tol, (dt, v, err)
    finally:
        for dt in (_capi.RPY_F32, _capi.RPY_F64):
            _select(op, w, d, dt, None)
