"""The signature entry point that allocates nothing
(rpy_stream_signature_batch_noalloc), through the C-ABI: its outputs equal
rpy_stream_signature_batch's bit for bit on two calls in a row, the second
smaller (the scratch keeps the first call's contents), on the length-sorted
specialised kernel, the unsorted one and the generic kernel's compensated
mode (whose running signatures live in the scratch); a batch whose
temporaries exceed the scratch is refused before any launch."""
import numpy as np
import pytest

from roughpy_amd import _capi, _device

pytestmark = pytest.mark.gpu


def _batch(torch, rng, n, lo, hi, w, dtype):
    lens = rng.integers(lo, hi + 1, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.maximum(lens, 1).repeat(lens))[:, None])
    return (torch.from_numpy(inc.astype(dtype)).cuda(), torch.from_numpy(offs).cuda(), n)


@pytest.mark.parametrize("w,d,code", [(5, 5, _capi.RPY_F32), (5, 5, _capi.RPY_F64),
                                      (5, 5, _capi.RPY_F32_COMPENSATED),
                                      (9, 3, _capi.RPY_F32_COMPENSATED),
                                      (9, 3, _capi.RPY_F64)])
def test_noalloc_matches_allocating_entry_point(gpu, w, d, code):
    torch = gpu
    L = _capi.lib()
    rng = np.random.default_rng(12345 + 60)
    dtype = np.float64 if code == _capi.RPY_F64 else np.float32
    tsize = int(L.rpy_tensor_size(w, d))
    # first a batch large enough to be length-sorted, then a small one
    for n, lo, hi in [(6000, 0, 300), (37, 1, 50)]:
        inc, offs, n = _batch(torch, rng, n, lo, hi, w, dtype)
        want = torch.empty((n, tsize), dtype=inc.dtype, device="cuda")
        got = torch.full((n, tsize), float("nan"), dtype=inc.dtype, device="cuda")
        st = _device.stream_handle()
        _capi.check(L.rpy_stream_signature_batch(want.data_ptr(), inc.data_ptr(),
                                                 offs.data_ptr(), n, w, d, code, st))
        _capi.check(L.rpy_stream_signature_batch_noalloc(got.data_ptr(), inc.data_ptr(),
                                                         offs.data_ptr(), n, w, d, code, st))
        torch.cuda.synchronize()
        assert torch.equal(got, want), (w, d, code, n)


def test_noalloc_refuses_batches_beyond_its_scratch(gpu):
    torch = gpu
    L = _capi.lib()
    n = 1_000_000
    offs = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    inc = torch.zeros((1, 5), dtype=torch.float32, device="cuda")
    out = torch.empty((1, 1), dtype=torch.float32, device="cuda")
    rc = L.rpy_stream_signature_batch_noalloc(out.data_ptr(), inc.data_ptr(), offs.data_ptr(),
                                              n, 5, 5, _capi.RPY_F32, _device.stream_handle())
    assert rc == _capi.RPY_ERR_SHAPE
    assert b"scratch" in L.rpy_last_error()
    # nothing was launched: the empty batch still answers
    assert L.rpy_stream_signature_batch_noalloc(out.data_ptr(), inc.data_ptr(), offs.data_ptr(),
                                                0, 5, 5, _capi.RPY_F32,
                                                _device.stream_handle()) == _capi.RPY_OK
