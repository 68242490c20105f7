"""The C-ABI entry points inside a HIP graph.

Each entry point is captured once with torch.cuda.graph (the library takes
the capturing stream), then the graph is replayed several times with new
inputs written into the captured buffers by eager work between replays; every
replay must equal the eager call on the same inputs.  No entry point records
a memset (counters and outputs are cleared by kernels), and temporaries are
stream-ordered allocations, which a graph records as allocation nodes.
"""
import numpy as np
import pytest

from roughpy_amd import _capi, _device

pytestmark = pytest.mark.gpu


def _replays(torch, refill, run, read, n_replays=4):
    """Capture run() into a graph, then for each replay: refill the inputs
    eagerly, replay, synchronise, and compare with an eager run."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()  # warm-up outside the capture (allocator pools, kernel loads)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    torch.cuda.synchronize()
    for i in range(n_replays):
        refill(i)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = read().clone()
        run()
        torch.cuda.synchronize()
        ref = read().clone()
        assert torch.equal(got, ref), f"replay {i} differs from the eager call"


def test_signature_batch_in_a_graph(gpu, rng):
    torch = gpu
    w, d, n = 5, 5, 64
    lens = rng.integers(10, 300, size=n)
    offs = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)).cuda()
    rows = int(lens.sum())
    inc = torch.zeros((rows, w), dtype=torch.float32, device="cuda")
    tsize = int(_capi.lib().rpy_tensor_size(w, d))
    out = torch.zeros((n, tsize), dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda")

    def refill(i):
        gen.manual_seed(100 + i)
        inc.copy_(torch.randn((rows, w), generator=gen, device="cuda") * 0.05)

    def run():
        _capi.check(_capi.lib().rpy_stream_signature_batch(
            out.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, w, d, _capi.RPY_F32,
            _device.stream_handle()))

    refill(99)
    _replays(torch, refill, run, lambda: out)


def test_fma_in_a_graph(gpu):
    torch = gpu
    from roughpy_amd.bases import TensorBasis
    w, d, batch = 4, 6, 256
    tb = TensorBasis(w, d)
    n = tb.size()
    a = torch.zeros((batch, n), device="cuda")
    b = torch.zeros((batch, n), device="cuda")
    c = torch.zeros((batch, n), device="cuda")
    c0 = torch.zeros((batch, n), device="cuda")
    gen = torch.Generator(device="cuda")
    B = _capi.DegreeBoundsC
    cs = tb.c_struct()

    def refill(i):
        gen.manual_seed(200 + i)
        a.copy_(torch.rand((batch, n), generator=gen, device="cuda") - 0.5)
        b.copy_(torch.rand((batch, n), generator=gen, device="cuda") - 0.5)
        c0.copy_(torch.rand((batch, n), generator=gen, device="cuda") - 0.5)

    def run():
        c.copy_(c0)
        _capi.check(_capi.lib().rpy_dense_ft_fma(
            c.data_ptr(), a.data_ptr(), b.data_ptr(), _capi.RPY_F32, cs, B(0, -1), B(0, -1),
            B(0, -1), batch, n, n, n, 1, 1, 1, _device.stream_handle()))

    refill(99)
    _replays(torch, refill, run, lambda: c)


def test_signature_backward_in_a_graph(gpu):
    torch = gpu
    w, d, n, L = 8, 4, 32, 64
    offs = torch.arange(0, n * L + 1, L, dtype=torch.int64, device="cuda")
    tsize = int(_capi.lib().rpy_tensor_size(w, d))
    inc = torch.zeros((n * L, w), device="cuda")
    G = torch.zeros((n, tsize), device="cuda")
    S = torch.zeros((n, tsize), device="cuda")
    grad = torch.zeros((n * L, w), device="cuda")
    gen = torch.Generator(device="cuda")
    lib = _capi.lib()

    def refill(i):
        gen.manual_seed(300 + i)
        inc.copy_(torch.randn((n * L, w), generator=gen, device="cuda") / 8)
        G.copy_(torch.randn((n, tsize), generator=gen, device="cuda"))

    def run():
        _capi.check(lib.rpy_stream_signature_batch(S.data_ptr(), inc.data_ptr(), offs.data_ptr(),
                                                   n, w, d, _capi.RPY_F32,
                                                   _device.stream_handle()))
        _capi.check(lib.rpy_stream_signature_vjp(grad.data_ptr(), inc.data_ptr(),
                                                 offs.data_ptr(), n, w, d, S.data_ptr(),
                                                 G.data_ptr(), _capi.RPY_F32,
                                                 _device.stream_handle()))

    refill(99)
    _replays(torch, refill, run, lambda: grad)


def test_dyadic_queries_in_a_graph(gpu, rng):
    torch = gpu
    from roughpy_amd import streams as rs
    w, d, n, r, J = 5, 5, 16, 10, 4
    lens = rng.integers(50, 900, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    params = np.concatenate([np.arange(L) / L for L in lens])
    inc = (rng.normal(size=(int(offs[-1]), w)) / 20).astype(np.float32)
    cache = rs.DyadicSignatureCache(inc, params, offs, w, d, resolution=r, levels=J)
    nq = 200
    sid = torch.zeros(nq, dtype=torch.int64, device="cuda")
    qa = torch.zeros(nq, dtype=torch.float64, device="cuda")
    qb = torch.zeros(nq, dtype=torch.float64, device="cuda")
    out = cache.query(sid, qa, qb)

    def refill(i):
        g = np.random.default_rng(400 + i)
        a = np.floor(g.uniform(0, 1, nq) * 2 ** r) / 2 ** r
        b = np.floor(g.uniform(0, 1, nq) * 2 ** r) / 2 ** r
        sid.copy_(torch.from_numpy(g.integers(0, n, size=nq)))
        qa.copy_(torch.from_numpy(np.minimum(a, b)))
        qb.copy_(torch.from_numpy(np.maximum(a, b)))

    def run():
        cache.query(sid, qa, qb, out=out)

    refill(99)
    _replays(torch, refill, run, lambda: out)
