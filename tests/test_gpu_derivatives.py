"""GPU parity of the derivatives.

* Forward mode: LieIncrementStream.signature_derivative against the
  reference goldens (tests/streams/test_lie_increment_path.py:219-287) and
  the oracle restatement of sig_derivative_impl.
* Reverse mode: the hand-written reverse-Chen VJP of batched signatures,
  checked against float64 central finite differences (step 1e-6) of the
  oracle signature, and the float32 gradient's error against the float64
  one stated.  Log-signature gradients (log series and t2l composed from
  differentiable kernels) are checked the same way.
"""
import json
import os

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import oracle as O
from roughpy_amd import _capi, _device
from roughpy_amd import autograd as ag
from roughpy_amd import streams as rs
from roughpy_amd.bases import TensorBasis

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))


def test_signature_derivative_goldens(gpu):
    g = GOLD["sig_derivative_w3_d2"]
    for depth, expected in [(1, [0.0, 0.0, 1.0, 0.0]), (2, g["expected"])]:
        p = rs.LieIncrementStream.from_increments(np.array([g["increment"]]),
                                                  indices=np.array([0.0]), width=3, depth=depth)
        d = p.signature_derivative(rs.RealInterval(0.0, 1.0), np.array(g["perturbation"]))
        assert_allclose(d.data, expected, atol=1e-15)


def test_signature_derivative_list_vs_oracle(gpu, rng):
    w, d, L = 3, 3, 20
    data = rng.normal(size=(L, w))
    idx = np.linspace(0.0, 1.0, L)
    p = rs.LieIncrementStream.from_increments(data, indices=idx, width=w, depth=d)
    so = O.IncrementStream(data, w, d, indices=idx)
    lsize = sum(O.witt_dim(w, k) for k in range(1, d + 1))
    perts = [(rs.RealInterval(0.0, 0.3), rng.uniform(0, 5, size=w)),
             (rs.RealInterval(0.3, 0.6), rng.uniform(0, 5, size=w)),
             (rs.RealInterval(0.6, 1.1), rng.uniform(0, 5, size=w))]
    got = p.signature_derivative(perts, 5).data
    pairs = []
    for ivl, pert in perts:
        E = so.signature(ivl.inf, ivl.sup, 5)
        lie = np.zeros(lsize)
        lie[:w] = pert
        pairs.append((E, O.lie_to_tensor(lie, w, d)))
    assert_allclose(got, O.signature_derivative(pairs, w, d), rtol=1e-12, atol=1e-12)


def _loss_oracle(inc, offs, G, w, d, s):
    S = O.signature(inc[offs[s]:offs[s + 1]], w, d)
    return float(np.dot(S, G[s]))


@pytest.mark.parametrize("w,d", [(2, 3), (3, 4), (5, 5), (8, 4)])
def test_signature_vjp_vs_finite_differences(gpu, rng, w, d):
    torch = gpu
    n = 6
    lens = rng.integers(1, 40, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.repeat(lens, lens))[:, None]
    tsize = O.tensor_size(w, d)
    G = rng.normal(size=(n, tsize))
    x = torch.from_numpy(inc).cuda().requires_grad_(True)
    S = ag.signature(x, torch.from_numpy(offs).cuda(), w, d)
    (S * torch.from_numpy(G).cuda()).sum().backward()
    grad = x.grad.cpu().numpy()
    h = 1e-6
    for _ in range(12):
        s = int(rng.integers(0, n))
        t = int(offs[s] + rng.integers(0, lens[s]))
        a = int(rng.integers(0, w))
        ip, im = inc.copy(), inc.copy()
        ip[t, a] += h
        im[t, a] -= h
        fd = (_loss_oracle(ip, offs, G, w, d, s) - _loss_oracle(im, offs, G, w, d, s)) / (2 * h)
        assert abs(grad[t, a] - fd) <= 1e-6 * max(1.0, abs(fd)), (s, t, a, grad[t, a], fd)


def test_signature_vjp_w8_d4_4096_streams(gpu):
    """Width 8, depth 4 (4681 coefficients), streams of length 256 with
    N(0, 1/256) increments; loss <G, S> with a fixed seeded G.  Float64
    gradient vs central differences on 64 sampled increments; float32
    gradient's error vs float64 stated."""
    torch = gpu
    rng = np.random.default_rng(12345 + 4)
    w, d, n, L = 8, 4, 4096, 256
    inc = rng.normal(0.0, np.sqrt(1.0 / L), size=(n * L, w))
    offs = np.arange(0, n * L + 1, L, dtype=np.int64)
    G = rng.normal(size=(n, O.tensor_size(w, d)))
    to = torch.from_numpy(offs).cuda()
    x64 = torch.from_numpy(inc).cuda().requires_grad_(True)
    (ag.signature(x64, to, w, d) * torch.from_numpy(G).cuda()).sum().backward()
    g64 = x64.grad.cpu().numpy()
    x32 = torch.from_numpy(inc.astype(np.float32)).cuda().requires_grad_(True)
    (ag.signature(x32, to, w, d) * torch.from_numpy(G.astype(np.float32)).cuda()).sum().backward()
    g32 = x32.grad.cpu().numpy()
    err32 = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"w8 d4 4096x256: float32 gradient max error / max |grad| = {err32:.2e}")
    assert err32 < 1e-3
    h = 1e-6
    worst = 0.0
    for _ in range(64):
        s = int(rng.integers(0, n))
        t = int(offs[s] + rng.integers(0, L))
        a = int(rng.integers(0, w))
        ip, im = inc.copy(), inc.copy()
        ip[t, a] += h
        im[t, a] -= h
        fd = (_loss_oracle(ip, offs, G, w, d, s) - _loss_oracle(im, offs, G, w, d, s)) / (2 * h)
        worst = max(worst, abs(g64[t, a] - fd) / max(1.0, abs(fd)))
    print(f"w8 d4 4096x256: float64 gradient vs central differences, worst {worst:.2e}")
    assert worst < 1e-6


def test_log_signature_gradient(gpu, rng):
    torch = gpu
    w, d = 3, 4
    n = 4
    lens = rng.integers(2, 15, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w)) / 3
    lsize = sum(O.witt_dim(w, k) for k in range(1, d + 1))
    Gl = rng.normal(size=(n, lsize))
    x = torch.from_numpy(inc).cuda().requires_grad_(True)
    ls = ag.log_signature(x, torch.from_numpy(offs).cuda(), w, d)
    for s in range(n):
        assert_allclose(ls[s].detach().cpu().numpy(),
                        O.log_signature(inc[offs[s]:offs[s + 1]], w, d), atol=1e-12)
    (ls * torch.from_numpy(Gl).cuda()).sum().backward()
    grad = x.grad.cpu().numpy()
    h = 1e-6
    for _ in range(10):
        s = int(rng.integers(0, n))
        t = int(offs[s] + rng.integers(0, lens[s]))
        a = int(rng.integers(0, w))
        ip, im = inc.copy(), inc.copy()
        ip[t, a] += h
        im[t, a] -= h
        fp = np.dot(O.log_signature(ip[offs[s]:offs[s + 1]], w, d), Gl[s])
        fm = np.dot(O.log_signature(im[offs[s]:offs[s + 1]], w, d), Gl[s])
        fd = (fp - fm) / (2 * h)
        assert abs(grad[t, a] - fd) <= 1e-6 * max(1.0, abs(fd)), (s, t, a)


def test_ft_mul_backward_matches_adjoints(gpu, rng):
    torch = gpu
    from roughpy_amd.bases import TensorBasis
    w, d = 3, 3
    tb = TensorBasis(w, d)
    n = tb.size()
    a = rng.normal(size=(2, n))
    b = rng.normal(size=(2, n))
    g = rng.normal(size=(2, n))
    ta = torch.from_numpy(a).cuda().requires_grad_(True)
    tbb = torch.from_numpy(b).cuda().requires_grad_(True)
    out = ag.ft_mul(ta, tbb, basis=tb)
    assert_allclose(out.detach().cpu().numpy(), O.ft_mul(a, b, w, d), atol=1e-13)
    (out * torch.from_numpy(g).cuda()).sum().backward()
    # brute force: d<g, a b>/da_i = <g, e_i b>
    E = np.eye(n)
    ga = np.array([[np.dot(g[r], O.ft_mul(E[i], b[r], w, d)) for i in range(n)] for r in range(2)])
    gb = np.array([[np.dot(g[r], O.ft_mul(a[r], E[i], w, d)) for i in range(n)] for r in range(2)])
    assert_allclose(ta.grad.cpu().numpy(), ga, atol=1e-12)
    assert_allclose(tbb.grad.cpu().numpy(), gb, atol=1e-12)


@pytest.mark.parametrize("w,d", [(8, 2), (8, 3), (8, 4), (4, 3), (4, 5), (4, 6), (2, 6), (2, 7),
                                 (3, 3), (5, 4)])
def test_signature_vjp_kernels_vs_host_lane_algebra(gpu, rng, w, d):
    """Both backward kernels (one wave per stream where the shape has a
    64-lane prefix split, the LDS workgroup kernel otherwise) against the
    host run of the lane algebra, which the CPU suite pins to central
    differences; float32 results within 1e-4 of the float64 ones."""
    from roughpy_amd import _capi
    torch = gpu
    L = _capi.lib()
    n = 5
    lens = rng.integers(1, 30, size=n)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w)) / 4
    tsize = O.tensor_size(w, d)
    G = rng.normal(size=(n, tsize))
    x = torch.from_numpy(inc).cuda().requires_grad_(True)
    (ag.signature(x, torch.from_numpy(offs).cuda(), w, d) * torch.from_numpy(G).cuda()).sum().backward()
    got = x.grad.cpu().numpy()
    x32 = torch.from_numpy(inc.astype(np.float32)).cuda().requires_grad_(True)
    (ag.signature(x32, torch.from_numpy(offs).cuda(), w, d)
     * torch.from_numpy(G.astype(np.float32)).cuda()).sum().backward()
    got32 = x32.grad.cpu().numpy()
    # the prefix length the float64 register kernel uses for the shape
    prefix64 = {(8, 2): 2, (8, 3): 2, (8, 4): 2, (4, 3): 3, (4, 5): 3, (4, 6): 3, (2, 6): 6,
                (2, 7): 6}
    emulated = (w, d) in prefix64
    for s in range(n):
        seg = np.ascontiguousarray(inc[offs[s]:offs[s + 1]])
        S = O.signature(seg, w, d)
        if emulated:
            ref = np.full(seg.shape, np.nan)
            _capi.check(L.rpy_debug_vjp_host_emulation(_capi.dblp(ref), _capi.dblp(seg), len(seg),
                                                        w, d, prefix64[(w, d)], _capi.dblp(S),
                                                        _capi.dblp(np.ascontiguousarray(G[s]))))
            assert_allclose(got[offs[s]:offs[s + 1]], ref, rtol=1e-11, atol=1e-11)
        ref64 = got[offs[s]:offs[s + 1]]
        scale = max(1.0, np.abs(ref64).max())
        assert np.abs(got32[offs[s]:offs[s + 1]] - ref64).max() <= 1e-4 * scale


def _log_loss_oracle(seg, g, w, d):
    return float(np.dot(O.log_signature(seg, w, d), g))


def test_log_signature_vjp_w8_d4_4096_streams(gpu):
    """Log-signatures at width 8, depth 4 (Lie dimension 1212), 4096 streams of
    length 256 with N(0, 1/256) increments; loss <G, logsig> with a fixed
    seeded G.  Float64 gradient vs float64 central differences (step 1e-6) of
    the oracle log-signature on 64 sampled increments; the float32 gradient's
    error against the float64 one stated."""
    torch = gpu
    rng = np.random.default_rng(12345 + 5)
    w, d, n, L = 8, 4, 4096, 256
    inc = rng.normal(0.0, np.sqrt(1.0 / L), size=(n * L, w))
    offs = np.arange(0, n * L + 1, L, dtype=np.int64)
    lsize = sum(O.witt_dim(w, k) for k in range(1, d + 1))
    Gl = rng.normal(size=(n, lsize))
    to = torch.from_numpy(offs).cuda()
    x64 = torch.from_numpy(inc).cuda().requires_grad_(True)
    ls64 = ag.log_signature(x64, to, w, d)
    pick = rng.choice(n, size=4, replace=False)
    for s in pick:
        ref = O.log_signature(inc[offs[s]:offs[s + 1]], w, d)
        assert_allclose(ls64[s].detach().cpu().numpy(), ref, rtol=0,
                        atol=1e-12 * max(1.0, np.abs(ref).max()))
    (ls64 * torch.from_numpy(Gl).cuda()).sum().backward()
    g64 = x64.grad.cpu().numpy()
    x32 = torch.from_numpy(inc.astype(np.float32)).cuda().requires_grad_(True)
    (ag.log_signature(x32, to, w, d) * torch.from_numpy(Gl).cuda()).sum().backward()
    g32 = x32.grad.cpu().numpy().astype(np.float64)
    err32 = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"log-signature w8 d4 4096x256: float32 gradient max error / max |grad| = {err32:.2e}")
    assert err32 < 1e-3
    h = 1e-6
    worst = 0.0
    for _ in range(64):
        s = int(rng.integers(0, n))
        r = int(rng.integers(0, L))
        a = int(rng.integers(0, w))
        ip = inc[offs[s]:offs[s + 1]].copy()
        im = ip.copy()
        ip[r, a] += h
        im[r, a] -= h
        fd = (_log_loss_oracle(ip, Gl[s], w, d) - _log_loss_oracle(im, Gl[s], w, d)) / (2 * h)
        worst = max(worst, abs(g64[offs[s] + r, a] - fd) / max(1.0, abs(fd)))
    print(f"log-signature w8 d4 4096x256: float64 gradient vs central differences, "
          f"worst {worst:.2e}")
    assert worst < 1e-6


@pytest.mark.parametrize("w,d,batch", [(3, 4, 6), (8, 4, 5), (5, 5, 3), (4, 6, 3), (2, 7, 4),
                                       (3, 8, 2), (1, 5, 2), (4, 1, 3)])
def test_ft_log_vjp_against_central_differences(gpu, w, d, batch):
    """rpy_dense_ft_log_vjp (reverse Horner chain of the log) against float64
    central differences (step 1e-6) of <H, log(S)> with the oracle's log, on
    sampled coefficients of every degree; (3,8) takes the global-workspace
    path (its float64 arrays, 197 KB, exceed LDS), the others run in LDS.
    The float32 call equals the float64 one to float32 rounding."""
    torch = gpu
    rng = np.random.default_rng(300 + 10 * w + d)
    tb = TensorBasis(w, d)
    n = tb.size()
    db = O.tensor_degree_begin(w, d)
    S = np.zeros((batch, n))
    S[:, 0] = 1.0
    for k in range(1, d + 1):
        S[:, db[k]:db[k + 1]] = rng.uniform(-1, 1, size=(batch, db[k + 1] - db[k])) / 2 ** k
    H = rng.normal(size=(batch, n))
    Sd = torch.from_numpy(S).cuda()
    Hd = torch.from_numpy(H).cuda()
    out = torch.full((batch, n), float("nan"), dtype=torch.float64, device="cuda")
    L = _capi.lib()
    _capi.check(L.rpy_dense_ft_log_vjp(out.data_ptr(), Sd.data_ptr(), Hd.data_ptr(),
                                       _capi.RPY_F64, tb.c_struct(), batch, n, n, n,
                                       _device.stream_handle()))
    got = out.cpu().numpy()
    assert (got[:, 0] == 0).all()
    h = 1e-6
    for e in range(batch):
        idx = [int(rng.integers(db[k], db[k + 1])) for k in range(1, d + 1) for _ in range(3)]
        for i in idx:
            sp, sm = S[e].copy(), S[e].copy()
            sp[i] += h
            sm[i] -= h
            fd = (np.dot(H[e], O.ft_log(sp, w, d)) - np.dot(H[e], O.ft_log(sm, w, d))) / (2 * h)
            assert abs(got[e, i] - fd) <= 1e-6 * max(1.0, abs(fd)), (e, i, got[e, i], fd)
    out32 = torch.full((batch, n), float("nan"), dtype=torch.float32, device="cuda")
    S32, H32 = Sd.float(), Hd.float()
    _capi.check(L.rpy_dense_ft_log_vjp(out32.data_ptr(), S32.data_ptr(), H32.data_ptr(),
                                       _capi.RPY_F32, tb.c_struct(), batch, n, n, n,
                                       _device.stream_handle()))
    S32d, H32d = S32.double(), H32.double()  # bound: the call reads them
    _capi.check(L.rpy_dense_ft_log_vjp(out.data_ptr(), S32d.data_ptr(), H32d.data_ptr(),
                                       _capi.RPY_F64, tb.c_struct(), batch, n, n, n,
                                       _device.stream_handle()))
    ref32 = out.cpu().numpy()
    assert_allclose(out32.cpu().numpy(), ref32, rtol=1e-6, atol=1e-6 * np.abs(ref32).max())


def test_ft_log_vjp_equals_the_composed_backward(gpu, rng):
    """The log cotangent kernel equals reverse mode through the Horner chain
    built from ft_mul autograd nodes (adjoint multiplication kernels), to
    float64 rounding."""
    torch = gpu
    w, d, batch = 4, 4, 7
    tb = TensorBasis(w, d)
    n = tb.size()
    S = rng.uniform(-0.3, 0.3, size=(batch, n))
    S[:, 0] = 1.0
    H = rng.normal(size=(batch, n))
    Sd = torch.from_numpy(S).cuda().requires_grad_(True)
    mask = torch.ones(n, dtype=torch.float64, device="cuda")
    mask[0] = 0.0
    Y = Sd * mask
    unit = torch.zeros((batch, n), dtype=torch.float64, device="cuda")
    unit[:, 0] = 1.0
    acc = torch.zeros_like(unit)
    for deg in range(d, 0, -1):
        acc = acc + unit * ((1.0 if deg % 2 else -1.0) / deg)
        acc = ag.ft_mul(acc, Y, basis=tb)
    Hd = torch.from_numpy(H).cuda()
    (acc * Hd).sum().backward()
    composed = Sd.grad.cpu().numpy()
    out = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    Sc = torch.from_numpy(S).cuda()
    _capi.check(_capi.lib().rpy_dense_ft_log_vjp(out.data_ptr(), Sc.data_ptr(), Hd.data_ptr(),
                                                 _capi.RPY_F64, tb.c_struct(), batch, n, n, n,
                                                 _device.stream_handle()))
    assert_allclose(out.cpu().numpy(), composed, rtol=1e-12, atol=1e-12 * np.abs(composed).max())
