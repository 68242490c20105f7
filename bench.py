"""Benchmark: batched signatures of uneven streams on one MI355X.

Workload (the reference's headline configuration): 16384 piecewise-linear
streams of width 5, lengths uniform in [64, 4096], increments N(0, 1/L) in
float32, depth-5 signatures (3906 coefficients), all in one persistent
launch.  A step is one pass of the signature kernel over the whole batch with
inputs already resident in HBM.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f32|f64]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, D, N_STREAMS, LMIN, LMAX = 5, 5, 16384, 64, 4096
SEED = 12345 + 3
# fused multiply-adds per increment of the depth-5 Chen step at width 5:
# sum_{m=1..D} (D - m + 1) W^m (shared Horner partials, unit level exploited)
FMA_PER_STEP = sum((D - m + 1) * W ** m for m in range(1, D + 1))
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X published vector peaks
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
PEAK_HBM_GBS = 8000.0


PROFILE = os.path.join(ROOT, "profiles", "chen_signature_w5_d5_16384streams.md")


def progress(msg):
    """One line per finished step on stderr, so a long run shows where it is."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(dtype):
    """HBM bytes per launch of the signature kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        for line in open(PROFILE):
            if line.startswith(f"traffic_bytes_{dtype} = "):
                return int(line.split("=")[1])
    except OSError:
        pass
    return None


def make_workload(seed=SEED):
    rng = np.random.default_rng(seed)
    lens = rng.integers(LMIN, LMAX + 1, size=N_STREAMS)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), W)).astype(np.float32)
    inc /= np.sqrt(np.repeat(lens, lens)).astype(np.float32)[:, None]
    return inc, offs, lens


def time_kernel(sig_fn, steps, warmup, torch):
    """Average ms per step with HIP events on the launch stream."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        sig_fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        sig_fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def cover_counts(A, B, r, J):
    """Per query: the number of pieces of the maximal dyadic cover of [A, B)
    (units of 2^-r) that are stored nodes (level <= J), vectorised over the
    queries (the same left-to-right greedy cover as the query kernel)."""
    A = A.astype(np.int64).copy()
    B = B.astype(np.int64)
    nodes = np.zeros(A.shape, dtype=np.int64)
    active = A < B
    while active.any():
        a = A[active]
        b = B[active]
        lg = np.full(a.shape, r, dtype=np.int64)
        nz = a != 0
        low = a[nz] & -a[nz]
        lg[nz] = np.log2(low).astype(np.int64)
        while True:
            over = a + (np.int64(1) << lg) > b
            if not over.any():
                break
            lg[over] -= 1
        nodes[np.flatnonzero(active)] += (r - lg) <= J
        A[active] = a + (np.int64(1) << lg)
        active = A < B
    return nodes


def dyadic_cache_leg(torch, inc, offs, lens, dinc32, doffs, n_queries=100000, levels=6,
                     resolution=12):
    """Dyadic cache of the same batch at resolution 12 (stored levels 0..6)
    and 100000 random clopen queries snapped to the 2^-12 grid."""
    from roughpy_amd import streams as rs
    rng = np.random.default_rng(SEED + 2)
    params = np.concatenate([np.arange(L, dtype=np.float64) / L for L in lens])
    dparams = torch.from_numpy(params).cuda()
    sid = torch.from_numpy(rng.integers(0, N_STREAMS, size=n_queries)).cuda()
    a = np.floor(rng.uniform(0, 1, size=n_queries) * 2 ** resolution) / 2 ** resolution
    b = np.floor(rng.uniform(0, 1, size=n_queries) * 2 ** resolution) / 2 ** resolution
    qa = torch.from_numpy(np.minimum(a, b)).cuda()
    qb = torch.from_numpy(np.maximum(a, b)).cuda()
    t0 = time.perf_counter()
    cache = rs.DyadicSignatureCache(dinc32, dparams, doffs, W, D, resolution=resolution,
                                    levels=levels)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    out = cache.query(sid, qa, qb)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(3):
        cache.query(sid, qa, qb, out=out)
    ev1.record()
    torch.cuda.synchronize()
    q_ms = ev0.elapsed_time(ev1) / 3
    t0 = time.perf_counter()
    cache.build()
    torch.cuda.synchronize()
    rebuild_ms = (time.perf_counter() - t0) * 1e3
    # algorithmic HBM bytes of the query batch: every stored node a query
    # multiplies is read once, every answer written once (raw fine-run rows
    # are < 1% and left out)
    tsize = int(cache.tsize)
    nodes = cover_counts(np.minimum(a, b) * 2 ** resolution, np.maximum(a, b) * 2 ** resolution,
                         resolution, levels)
    q_bytes = float((nodes.sum() + n_queries) * tsize * 4)
    res = {"levels_stored": levels, "resolution": resolution, "dtype": "f32",
           "cache_bytes": int(cache.cache.numel()) * 4, "build_ms": rebuild_ms,
           "first_build_ms": build_ms, "queries": n_queries, "query_ms": q_ms,
           "queries_per_s": n_queries / (q_ms * 1e-3),
           "query_nodes_mean": float(nodes.mean()),
           "query_hbm_gb_s": q_bytes / (q_ms * 1e-3) / 1e9,
           "query_hbm_frac": q_bytes / (q_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
    del cache, out
    torch.cuda.empty_cache()
    return res


def level_rel_err(x, ref, width, depth):
    """Max over rows and degrees of max|x - ref| / max|ref| within a degree."""
    # degree k starts at 1 + w + ... + w^(k-1)
    db = [sum(width ** j for j in range(k)) for k in range(depth + 2)]
    err = 0.0
    for k in range(depth + 1):
        a = x[:, db[k]:db[k + 1]].double()
        b = ref[:, db[k]:db[k + 1]].double()
        num = (a - b).abs().amax(dim=1)
        den = b.abs().amax(dim=1).clamp_min(1e-300)
        err = max(err, float((num / den).max()))
    return err


def fma_leg(torch, steps, warmup):
    """Batched dense free-tensor FMA C += A (x) B, width 4, depth 6 (5461
    coefficients), batch 65536; A, B, C with unit scalar level and degree-k
    coefficients uniform[-1, 1] / k!.  Errors are of one FMA from the same C.
    HBM-bound: algorithmic traffic per tensor triple is read A, B, C and
    write C."""
    from math import factorial

    from roughpy_amd import _capi, _device
    from roughpy_amd.bases import TensorBasis
    w, d, batch = 4, 6, 65536
    tb = TensorBasis(w, d)
    n = tb.size()
    db = tb.degree_begin
    cs = tb.c_struct()
    scale = torch.empty(n, dtype=torch.float32, device="cuda")
    for k in range(d + 1):
        scale[db[k]:db[k + 1]] = 1.0 / factorial(k)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 20)
    A32 = (torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1) * scale
    B32 = (torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1) * scale
    C0 = (torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1) * scale
    A32[:, 0] = 1.0
    B32[:, 0] = 1.0
    C0[:, 0] = 1.0
    L = _capi.lib()
    Bd = _capi.DegreeBoundsC
    fma_per = sum((m + 1) * w ** m for m in range(d + 1))
    res = {"workload": f"C += A (x) B, width {w}, depth {d} ({n} coefficients), batch {batch}",
           "fma_per_product": fma_per}
    outs = {}
    for name, code, A, B in (("f64", _capi.RPY_F64, A32.double(), B32.double()),
                             ("f32", _capi.RPY_F32, A32, B32),
                             ("f32_compensated", _capi.RPY_F32_COMPENSATED, A32, B32)):
        C = C0.to(A.dtype, copy=True)

        def fn():
            _capi.check(L.rpy_dense_ft_fma(C.data_ptr(), A.data_ptr(), B.data_ptr(), code,
                                           cs, Bd(0, -1), Bd(0, -1), Bd(0, -1),
                                           batch, n, n, n, 1, 1, 1, _device.stream_handle()))
        try:
            fn()
        except NotImplementedError:
            continue
        torch.cuda.synchronize()
        outs[name] = C.clone()
        ms = time_kernel(fn, steps, warmup, torch)
        es = A.element_size()
        gbs = 4.0 * n * es * batch / (ms * 1e-3) / 1e9
        res[name] = {"ms_per_step": ms, "products_per_s": batch / (ms * 1e-3),
                     "hbm_gb_s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
                     "tflops": 2.0 * fma_per * batch / (ms * 1e-3) / 1e12}
        del C
    for name in ("f32", "f32_compensated"):
        if name in outs:
            res[name]["rel_err_vs_f64"] = level_rel_err(outs[name], outs["f64"], w, d)
    del outs, A32, B32, C0
    torch.cuda.empty_cache()
    return res


def hbm_leg(torch, steps, warmup):
    """What plain device copies reach on this GPU's HBM, beside the 8 TB/s
    nominal figure the HBM rooflines are quoted against: torch's copy_
    (read and write, 1:1) and fill_ (write only) over 4 GiB."""
    nbytes = 4 << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    copy_ms = time_kernel(lambda: b.copy_(a), steps, warmup, torch)
    fill_ms = time_kernel(lambda: b.fill_(2.0), steps, warmup, torch)
    del a, b
    torch.cuda.empty_cache()
    return {"workload": "4 GiB float32, torch copy_ and fill_",
            "copy_tb_s": 2 * nbytes / (copy_ms * 1e-3) / 1e12,
            "fill_tb_s": nbytes / (fill_ms * 1e-3) / 1e12,
            "nominal_tb_s": PEAK_HBM_GBS / 1e3}


def dense_leg(torch, steps, warmup):
    """The other dense operations of the path on a batch of 16384 tensors of
    width 5, depth 5 (3906 coefficients; Lie dimension 829), inputs resident
    in HBM: in-place product, exp, fmexp, log, antipode, adjoint of left
    multiplication, lie_to_tensor and tensor_to_lie.  Each is HBM-bound; the
    algorithmic bytes per element are the operands read plus the result
    written.  `ms` times the C-ABI call alone on preallocated outputs (HIP
    events on the launch stream); `api_ms` times the same operation through
    the Python mirror of roughpy.compute, which allocates its result on every
    call, so the difference is the host and allocation overhead of the API."""
    from math import factorial

    from roughpy_amd import _capi, _device
    from roughpy_amd import compute as rc
    from roughpy_amd.bases import LieBasis, TensorBasis
    w, d, batch = 5, 5, 16384
    tb, lb = TensorBasis(w, d), LieBasis(w, d)
    n, nl = tb.size(), lb.size()
    db = tb.degree_begin
    cs = tb.c_struct()
    L = _capi.lib()
    st = _device.stream_handle()
    Bd = _capi.DegreeBoundsC
    full = Bd(0, -1)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 80)
    scale = torch.empty(n, device="cuda")
    for k in range(d + 1):
        scale[db[k]:db[k + 1]] = 1.0 / factorial(k)
    res = {"workload": f"batch {batch}, width {w}, depth {d} ({n} coefficients, Lie dimension "
                       f"{nl})"}
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        code = _capi.RPY_F64 if dt == torch.float64 else _capi.RPY_F32
        A = ((torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1) * scale).to(dt)
        B = ((torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1) * scale).to(dt)
        A[:, 0] = 1
        B[:, 0] = 1
        X = B.clone()
        X[:, 0] = 0
        Lz = (torch.rand((batch, nl), generator=gen, device="cuda") * 2 - 1).to(dt)
        O = torch.empty((batch, n), dtype=dt, device="cuda")
        OL = torch.empty((batch, nl), dtype=dt, device="cuda")
        l2t = lb.get_l2t_matrix(np.dtype("float64" if name == "f64" else "float32"))
        t2l = lb.get_t2l_matrix(np.dtype("float64" if name == "f64" else "float32"))
        es = A.element_size()
        P = lambda t: t.data_ptr()  # noqa: E731
        kernels = {
            "inplace_mul": (lambda: L.rpy_dense_ft_inplace_mul(P(A), P(B), code, cs, full, full,
                                                               batch, n, n, 1, 1, st), 3 * n),
            "exp": (lambda: L.rpy_dense_ft_exp(P(O), P(X), code, cs, full, full, batch, n, n, 1,
                                               1, st), 2 * n),
            "fmexp": (lambda: L.rpy_dense_ft_fmexp(P(O), P(B), P(X), code, cs, full, full, full,
                                                   batch, n, n, n, 1, 1, 1, st), 3 * n),
            "log": (lambda: L.rpy_dense_ft_log(P(O), P(B), code, cs, full, full, batch, n, n, 1,
                                               1, st), 2 * n),
            "antipode": (lambda: L.rpy_dense_ft_antipode(P(O), P(B), code, cs, full, full, batch,
                                                         n, n, 1, 1, st), 2 * n),
            "adjoint_left_mul": (lambda: L.rpy_dense_ft_adjoint_left_mul(
                P(O), P(B), P(X), code, cs, full, full, full, batch, n, n, n, 1, 1, 1, st),
                3 * n),
            "lie_to_tensor": (lambda: L.rpy_dense_sparse_map(P(O), P(Lz), l2t.device_struct(),
                                                             code, 1.0, batch, n, nl, 1, 1, st),
                              nl + n),
            "tensor_to_lie": (lambda: L.rpy_dense_sparse_map(P(OL), P(B), t2l.device_struct(),
                                                             code, 1.0, batch, nl, n, 1, 1, st),
                              n + nl),
        }
        api = {
            "inplace_mul": lambda: rc.ft_inplace_mul(rc.FreeTensor(A, tb), rc.FreeTensor(B, tb)),
            "exp": lambda: rc.ft_exp(rc.FreeTensor(X, tb)),
            "fmexp": lambda: rc.ft_fmexp(rc.FreeTensor(B, tb), rc.FreeTensor(X, tb)),
            "log": lambda: rc.ft_log(rc.FreeTensor(B, tb)),
            "antipode": lambda: rc.antipode(rc.FreeTensor(B, tb)),
            "adjoint_left_mul": lambda: rc.ft_adjoint_left_mul(rc.FreeTensor(B, tb),
                                                               rc.ShuffleTensor(X, tb)),
            "lie_to_tensor": lambda: rc.lie_to_tensor(rc.Lie(Lz, lb), tb),
            "tensor_to_lie": lambda: rc.tensor_to_lie(rc.FreeTensor(B, tb), lb),
        }
        out = {}
        for op, (fn, vals) in kernels.items():
            def call(fn=fn):
                _capi.check(fn())
            ms = time_kernel(call, steps, warmup, torch)
            api_ms = time_kernel(api[op], steps, warmup, torch)
            gbs = float(vals) * es * batch / (ms * 1e-3) / 1e9
            out[op] = {"ms": ms, "hbm_gb_s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
                       "api_ms": api_ms}
        res[name] = out
        del A, B, X, Lz, O, OL
    torch.cuda.empty_cache()
    return res


def log_signature_leg(torch, steps, warmup):
    """Log-signatures and their gradient at width 8, depth 4 (4681 tensor
    coefficients, Lie dimension 1212), 4096 streams of length 256 with
    N(0, 1/256) increments, loss <G, logsig> with a seeded G: forward (the
    batched log-signature entry point) and backward (autograd through the
    library's kernels), float32 and float64."""
    from roughpy_amd import _capi, _device
    from roughpy_amd import autograd as ag
    from roughpy_amd.bases import LieBasis
    w, d, n, L_ = 8, 4, 4096, 256
    rng = np.random.default_rng(SEED + 32)
    inc32 = torch.from_numpy(rng.normal(0.0, np.sqrt(1.0 / L_), size=(n * L_, w))
                             .astype(np.float32)).cuda()
    offs = torch.arange(0, n * L_ + 1, L_, dtype=torch.int64, device="cuda")
    lb = LieBasis(w, d)
    t2l = lb.get_t2l_matrix(np.float64)
    G = torch.from_numpy(rng.normal(size=(n, lb.size()))).cuda()
    res = {"workload": f"log-signatures, width {w}, depth {d} (Lie dimension {lb.size()}), "
                       f"{n} streams of length {L_}"}
    lib = _capi.lib()
    for name, code, inc in (("f64", _capi.RPY_F64, inc32.double()), ("f32", _capi.RPY_F32, inc32)):
        out = torch.empty((n, lb.size()), dtype=torch.float64, device="cuda")

        def fwd():
            _capi.check(lib.rpy_stream_log_signature_batch(
                out.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, w, d, code,
                t2l.device_struct(), _device.stream_handle()))

        x = inc.clone().requires_grad_(True)

        def fwd_bwd():
            x.grad = None
            (ag.log_signature(x, offs, w, d) * G).sum().backward()

        fwd_ms = time_kernel(fwd, steps, warmup, torch)
        both_ms = time_kernel(fwd_bwd, max(1, steps // 2), 1, torch)
        res[name] = {"forward_ms": fwd_ms, "forward_backward_ms": both_ms,
                     "gradients_per_s": n / (both_ms * 1e-3)}
        del out, x
    del inc32, G
    torch.cuda.empty_cache()
    return res


def headline_backward_leg(torch, dinc32, doffs, steps, warmup):
    """Reverse-Chen backward of the headline batch (16384 streams of width 5,
    lengths U{64..4096}, depth 5): loss <G, S> with a seeded random G, the
    same four Chen-step equivalents per increment as the width-8 leg."""
    from roughpy_amd import _capi, _device
    w, d = W, D
    n = int(doffs.shape[0]) - 1
    rows = int(dinc32.shape[0])
    lib = _capi.lib()
    tsize = int(lib.rpy_tensor_size(w, d))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 41)
    G32 = torch.randn((n, tsize), generator=gen, device="cuda")
    fma_step = sum((d - m + 1) * w ** m for m in range(1, d + 1))
    res = {"workload": f"backward of the headline batch: {n} streams, width {w}, depth {d}, "
                       f"{rows} increments"}
    for name, code, inc, G in (("f32", _capi.RPY_F32, dinc32, G32),
                               ("f64", _capi.RPY_F64, dinc32.double(), G32.double())):
        S = torch.empty((n, tsize), dtype=inc.dtype, device="cuda")
        grad = torch.empty_like(inc)
        _capi.check(lib.rpy_stream_signature_batch(S.data_ptr(), inc.data_ptr(), doffs.data_ptr(),
                                                   n, w, d, code, _device.stream_handle()))

        def bwd():
            _capi.check(lib.rpy_stream_signature_vjp(grad.data_ptr(), inc.data_ptr(),
                                                     doffs.data_ptr(), n, w, d, S.data_ptr(),
                                                     G.data_ptr(), code, _device.stream_handle()))
        ms = time_kernel(bwd, steps, warmup, torch)
        tf = 2.0 * 4 * fma_step * rows / (ms * 1e-3) / 1e12
        res[name] = {"backward_ms": ms, "gradients_per_s": n / (ms * 1e-3),
                     "backward_tflops": tf, "backward_frac": tf / PEAK_TFLOPS[name]}
        del S, grad
    torch.cuda.empty_cache()
    return res


def backward_leg(torch, steps, warmup):
    """Signature forward + reverse-Chen backward, width 8, depth 4 (4681
    coefficients), 4096 streams of length 256, N(0, 1/256) increments, loss
    <G, S> with a seeded random G.  Algorithmic FLOPs of the backward: four
    Chen-step equivalents per increment (prefix recovery by exp(-x), partial
    recomputation, and the two adjoint FMAs of every forward FMA)."""
    from roughpy_amd import _capi, _device
    w, d, n, L_ = 8, 4, 4096, 256
    rng = np.random.default_rng(SEED + 30)
    inc32 = torch.from_numpy(rng.normal(0.0, np.sqrt(1.0 / L_), size=(n * L_, w))
                             .astype(np.float32)).cuda()
    offs = torch.arange(0, n * L_ + 1, L_, dtype=torch.int64, device="cuda")
    tsize = int(_capi.lib().rpy_tensor_size(w, d))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 31)
    G32 = torch.randn((n, tsize), generator=gen, device="cuda")
    fma_step = sum((d - m + 1) * w ** m for m in range(1, d + 1))
    lib = _capi.lib()
    res = {"workload": f"signature + backward, width {w}, depth {d} ({tsize} coefficients), "
                       f"{n} streams of length {L_}", "fma_per_step": fma_step}
    grads = {}
    for name, code, inc, G in (("f64", _capi.RPY_F64, inc32.double(), G32.double()),
                               ("f32", _capi.RPY_F32, inc32, G32)):
        S = torch.empty((n, tsize), dtype=inc.dtype, device="cuda")
        grad = torch.zeros_like(inc)

        def fwd():
            _capi.check(lib.rpy_stream_signature_batch(S.data_ptr(), inc.data_ptr(),
                                                       offs.data_ptr(), n, w, d, code,
                                                       _device.stream_handle()))

        def bwd():
            _capi.check(lib.rpy_stream_signature_vjp(grad.data_ptr(), inc.data_ptr(),
                                                     offs.data_ptr(), n, w, d, S.data_ptr(),
                                                     G.data_ptr(), code, _device.stream_handle()))
        fwd_ms = time_kernel(fwd, steps, warmup, torch)
        bwd_ms = time_kernel(bwd, max(1, steps // 2), 1, torch)
        grads[name] = grad.clone()
        peak = PEAK_TFLOPS[name]
        f_tf = 2.0 * fma_step * n * L_ / (fwd_ms * 1e-3) / 1e12
        b_tf = 2.0 * 4 * fma_step * n * L_ / (bwd_ms * 1e-3) / 1e12
        res[name] = {"forward_ms": fwd_ms, "backward_ms": bwd_ms,
                     "forward_tflops": f_tf, "forward_frac": f_tf / peak,
                     "backward_tflops": b_tf, "backward_frac": b_tf / peak,
                     "gradients_per_s": n / ((fwd_ms + bwd_ms) * 1e-3)}
        del S, grad
    g64, g32 = grads["f64"], grads["f32"].double()
    res["f32"]["grad_rel_err_vs_f64"] = float((g32 - g64).abs().max() / g64.abs().max())
    del grads, inc32, G32
    torch.cuda.empty_cache()
    return res


def gram_leg(torch, steps, warmup):
    """Signature kernel by signature dot: 2048 streams, width 6, length 512,
    N(0, 1/512) increments, depth-5 float32 signatures (9331 coefficients),
    then the 2048 x 2048 Gram matrix on MFMA (bf16 hi/lo split, three passes,
    upper-triangle tiles).  Error vs a float64 GEMM of the same signatures."""
    from roughpy_amd import _capi, _device
    w, d, n, L_ = 6, 5, 2048, 512
    rng = np.random.default_rng(SEED + 40)
    inc = torch.from_numpy(rng.normal(0.0, np.sqrt(1.0 / L_), size=(n * L_, w))
                           .astype(np.float32)).cuda()
    offs = torch.arange(0, n * L_ + 1, L_, dtype=torch.int64, device="cuda")
    lib = _capi.lib()
    k = int(lib.rpy_tensor_size(w, d))
    S = torch.empty((n, k), dtype=torch.float32, device="cuda")
    Gm = torch.empty((n, n), dtype=torch.float32, device="cuda")

    def sig():
        _capi.check(lib.rpy_stream_signature_batch(S.data_ptr(), inc.data_ptr(), offs.data_ptr(),
                                                   n, w, d, _capi.RPY_F32,
                                                   _device.stream_handle()))

    def gram():
        _capi.check(lib.rpy_signature_gram(Gm.data_ptr(), S.data_ptr(), n, k, k, n,
                                           _device.stream_handle()))
    sig_ms = time_kernel(sig, steps, warmup, torch)
    gram_ms = time_kernel(gram, steps, warmup, torch)
    S64 = S.double()
    ref = S64 @ S64.T
    dg = torch.sqrt(torch.outer(torch.diagonal(ref), torch.diagonal(ref)))
    rel = float(((Gm.double() - ref).abs() / dg).max())
    fma_step = sum((d - m + 1) * w ** m for m in range(1, d + 1))
    sig_tf = 2.0 * fma_step * n * L_ / (sig_ms * 1e-3) / 1e12
    # algorithmic work is one n x n x k GEMM; the three bf16 passes and the
    # upper-triangle tiling are the kernel's own business
    gemm_flops = 2.0 * n * n * k
    gram_tf = gemm_flops / (gram_ms * 1e-3) / 1e12
    res = {"workload": f"{n} signatures, width {w}, depth {d} ({k} coefficients), length {L_}, "
                       f"then the {n} x {n} Gram matrix",
           "signature_ms": sig_ms, "signature_tflops": sig_tf,
           "signature_frac": sig_tf / PEAK_TFLOPS["f32"],
           "gram_ms": gram_ms, "gram_bf16_tflops": gram_tf,
           "gram_frac": gram_tf / PEAK_BF16_TFLOPS,
           "gram_max_rel_err_vs_f64_gemm": rel}
    del S, Gm, S64, ref, inc
    torch.cuda.empty_cache()
    return res


def shuffle_leg(torch, steps, warmup):
    """Shuffle product C += A (shuffle) B, width 4, depth 6 (5461 coefficients),
    batch 16384, uniform[-1, 1] coefficients.  Algorithmic work: every output
    word of degree n sums 2^n products, 2 * sum_n (2W)^n flop per element;
    traffic read A, B, C and write C."""
    from roughpy_amd import _capi, _device
    from roughpy_amd.bases import TensorBasis
    w, d, batch = 4, 6, 16384
    tb = TensorBasis(w, d)
    n = tb.size()
    cs = tb.c_struct()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED + 60)
    A32 = torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1
    B32 = torch.rand((batch, n), generator=gen, device="cuda") * 2 - 1
    terms = sum((2 * w) ** k for k in range(d + 1))
    res = {"workload": f"C += A (shuffle) B, width {w}, depth {d} ({n} coefficients), "
                       f"batch {batch}", "products_per_element": terms}
    lib = _capi.lib()
    for name, code, A, B in (("f64", _capi.RPY_F64, A32.double(), B32.double()),
                             ("f32", _capi.RPY_F32, A32, B32)):
        C = torch.zeros_like(A)

        def fn():
            _capi.check(lib.rpy_dense_st_fma(C.data_ptr(), A.data_ptr(), B.data_ptr(), code, cs,
                                             batch, n, n, n, 1, 1, 1, _device.stream_handle()))
        ms = time_kernel(fn, steps, warmup, torch)
        tf = 2.0 * terms * batch / (ms * 1e-3) / 1e12
        gbs = 4.0 * n * A.element_size() * batch / (ms * 1e-3) / 1e9

        def adj():
            _capi.check(lib.rpy_dense_st_adj_mul(C.data_ptr(), A.data_ptr(), B.data_ptr(), code,
                                                 cs, batch, n, n, n, 1, 1, 1,
                                                 _device.stream_handle()))
        ms_adj = time_kernel(adj, max(1, steps // 3), 1, torch)
        res[name] = {"ms_per_step": ms, "tflops": tf, "valu_frac": tf / PEAK_TFLOPS[name],
                     "hbm_gb_s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
                     "adjoint_ms_per_step": ms_adj,
                     "adjoint_tflops": 2.0 * terms * batch / (ms_adj * 1e-3) / 1e12}
        del C
    del A32, B32
    torch.cuda.empty_cache()
    return res


def simplify_leg(torch, steps):
    """Stream.simplify of one stream of 2^20 unit-spaced float32 increments
    (width 5, depth 5, N(0, 1/64) entries) over a partition into 16384 equal
    pieces of 64 increments: one native row selection, one gather on the GPU
    and one log-signature launch (Chen steps, log, Hall projection).  Reports
    the end-to-end rate and the log-signature launch alone, whose arithmetic
    is the headline's Chen-step count per increment."""
    import time

    from roughpy_amd import _capi, _device
    from roughpy_amd import streams as rs
    from roughpy_amd.bases import LieBasis
    n, pieces = 1 << 20, 16384
    rng = np.random.default_rng(SEED + 70)
    inc = (rng.normal(size=(n, W)) / 8.0).astype(np.float32)
    st = rs.LieIncrementStream.from_increments(inc, depth=D)
    part = rs.Partition(rs.RealInterval(0.0, float(n)),
                        np.arange(1, pieces, dtype=np.float64) * (n // pieces))
    st.simplify(part, 0)  # warm-up: device rows, bases, maps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st.simplify(part, 0)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    mids = part.intermediates()
    idx, offs = st._select_batch(np.array([0.0] + mids), np.array(mids + [float(n)]), 0)
    rows = st._dev_rows.index_select(0, torch.from_numpy(idx).cuda())
    doffs = torch.from_numpy(offs).cuda()
    lb = LieBasis(W, D)
    t2l = lb.get_t2l_matrix(np.float64)
    out = torch.empty((pieces, lb.size()), dtype=torch.float64, device="cuda")
    lib = _capi.lib()

    def fn():
        _capi.check(lib.rpy_stream_log_signature_batch(
            out.data_ptr(), rows.data_ptr(), doffs.data_ptr(), pieces, W, D, _capi.RPY_F32,
            t2l.device_struct(), _device.stream_handle()))
    ms_k = time_kernel(fn, steps, 1, torch)
    tf = 2.0 * FMA_PER_STEP * n / (ms_k * 1e-3) / 1e12
    return {"workload": f"simplify a stream of {n} increments (width {W}, depth {D}, float32) "
                        f"into {pieces} pieces of {n // pieces}",
            "ms_per_call": ms, "pieces_per_s": pieces / (ms * 1e-3),
            "log_signature_launch_ms": ms_k, "launch_share": ms_k / ms,
            "launch_tflops": tf, "launch_valu_frac": tf / PEAK_TFLOPS["f32"]}


def serialize_leg(torch, inc, offs, lens, n_streams=256, levels=6, resolution=12):
    """A dyadic cache of the first 256 streams of the batch (stored levels
    0..6, float32) written to and read back from the binary cache format in
    host memory: dump = device-to-host copy plus framing, load = parsing plus
    host-to-device copy, set beside rebuilding the same cache on the GPU."""
    from roughpy_amd import serialization as ser
    from roughpy_amd import streams as rs
    rows = int(offs[n_streams])
    params = np.concatenate([np.arange(L, dtype=np.float64) / L for L in lens[:n_streams]])
    cache = rs.DyadicSignatureCache(inc[:rows], params, offs[:n_streams + 1], W, D,
                                    resolution=resolution, levels=levels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cache.build()
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    blob = ser.dumps(cache)
    dump_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    back = ser.loads(blob)
    torch.cuda.synchronize()
    load_ms = (time.perf_counter() - t0) * 1e3
    same = bool(torch.equal(back.cache, cache.cache))
    gb = len(blob) / 1e9
    del back, cache
    torch.cuda.empty_cache()
    return {"workload": f"dyadic cache of {n_streams} streams, width {W}, depth {D}, levels "
                        f"0..{levels}, float32, in host memory",
            "bytes": len(blob), "dump_ms": dump_ms, "dump_gb_s": gb / (dump_ms * 1e-3),
            "load_ms": load_ms, "load_gb_s": gb / (load_ms * 1e-3), "rebuild_ms": build_ms,
            "bit_identical": same}


def graph_leg(torch, reps=200):
    """A launch-bound multi-launch call: float64 log-signatures of 64 short
    streams (width 5, depth 5, lengths 2..16) from entry points that
    allocate nothing -- length keys and the persistent signature launch
    (rpy_stream_signature_batch_noalloc), the tensor log and the Hall
    projection on preallocated buffers -- timed as eager calls and as
    replays of one HIP graph captured once."""
    from roughpy_amd import _capi, _device
    from roughpy_amd.bases import LieBasis, TensorBasis
    w, d, n = 5, 5, 64
    rng = np.random.default_rng(SEED + 50)
    lens = rng.integers(2, 17, size=n)
    offs = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)).cuda()
    inc = torch.from_numpy(rng.normal(size=(int(lens.sum()), w)) * 0.1).cuda()
    tb, lb = TensorBasis(w, d), LieBasis(w, d)
    t2l = lb.get_t2l_matrix(np.float64)
    tsize = tb.size()
    sig = torch.empty((n, tsize), dtype=torch.float64, device="cuda")
    lg = torch.empty((n, tsize), dtype=torch.float64, device="cuda")
    out = torch.empty((n, lb.size()), dtype=torch.float64, device="cuda")
    lib = _capi.lib()
    cs = tb.c_struct()
    full = _capi.DegreeBoundsC(0, -1)

    def run():
        st = _device.stream_handle()
        _capi.check(lib.rpy_stream_signature_batch_noalloc(
            sig.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, w, d, _capi.RPY_F64, st))
        _capi.check(lib.rpy_dense_ft_log(lg.data_ptr(), sig.data_ptr(), _capi.RPY_F64, cs, full,
                                         full, n, tsize, tsize, 1, 1, st))
        out.zero_()
        _capi.check(lib.rpy_dense_sparse_map(out.data_ptr(), lg.data_ptr(), t2l.device_struct(),
                                             _capi.RPY_F64, 1.0, n, lb.size(), tsize, 1, 1, st))
    eager_ms = time_kernel(run, reps, 5, torch)
    ref = out.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    graph_ms = time_kernel(g.replay, reps, 5, torch)
    same = bool(torch.equal(out, ref))
    return {"workload": f"float64 log-signatures of {n} streams, width {w}, depth {d}, lengths "
                        "2..16: length keys, signature, log, clearing and Hall projection launches",
            "eager_us_per_call": eager_ms * 1e3, "graph_us_per_replay": graph_ms * 1e3,
            "replay_equals_eager": same}


def ref_cpu_rates(w, d, ops, target_seconds=0.4):
    """The reference's own roughpy_compute kernels (oracle/_ref/libref.so,
    compiled from the reference headers) on one host core, float64, over a
    bounded sample of tensors of the leg's shape: elements per second for
    each named operation.  ops: {label: (harness function, operand count)}
    where the harness function takes (out, operands..., width, depth, n)."""
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    size = O.tensor_size(w, d)
    rng = np.random.default_rng(SEED + 90)
    out = {}
    for label, (name, n_in) in ops.items():
        f = getattr(R, f"{name}_f64")

        def run(n):
            ins = [np.ascontiguousarray(rng.uniform(-1, 1, size=(n, size)) / 4)
                   for _ in range(n_in)]
            for a in ins:
                a[:, 0] = 1.0
            o = np.zeros((n, size))
            t0 = time.perf_counter()
            f(o.ctypes.data, *[a.ctypes.data for a in ins], w, d, n)
            return time.perf_counter() - t0

        dt = run(2)
        n = int(max(2, min(4096, target_seconds / max(dt / 2, 1e-9))))
        dt = run(n)
        out[label] = {"value": n / dt, "unit": "elements/s", "cores": 1, "kind": "reference",
                      "sample": f"{n} float64 tensors, {dt:.2f} s"}
    return out


def host_cpu():
    """CPU model and the number of logical cores this process may use."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    try:  # lscpu may print only the family; /proc/cpuinfo names the part
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                full = line.split(":", 1)[1].strip()
                if full and (model is None or len(full) > len(model)):
                    model = full
                break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline(inc, offs, target_seconds=8.0):
    """Reference CPU path on bounded samples of the same workload: the
    reference's own roughpy_compute ft_fmexp per increment (oracle/_ref/
    libref.so, built from the reference headers) -- on one thread, as
    roughpy.compute runs (it holds the GIL), and with one stream per thread
    on every host core this process may use, as GIL-free Stream.signature
    calls would.  About target_seconds of wall time each."""
    from oracle import oracle as O
    use_ref = O.ref_lib() is not None
    model, cores = host_cpu()
    rng = np.random.default_rng(SEED + 1)

    def run(sample, threads):
        lens = offs[1:][sample] - offs[:-1][sample]
        sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        rows = np.concatenate([inc[offs[s]:offs[s + 1]] for s in sample])
        t0 = time.perf_counter()
        O.signature_batch(rows, sub_off, W, D, nthreads=threads, use_ref=use_ref)
        return time.perf_counter() - t0

    def rate(threads):
        calib = rng.choice(N_STREAMS, size=max(4, 2 * threads), replace=False)
        per_stream = run(calib, threads) / len(calib)
        n = int(min(N_STREAMS, max(len(calib), target_seconds / max(per_stream, 1e-9))))
        sample = rng.choice(N_STREAMS, size=n, replace=False)
        dt = run(sample, threads)
        return n / dt, n, dt

    v1, n1, t1 = rate(1)
    vn, nn, tn = rate(cores)
    kind = "reference" if use_ref else "port"
    return {
        "value": vn,
        "unit": "signatures/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{nn} of the {N_STREAMS} streams (same length distribution), float32 data, "
                  f"{tn:.1f} s wall on {cores} threads (one stream per thread)",
        "cpu_model": model,
        "single_thread": {"value": v1, "unit": "signatures/s", "cores": 1, "kind": kind,
                          "sample": f"{n1} streams, {t1:.1f} s on one thread"},
    }


def compensated_arithmetic():
    """What the compensated mode of the headline shape runs."""
    from roughpy_amd import tune
    from roughpy_amd import _capi
    v, _ = tune.config(_capi.RPY_OP_SIGNATURE, W, D, _capi.RPY_F32_COMPENSATED)
    if v[0] == _capi.RPY_KIND_BLOCK_MERGE:
        return (f"float32 segments of {v[2]} rows (at most 16 per stream) on the compensated "
                "lanes, merged per stream with Kahan sums")
    if v[0] in (_capi.RPY_KIND_LANE_CHUNK, _capi.RPY_KIND_STREAM_GROUP):
        return (f"float32 Chen steps ({tune.describe(v)}), signature levels 1 and 2 "
                "accumulated with Kahan's compensated summation")
    return "float64 arithmetic on float32 data, each coefficient rounded once"


def block_merge_leg(torch, dinc32, doffs, s64, steps, warmup):
    """The compensated mode's block-merge variant: segments of T rows on the
    compensated lanes, merged per stream with Kahan sums."""
    from roughpy_amd import _capi, _device, tune
    L = _capi.lib()
    n = int(doffs.shape[0]) - 1
    out = torch.empty((n, int(L.rpy_tensor_size(W, D))), dtype=torch.float32, device="cuda")
    res = {}
    for T in (256, 512):
        v = (_capi.RPY_KIND_BLOCK_MERGE, 0, T, 0)
        tune.select(_capi.RPY_OP_SIGNATURE, W, D, _capi.RPY_F32_COMPENSATED, v)
        try:
            def fn():
                _capi.check(L.rpy_stream_signature_batch(out.data_ptr(), dinc32.data_ptr(),
                                                         doffs.data_ptr(), n, W, D,
                                                         _capi.RPY_F32_COMPENSATED,
                                                         _device.stream_handle()))
            ms = time_kernel(fn, steps, warmup, torch)
        finally:
            tune.select(_capi.RPY_OP_SIGNATURE, W, D, _capi.RPY_F32_COMPENSATED, None)
        res[f"segments_{T}"] = {"ms_per_step": ms, "value": n / (ms * 1e-3),
                                "rel_err_vs_f64": level_rel_err(out, s64, W, D)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs",
                    default="dyadic,hbm,fma,dense,backward,logsig,gram,shuffle,simplify,"
                            "serialize,graph",
                    help="comma-separated secondary workloads to run after the headline "
                         "(dyadic, hbm, fma, dense, backward, logsig, gram, shuffle, simplify, "
                         "serialize, graph; empty for none)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device-to-device copy, the reference "
                         "dispatch for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the streams shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    import torch

    from roughpy_amd import _capi
    from roughpy_amd import _device

    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    inc, offs, lens = make_workload()
    total_incr = int(offs[-1])
    dinc32 = torch.from_numpy(inc).cuda()
    dinc64 = dinc32.double()
    doffs = torch.from_numpy(offs).cuda()
    tsize = int(_capi.lib().rpy_tensor_size(W, D))
    L = _capi.lib()

    def make_fn(dtype):
        out = torch.empty((N_STREAMS, tsize),
                          dtype=torch.float64 if dtype == "f64" else torch.float32, device="cuda")
        src = dinc64 if dtype == "f64" else dinc32
        code = {"f32": _capi.RPY_F32, "f64": _capi.RPY_F64,
                "f32c": _capi.RPY_F32_COMPENSATED}[dtype]

        def fn():
            _capi.check(L.rpy_stream_signature_batch(out.data_ptr(), src.data_ptr(),
                                                     doffs.data_ptr(), N_STREAMS, W, D, code,
                                                     _device.stream_handle()))
        return fn, out

    results = {}
    for dt in ("f32", "f64", "f32c"):
        fn, out = make_fn(dt)
        ms = time_kernel(fn, args.steps if dt == args.dtype else max(2, args.steps // 2),
                         args.warmup, torch)
        results[dt] = (ms, out)
        progress(f"signatures {dt}: {ms:.3f} ms")

    ms, out = results[args.dtype]
    flops = 2.0 * FMA_PER_STEP * total_incr
    achieved = flops / (ms * 1e-3) / 1e12
    s32 = results["f32"][1]
    s64 = results["f64"][1]
    # per-level normwise relative error of the float32 kernel vs the float64 kernel
    err = level_rel_err(s32, s64, W, D)

    line = {
        "metric": "signatures/sec, w=5 N=5 batched; % VALU/HBM roofline; fp32 vs fp64 rel-err",
        "value": N_STREAMS / (ms * 1e-3),
        "unit": "signatures/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded N(0,1/L) increments, L ~ U{64..4096})",
        "config": {
            "workload": "batched signatures of 16384 uneven streams, width 5, depth 5 (3906 "
                        "coefficients), one persistent launch",
            "n_streams": N_STREAMS, "width": W, "depth": D, "lengths": [LMIN, LMAX],
            "total_increments": total_incr,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_TFLOPS[args.dtype],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS[args.dtype],
            "traffic": profiled_traffic(args.dtype),
        },
        "f64_kernel": {"ms_per_step": results["f64"][0],
                       "value": N_STREAMS / (results["f64"][0] * 1e-3),
                       "tflops": flops / (results["f64"][0] * 1e-3) / 1e12},
        "f32_kernel": {"ms_per_step": results["f32"][0],
                       "value": N_STREAMS / (results["f32"][0] * 1e-3),
                       "tflops": flops / (results["f32"][0] * 1e-3) / 1e12},
        "rel_err_f32_vs_f64": err,
        "f32_compensated_kernel": {
            "ms_per_step": results["f32c"][0],
            "value": N_STREAMS / (results["f32c"][0] * 1e-3),
            "speedup_vs_f64": results["f64"][0] / results["f32c"][0],
            "arithmetic": compensated_arithmetic(),
            "rel_err_vs_f64": level_rel_err(results["f32c"][1], s64, W, D)},
    }
    line["f32_compensated_kernel"]["block_merge"] = block_merge_leg(
        torch, dinc32, doffs, s64, max(2, args.steps // 2), 1)
    line["f32_compensated_kernel"]["block_merge"]["speedup_vs_f64"] = {
        k: results["f64"][0] / v["ms_per_step"]
        for k, v in line["f32_compensated_kernel"]["block_merge"].items()}
    progress("compensated block merge done")
    legs = [x for x in args.legs.split(",") if x]
    if "dyadic" in legs:
        line["dyadic_cache"] = dyadic_cache_leg(torch, inc, offs, lens, dinc32, doffs)
        progress("dyadic_cache done")
    if "backward" in legs:
        del dinc64, s64
        torch.cuda.empty_cache()
        line["backward_w5_d5"] = headline_backward_leg(torch, dinc32, doffs, max(2, args.steps // 4), 1)
        progress("backward_w5_d5 done")
    del results, s32, out
    if "backward" not in legs:
        del dinc64, s64
    torch.cuda.empty_cache()
    if "hbm" in legs:
        line["hbm_copy"] = hbm_leg(torch, max(3, args.steps // 2), 1)
        progress("hbm_copy done")
    if "fma" in legs:
        line["fma_w4_d6"] = fma_leg(torch, max(3, args.steps // 2), 1)
        progress("fma_w4_d6 done")
    if "dense" in legs:
        line["dense_w5_d5"] = dense_leg(torch, max(3, args.steps // 2), 1)
        progress("dense_w5_d5 done")
    if "backward" in legs:
        line["backward_w8_d4"] = backward_leg(torch, max(3, args.steps // 2), 1)
        progress("backward_w8_d4 done")
    if "logsig" in legs:
        line["log_signature_w8_d4"] = log_signature_leg(torch, max(3, args.steps // 2), 1)
        progress("log_signature_w8_d4 done")
    if "gram" in legs:
        line["gram_w6_d5"] = gram_leg(torch, max(3, args.steps // 2), 1)
        progress("gram_w6_d5 done")
    if "shuffle" in legs:
        line["shuffle_w4_d6"] = shuffle_leg(torch, max(3, args.steps // 2), 1)
        progress("shuffle_w4_d6 done")
    if "simplify" in legs:
        line["simplify_w5_d5"] = simplify_leg(torch, max(3, args.steps // 2))
        progress("simplify_w5_d5 done")
    if "serialize" in legs:
        line["serialize_cache"] = serialize_leg(torch, inc, offs, lens)
        progress("serialize_cache done")
    if "graph" in legs:
        line["graph_replay"] = graph_leg(torch)
        progress("graph_replay done")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(inc, offs)
        progress("cpu_baseline done")
        # the reference CPU kernels beside the secondary legs (one core, float64)
        if "fma_w4_d6" in line:
            line["fma_w4_d6"]["cpu_reference"] = ref_cpu_rates(
                4, 6, {"ft_fma": ("ref_ft_fma", 2)})
        if "dense_w5_d5" in line:
            line["dense_w5_d5"]["cpu_reference"] = ref_cpu_rates(5, 5, {
                "inplace_mul": ("ref_ft_inplace_mul", 1), "exp": ("ref_ft_exp", 1),
                "fmexp": ("ref_ft_fmexp", 2), "log": ("ref_ft_log", 1),
                "antipode": ("ref_antipode", 1), "adjoint_left_mul": ("ref_adj_lmul", 2)})
        if "shuffle_w4_d6" in line:
            line["shuffle_w4_d6"]["cpu_reference"] = ref_cpu_rates(4, 6, {
                "st_fma": ("ref_st_fma", 2), "st_adj_mul": ("ref_st_adj_mul", 2)})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
