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


PROFILE = os.path.join(ROOT, "profiles", "chen_signature_w5_d5_16384streams.md")


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
    res = {"levels_stored": levels, "resolution": resolution, "dtype": "f32",
           "cache_bytes": int(cache.cache.numel()) * 4, "build_ms": rebuild_ms,
           "first_build_ms": build_ms, "queries": n_queries, "query_ms": q_ms,
           "queries_per_s": n_queries / (q_ms * 1e-3)}
    del cache, out
    torch.cuda.empty_cache()
    return res


def cpu_baseline(inc, offs, target_seconds=10.0):
    """Reference CPU path on a bounded sample: the reference's own
    roughpy_compute ft_fmexp per increment (oracle/_ref/libref.so, built from
    the reference headers), one stream per thread across the host cores."""
    from oracle import oracle as O
    use_ref = O.ref_lib() is not None
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(SEED + 1)

    def run(sample):
        lens = offs[1:][sample] - offs[:-1][sample]
        sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        rows = np.concatenate([inc[offs[s]:offs[s + 1]] for s in sample])
        t0 = time.perf_counter()
        O.signature_batch(rows, sub_off, W, D, nthreads=threads, use_ref=use_ref)
        return time.perf_counter() - t0

    calib = rng.choice(N_STREAMS, size=2 * threads, replace=False)
    dt = run(calib)
    per_stream = dt / len(calib)
    n = int(min(N_STREAMS, max(len(calib), target_seconds / max(per_stream, 1e-9))))
    sample = rng.choice(N_STREAMS, size=n, replace=False)
    dt = run(sample)
    return {
        "value": n / dt,
        "unit": "signatures/s",
        "cores": threads,
        "kind": "reference" if use_ref else "port",
        "sample": f"{n} of the {N_STREAMS} streams (same length distribution), float32 data, "
                  f"{dt:.1f} s wall on {threads} threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device-to-device copy, the reference "
                         "dispatch for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the streams shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
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
        out = torch.empty((N_STREAMS, tsize), dtype=torch.float32 if dtype == "f32" else torch.float64,
                          device="cuda")
        src = dinc32 if dtype == "f32" else dinc64
        code = _capi.RPY_F32 if dtype == "f32" else _capi.RPY_F64

        def fn():
            _capi.check(L.rpy_stream_signature_batch(out.data_ptr(), src.data_ptr(),
                                                     doffs.data_ptr(), N_STREAMS, W, D, code,
                                                     _device.stream_handle()))
        return fn, out

    results = {}
    for dt in ("f32", "f64"):
        fn, out = make_fn(dt)
        ms = time_kernel(fn, args.steps if dt == args.dtype else max(2, args.steps // 2),
                         args.warmup, torch)
        results[dt] = (ms, out)

    ms, out = results[args.dtype]
    flops = 2.0 * FMA_PER_STEP * total_incr
    achieved = flops / (ms * 1e-3) / 1e12
    s32 = results["f32"][1].double()
    s64 = results["f64"][1]
    # per-level normwise relative error of the float32 kernel vs the float64 kernel
    from oracle import oracle as O
    db = O.tensor_degree_begin(W, D)
    err = 0.0
    for k in range(D + 1):
        a = s32[:, db[k]:db[k + 1]]
        b = s64[:, db[k]:db[k + 1]]
        num = (a - b).abs().amax(dim=1)
        den = b.abs().amax(dim=1).clamp_min(1e-300)
        err = max(err, float((num / den).max()))

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
    }
    line["dyadic_cache"] = dyadic_cache_leg(torch, inc, offs, lens, dinc32, doffs)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(inc, offs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
