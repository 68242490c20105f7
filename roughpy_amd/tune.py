"""Autotuner of the signature kernels: times every compiled variant of the
signature forward and backward on the tuned shapes and writes the table the
library's dispatch compiles in.

    python -m roughpy_amd.tune [--ops sig,vjp] [--shapes 5x5,8x4,6x5,4x6]
                               [--write [PATH]] [--report PATH]

For each (op, width, depth, dtype) the variants listed by rpy_tune_variants
(kernel kind, prefix length P, rows per LDS block, waves per EU) run on a
seeded synthetic batch of that shape's workload through the C-ABI with the
variant forced by rpy_tune_select; each time is the best of three means of
--steps calls.  A variant counts only if its result agrees with the current
float64 variant's result: within 1e-12 (float64) or 1e-5 (float32,
compensated float32), max over degrees of max|x - ref| / max|ref|, or, for
the float32 modes, within twice the current variant's own error (the float32
rounding of a long stream can exceed 1e-5 whatever the variant); backward:
relative to the largest gradient entry, 1e-10 and 1e-4.  The fastest
agreeing variant becomes the shape's row in roughpy_amd/csrc/
signature_tuned.inc (--write, or --write PATH) when it beats the variant
the library runs now (its tuned row, else the built-in default) by more
than 3%, else that variant's row is written; rebuild the library
(python -m roughpy_amd.build) to compile it in.  --report writes every
timing as a markdown table.

Workloads (increments N(0, 1/L) per channel, seeded):
  5x5  16384 streams, lengths uniform in [64, 4096]
  8x4  4096 streams of length 256
  6x5  2048 streams of length 512
  4x6  4096 streams, lengths uniform in [64, 1024]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _capi, _device

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "csrc", "signature_tuned.inc")

WORKLOADS = {
    (5, 5): dict(n=16384, lmin=64, lmax=4096),
    (8, 4): dict(n=4096, lmin=256, lmax=256),
    (6, 5): dict(n=2048, lmin=512, lmax=512),
    (4, 6): dict(n=4096, lmin=64, lmax=1024),
}
DTYPES = {"f32": _capi.RPY_F32, "f64": _capi.RPY_F64, "f32c": _capi.RPY_F32_COMPENSATED}
DTYPE_NAMES = {v: k for k, v in DTYPES.items()}
DTYPE_ENUM = {_capi.RPY_F32: "RPY_F32", _capi.RPY_F64: "RPY_F64",
              _capi.RPY_F32_COMPENSATED: "RPY_F32_COMPENSATED"}
KIND_NAMES = {0: "generic", 1: "lane-chunk", 2: "stream-group", 3: "wave", 4: "block-merge",
              5: "stream-table"}
SIG_TOL = {_capi.RPY_F64: 1e-12, _capi.RPY_F32: 1e-5, _capi.RPY_F32_COMPENSATED: 1e-5}
VJP_TOL = {_capi.RPY_F64: 1e-10, _capi.RPY_F32: 1e-4}


def variants(op, w, d, dtype):
    L = _capi.lib()
    n = L.rpy_tune_variants(op, w, d, dtype, None, 0)
    arr = (_capi.KernelVariantC * n)()
    L.rpy_tune_variants(op, w, d, dtype, arr, n)
    return [a.astuple() for a in arr]


def select(op, w, d, dtype, v):
    """Forces variant v (a 4-tuple) for the shape, or clears with None."""
    L = _capi.lib()
    if v is None:
        _capi.check(L.rpy_tune_select(op, w, d, dtype, None))
    else:
        c = _capi.KernelVariantC(*v)
        _capi.check(L.rpy_tune_select(op, w, d, dtype, ctypes.byref(c)))


def config(op, w, d, dtype):
    """(variant tuple, source) of the shape: source 0 default, 1 tuned, 2 selected."""
    L = _capi.lib()
    c = _capi.KernelVariantC()
    src = ctypes.c_int32()
    _capi.check(L.rpy_tune_config(op, w, d, dtype, ctypes.byref(c), ctypes.byref(src)))
    return c.astuple(), src.value


def describe(v):
    kind, p, blk, wpe = v
    s = f"{KIND_NAMES.get(kind, kind)}"
    if kind:
        s += f" P={p}"
    if blk:
        s += f" block={blk}"
    if wpe:
        s += f" waves/EU={wpe}"
    return s


def workload(w, d, seed=2024):
    spec = WORKLOADS[(w, d)]
    rng = np.random.default_rng(seed + 100 * w + d)
    lens = rng.integers(spec["lmin"], spec["lmax"] + 1, size=spec["n"])
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = rng.normal(size=(int(offs[-1]), w)) / np.sqrt(np.repeat(lens, lens))[:, None]
    return inc, offs


def level_err(x, ref, w, d):
    db = [sum(w ** j for j in range(k)) for k in range(d + 2)]
    err = 0.0
    for k in range(d + 1):
        a = x[:, db[k]:db[k + 1]].double()
        b = ref[:, db[k]:db[k + 1]].double()
        num = (a - b).abs().amax(dim=1)
        den = b.abs().amax(dim=1).clamp_min(1e-300)
        err = max(err, float((num / den).max()))
    return err


def time_fn(torch, fn, steps, warmup, budget_ms=None):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(torch.cuda.current_stream())
    fn()
    stop.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    first = start.elapsed_time(stop)
    if budget_ms is not None and first > budget_ms:
        return first  # far slower than the best so far: one run is enough
    best = first
    for _ in range(3):
        start.record(torch.cuda.current_stream())
        for _ in range(steps):
            fn()
        stop.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        best = min(best, start.elapsed_time(stop) / steps)
    return best


def tune_signature(torch, w, d, dtypes, steps, warmup, log):
    L = _capi.lib()
    inc, offs = workload(w, d)
    n = len(offs) - 1
    tsize = int(L.rpy_tensor_size(w, d))
    d_offs = torch.from_numpy(offs).cuda()
    d64 = torch.from_numpy(inc).cuda()
    d32 = d64.float()
    op = _capi.RPY_OP_SIGNATURE
    # float64 reference: the current float64 variant
    select(op, w, d, _capi.RPY_F64, None)
    ref = torch.empty((n, tsize), dtype=torch.float64, device="cuda")
    _capi.check(L.rpy_stream_signature_batch(ref.data_ptr(), d64.data_ptr(), d_offs.data_ptr(), n,
                                             w, d, _capi.RPY_F64, _device.stream_handle()))
    rows = []
    for dt in dtypes:
        src = d64 if dt == _capi.RPY_F64 else d32
        out = torch.empty((n, tsize), dtype=src.dtype, device="cuda")

        def run():
            _capi.check(L.rpy_stream_signature_batch(out.data_ptr(), src.data_ptr(),
                                                     d_offs.data_ptr(), n, w, d, dt,
                                                     _device.stream_handle()))
        best = None
        # the current variant's own error bounds the others' in the float32 modes
        select(op, w, d, dt, None)
        out.fill_(float("nan"))
        run()
        tol = SIG_TOL[dt]
        if dt != _capi.RPY_F64:
            tol = max(tol, 2.0 * level_err(out, ref, w, d))
        for v in variants(op, w, d, dt):
            select(op, w, d, dt, v)
            out.fill_(float("nan"))
            ms = time_fn(torch, run, steps, warmup,
                         budget_ms=None if best is None else 10 * best)
            err = level_err(out, ref, w, d)
            ok = err <= tol
            rows.append(dict(op="signature", width=w, depth=d, dtype=DTYPE_NAMES[dt],
                             variant=list(v), name=describe(v), ms=ms, err=err, ok=ok))
            log(f"signature {w}x{d} {DTYPE_NAMES[dt]:4s} {describe(v):32s} {ms:9.3f} ms "
                f"err {err:.2e}{'' if ok else '  REJECTED'}")
            if ok and (best is None or ms < best):
                best = ms
        select(op, w, d, dt, None)
        del out
    del ref, d64, d32
    torch.cuda.empty_cache()
    return rows


def tune_vjp(torch, w, d, dtypes, steps, warmup, log):
    L = _capi.lib()
    spec = WORKLOADS[(w, d)]
    n, length = min(spec["n"], 4096), 256
    rng = np.random.default_rng(7000 + 10 * w + d)
    inc = rng.normal(size=(n * length, w)) / np.sqrt(length)
    offs = np.arange(0, n * length + 1, length, dtype=np.int64)
    tsize = int(L.rpy_tensor_size(w, d))
    G = rng.normal(size=(n, tsize))
    d_offs = torch.from_numpy(offs).cuda()
    op = _capi.RPY_OP_SIGNATURE_VJP
    ref = None
    rows = []
    for dt in sorted(dtypes, key=lambda t: t != _capi.RPY_F64):  # float64 first
        tdt = torch.float64 if dt == _capi.RPY_F64 else torch.float32
        x = torch.from_numpy(inc).to(device="cuda", dtype=tdt)
        g = torch.from_numpy(G).to(device="cuda", dtype=tdt)
        S = torch.empty((n, tsize), dtype=tdt, device="cuda")
        _capi.check(L.rpy_stream_signature_batch(S.data_ptr(), x.data_ptr(), d_offs.data_ptr(),
                                                 n, w, d, dt, _device.stream_handle()))
        grad = torch.empty_like(x)

        def run():
            _capi.check(L.rpy_stream_signature_vjp(grad.data_ptr(), x.data_ptr(),
                                                   d_offs.data_ptr(), n, w, d, S.data_ptr(),
                                                   g.data_ptr(), dt, _device.stream_handle()))
        if ref is None:
            select(op, w, d, dt, None)
            run()
            ref = grad.double().clone()
        best = None
        for v in variants(op, w, d, dt):
            select(op, w, d, dt, v)
            grad.fill_(float("nan"))
            ms = time_fn(torch, run, steps, warmup,
                         budget_ms=None if best is None else 10 * best)
            err = float((grad.double() - ref).abs().max() / ref.abs().max())
            ok = err <= VJP_TOL[dt]
            rows.append(dict(op="signature_vjp", width=w, depth=d, dtype=DTYPE_NAMES[dt],
                             variant=list(v), name=describe(v), ms=ms, err=err, ok=ok))
            log(f"vjp       {w}x{d} {DTYPE_NAMES[dt]:4s} {describe(v):32s} {ms:9.3f} ms "
                f"err {err:.2e}{'' if ok else '  REJECTED'}")
            if ok and (best is None or ms < best):
                best = ms
        select(op, w, d, dt, None)
        del x, g, S, grad
    torch.cuda.empty_cache()
    return rows


def winners(rows, margin=0.03):
    """Fastest agreeing variant per shape; the current choice stays unless
    another variant is more than `margin` faster (timing noise is a few
    percent)."""
    best, cur = {}, {}
    for r in rows:
        if not r["ok"]:
            continue
        key = (r["op"], r["width"], r["depth"], r["dtype"])
        if r.get("current"):
            cur[key] = r
        if key not in best or r["ms"] < best[key]["ms"]:
            best[key] = r
    for key, r in cur.items():
        if best[key]["ms"] > (1.0 - margin) * r["ms"]:
            best[key] = r
    return best


def table_text(best, rows):
    ops = {"signature": "RPY_OP_SIGNATURE", "signature_vjp": "RPY_OP_SIGNATURE_VJP"}
    lines = ["// Tuned kernel variants: {op, width, depth, dtype, {kind, prefix, block, "
             "waves_per_eu}}.",
             "// Written by `python -m roughpy_amd.tune --write` from timings on MI355X; "
             "each row's",
             "// comment gives the winner's time and that of the variant the library ran "
             "before, on the same workload."]
    for key in sorted(best):
        r = best[key]
        op, w, d, dt = key
        prev = [x for x in rows if (x["op"], x["width"], x["depth"], x["dtype"]) == key
                and x.get("current")]
        dms = f", previous choice {prev[0]['ms']:.3f} ms" if prev else ""
        k, p, blk, wpe = r["variant"]
        dt_enum = DTYPE_ENUM[DTYPES[dt]]
        lines.append(f"        {{{ops[op]}, {w}, {d}, {dt_enum}, {{{k}, {p}, {blk}, {wpe}}}}},"
                     f"  // {r['name']}: {r['ms']:.3f} ms{dms}")
    return "\n".join(lines) + "\n"


def report_text(rows, best, argv):
    lines = ["# Signature kernel autotuning on MI355X", "",
             f"Command: `python -m roughpy_amd.tune {' '.join(argv)}`", "",
             "Workloads: " + "; ".join(
                 f"{w}x{d}: {s['n']} streams, lengths {s['lmin']}..{s['lmax']}"
                 for (w, d), s in WORKLOADS.items()) + " (backward: length 256, at most "
             "4096 streams).", "",
             "Error: max over degrees of max|x - ref| / max|ref| against the current "
             "float64 variant on the float64 increments (the float32 modes read them "
             "rounded to float32); backward: max error over the largest gradient entry "
             "against the float64 backward.", "",
             "| op | shape | dtype | variant | ms | error | chosen |",
             "|---|---|---|---|---|---|---|"]
    for r in rows:
        key = (r["op"], r["width"], r["depth"], r["dtype"])
        chosen = "**yes**" if best.get(key) is r else ""
        flag = "" if r["ok"] else " (rejected)"
        lines.append(f"| {r['op']} | {r['width']}x{r['depth']} | {r['dtype']} | {r['name']}"
                     f"{' (current)' if r.get('current') else ''} | {r['ms']:.3f} | "
                     f"{r['err']:.1e}{flag} | {chosen} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--ops", default="sig,vjp")
    ap.add_argument("--shapes", default="5x5,8x4,6x5,4x6")
    ap.add_argument("--dtypes", default="f32,f64,f32c")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--write", nargs="?", const=TABLE, default=None, metavar="PATH",
                    help=f"write the table (default path {TABLE})")
    ap.add_argument("--report", default=None, help="markdown table of every timing")
    ap.add_argument("--json", default=None, help="every timing as JSON")
    args = ap.parse_args(argv)
    torch = _device.require_gpu()
    torch.cuda.set_device(0)

    def log(msg):
        print(msg, flush=True)

    shapes = [tuple(int(x) for x in s.split("x")) for s in args.shapes.split(",") if s]
    dtypes = [DTYPES[s] for s in args.dtypes.split(",") if s]
    ops = set(args.ops.split(","))
    rows = []
    t0 = time.time()
    for (w, d) in shapes:
        if "sig" in ops:
            now = {dt: config(_capi.RPY_OP_SIGNATURE, w, d, dt)[0] for dt in dtypes}
            for r in tune_signature(torch, w, d, dtypes, args.steps, args.warmup, log):
                r["current"] = tuple(r["variant"]) == now[DTYPES[r["dtype"]]]
                rows.append(r)
        if "vjp" in ops:
            vd = [dt for dt in dtypes if dt != _capi.RPY_F32_COMPENSATED]
            now = {dt: config(_capi.RPY_OP_SIGNATURE_VJP, w, d, dt)[0] for dt in vd}
            for r in tune_vjp(torch, w, d, vd, args.steps, args.warmup, log):
                r["current"] = tuple(r["variant"]) == now[DTYPES[r["dtype"]]]
                rows.append(r)
    log(f"tuned in {time.time() - t0:.1f} s")
    best = winners(rows)
    if args.write:
        with open(args.write, "w") as f:
            f.write(table_text(best, rows))
        log(f"wrote {args.write}")
    if args.report:
        with open(args.report, "w") as f:
            f.write(report_text(rows, best, argv))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f, indent=1)


if __name__ == "__main__":
    main()
