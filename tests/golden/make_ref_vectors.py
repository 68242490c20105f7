"""Writes the fixtures made from the reference's own roughpy_compute kernels,
run in this container by the reference harness oracle/_ref/libref.so (built
from the reference headers by oracle/Makefile):

* tests/golden/ref_vectors.npz -- full input/output vectors at small shapes
  (w, d) in SHAPES; the GPU tests compare against them on a machine where the
  reference tree is absent;
* tests/golden/ref_digests.json -- at the benchmark shapes in BASELINE_SHAPES
  (width 4 depth 6, width 5 depth 5, width 8 depth 4, width 6 depth 5), whose
  vectors are too large to commit: for every kernel output, the SHA-256 of
  its bytes and a few sampled coefficients.  The inputs are regenerated
  bit for bit by baseline_inputs(), so a test can rerun the oracle on them and
  check that it reproduces the reference kernels exactly.

Run:  make -C oracle && python tests/golden/make_ref_vectors.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SHAPES = [(2, 3), (3, 4), (4, 3)]
BASELINE_SHAPES = [(4, 6), (5, 5), (8, 4), (6, 5)]
# stream lengths of the signature vectors at the benchmark shapes
BASELINE_LENGTHS = [0, 7, 40, 1]
# output coefficients sampled into the digest file (besides the hash)
N_SAMPLES = 8


def _rand_tensor(rng, w, d, n, dtype, unit=None, scale_by_factorial=True):
    db = O.tensor_degree_begin(w, d)
    x = rng.uniform(-1.0, 1.0, size=(n, int(db[-1])))
    if scale_by_factorial:
        f = 1.0
        for k in range(1, d + 1):
            f *= k
            x[:, db[k]:db[k + 1]] /= f
    if unit is not None:
        x[:, 0] = unit
    return x.astype(dtype)


def baseline_inputs(w, d, dtype):
    """Deterministic inputs at a benchmark shape (one tensor per operand; a
    batch of four short streams for the signature): a, b with unit scalar
    level, c general, x with zero scalar level, levels scaled by 1/k!."""
    sfx = "f64" if np.dtype(dtype) == np.float64 else "f32"
    seed = 12345 + 1000 * w + 10 * d + (0 if sfx == "f64" else 1)
    rng = np.random.default_rng(seed)
    a = _rand_tensor(rng, w, d, 1, dtype, unit=1.0)
    b = _rand_tensor(rng, w, d, 1, dtype, unit=1.0)
    c = _rand_tensor(rng, w, d, 1, dtype)
    x = _rand_tensor(rng, w, d, 1, dtype, unit=0.0)
    lens = np.array(BASELINE_LENGTHS, dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    inc = (rng.normal(size=(int(offs[-1]), w)) / 4.0).astype(dtype)
    return dict(a=a, b=b, c=c, x=x, sig_increments=inc, sig_offsets=offs)


def reference_outputs(R, w, d, g):
    """Every kernel output of the reference harness on the inputs g."""
    import ctypes
    sfx = "f64" if g["a"].dtype == np.float64 else "f32"
    p = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    a, b, c, x = g["a"], g["b"], g["c"], g["x"]
    n = a.shape[0]
    fma = c.copy()
    getattr(R, f"ref_ft_fma_{sfx}")(p(fma), p(a), p(b), w, d, n)
    imul = a.copy()
    getattr(R, f"ref_ft_inplace_mul_{sfx}")(p(imul), p(b), w, d, n)
    ex = np.zeros_like(a)
    getattr(R, f"ref_ft_exp_{sfx}")(p(ex), p(x), w, d, n)
    fmx = np.zeros_like(a)
    getattr(R, f"ref_ft_fmexp_{sfx}")(p(fmx), p(a), p(x), w, d, n)
    lg = np.zeros_like(a)
    getattr(R, f"ref_ft_log_{sfx}")(p(lg), p(a), w, d, n)
    ap = np.zeros_like(a)
    getattr(R, f"ref_antipode_{sfx}")(p(ap), p(c), w, d, n)
    adj = np.zeros_like(a)
    getattr(R, f"ref_adj_lmul_{sfx}")(p(adj), p(a), p(c), w, d, n)
    st = c.copy()
    getattr(R, f"ref_st_fma_{sfx}")(p(st), p(a), p(b), w, d, n)
    st_adj = np.zeros_like(a)
    getattr(R, f"ref_st_adj_mul_{sfx}")(p(st_adj), p(a), p(c), w, d, n)
    sig = O.signature_batch(g["sig_increments"], g["sig_offsets"], w, d, use_ref=True)
    return dict(fma=fma, inplace_mul=imul, exp=ex, fmexp=fmx, log=lg, antipode=ap,
                adj_lmul=adj, st_fma=st, st_adj_mul=st_adj, signature=sig)


def digest(arr):
    """SHA-256 of the array's bytes (C order, native little-endian)."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def sample_indices(size, n=N_SAMPLES):
    """Fixed coefficient indices spread over a flattened output."""
    return [int(i) for i in np.linspace(0, size - 1, n).round().astype(np.int64)]


def write_digests(R, path):
    import json
    out = {}
    for w, d in BASELINE_SHAPES:
        for dtype, sfx in ((np.float64, "f64"), (np.float32, "f32")):
            g = baseline_inputs(w, d, dtype)
            for name, arr in reference_outputs(R, w, d, g).items():
                flat = arr.ravel()
                idx = sample_indices(flat.size)
                out[f"w{w}_d{d}_{sfx}/{name}"] = {
                    "sha256": digest(arr), "shape": list(arr.shape),
                    "samples": {str(i): float(flat[i]) for i in idx}}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"wrote {path} ({len(out)} digests)")


def main(path=os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")):
    R = O.ref_lib()
    if R is None:
        raise SystemExit("reference harness not built: run `make -C oracle` first")
    rng = np.random.default_rng(12345)
    out = {}
    n = 3
    for w, d in SHAPES:
        for dtype, sfx in ((np.float64, "f64"), (np.float32, "f32")):
            key = f"w{w}_d{d}_{sfx}"
            a = _rand_tensor(rng, w, d, n, dtype, unit=1.0)
            b = _rand_tensor(rng, w, d, n, dtype, unit=1.0)
            c = _rand_tensor(rng, w, d, n, dtype)
            x = _rand_tensor(rng, w, d, n, dtype, unit=0.0)
            p = lambda arr: arr.ctypes.data_as(__import__("ctypes").c_void_p)  # noqa: E731
            fma = c.copy()
            getattr(R, f"ref_ft_fma_{sfx}")(p(fma), p(a), p(b), w, d, n)
            imul = a.copy()
            getattr(R, f"ref_ft_inplace_mul_{sfx}")(p(imul), p(b), w, d, n)
            ex = np.zeros_like(a)
            getattr(R, f"ref_ft_exp_{sfx}")(p(ex), p(x), w, d, n)
            fmx = np.zeros_like(a)
            getattr(R, f"ref_ft_fmexp_{sfx}")(p(fmx), p(a), p(x), w, d, n)
            lg = np.zeros_like(a)
            getattr(R, f"ref_ft_log_{sfx}")(p(lg), p(a), w, d, n)
            ap = np.zeros_like(a)
            getattr(R, f"ref_antipode_{sfx}")(p(ap), p(c), w, d, n)
            adj = np.zeros_like(a)
            getattr(R, f"ref_adj_lmul_{sfx}")(p(adj), p(a), p(c), w, d, n)
            st = c.copy()
            getattr(R, f"ref_st_fma_{sfx}")(p(st), p(a), p(b), w, d, n)
            st_adj = np.zeros_like(a)
            getattr(R, f"ref_st_adj_mul_{sfx}")(p(st_adj), p(a), p(c), w, d, n)
            for name, arr in dict(a=a, b=b, c=c, x=x, fma=fma, inplace_mul=imul, exp=ex,
                                  fmexp=fmx, log=lg, antipode=ap, adj_lmul=adj, st_fma=st,
                                  st_adj_mul=st_adj).items():
                out[f"{key}/{name}"] = arr
            lens = np.array([0, 5, 17, 1], dtype=np.int64)
            offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            inc = (rng.normal(size=(int(offs[-1]), w)) / 4.0).astype(dtype)
            sig = O.signature_batch(inc, offs, w, d, use_ref=True)
            out[f"{key}/sig_increments"] = inc
            out[f"{key}/sig_offsets"] = offs
            out[f"{key}/signature"] = sig
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")
    write_digests(R, os.path.join(os.path.dirname(path), "ref_digests.json"))


if __name__ == "__main__":
    main()
