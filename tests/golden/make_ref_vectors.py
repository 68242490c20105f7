"""Writes tests/golden/ref_vectors.npz: input/output vectors of the
reference's own roughpy_compute kernels, produced in this container by the
reference harness oracle/_ref/libref.so (built from the reference headers by
oracle/Makefile).  The GPU tests compare against these vectors on a machine
where the reference tree is absent.

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


if __name__ == "__main__":
    main()
