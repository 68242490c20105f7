"""Tensor and Lie bases, mirroring roughpy.compute.TensorBasis / LieBasis.

Reference types: roughpy/compute/_src/tensor_basis.c:12-370 (TensorBasis:
width, depth, degree_begin, size(), truncate, equality and hashing) and
roughpy/compute/_src/lie_basis.c (LieBasis: greedy Hall set, degree_begin,
data, size(), get_l2t_matrix / get_t2l_matrix cached per dtype).  The Hall
set and both sparse maps are built by the native library on the host.
"""
from __future__ import annotations

import numpy as np

from . import _capi


# (l2t or t2l, width, depth, Hall data) -> float64 SparseMatrix
_MAP_CACHE = {}

# 64-bit FNV-1a parameters (the published FNV offset basis and prime)
_FNV_OFFSET_64 = 0xCBF29CE484222325
_FNV_PRIME_64 = 0x00000100000001B3
_U64 = (1 << 64) - 1


def fnv1a64(data: bytes, state: int = _FNV_OFFSET_64) -> int:
    """64-bit FNV-1a of `data` continuing from `state`: per byte, xor it into
    the state, then multiply by the FNV prime modulo 2^64."""
    for byte in data:
        state = ((state ^ byte) * _FNV_PRIME_64) & _U64
    return state


class TensorBasis:
    """Dense tensor basis: level k occupies [degree_begin[k], degree_begin[k+1])."""

    __slots__ = ("_width", "_depth", "_degree_begin")

    def __init__(self, width: int, depth: int, degree_begin=None):
        width, depth = int(width), int(depth)
        if width < 1:
            raise ValueError("width must be positive")
        if depth < 0:
            raise ValueError("depth must be non-negative")
        db = np.zeros(depth + 2, dtype=np.intp)
        for k in range(depth + 1):
            db[k + 1] = 1 + width * db[k]
        if degree_begin is not None:
            given = np.asarray(degree_begin, dtype=np.intp)
            if given.shape != db.shape or not np.array_equal(given, db):
                raise ValueError("degree_begin does not describe the dense tensor layout")
        self._width, self._depth = width, depth
        db.setflags(write=False)
        self._degree_begin = db

    @property
    def width(self) -> int:
        return self._width

    @property
    def depth(self) -> int:
        return self._depth

    @property
    def degree_begin(self) -> np.ndarray:
        return self._degree_begin

    def size(self) -> int:
        return int(self._degree_begin[self._depth + 1])

    def truncate(self, new_depth: int) -> "TensorBasis":
        """The basis of the same width truncated at new_depth, which must be
        in [0, depth) (roughpy/compute/_src/tensor_basis.c:173-203)."""
        new_depth = int(new_depth)
        if new_depth < 0:
            raise ValueError("new_depth must be non-negative")
        if new_depth >= self._depth:
            raise ValueError("new_depth must be less than depth")
        return TensorBasis(self._width, new_depth)

    def __eq__(self, other):
        if not isinstance(other, TensorBasis):
            return NotImplemented
        return (other._width == self._width and other._depth == self._depth
                and np.array_equal(other._degree_begin, self._degree_begin))

    def __hash__(self):
        """64-bit FNV-1a over b"TensorBasis", the width and depth (int32) and
        the depth + 2 degree_begin entries (pointer-sized integers), as
        tensor_basis_hash computes it (roughpy/compute/_src/tensor_basis.c:
        144-171); -1 is reported as -2."""
        payload = (b"TensorBasis" + np.int32(self._width).tobytes()
                   + np.int32(self._depth).tobytes()
                   + np.ascontiguousarray(self._degree_begin, dtype=np.intp).tobytes())
        state = fnv1a64(payload)
        h = state - (1 << 64) if state >= (1 << 63) else state
        return -2 if h == -1 else h

    def __repr__(self):
        return f"TensorBasis({self._width}, {self._depth})"

    def c_struct(self) -> _capi.TensorBasisC:
        db64 = np.ascontiguousarray(self._degree_begin, dtype=np.int64)
        s = _capi.TensorBasisC(self._width, self._depth, _capi.i64p(db64))
        s._keep = db64
        return s


class SparseMatrix:
    """Host CSR matrix (rows x cols) with a per-device cached copy."""

    def __init__(self, rows, cols, offsets, indices, data):
        self.rows, self.cols = int(rows), int(cols)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        self.data = np.ascontiguousarray(data)
        self._dev = None

    @property
    def nnz(self) -> int:
        return int(self.indices.shape[0])

    @property
    def shape(self):
        return (self.rows, self.cols)

    def to_dense(self) -> np.ndarray:
        m = np.zeros((self.rows, self.cols), dtype=self.data.dtype)
        for r in range(self.rows):
            for p in range(self.offsets[r], self.offsets[r + 1]):
                m[r, self.indices[p]] += self.data[p]
        return m

    def device_struct(self):
        """SparseMatrixC pointing at device copies (kept alive on self)."""
        from . import _device
        if self._dev is None:
            t = _device.require_gpu()
            offs = t.from_numpy(self.offsets).cuda()
            idx = t.from_numpy(self.indices).cuda()
            dat = t.from_numpy(self.data).cuda()
            s = _capi.SparseMatrixC(self.rows, self.cols, self.nnz, offs.data_ptr(),
                                    idx.data_ptr(), dat.data_ptr(), 1)
            self._dev = (s, offs, idx, dat)
        return self._dev[0]


def _lie_struct(width, depth, db, data):
    s = _capi.LieBasisC(width, depth, _capi.i64p(db), _capi.i64p(data))
    s._keep = (db, data)
    return s


class LieBasis:
    """Hall basis of the free Lie algebra (keys 1..size, row 0 = (0, 0)).

    LieBasis(width, depth) builds the greedy Hall set; a user Hall set may be
    given with degree_begin and data (checked like check_lie_basis_data).
    """

    def __init__(self, width: int, depth: int, degree_begin=None, data=None):
        width, depth = int(width), int(depth)
        if width < 1 or depth < 0:
            raise ValueError("invalid width/depth")
        L = _capi.lib()
        if degree_begin is None and data is None:
            n = int(L.rpy_lie_basis_size(width, depth))
            db = np.zeros(depth + 2, dtype=np.int64)
            dat = np.zeros((n + 1, 2), dtype=np.int64)
            _capi.check(L.rpy_lie_basis_build(width, depth, _capi.i64p(db), _capi.i64p(dat)))
        else:
            if degree_begin is None or data is None:
                raise ValueError("degree_begin and data must be given together")
            db = np.ascontiguousarray(degree_begin, dtype=np.int64)
            dat = np.ascontiguousarray(data, dtype=np.int64)
            if db.shape != (depth + 2,) or dat.ndim != 2 or dat.shape[1] != 2:
                raise ValueError("degree_begin must have depth+2 entries and data shape (n+1, 2)")
            if dat.shape[0] < db[depth + 1]:
                raise ValueError("data is shorter than degree_begin requires")
            _capi.check(L.rpy_check_lie_basis_data(_lie_struct(width, depth, db, dat)))
        self._width, self._depth = width, depth
        self._degree_begin = db.astype(np.intp)
        self._data = dat.astype(np.intp)
        self._degree_begin.setflags(write=False)
        self._data.setflags(write=False)
        self._l2t = {}
        self._t2l = {}

    @property
    def width(self) -> int:
        return self._width

    @property
    def depth(self) -> int:
        return self._depth

    @property
    def degree_begin(self) -> np.ndarray:
        return self._degree_begin

    @property
    def data(self) -> np.ndarray:
        return self._data

    def size(self) -> int:
        return int(self._degree_begin[self._depth + 1]) - 1

    def degree(self, key: int) -> int:
        for d in range(self._depth + 1):
            if self._degree_begin[d] <= key < self._degree_begin[d + 1]:
                return d
        raise IndexError(key)

    def parents(self, key: int):
        return int(self._data[key, 0]), int(self._data[key, 1])

    def key_to_str(self, key: int) -> str:
        a, b = self.parents(key)
        if a == 0:
            return str(b)
        return f"[{self.key_to_str(a)},{self.key_to_str(b)}]"

    def __iter__(self):
        return iter(range(1, self.size() + 1))

    def __eq__(self, other):
        return (isinstance(other, LieBasis) and other._width == self._width
                and other._depth == self._depth
                and np.array_equal(other._degree_begin, self._degree_begin)
                and np.array_equal(other._data[: self.size() + 1], self._data[: self.size() + 1]))

    def __hash__(self):
        return hash(("LieBasis", self._width, self._depth))

    def __repr__(self):
        return f"LieBasis(width={self._width}, depth={self._depth})"

    def _c(self):
        return _lie_struct(self._width, self._depth,
                           np.ascontiguousarray(self._degree_begin, dtype=np.int64),
                           np.ascontiguousarray(self._data, dtype=np.int64))

    def _build(self, which: str) -> SparseMatrix:
        L = _capi.lib()
        fn = L.rpy_l2t_matrix if which == "l2t" else L.rpy_t2l_matrix
        c = self._c()
        nnz = np.zeros(1, dtype=np.int64)
        _capi.check(fn(c, _capi.i64p(nnz), None, None, None))
        tdim = int(_capi.lib().rpy_tensor_size(self._width, self._depth))
        rows, cols = (tdim, self.size()) if which == "l2t" else (self.size(), tdim)
        offs = np.zeros(rows + 1, dtype=np.int64)
        idx = np.zeros(int(nnz[0]), dtype=np.int64)
        dat = np.zeros(int(nnz[0]), dtype=np.float64)
        _capi.check(fn(c, _capi.i64p(nnz), _capi.i64p(offs), _capi.i64p(idx), _capi.dblp(dat)))
        return SparseMatrix(rows, cols, offs, idx, dat)

    def _get(self, which: str, dtype) -> SparseMatrix:
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise TypeError("only floating point data types are supported")
        cache = self._l2t if which == "l2t" else self._t2l
        if dt not in cache:
            base = cache.get(np.dtype(np.float64))
            if base is None:
                # bases with the same Hall data share one matrix (and its
                # device copy), so a fresh LieBasis does not rebuild it
                key = (which, self._width, self._depth,
                       np.ascontiguousarray(self._degree_begin, dtype=np.int64).tobytes(),
                       np.ascontiguousarray(self._data[: self.size() + 1],
                                            dtype=np.int64).tobytes())
                base = _MAP_CACHE.get(key)
                if base is None:
                    base = self._build(which)
                    _MAP_CACHE[key] = base
                cache[np.dtype(np.float64)] = base
            if dt != np.dtype(np.float64):
                cache[dt] = SparseMatrix(base.rows, base.cols, base.offsets, base.indices,
                                         base.data.astype(dt))
        return cache[dt]

    def get_l2t_matrix(self, dtype=np.float64) -> SparseMatrix:
        return self._get("l2t", dtype)

    def get_t2l_matrix(self, dtype=np.float64) -> SparseMatrix:
        return self._get("t2l", dtype)


def check_lie_basis_data(data, degree_begin, width, depth, total_order=None, basis_major=0,
                         raise_on_fail=False) -> bool:
    """Validity of user Hall-set data (lie_basis.c:1651-1682); Bourbaki major."""
    db = np.ascontiguousarray(degree_begin, dtype=np.int64)
    dat = np.ascontiguousarray(data, dtype=np.int64)
    L = _capi.lib()
    rc = L.rpy_check_lie_basis_data(_lie_struct(int(width), int(depth), db, dat))
    if rc != _capi.RPY_OK:
        if raise_on_fail:
            _capi.check(rc)
        return False
    return True
