"""The main library's algebra objects, on the GPU kernels.

FreeTensor, ShuffleTensor and Lie are truncated elements of width w and
depth D with the Python behaviour of roughpy's objects (roughpy/_roughpy.pyi:
FreeTensor :512-560, ShuffleTensor :562-620, Lie :444-498): construction
from a scalar (free and shuffle tensors only), a sequence or buffer of
coefficients, or (index, value) pairs / a dict; width and depth from the
arguments, a Context, or the data (a flat sequence of 1 + w values is a
degree-one element of width w); arithmetic, in-place arithmetic and the
fused forms (add_mul, sub_scal_div, ...); ``str``/``repr`` in the reference
formats.  Context (:389-416) and get_context (:419-423) bind a width, depth
and scalar type and provide cbh, compute_signature, to_logsignature,
lie_to_tensor, tensor_to_lie and zero_lie
(roughpy/src/algebra/context.cpp:114-262).

Products (the free tensor product, the shuffle product, the Lie bracket),
exp, log, fmexp, the antipode, the Lie/tensor maps, CBH and signatures run
on the GPU library; sums and scalings act on the host coefficients.  The
coefficients of an element are a dense host array over the whole truncated
basis; ``size()`` is the number of coefficients it was given (the full
dimension for computed results), ``degree()`` the highest degree holding a
non-zero coefficient.
"""
from __future__ import annotations

import numbers
from typing import Any, Optional, Sequence

import numpy as np

from . import compute as rc
from .bases import LieBasis, TensorBasis

__all__ = [
    "ScalarType", "DPReal", "SPReal", "VectorType", "TensorKey", "FreeTensor", "ShuffleTensor",
    "Lie", "Context", "get_context",
]


class ScalarType:
    """Coefficient type of an algebra (roughpy's ScalarTypeMeta)."""

    def __init__(self, name: str, dtype):
        self.name = name
        self.dtype = np.dtype(dtype)

    def __repr__(self):
        return self.name

    __str__ = __repr__


DPReal = ScalarType("DPReal", np.float64)
SPReal = ScalarType("SPReal", np.float32)


class VectorType:
    """Storage of an element (roughpy.VectorType); the values are held
    densely whichever is named."""
    DenseVector = "DenseVector"
    SparseVector = "SparseVector"


def _scalar_type(coeffs, default=DPReal) -> ScalarType:
    if coeffs is None:
        return default
    if isinstance(coeffs, ScalarType):
        return coeffs
    if isinstance(coeffs, str):
        name = coeffs.lower()
        if name in ("dpreal", "double", "float64", "f64"):
            return DPReal
        if name in ("spreal", "float", "float32", "f32"):
            return SPReal
        raise ValueError(f"unknown scalar type {coeffs!r}")
    dt = np.dtype(coeffs)
    if dt == np.float64:
        return DPReal
    if dt == np.float32:
        return SPReal
    raise TypeError(f"unsupported scalar type {coeffs!r}")


def _tensor_db(width: int, depth: int):
    db = [0]
    for _ in range(depth + 1):
        db.append(1 + width * db[-1])
    return db


class TensorKey:
    """A word of letters 1..width (roughpy.TensorKey, _roughpy.pyi:338-348)."""

    def __init__(self, *letters: int, width: int, depth: int):
        if len(letters) == 1 and isinstance(letters[0], (list, tuple)):
            letters = tuple(letters[0])
        if len(letters) > depth:
            raise ValueError("word longer than the depth")
        if any(not 1 <= int(a) <= width for a in letters):
            raise ValueError("letters must lie in 1..width")
        self.letters = tuple(int(a) for a in letters)
        self.width, self.max_degree = int(width), int(depth)

    def degree(self) -> int:
        return len(self.letters)

    def index(self) -> int:
        db = _tensor_db(self.width, self.max_degree)
        idx = 0
        for a in self.letters:
            idx = idx * self.width + (a - 1)
        return db[len(self.letters)] + idx

    def __str__(self):
        return "(" + ",".join(str(a) for a in self.letters) + ")"

    def __repr__(self):
        return f"TensorKey{str(self)}"

    def __eq__(self, other):
        return (isinstance(other, TensorKey) and self.letters == other.letters
                and self.width == other.width)

    def __hash__(self):
        return hash((self.letters, self.width))


def _word_of(index: int, width: int, depth: int):
    db = _tensor_db(width, depth)
    for k in range(depth + 1):
        if index < db[k + 1]:
            rest, letters = index - db[k], []
            for _ in range(k):
                letters.append(rest % width + 1)
                rest //= width
            return tuple(reversed(letters))
    raise IndexError(index)


def _fmt(v) -> str:
    return f"{float(v):g}"


class _HostDLPack:
    """An object exporting __dlpack__ read as host memory (producers that
    define no __dlpack_device__ are taken to be on the CPU)."""

    def __init__(self, obj):
        self._obj = obj

    def __dlpack__(self, stream=None, **kwargs):
        return self._obj.__dlpack__(stream=stream)

    def __dlpack_device__(self):
        dev = getattr(self._obj, "__dlpack_device__", None)
        return dev() if dev is not None else (1, 0)


class _IteratorItem:
    def __init__(self, key, value):
        self._key, self._value = key, value

    def key(self):
        return self._key

    def value(self):
        return self._value


class _Element:
    """Shared behaviour of the three element types: a dense coefficient
    array `_data` over the basis of (width, depth), the stored size and the
    scalar type."""

    _kind = "element"

    def __init__(self, data: Any = None, *, width: Optional[int] = None,
                 depth: Optional[int] = None, ctx: Optional["Context"] = None,
                 dtype=None, coeffs=None, vector_type=None, **kwargs):
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        if ctx is not None:
            width = ctx.width if width is None else width
            depth = ctx.depth if depth is None else depth
            stype = ctx.ctype if (dtype is None and coeffs is None) else None
        else:
            stype = None
        if isinstance(data, str):
            raise ValueError(f"unexpected string as {self._kind} data")
        storage = VectorType.DenseVector
        pairs = None
        values = None
        if data is None:
            values = np.zeros(0)
        elif isinstance(data, dict):
            pairs, storage = list(data.items()), VectorType.SparseVector
        elif isinstance(data, tuple) and len(data) == 2 and not isinstance(
                data[1], (list, tuple, np.ndarray)) and (
                isinstance(data[0], (TensorKey, numbers.Integral))):
            pairs, storage = [data], VectorType.SparseVector
        elif isinstance(data, (list, tuple)) and data and all(
                isinstance(x, tuple) and len(x) == 2 for x in data):
            pairs, storage = list(data), VectorType.SparseVector
        elif isinstance(data, (list, tuple)) and any(isinstance(x, dict) for x in data):
            raise ValueError("a sequence of dictionaries is not one element")
        elif isinstance(data, (list, tuple)) and any(isinstance(x, str) for x in data):
            raise ValueError("unexpected string in key-scalar data")
        elif isinstance(data, numbers.Number):
            values = self._from_scalar(data)
        else:
            if hasattr(data, "__dlpack__") and not isinstance(data, np.ndarray):
                data = np.from_dlpack(_HostDLPack(data))
            values = np.asarray(data)
            if values.ndim != 1:
                raise ValueError(f"{self._kind} data must be one-dimensional")
        if values is not None and values.dtype.kind in "iub":
            values = values.astype(np.float64)
        if stype is None:
            guess = values.dtype if values is not None and values.dtype.kind == "f" else None
            stype = _scalar_type(dtype if dtype is not None else coeffs,
                                 _scalar_type(guess) if guess is not None else DPReal)
        if vector_type is not None:
            storage = vector_type
        if pairs is not None:
            if width is None:
                raise ValueError("key-value data needs a width")
        if width is None:
            width = self._deduce_width(len(values))
        width = int(width)
        if depth is None:
            depth = max(2, self._degree_of_size(width, len(values) if values is not None else 0,
                                                 64))
        depth = int(depth)
        self._setup(width, depth, stype)
        dim = self.dimension()
        full = np.zeros(dim, dtype=stype.dtype)
        if pairs is not None:
            for k, v in pairs:
                idx = k.index() if isinstance(k, TensorKey) else int(k)
                if isinstance(v, str):
                    raise ValueError("unexpected string in key-scalar data")
                if not 0 <= idx < dim:
                    raise ValueError(f"key {k} outside the basis")
                full[idx] = v
            self._size = dim
        else:
            if len(values) > dim:
                raise ValueError(f"{len(values)} coefficients do not fit a {self._kind} of "
                                 f"width {width} and depth {depth} ({dim})")
            full[:len(values)] = values
            self._size = len(values)
        self._data = full
        self.storage_type = storage

    # -- construction helpers (overridden per element type) -----------------
    def _from_scalar(self, v):
        return np.array([float(v)])

    def _deduce_width(self, n: int) -> int:
        if n < 2:
            raise ValueError(f"cannot deduce the width of a {self._kind} from {n} values")
        return n - 1

    @staticmethod
    def _degree_of_size(width, n, cap):
        if n <= 1:
            return 0
        db = _tensor_db(width, cap)
        for k in range(cap + 1):
            if n <= db[k + 1]:
                return k
        return cap

    def _setup(self, width, depth, stype):
        self.width, self.max_degree, self.dtype = width, depth, stype

    def _new(self, data, size=None):
        obj = self.__class__.__new__(self.__class__)
        obj._setup(self.width, self.max_degree, self.dtype)
        obj._data = np.ascontiguousarray(data, dtype=self.dtype.dtype)
        obj._size = self.dimension() if size is None else size
        obj.storage_type = self.storage_type
        return obj

    def _check(self, other):
        if not isinstance(other, self.__class__):
            raise TypeError(f"expected a {self.__class__.__name__}")
        if other.width != self.width:
            raise ValueError("elements of different widths")
        if other.max_degree != self.max_degree:
            raise ValueError("elements of different depths")

    # -- sizes ----------------------------------------------------------------
    def size(self) -> int:
        return int(self._size)

    def dimension(self) -> int:
        raise NotImplementedError

    def _level_begin(self):
        raise NotImplementedError

    def degree(self) -> int:
        nz = np.flatnonzero(self._data)
        if nz.size == 0:
            return 0
        lb = self._level_begin()
        last = int(nz[-1])
        for k in range(len(lb) - 1):
            if last < lb[k + 1]:
                return k
        return self.max_degree

    def __array__(self, dtype=None, copy=None):
        out = np.array(self._data, copy=True)
        return out if dtype is None else out.astype(dtype)

    # -- linear structure -----------------------------------------------------
    def __neg__(self):
        return self._new(-self._data, self._size)

    def __add__(self, other):
        self._check(other)
        return self._new(self._data + other._data, max(self._size, other._size))

    def __sub__(self, other):
        self._check(other)
        return self._new(self._data - other._data, max(self._size, other._size))

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return self._new(self._data * other, self._size)
        return NotImplemented

    def __truediv__(self, other):
        if isinstance(other, numbers.Number):
            return self._new(self._data / other, self._size)
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return self._new(self._data * other, self._size)
        self._check(other)
        return self._product(other)

    def _product(self, other):
        raise NotImplementedError

    def _assign(self, res):
        self._data, self._size = res._data, res._size
        return self

    def __iadd__(self, other):
        return self._assign(self + other)

    def __isub__(self, other):
        return self._assign(self - other)

    def __imul__(self, other):
        return self._assign(self * other)

    def __itruediv__(self, other):
        return self._assign(self / other)

    def add_scal_mul(self, other, scalar):
        return self._assign(self + other * scalar)

    def sub_scal_mul(self, other, scalar):
        return self._assign(self - other * scalar)

    def add_scal_div(self, other, scalar):
        return self._assign(self + other / scalar)

    def sub_scal_div(self, other, scalar):
        return self._assign(self - other / scalar)

    def add_mul(self, lhs, rhs):
        return self._assign(self + lhs * rhs)

    def sub_mul(self, lhs, rhs):
        return self._assign(self - lhs * rhs)

    def mul_smul(self, other, scalar):
        return self._assign(self * (other * scalar))

    def mul_sdiv(self, other, scalar):
        return self._assign(self * (other / scalar))

    # -- comparison and display -------------------------------------------------
    def __eq__(self, other):
        if not isinstance(other, self.__class__):
            return NotImplemented
        return (self.width == other.width and self.max_degree == other.max_degree
                and np.array_equal(self._data, other._data))

    def __ne__(self, other):
        eq = self.__eq__(other)
        return eq if eq is NotImplemented else not eq

    __hash__ = None

    def __repr__(self):
        return (f"{self.__class__.__name__}(width={self.width}, depth={self.max_degree}, "
                f"ctype={self.dtype})")

    def _key_str(self, i: int) -> str:
        raise NotImplementedError

    def __str__(self):
        terms = [f"{_fmt(v)}{self._key_str(i)}" for i, v in enumerate(self._data) if v != 0]
        return "{ " + " ".join(terms) + " }"

    def __getitem__(self, item):
        return self._data[self._index_of(item)]

    def _index_of(self, item) -> int:
        if isinstance(item, numbers.Integral):
            return int(item)
        raise TypeError(f"cannot index a {self.__class__.__name__} with {item!r}")

    def __reduce__(self):
        return (_rebuild, (self.__class__, self.width, self.max_degree, self.dtype.name,
                           self._data, self._size, self.storage_type))


def _rebuild(cls, width, depth, sname, data, size, storage):
    obj = cls.__new__(cls)
    obj._setup(width, depth, _scalar_type(sname))
    obj._data, obj._size, obj.storage_type = np.asarray(data), size, storage
    return obj


class _TensorElement(_Element):
    def dimension(self) -> int:
        return self._basis().size()

    def _basis(self) -> TensorBasis:
        return TensorBasis(self.width, self.max_degree)

    def _level_begin(self):
        return _tensor_db(self.width, self.max_degree)

    def _key_str(self, i):
        return "(" + ",".join(str(a) for a in _word_of(i, self.width, self.max_degree)) + ")"

    def _index_of(self, item):
        if isinstance(item, TensorKey):
            return item.index()
        if isinstance(item, (tuple, list)):
            return TensorKey(*item, width=self.width, depth=self.max_degree).index()
        return super()._index_of(item)

    def __iter__(self):
        for i, v in enumerate(self._data):
            if v != 0:
                yield _IteratorItem(TensorKey(*_word_of(i, self.width, self.max_degree),
                                              width=self.width, depth=self.max_degree), v)


class FreeTensor(_TensorElement):
    """Truncated free tensor (roughpy.FreeTensor)."""

    _kind = "tensor"

    def _ft(self):
        return rc.DenseFreeTensor(self._data, self._basis())

    def _product(self, other):
        return self._new(rc.ft_mul(self._ft(), other._ft()).data)

    def exp(self) -> "FreeTensor":
        """The truncated series sum_{i <= D} X^i / i!, Horner's form of the
        main library's exp (vendored/libalgebra_lite/free_tensor.h:856-871),
        which keeps X's scalar term.  Without a scalar term it is
        free_tensor_exp.hpp:13-37, one fused launch; with one, D products."""
        if self._data[0] == 0:
            return self._new(rc.ft_exp(self._ft()).data)
        one = np.zeros_like(self._data)
        one[0] = 1
        acc = one.copy()
        for i in range(self.max_degree, 0, -1):
            acc = np.asarray(rc.ft_mul(rc.DenseFreeTensor(acc, self._basis()),
                                       rc.DenseFreeTensor(self._data / i, self._basis())).data)
            acc = acc + one
        return self._new(acc)

    def log(self) -> "FreeTensor":
        """log(X) truncated, taking X[0] = 1 (free_tensor_log.hpp:12-40)."""
        return self._new(rc.ft_log(self._ft()).data)

    def antipode(self) -> "FreeTensor":
        return self._new(rc.antipode(self._ft()).data)

    def fmexp(self, other: "FreeTensor") -> "FreeTensor":
        """self <- self exp(other) in place (free_tensor_fmexp.hpp:15-40)."""
        self._check(other)
        return self._assign(self._new(rc.ft_fmexp(self._ft(), other._ft()).data))


class ShuffleTensor(_TensorElement):
    """Truncated shuffle tensor (roughpy.ShuffleTensor); its product is the
    shuffle product."""

    _kind = "shuffle tensor"

    def _product(self, other):
        a = rc.DenseShuffleTensor(self._data, self._basis())
        b = rc.DenseShuffleTensor(other._data, self._basis())
        return self._new(rc.st_mul(a, b).data)


class Lie(_Element):
    """Truncated Lie element in the Hall basis (roughpy.Lie); its product
    is the Lie bracket."""

    _kind = "Lie element"

    def _from_scalar(self, v):
        raise TypeError("a Lie element cannot be made from a scalar")

    def _deduce_width(self, n: int) -> int:
        if n < 1:
            raise ValueError("cannot deduce the width of an empty Lie element")
        return n

    @staticmethod
    def _degree_of_size(width, n, cap):
        if n <= width:
            return 1 if n else 0
        size = width
        for k in range(2, cap + 1):
            size = LieBasis(width, k).size()
            if n <= size:
                return k
        return cap

    def _basis(self) -> LieBasis:
        return LieBasis(self.width, self.max_degree)

    def dimension(self) -> int:
        return self._basis().size()

    def _level_begin(self):
        db = self._basis().degree_begin
        # key k (1-based) sits at coefficient k - 1
        return [0] + [int(x) - 1 for x in db[1:]]

    def _key_str(self, i):
        return "(" + self._basis().key_to_str(i + 1) + ")"

    def __iter__(self):
        b = self._basis()
        for i, v in enumerate(self._data):
            if v != 0:
                yield _IteratorItem(b.key_to_str(i + 1), v)

    def _lie(self):
        return rc.DenseLie(self._data, self._basis())

    def _product(self, other):
        """[self, other] = t2l(x y - y x) with x, y the tensor images."""
        tb = TensorBasis(self.width, self.max_degree)
        x = rc.lie_to_tensor(self._lie(), tb)
        y = rc.lie_to_tensor(other._lie(), tb)
        xy = rc.ft_mul(x, y).data
        yx = rc.ft_mul(y, x).data
        com = rc.DenseFreeTensor(np.asarray(xy) - np.asarray(yx), tb)
        return self._new(rc.tensor_to_lie(com, self._basis()).data)


class Context:
    """Width, depth and scalar type of an algebra (roughpy.Context,
    roughpy/src/algebra/context.cpp)."""

    def __init__(self, width: int, depth: int, ctype: ScalarType = DPReal):
        if width < 1 or depth < 1:
            raise ValueError("width and depth must be positive")
        self.width, self.depth, self.ctype = int(width), int(depth), ctype
        self.lie_basis = LieBasis(self.width, self.depth)
        self.tensor_basis = TensorBasis(self.width, self.depth)

    def __repr__(self):
        return f"Context(width={self.width}, depth={self.depth}, ctype={self.ctype})"

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        return None

    def _degree(self, degree):
        if degree is None or degree < 0:
            return self.depth
        if degree > self.depth:
            raise ValueError("the requested degree exceeds the maximum degree for this basis")
        return int(degree)

    def lie_size(self, degree: Optional[int] = None) -> int:
        d = self._degree(degree)
        return 0 if d == 0 else LieBasis(self.width, d).size()

    def tensor_size(self, degree: Optional[int] = None) -> int:
        return _tensor_db(self.width, self._degree(degree))[self._degree(degree) + 1]

    def zero_lie(self, vtype=None) -> Lie:
        out = Lie(None, width=self.width, depth=self.depth, ctx=self)
        out.storage_type = vtype or VectorType.SparseVector
        return out

    def lie_to_tensor(self, lie: Lie) -> FreeTensor:
        if not isinstance(lie, Lie):
            raise TypeError("expected a Lie object")
        res = rc.lie_to_tensor(rc.DenseLie(lie._data, LieBasis(lie.width, lie.max_degree)),
                               self.tensor_basis)
        return self._tensor(res.data)

    def tensor_to_lie(self, tensor: FreeTensor) -> Lie:
        if not isinstance(tensor, FreeTensor):
            raise TypeError("expected a FreeTensor object")
        res = rc.tensor_to_lie(rc.DenseFreeTensor(tensor._data, TensorBasis(
            tensor.width, tensor.max_degree)), self.lie_basis)
        return self._lie(res.data)

    def to_logsignature(self, signature: FreeTensor) -> Lie:
        """tensor_to_lie(signature.log()) (context.cpp:213-228)."""
        if not isinstance(signature, FreeTensor):
            raise TypeError("expected a FreeTensor object")
        return self.tensor_to_lie(signature.log())

    def cbh(self, lies: Sequence[Lie], vec_type=None) -> Lie:
        """log(exp(l_1) ... exp(l_n)) in the context's truncation; an empty
        sequence gives zero and one element comes back unchanged
        (Context::cbh, algebra/src/context.cpp:187-216)."""
        lies = list(lies)
        if not lies:
            return self.zero_lie(vec_type)
        for lie in lies:
            if not isinstance(lie, Lie) or lie.width != self.width:
                raise ValueError("cbh takes Lie elements of the context's width")
        if len(lies) == 1:
            return self._lie(self._fit_lie(lies[0]))
        tb = self.tensor_basis
        acc = np.zeros(tb.size(), dtype=self.ctype.dtype)
        acc[0] = 1
        for lie in lies:
            x = rc.lie_to_tensor(rc.DenseLie(self._fit_lie(lie), self.lie_basis), tb).data
            acc = rc.ft_fmexp(rc.DenseFreeTensor(acc, tb), rc.DenseFreeTensor(x, tb)).data
        return self.to_logsignature(self._tensor(acc))

    def compute_signature(self, data: Any, vtype=None) -> FreeTensor:
        """Signature of the piecewise-linear path with the rows of `data` as
        increments: the Chen product of exp(row) (Context::signature through
        compute_signature, algebra/src/lite_context.h:545-563)."""
        from .streams import signature_batch
        arr = np.asarray(data, dtype=self.ctype.dtype)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1) if arr.size else arr.reshape(0, self.width)
        if arr.ndim != 2 or arr.shape[1] != self.width:
            raise ValueError(f"signature data must be rows of {self.width} increments")
        if arr.shape[0] == 0:
            unit = np.zeros(self.tensor_basis.size(), dtype=self.ctype.dtype)
            unit[0] = 1
            return self._tensor(unit)
        sig = signature_batch(np.ascontiguousarray(arr), np.array([0, arr.shape[0]]),
                              self.width, self.depth, dtype=self.ctype.dtype)
        return self._tensor(sig[0])

    # -- helpers ----------------------------------------------------------------
    def _fit_lie(self, lie: Lie) -> np.ndarray:
        """A Lie element's coefficients in this context's basis (truncated or
        zero-padded: the Hall bases of one width agree on common degrees)."""
        n = self.lie_basis.size()
        out = np.zeros(n, dtype=self.ctype.dtype)
        m = min(n, lie._data.shape[0])
        out[:m] = lie._data[:m]
        return out

    def _tensor(self, data) -> FreeTensor:
        obj = FreeTensor.__new__(FreeTensor)
        obj._setup(self.width, self.depth, self.ctype)
        obj._data = np.ascontiguousarray(np.asarray(data), dtype=self.ctype.dtype)
        obj._size = self.tensor_basis.size()
        obj.storage_type = VectorType.DenseVector
        return obj

    def _lie(self, data) -> Lie:
        obj = Lie.__new__(Lie)
        obj._setup(self.width, self.depth, self.ctype)
        obj._data = np.ascontiguousarray(np.asarray(data), dtype=self.ctype.dtype)
        obj._size = self.lie_basis.size()
        obj.storage_type = VectorType.DenseVector
        return obj


_contexts = {}


def get_context(width: int, depth: int, coeffs=None) -> Context:
    """The context of (width, depth, scalar type), one object per key
    (roughpy.get_context)."""
    stype = _scalar_type(coeffs)
    key = (int(width), int(depth), stype.name)
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = _contexts[key] = Context(int(width), int(depth), stype)
    return ctx
