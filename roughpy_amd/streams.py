"""Streams on the GPU: LieIncrementStream and batched signatures.

* ``LieIncrementStream.from_increments(...)`` mirrors the reference
  constructor (roughpy/src/streams/lie_increment_stream.cpp:73-245) and its
  ``signature`` / ``log_signature`` queries mirror Stream.signature /
  log_signature (roughpy/src/streams/stream.cpp:184-276, arguments as
  signature_arguments.cpp:56-241): interval or (inf, sup), optional
  resolution (int, or float converted by param_to_resolution), depth=.
* ``signature_batch`` / ``log_signature_batch`` are the batched entry points
  over many uneven streams (CSR offsets), one persistent GPU launch each.

A signature query at the stream's resolution reads the stream's dyadic
signature cache (built on the GPU on the first such query and kept); other
queries, and log-signatures, decide on the host which increments they cover
(rpy_stream_select_rows, native) and run one Chen-product launch over them.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from . import _capi, _device
from .bases import LieBasis, TensorBasis
from .compute import DenseFreeTensor, DenseLie

__all__ = [
    "RealInterval", "Partition", "LieIncrementStream", "PiecewiseAbelianStream",
    "DyadicSignatureCache", "signature_batch", "log_signature_batch",
    "param_to_resolution", "py_param_to_resolution", "to_dyadic_intervals", "cache_root",
]


def param_to_resolution(arg: float) -> int:
    return int(_capi.lib().rpy_param_to_resolution(float(arg)))


def to_dyadic_intervals(inf: float, sup: float, tol: int):
    """Clopen [inf, sup) as maximal aligned dyadic pieces (k, n), left to right."""
    L = _capi.lib()
    import ctypes
    count = ctypes.c_int32(0)
    rc = L.rpy_to_dyadic_intervals(float(inf), float(sup), int(tol), None, None, 0,
                                   ctypes.byref(count))
    if rc not in (_capi.RPY_OK, _capi.RPY_ERR_SHAPE):
        _capi.check(rc)
    n = count.value
    ks = np.zeros(max(n, 1), dtype=np.int64)
    ns = np.zeros(max(n, 1), dtype=np.int32)
    _capi.check(L.rpy_to_dyadic_intervals(float(inf), float(sup), int(tol), _capi.i64p(ks),
                                           ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                           n, ctypes.byref(count)))
    return [(int(ks[i]), int(ns[i])) for i in range(n)]


class RealInterval:
    """Clopen real interval [inf, sup)."""

    __slots__ = ("inf", "sup")

    def __init__(self, inf: float, sup: float):
        self.inf, self.sup = float(inf), float(sup)

    def __repr__(self):
        return f"RealInterval({self.inf}, {self.sup})"


def _bounds_of(interval):
    """(inf, sup) of a RealInterval, a (inf, sup) pair or an object with
    inf/sup attributes or methods."""
    if isinstance(interval, RealInterval):
        return interval.inf, interval.sup
    if isinstance(interval, (tuple, list)):
        return float(interval[0]), float(interval[1])
    inf_ = interval.inf() if callable(interval.inf) else interval.inf
    sup_ = interval.sup() if callable(interval.sup) else interval.sup
    return float(inf_), float(sup_)


class Partition(RealInterval):
    """A clopen interval [inf, sup) cut at sorted intermediate points, with
    the behaviour of intervals::Partition (intervals/src/partition.cpp:44-250)
    and its Python binding (roughpy/src/intervals/partition.cpp:51-130).

    Intermediates outside [inf, sup) or equal to an end point are dropped;
    ``p[i]`` is the i-th sub-interval, left to right."""

    __slots__ = ("_mids",)

    def __init__(self, interval, intermediates=()):
        inf_, sup_ = _bounds_of(interval)
        super().__init__(inf_, sup_)
        kept = (float(m) for m in intermediates)
        self._mids = sorted(m for m in kept if inf_ < m < sup_)

    @classmethod
    def _raw(cls, inf_, sup_, mids):
        obj = cls.__new__(cls)
        RealInterval.__init__(obj, inf_, sup_)
        obj._mids = list(mids)
        return obj

    def _points(self):
        """inf, the intermediates and sup, in order."""
        return [self.inf, *self._mids, self.sup]

    def __len__(self):
        return 1 + len(self._mids)

    def __getitem__(self, i: int) -> RealInterval:
        if not 0 <= i <= len(self._mids):
            raise IndexError("partition index out of range")
        pts = self._points()
        return RealInterval(pts[i], pts[i + 1])

    def intermediates(self):
        return list(self._mids)

    def refine_midpoints(self) -> "Partition":
        """Every sub-interval [a, b) cut at (b + a) / 2."""
        pts = self._points()
        mids = []
        for a, b in zip(pts[:-1], pts[1:]):
            mids.append((b + a) / 2.0)
            mids.append(b)
        return Partition._raw(self.inf, self.sup, mids[:-1])

    def mesh(self) -> float:
        """The length of the shortest sub-interval (what partition.cpp:75-91
        computes)."""
        pts = self._points()
        return min(abs(b - a) for a, b in zip(pts[:-1], pts[1:]))

    def insert_intermediate(self, point: float) -> None:
        point = float(point)
        if not self.inf <= point < self.sup:
            raise ValueError("provided intermediate does not lie within the interval")
        import bisect
        bisect.insort_left(self._mids, point)

    def merge(self, other: "Partition") -> "Partition":
        """The partition of the hull [min inf, max sup) of two partitions.

        Overlapping or touching intervals: the cut points are every
        intermediate of either partition, the larger inf when the infs
        differ, and the smaller sup when the sups differ and it is not that
        inf; a value cut in both (or a cut that is also such an end point)
        appears as often as it does in the one that cuts it most.  Disjoint
        intervals: the two end points facing the gap only -- the reference
        drops both partitions' intermediates then
        (intervals/src/partition.cpp:146-161)."""
        lo, hi = min(self.inf, other.inf), max(self.sup, other.sup)
        if self.sup < other.inf or other.sup < self.inf:
            left, right = (self, other) if self.sup < other.inf else (other, self)
            cuts = [v for v, nonempty in ((left.sup, left.sup != left.inf),
                                          (right.inf, right.inf != right.sup)) if nonempty]
            return Partition._raw(lo, hi, cuts)
        from collections import Counter
        mult = Counter(self._mids)
        for v, c in Counter(other._mids).items():
            mult[v] = max(mult[v], c)
        inner_inf, inner_sup = max(self.inf, other.inf), min(self.sup, other.sup)
        if self.inf != other.inf:
            mult[inner_inf] = max(mult[inner_inf], 1)
        if self.sup != other.sup and inner_sup != inner_inf:
            mult[inner_sup] = max(mult[inner_sup], 1)
        cuts = [v for v in sorted(mult) for _ in range(mult[v])]
        return Partition._raw(lo, hi, cuts)

    def __eq__(self, other):
        return (isinstance(other, Partition) and self.inf == other.inf
                and self.sup == other.sup and self._mids == other._mids)

    def __hash__(self):
        return hash((self.inf, self.sup, tuple(self._mids)))

    def __reduce__(self):
        return (Partition._raw, (self.inf, self.sup, list(self._mids)))

    def __repr__(self):
        return f"Partition({self.inf}, {self.sup}, {self._mids})"

    def __str__(self):
        return "".join(f"[{self[i].inf}, {self[i].sup})" for i in range(len(self)))


def _cbh_rows(lies, tb: TensorBasis, lb: LieBasis):
    """Context::cbh over the rows of each group (algebra/src/context.cpp:
    187-196): log(prod_j exp(l_j)) in the tensor algebra, mapped back to the
    Lie basis.  lies: (T, M, lie_dim) GPU tensor, T factors for each of M
    groups (zero rows are unit factors).  Returns (M, lie_dim)."""
    from . import compute as rc
    t = _device.torch()
    T, M = int(lies.shape[0]), int(lies.shape[1])
    coll = t.zeros((M, tb.size()), dtype=lies.dtype, device=lies.device)
    coll[:, 0] = 1
    for j in range(T):
        X = rc.lie_to_tensor(rc.Lie(lies[j].contiguous(), lb), tb).data
        coll = rc.ft_fmexp(rc.FreeTensor(coll, tb), rc.FreeTensor(X, tb)).data
    logc = rc.ft_log(rc.FreeTensor(coll, tb)).data
    return rc.tensor_to_lie(rc.FreeTensor(logc, tb), lb).data


def py_param_to_resolution(accuracy: float) -> int:
    """-min(5, e - 1) for accuracy = m 2^e with 0.5 <= |m| < 1: the Python
    bindings' conversion of an accuracy to a resolution, used for the
    resolution from_increments infers and for float resolution arguments
    (roughpy/src/args/kwargs_to_path_metadata.cpp:224-229; the streams
    library's own param_to_resolution caps at 0 instead of 5)."""
    _, e = math.frexp(float(accuracy))
    return -min(5, e - 1)


def _resolution_arg(res) -> Optional[int]:
    if res is None:
        return None
    if isinstance(res, bool):
        raise TypeError("resolution should be either float or int")
    if isinstance(res, (int, np.integer)):
        return int(res)
    if isinstance(res, (float, np.floating)):
        return py_param_to_resolution(float(res))
    raise TypeError("resolution should be either float or int")


_VTYPES = ("dense", "sparse")


def _check_vtype(vtype):
    """vtype as the reference takes it (algebra::VectorType, Dense or Sparse):
    a name, or anything whose name attribute is one."""
    if vtype is None:
        return
    name = getattr(vtype, "name", vtype)
    if not isinstance(name, str) or name.lower() not in _VTYPES:
        raise TypeError(f"vtype must be one of {_VTYPES}, got {vtype!r}")


def _schema_increment_channels(schema, channel_types):
    """Number of channels a schema or channel_types argument describes when
    every channel is an increment (the only kind a LieIncrementStream of
    this path carries), None when neither is given or the channels are
    homogeneous increments of the data's width."""
    spec = schema if schema is not None else channel_types
    if spec is None:
        return None
    if isinstance(spec, str):
        if spec.lower() == "increment":
            return None
        raise NotImplementedError(f"channel type {spec!r} is not supported")
    items = list(spec.values()) if isinstance(spec, dict) else list(spec)
    for it in items:
        name = getattr(it, "name", it)
        if not isinstance(name, str) or name.lower() != "increment":
            raise NotImplementedError(
                f"only increment channels are supported, got {it!r}")
    return len(items)


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float64:
        return _capi.RPY_F64
    if dt == np.float32:
        return _capi.RPY_F32
    raise TypeError(f"unsupported dtype {dt}")


def _rows_of(inc, width: int):
    """(rows, width) view of the increments: a 2-d array must have `width`
    columns; a 1-d one is read as consecutive rows."""
    if inc.ndim == 2:
        if int(inc.shape[1]) != int(width):
            raise ValueError(f"increments have {int(inc.shape[1])} columns, width is {width}")
        return inc
    if inc.ndim == 1 and int(inc.shape[0]) % int(width) == 0:
        return inc.reshape(-1, int(width))
    raise ValueError(f"increments of shape {tuple(inc.shape)} are not rows of width {width}")


def _check_host_offsets(offsets, rows: int):
    """CSR offsets given on the host are checked before any launch: they
    start at or above 0, never decrease and end at or below the row count.
    Offsets already on the device are taken as they are (the caller
    guarantees the same)."""
    if _device.is_device_array(offsets):
        return
    off = np.asarray(offsets, dtype=np.int64).reshape(-1)
    if off.shape[0] < 1:
        raise ValueError("offsets must hold at least one entry")
    if off[0] < 0 or np.any(np.diff(off) < 0) or off[-1] > rows:
        raise ValueError("offsets must be non-decreasing, start at 0 or above and end at "
                         f"or below the number of increment rows ({rows})")


def signature_batch(increments, offsets, width: int, depth: int, dtype=None,
                    compensated: bool = False):
    """Signatures of uneven streams.

    increments: (rows, width) array (NumPy or GPU tensor); offsets: (n+1,)
    int64 CSR offsets, stream s owning rows offsets[s]:offsets[s+1].
    Returns an (n, tensor_size) GPU tensor (NumPy if increments is NumPy).
    compensated (float32 data only): the float32 kernels with signature
    levels 1 and 2 accumulated by Kahan's compensated summation (smaller
    error than plain float32 on long streams); shapes without a fixed-shape
    kernel use float64 arithmetic on the float32 data, rounded once.
    """
    dt = np.dtype(dtype) if dtype is not None else _device.np_dtype(increments)
    if dt not in (np.float32, np.float64):
        dt = np.dtype(np.float64)
    code = _dtype_code(dt)
    if compensated:
        if dt != np.float32:
            raise TypeError("compensated signatures take float32 data")
        code = _capi.RPY_F32_COMPENSATED
    inc = _rows_of(_device.to_device(increments, dt), width)
    _check_host_offsets(offsets, int(inc.shape[0]))
    off = _device.to_device(np.asarray(offsets, dtype=np.int64) if isinstance(offsets, (list, tuple))
                            else offsets)
    t = _device.torch()
    off = off.to(t.int64).contiguous()
    n = int(off.shape[0]) - 1
    if n < 0:
        raise ValueError("offsets must hold at least one entry")
    tsize = int(_capi.lib().rpy_tensor_size(int(width), int(depth)))
    out = _device.empty((n, tsize), dt)
    _capi.check(_capi.lib().rpy_stream_signature_batch(
        out.data_ptr(), inc.data_ptr(), off.data_ptr(), n, int(width), int(depth), code,
        _device.stream_handle()))
    return _device.to_host(out) if isinstance(increments, np.ndarray) else out


def log_signature_batch(increments, offsets, width: int, depth: int, dtype=None,
                        lie_basis: Optional[LieBasis] = None):
    """Log-signatures in the Hall basis (float64), one row per stream."""
    dt = np.dtype(dtype) if dtype is not None else _device.np_dtype(increments)
    if dt not in (np.float32, np.float64):
        dt = np.dtype(np.float64)
    code = _dtype_code(dt)
    lie_basis = lie_basis or LieBasis(width, depth)
    t2l = lie_basis.get_t2l_matrix(np.float64)
    inc = _rows_of(_device.to_device(increments, dt), width)
    _check_host_offsets(offsets, int(inc.shape[0]))
    t = _device.torch()
    off = _device.to_device(offsets if _device.is_device_array(offsets)
                            else np.asarray(offsets, dtype=np.int64)).to(t.int64).contiguous()
    n = int(off.shape[0]) - 1
    out = _device.empty((n, lie_basis.size()), np.float64)
    _capi.check(_capi.lib().rpy_stream_log_signature_batch(
        out.data_ptr(), inc.data_ptr(), off.data_ptr(), n, int(width), int(depth), code,
        t2l.device_struct(), _device.stream_handle()))
    return _device.to_host(out) if isinstance(increments, np.ndarray) else out


def signature_gram(signatures):
    """Gram matrix of dense signatures, G[i, j] = <S_i, S_j> (float32).

    The signature kernel by signature dot
    (examples/signature-kernel-by-signature-dot.py:15-23) for a batch: one
    MFMA GEMM with bf16 hi/lo split operands and float32 accumulation.
    signatures: (n, k) float32 (NumPy or GPU tensor).
    """
    S = _device.to_device(signatures, np.float32)
    if S.ndim != 2:
        raise ValueError("signatures must be a 2-d array (n, tensor_size)")
    n, k = int(S.shape[0]), int(S.shape[1])
    G = _device.empty((n, n), np.float32)
    _capi.check(_capi.lib().rpy_signature_gram(G.data_ptr(), S.data_ptr(), n, k, k, n,
                                               _device.stream_handle()))
    return _device.to_host(G) if isinstance(signatures, np.ndarray) else G


def cache_root(inf: float, sup: float, resolution: int):
    """Root [o, o + 2^k) of a dyadic cache over [inf, sup) at `resolution`:
    o is inf rounded down to the resolution grid (a multiple of
    2^-resolution, so every stored node is a union of resolution pieces) and
    2^k the smallest power of two reaching sup.  Returns (o, k)."""
    inf, sup = float(inf), float(sup)
    if not (math.isfinite(inf) and math.isfinite(sup)) or not inf < sup:
        raise ValueError("a cache support must be a finite interval [inf, sup) with inf < sup")
    o = math.ldexp(math.floor(math.ldexp(inf, resolution)), -resolution)
    k = math.ceil(math.log2(sup - o))
    while o + math.ldexp(1.0, k) < sup:
        k += 1
    return o, int(k)


class DyadicSignatureCache:
    """Dyadic signature cache of many streams, built on the GPU.

    increments: (rows, width); params: (rows,) float64, strictly increasing
    within each stream; offsets: (n_streams + 1,) CSR offsets.  ``support``
    (inf, sup) is the parameter range the cache covers (default [0, 1)); it
    is held in a root [o, o + 2^k) (cache_root: o is inf rounded down to the
    resolution grid), the node of power 0, and nodes of powers 0..levels
    below it are stored (level-major, power j meaning width 2^(k - j); all
    are unions of resolution pieces).  Rows outside the root are never
    covered.  ``query`` answers clopen interval queries at ``resolution``
    with the semantics of Stream.signature(RealInterval(inf, sup),
    resolution) for streams of that resolution
    (streams/src/stream.cpp:107-171): the Chen product over the rows whose
    param lies in [inf, sup) with both ends rounded down to the resolution
    grid, at any parameter range as the reference's caching layer answers
    (streams/src/dyadic_caching_layer.cpp:86-130).  The device kernels work
    on [0, 1): params and query bounds are mapped by t -> (t - o) 2^-k and
    the resolution becomes resolution + k; params whose map is not exact
    are refused (ValueError).
    """

    def __init__(self, increments, params, offsets, width: int, depth: int,
                 resolution: int = 12, levels: int = 6, dtype=None, support=None):
        dt = np.dtype(dtype) if dtype is not None else _device.np_dtype(increments)
        if dt not in (np.float32, np.float64):
            dt = np.dtype(np.float64)
        self.dtype = dt
        self.width, self.depth = int(width), int(depth)
        self.resolution, self.levels = int(resolution), int(levels)
        t = _device.torch()
        self._inc = _rows_of(_device.to_device(increments, dt), self.width)
        _check_host_offsets(offsets, int(self._inc.shape[0]))
        self._set_root(support)
        self._params = self._to_unit(_device.to_device(params).to(t.float64).contiguous())
        if int(self._params.shape[0]) != int(self._inc.shape[0]):
            raise ValueError("params must hold one value per increment row")
        self._offs = _device.to_device(np.asarray(offsets, dtype=np.int64)
                                       if isinstance(offsets, (list, tuple, np.ndarray))
                                       else offsets).to(t.int64).contiguous()
        self.n_streams = int(self._offs.shape[0]) - 1
        L = _capi.lib()
        n = int(L.rpy_dyadic_cache_elements(self.n_streams, self.width, self.depth, self.levels))
        if n < 0:
            raise ValueError("invalid cache shape")
        self.tsize = int(L.rpy_tensor_size(self.width, self.depth))
        self.cache = _device.empty((n,), dt)
        self.build()

    def _set_root(self, support):
        if support is None:
            self.origin, self.root_power = 0.0, 0
        else:
            inf_, sup_ = _bounds_of(support)
            self.origin, self.root_power = cache_root(inf_, sup_, self.resolution)
        # resolution of the device kernels, on the root mapped to [0, 1)
        self._res_unit = self.resolution + self.root_power
        if not 0 <= self._res_unit <= 30:
            raise ValueError(f"resolution {self.resolution} on a root of width "
                             f"2^{self.root_power} is outside what the cache supports "
                             "(0 <= resolution + log2(root width) <= 30)")
        if not 0 <= self.levels <= min(self._res_unit, 24):
            raise ValueError("levels must lie in 0..min(resolution + log2(root width), 24)")

    def _to_unit(self, params):
        """Params (device float64) on the root mapped to [0, 1); the map is
        exact, and params outside the root stay outside [0, 1)."""
        if self.root_power == 0 and self.origin == 0.0:
            return params
        scale = math.ldexp(1.0, -self.root_power)
        unit = (params - self.origin) * scale
        back = unit * math.ldexp(1.0, self.root_power) + self.origin
        inside = (unit >= 0) & (unit < 1)
        if bool(((back != params) & inside).any()):
            raise ValueError("params are not exactly representable on the cache's root")
        return unit

    def params(self):
        """The stream parameters in their own units (device float64)."""
        if self.root_power == 0 and self.origin == 0.0:
            return self._params
        return self._params * math.ldexp(1.0, self.root_power) + self.origin

    @classmethod
    def _from_state(cls, increments, params, offsets, nodes, width, depth, resolution, levels,
                    dtype, origin=0.0, root_power=0):
        """A cache from stored nodes, without rebuilding them (params in
        their own units)."""
        obj = cls.__new__(cls)
        t = _device.torch()
        obj.dtype = np.dtype(dtype)
        obj.width, obj.depth = int(width), int(depth)
        obj.resolution, obj.levels = int(resolution), int(levels)
        obj.origin, obj.root_power = float(origin), int(root_power)
        obj._res_unit = obj.resolution + obj.root_power
        obj._inc = _device.to_device(np.ascontiguousarray(increments), obj.dtype).reshape(
            -1, obj.width)
        obj._params = obj._to_unit(_device.to_device(np.ascontiguousarray(params,
                                                                         dtype=np.float64)))
        obj._offs = _device.to_device(np.ascontiguousarray(offsets, dtype=np.int64)).to(t.int64)
        obj.n_streams = int(obj._offs.shape[0]) - 1
        L = _capi.lib()
        n = int(L.rpy_dyadic_cache_elements(obj.n_streams, obj.width, obj.depth, obj.levels))
        if n != int(np.asarray(nodes).shape[0]):
            raise ValueError("stored cache size does not match its shape")
        obj.tsize = int(L.rpy_tensor_size(obj.width, obj.depth))
        obj.cache = _device.to_device(np.ascontiguousarray(nodes), obj.dtype)
        return obj

    def dump(self, path) -> None:
        """Writes the cache (nodes and the stream data they came from) to a
        file; the counterpart of DyadicCachingLayer::dump_cache
        (streams/src/dyadic_caching_layer.cpp:183-201)."""
        from . import serialization
        serialization.dump(self, path)

    @classmethod
    def load(cls, path) -> "DyadicSignatureCache":
        """Reads a dumped cache back into device memory without rebuilding
        it (DyadicCachingLayer::load_cache, dyadic_caching_layer.cpp:168-181)."""
        from . import serialization
        obj = serialization.load(path)
        if not isinstance(obj, cls):
            raise ValueError("file does not hold a dyadic signature cache")
        return obj

    def __reduce__(self):
        from . import serialization
        return (serialization.loads, (serialization.dumps(self),))

    def build(self):
        _capi.check(_capi.lib().rpy_dyadic_cache_build(
            self.cache.data_ptr(), self._inc.data_ptr(), self._params.data_ptr(),
            self._offs.data_ptr(), self.n_streams, self.width, self.depth, self._res_unit,
            self.levels, _dtype_code(self.dtype), _device.stream_handle()))

    def node(self, power: int, stream: int, k: int):
        """Stored signature of stream `stream` over the k-th piece of power
        `power` below the root: [o + k w, o + (k+1) w) with o = self.origin
        and w = 2^(root_power - power) (for the default [0, 1) root:
        [k 2^-power, (k+1) 2^-power))."""
        if not (0 <= power <= self.levels and 0 <= k < (1 << power)):
            raise IndexError("node outside the stored levels")
        if not 0 <= stream < self.n_streams:
            raise IndexError(f"stream {stream} outside 0..{self.n_streams - 1}")
        off = ((1 << power) - 1) * self.n_streams + stream * (1 << power) + k
        return self.cache[off * self.tsize:(off + 1) * self.tsize]

    def query(self, stream_ids, inf, sup, out=None):
        """Signatures of [inf[q], sup[q]) on stream stream_ids[q].

        Stream ids given on the host must lie in 0..n_streams-1 (IndexError
        otherwise); ids already on the device are not checked, and one
        outside that range answers the unit tensor, like an empty query.
        `out`, when given, must be a contiguous device tensor of shape
        (queries, tensor_size) and the cache's dtype."""
        t = _device.torch()
        if not _device.is_device_array(stream_ids):
            ids = np.asarray(stream_ids, dtype=np.int64).reshape(-1)
            if ids.size and (ids.min() < 0 or ids.max() >= self.n_streams):
                raise IndexError(f"stream ids must lie in 0..{self.n_streams - 1}")
        sid = _device.to_device(np.asarray(stream_ids, dtype=np.int64)
                                if not _device.is_device_array(stream_ids) else stream_ids)
        sid = sid.to(t.int64).contiguous()
        qi = _device.to_device(np.asarray(inf, dtype=np.float64)
                               if not _device.is_device_array(inf) else inf).to(t.float64).contiguous()
        qs = _device.to_device(np.asarray(sup, dtype=np.float64)
                               if not _device.is_device_array(sup) else sup).to(t.float64).contiguous()
        if self.root_power != 0 or self.origin != 0.0:
            # clip to the root, then map to [0, 1)
            w = math.ldexp(1.0, self.root_power)
            scale = math.ldexp(1.0, -self.root_power)
            qi = (qi.clamp(min=self.origin, max=self.origin + w) - self.origin) * scale
            qs = (qs.clamp(min=self.origin, max=self.origin + w) - self.origin) * scale
        nq = int(sid.shape[0])
        if qi.shape[0] != nq or qs.shape[0] != nq:
            raise ValueError("stream_ids, inf and sup must have the same length")
        if out is None:
            out = _device.empty((nq, self.tsize), self.dtype)
        else:
            if not _device.is_device_array(out):
                raise TypeError("out must be a device tensor")
            if tuple(out.shape) != (nq, self.tsize):
                raise ValueError(f"out must have shape {(nq, self.tsize)}, "
                                 f"not {tuple(out.shape)}")
            if _device.np_dtype(out) != self.dtype:
                raise TypeError(f"out must have dtype {self.dtype}, not {_device.np_dtype(out)}")
            if not out.is_contiguous() or out.device != self.cache.device:
                raise ValueError("out must be contiguous and on the cache's device")
        _capi.check(_capi.lib().rpy_dyadic_cache_query(
            out.data_ptr(), self.cache.data_ptr(), self._inc.data_ptr(), self._params.data_ptr(),
            self._offs.data_ptr(), self.n_streams, self.width, self.depth, self._res_unit,
            self.levels, sid.data_ptr(), qi.data_ptr(), qs.data_ptr(), nq,
            _dtype_code(self.dtype), _device.stream_handle()))
        return out


class LieIncrementStream:
    """A stream given by its increments at parameter values (indices)."""

    def __init__(self, params, rows, width, depth, resolution, support, dtype):
        self._params = params          # strictly increasing float64
        self._rows = rows              # (n, width) in dtype
        self.width = int(width)
        self.depth = int(depth)
        self.resolution = int(resolution)
        self.support = support         # (inf, sup)
        self.dtype = np.dtype(dtype)
        self._dev_rows = None
        self._sel_buf = None  # row-selection indices (pinned), reused by simplify
        self._sel_pinned = None
        self._cache = None    # dyadic signature cache: None not built, False unusable

    @classmethod
    def from_increments(cls, data, indices=None, *, width=None, depth=None, resolution=None,
                        dtype=None, coeffs=None, support=None, ctx=None, vtype=None,
                        include_time=False, schema=None, channel_types=None, **kwargs):
        """Mirror of LieIncrementStream.from_increments
        (roughpy/src/streams/lie_increment_stream.cpp:73-245, keywords as
        roughpy/src/args/kwargs_to_path_metadata.cpp:42-219).

        * Without ctx, width defaults to the data's column count and depth
          must be given: otherwise ValueError (:103-109).
        * indices: None gives 0..N-1 with resolution 0; an int names a data
          column read as cumulative parameter steps, the column staying part
          of the increments (:133-168); an array or sequence gives the
          parameters.  Given indices without a resolution use
          py_param_to_resolution(min index) + 1 (the reference's
          running-minimum quirk, :193-206).
        * include_time appends a channel holding each row's parameter
          increment (index minus the previous row's index, starting from 0;
          streams/src/lie_increment_stream.cpp:135-159), so the width grows
          by one.
        * support replaces the support [min, max + 2^-resolution)
          (Stream::restrict_to, streams/src/stream.cpp:73-80); it must be a
          RealInterval.
        * vtype ("dense" or "sparse") selects the reference's vector storage
          and leaves the values unchanged; schema / channel_types are
          accepted when they describe only increment channels.
        On a repeated index the first row is kept.
        """
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        if ctx is not None:
            width = width if width is not None else ctx.width
            depth = depth if depth is not None else ctx.depth
            if dtype is None:
                dtype = getattr(ctx, "dtype", None)
        _check_vtype(vtype)
        if isinstance(data, (list, tuple)) and data and hasattr(data[0], "data"):
            data = np.stack([np.asarray(getattr(d, "data")) for d in data])
        arr = np.asarray(data)
        if dtype is None:
            dtype = coeffs
        if dtype is None:
            dtype = np.float32 if arr.dtype == np.float32 else np.float64
        dtype = np.dtype(dtype)
        if arr.size == 0:
            arr = np.zeros((0, int(width or 0)), dtype=dtype)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1)
        n = arr.shape[0]
        n_schema = _schema_increment_channels(schema, channel_types)
        if width is None:
            width = n_schema if n_schema is not None else arr.shape[1]
        width = int(width)
        if n_schema is not None and n_schema != width:
            raise ValueError("specified width does not match the schema width")
        if width == 0 or not depth:
            raise ValueError("either ctx or both width and depth must be specified")
        depth = int(depth)
        if arr.shape[1] < width:
            raise ValueError(f"data has {arr.shape[1]} columns, width {width} requested")
        if indices is None:
            idx = np.arange(n, dtype=np.float64)
        elif isinstance(indices, (int, np.integer)) and not isinstance(indices, bool):
            icol = int(indices)
            if icol < 0 or icol >= arr.shape[1]:
                raise RuntimeError("cannot find index column in provided data")
            # running sum of the column (the first row's value is its index)
            idx = np.cumsum(arr[:, icol].astype(np.float64))
        elif isinstance(indices, (np.ndarray, list, tuple)) or hasattr(indices, "__array__"):
            idx = np.asarray(indices, dtype=np.float64).reshape(-1)
        else:
            raise TypeError("unexpected type provided to 'indices' argument")
        if idx.shape[0] == 0:
            # empty indices are the default ones (:170-175)
            idx = np.arange(n, dtype=np.float64)
            indices = None
        if indices is None:
            res = 0  # default indices force resolution 0 whatever was asked
        else:
            if idx.shape[0] != n:
                raise ValueError("mismatch between number of rows in data and number of indices")
            res = resolution
            if res is None and n:
                min_diff = math.inf
                for v in idx:  # `previous` stays 0 in the reference
                    if v - 0.0 < min_diff:
                        min_diff = v - 0.0
                res = py_param_to_resolution(min_diff) + 1
        res = int(res if res is not None else 0)
        rows = np.ascontiguousarray(arr[:, :width], dtype=dtype)
        if include_time:
            prev = np.concatenate([[0.0], idx[:-1]]) if n else idx
            rows = np.concatenate([rows, (idx - prev).astype(dtype)[:, None]], axis=1)
            width += 1
        # first occurrence wins on a repeated index; rows sorted by index
        uniq, first = np.unique(idx, return_index=True)
        params = uniq.astype(np.float64)
        rows = np.ascontiguousarray(rows[first]) if n else rows
        if n:
            sup_support = (float(idx.min()), float(idx.max()) + math.ldexp(1.0, -res))
        else:
            sup_support = (0.0, 0.0)
        if support is not None:
            if not isinstance(support, RealInterval):
                raise TypeError("Support must be an Interval")
            sup_support = (support.inf, support.sup)
        return cls(params, rows, width, depth, res, sup_support, dtype)

    def __reduce__(self):
        """Pickling stores the stream in the binary stream format, as the
        reference pickles a stream's serialised archive
        (roughpy/src/streams/stream.cpp:498-519)."""
        from . import serialization
        return (serialization.loads, (serialization.dumps(self),))

    # -- queries -------------------------------------------------------------
    def _parse(self, interval_or_inf, sup, resolution):
        res = _resolution_arg(resolution)
        if interval_or_inf is None:
            return self.support, res
        if isinstance(interval_or_inf, RealInterval):
            if sup is not None:
                res = _resolution_arg(sup)
            return (interval_or_inf.inf, interval_or_inf.sup), res
        if isinstance(interval_or_inf, (int, float, np.integer, np.floating)) and not isinstance(
                interval_or_inf, bool):
            if sup is None:
                if isinstance(interval_or_inf, (float, np.floating)):
                    raise TypeError("expecting integer resolution")
                return self.support, int(interval_or_inf)
            inf_, sup_ = float(interval_or_inf), float(sup)
            if inf_ > sup_:
                raise ValueError("inf must not be larger than sup")
            return (inf_, sup_), res
        if hasattr(interval_or_inf, "inf") and hasattr(interval_or_inf, "sup"):
            inf_ = interval_or_inf.inf() if callable(interval_or_inf.inf) else interval_or_inf.inf
            sup_ = interval_or_inf.sup() if callable(interval_or_inf.sup) else interval_or_inf.sup
            if sup is not None:
                res = _resolution_arg(sup)
            return (float(inf_), float(sup_)), res
        raise TypeError("unexpected type for argument interval_or_inf")

    def _refine(self, interval, resolution):
        """(a, b, query resolution) of a query clipped to the support
        (Stream::refine_interval, streams/src/stream.cpp:52-68), or None when
        nothing is left."""
        lo, hi = self.support
        inf_, sup_ = interval
        if sup_ < lo or inf_ > hi:
            return None
        a, b = max(inf_, lo), min(sup_, hi)
        if b - a == 0.0:
            return None
        q_res = resolution if resolution is not None else max(
            self.resolution, param_to_resolution(b - a) + 2)
        return a, b, int(q_res)

    def _query_cache(self):
        """The stream's dyadic signature cache, built on first use and kept:
        queries at the stream's resolution read it, as the reference's
        streams memoise their dyadic pieces (DyadicCachingLayer,
        streams/src/dyadic_caching_layer.cpp:86-130).  The finest stored
        nodes hold about 32 rows.  None when the stream has no rows or its
        support and resolution do not fit the cache's range."""
        if self._cache is None:
            self._cache = False
            n = int(self._params.shape[0])
            lo, hi = self.support
            if n and hi > lo:
                try:
                    _, k = cache_root(lo, hi, self.resolution)
                    levels = int(math.log2(n / 32)) if n >= 64 else 0
                    levels = max(0, min(levels, self.resolution + k, 20))
                    self._cache = DyadicSignatureCache(
                        self._rows, self._params, np.array([0, n], dtype=np.int64), self.width,
                        self.depth, resolution=self.resolution, levels=levels, dtype=self.dtype,
                        support=(lo, hi))
                except ValueError:
                    self._cache = False
        return self._cache or None

    def _select(self, interval, resolution):
        """Row indices the query covers (Stream::refine_interval + dyadic plan)."""
        q = self._refine(interval, resolution)
        if q is None:
            return np.zeros(0, dtype=np.int64)
        a, b, q_res = q
        L = _capi.lib()
        n = self._params.shape[0]
        params = np.ascontiguousarray(self._params, dtype=np.float64)
        nr = np.zeros(1, dtype=np.int64)
        cap = 4 * 64 + 8
        ranges = np.zeros(2 * cap, dtype=np.int64)
        _capi.check(L.rpy_stream_select_rows(_capi.dblp(params), n, self.resolution, a, b,
                                             int(q_res), _capi.i64p(ranges), cap,
                                             _capi.i64p(nr)))
        sel = [np.arange(ranges[2 * i], ranges[2 * i + 1]) for i in range(int(nr[0]))]
        return np.concatenate(sel).astype(np.int64) if sel else np.zeros(0, dtype=np.int64)

    def signature(self, interval_or_inf=None, sup=None, resolution=None, *, depth=None,
                  dtype=None, ctx=None) -> DenseFreeTensor:
        interval, res = self._parse(interval_or_inf, sup, resolution)
        depth = int(depth if depth is not None else (ctx.depth if ctx is not None else self.depth))
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        basis = TensorBasis(self.width, depth)
        q = self._refine(interval, res)
        if q is not None and q[2] == self.resolution and depth == self.depth and dt == self.dtype:
            cache = self._query_cache()
            if cache is not None:
                out = cache.query(np.zeros(1, dtype=np.int64), np.array([q[0]]),
                                  np.array([q[1]]))
                return DenseFreeTensor(_device.to_host(out)[0], basis)
        rows = self._rows[self._select(interval, res)]
        if rows.shape[0] == 0:
            data = np.zeros(basis.size(), dtype=dt)
            data[0] = 1
            return DenseFreeTensor(data, basis)
        offs = np.array([0, rows.shape[0]], dtype=np.int64)
        sig = signature_batch(np.ascontiguousarray(rows, dtype=dt), offs, self.width, depth,
                              dtype=dt)
        return DenseFreeTensor(sig[0], basis)

    def signature_derivative(self, interval_or_perturbations, perturbation=None,
                             resolution=None, *, depth=None, ctx=None) -> DenseFreeTensor:
        """Directional derivative of the signature (forward mode), mirror of
        Stream.signature_derivative (streams/src/stream.cpp:173-230) and
        LiteContext::sig_derivative_impl (algebra/src/lite_context.h:
        580-636): for pairs (interval_i, perturbation_i),
            result = result (x) E_i + E_i (x) dexp_{X_i}(P_i),
        E_i the interval's signature, X_i = log E_i, P_i = l2t(perturbation_i),
        dexp_X(P) = sum_{d>=0} (-1)^d ad_X^d(P) / (d+1)!.
        Forms: (interval, perturbation[, resolution]) or
        ([(interval, perturbation), ...][, resolution])."""
        from . import compute as rc
        if perturbation is None or isinstance(perturbation, (int, float, np.integer,
                                                              np.floating)):
            pairs = list(interval_or_perturbations)
            if perturbation is not None:
                resolution = perturbation
        else:
            pairs = [(interval_or_perturbations, perturbation)]
        res = _resolution_arg(resolution)
        depth = int(depth if depth is not None else (ctx.depth if ctx is not None else self.depth))
        tb = TensorBasis(self.width, depth)
        lb = LieBasis(self.width, depth)
        t = _device.require_gpu()

        def T(x):
            return rc.FreeTensor(x, tb)

        def mul(a, b):
            return rc.ft_mul(T(a), T(b)).data

        result = t.zeros((1, tb.size()), dtype=t.float64, device="cuda")
        for ivl, pert in pairs:
            if isinstance(ivl, RealInterval):
                interval = (ivl.inf, ivl.sup)
            elif isinstance(ivl, (tuple, list)):
                interval = (float(ivl[0]), float(ivl[1]))
            else:
                interval = (float(ivl.inf() if callable(ivl.inf) else ivl.inf),
                            float(ivl.sup() if callable(ivl.sup) else ivl.sup))
            rows = self._rows[self._select(interval, res)]
            if rows.shape[0]:
                offs = np.array([0, rows.shape[0]], dtype=np.int64)
                E = signature_batch(t.from_numpy(np.ascontiguousarray(rows, dtype=np.float64)).cuda(),
                                    t.from_numpy(offs).cuda(), self.width, depth,
                                    dtype=np.float64)
            else:
                E = t.zeros((1, tb.size()), dtype=t.float64, device="cuda")
                E[0, 0] = 1.0
            X = rc.ft_log(T(E)).data
            pdata = np.asarray(getattr(pert, "data", pert), dtype=np.float64).reshape(1, -1)
            lie = np.zeros((1, lb.size()))
            n = min(lb.size(), pdata.shape[1])
            lie[0, :n] = pdata[0, :n]
            P = rc.lie_to_tensor(rc.Lie(t.from_numpy(lie).cuda(), lb), tb).data
            ad = mul(X, P) - mul(P, X)
            series = P.clone()
            fact = 1.0
            for d in range(1, depth + 1):
                fact *= d + 1
                series = series + ad * (((-1.0) ** d) / fact)
                ad = mul(X, ad) - mul(ad, X)
            result = mul(result, E) + mul(E, series)
        return DenseFreeTensor(_device.to_host(result)[0], tb)

    def simplify(self, partition: Partition, resolution=None, *, depth=None,
                 ctx=None) -> "PiecewiseAbelianStream":
        """Stream.simplify (streams/src/stream.cpp:232-258, Python binding
        roughpy/src/streams/stream.cpp:406-469): the log-signature of every
        sub-interval of the partition at `resolution` (default: the stream's)
        becomes a piece of a PiecewiseAbelianStream.  All pieces are one
        batched log-signature launch over the covered rows."""
        res = self.resolution if resolution is None else _resolution_arg(resolution)
        depth = int(depth if depth is not None else (ctx.depth if ctx is not None else self.depth))
        n = len(partition)
        mids = np.fromiter(partition._mids, dtype=np.float64, count=n - 1)
        infs = np.concatenate([[float(partition.inf)], mids])
        sups = np.concatenate([mids, [float(partition.sup)]])
        t = _device.require_gpu()
        # the index buffer is pinned host memory kept between calls: a fresh
        # pageable one costs a page fault per 4 KiB on first write and a
        # staged copy to the device, together more than the selection itself
        if self._sel_buf is None:
            self._sel_pinned = t.empty(self._params.shape[0] + 16, dtype=t.int64, pin_memory=True)
            self._sel_buf = self._sel_pinned.numpy()
        idx, offs = self._select_batch(infs, sups, res, self._sel_buf)
        if self._dev_rows is None:
            self._dev_rows = _device.to_device(np.ascontiguousarray(self._rows), self.dtype)
        # a blocking copy: the buffer is rewritten by the next call.  The
        # selection may have needed a larger array of its own; upload that one
        if idx.base is self._sel_buf or idx is self._sel_buf:
            dev_idx = self._sel_pinned[:idx.shape[0]].to(self._dev_rows.device)
        else:
            dev_idx = _device.to_device(np.ascontiguousarray(idx)).to(self._dev_rows.device)
        rows = self._dev_rows.index_select(0, dev_idx)
        lb = LieBasis(self.width, depth)
        lies = log_signature_batch(rows, offs, self.width, depth, dtype=self.dtype, lie_basis=lb)
        del n
        return PiecewiseAbelianStream(infs, sups, lies, self.width, depth, res, self.support)

    def _select_batch(self, infs, sups, resolution, buf=None):
        """Row indices (CSR) that each query [infs[q], sups[q]) covers at
        `resolution`, one native call.  `buf` (int64, at least the stream's
        length + 16) is filled instead of a new array, and the returned
        indices are a view of it."""
        L = _capi.lib()
        params = np.ascontiguousarray(self._params, dtype=np.float64)
        nq = int(infs.shape[0])
        offs = np.zeros(nq + 1, dtype=np.int64)
        lo, hi = self.support

        def call(idx):
            return L.rpy_stream_select_rows_batch(
                _capi.dblp(params), params.shape[0], self.resolution, lo, hi, _capi.dblp(infs),
                _capi.dblp(sups), nq, int(resolution), _capi.i64p(offs),
                None if idx is None else _capi.i64p(idx), 0 if idx is None else idx.shape[0])

        # disjoint queries (a partition) select each row at most once, so one
        # pass into a buffer of the stream's length usually suffices; more rows
        # than that take a counting pass first
        need = params.shape[0] + 16
        idx = buf if buf is not None and buf.shape[0] >= need else np.empty(need, dtype=np.int64)
        rc = call(idx)
        if rc == _capi.RPY_ERR_SHAPE:
            _capi.check(call(None))
            idx = np.empty(max(1, int(offs[-1])), dtype=np.int64)
            rc = call(idx)
        _capi.check(rc)
        return idx[:int(offs[-1])], offs

    def log_signature(self, interval_or_inf=None, sup=None, resolution=None, *, depth=None,
                      dtype=None, ctx=None) -> DenseLie:
        interval, res = self._parse(interval_or_inf, sup, resolution)
        depth = int(depth if depth is not None else (ctx.depth if ctx is not None else self.depth))
        rows = self._rows[self._select(interval, res)]
        lb = LieBasis(self.width, depth)
        if rows.shape[0] == 0:
            return DenseLie(np.zeros(lb.size(), dtype=np.float64), lb)
        offs = np.array([0, rows.shape[0]], dtype=np.int64)
        ls = log_signature_batch(np.ascontiguousarray(rows), offs, self.width, depth,
                                 dtype=self.dtype if dtype is None else dtype, lie_basis=lb)
        return DenseLie(ls[0], lb)


class PiecewiseAbelianStream:
    """A stream given by Lie pieces on consecutive intervals, mirror of
    PiecewiseAbelianStream (streams/src/piecewise_abelian_stream.cpp:
    43-189) queried through Stream (streams/src/stream.cpp:54-171).

    A query [a, b) is trimmed to the support and dissected into dyadic
    intervals at the resolution; on each dyadic interval [c, d) the pieces
    it meets are taken in order -- a piece inside [c, d) whole, any other
    overlapping piece [p, q) scaled by (q - c) / (q - p) (the reference's
    branch order, piecewise_abelian_stream.cpp:170-181, sends partial
    overlaps on either side there) -- and combined by CBH; zero results are
    dropped and the dyadic results combined by CBH again.  Every exponential,
    product, logarithm and basis map is a batched GPU launch.
    """

    def __init__(self, infs, sups, lies, width: int, depth: int, resolution: int,
                 support):
        self._inf = np.asarray(infs, dtype=np.float64).reshape(-1)
        self._sup = np.asarray(sups, dtype=np.float64).reshape(-1)
        t = _device.require_gpu()
        self._lies = _device.to_device(lies, np.float64).reshape(len(self._inf), -1)
        self.width = int(width)
        self.depth = int(depth)
        self.resolution = int(resolution)
        self.support = (float(support[0]), float(support[1]))
        self._tb = TensorBasis(self.width, self.depth)
        self._lb = LieBasis(self.width, self.depth)
        if int(self._lies.shape[1]) != self._lb.size():
            raise ValueError("piece data does not match the Lie basis size")
        del t

    def __reduce__(self):
        from . import serialization
        return (serialization.loads, (serialization.dumps(self),))

    @property
    def pieces(self):
        """[(RealInterval, Lie coefficients as NumPy)] in order."""
        host = _device.to_host(self._lies)
        return [(RealInterval(a, b), host[i]) for i, (a, b) in enumerate(zip(self._inf, self._sup))]

    def _terms(self, c: float, d: float):
        """(piece, multiplier) pairs for the dyadic interval [c, d).  Pieces are
        consecutive, so the scan starts at the first piece ending after c."""
        out = []
        start = int(np.searchsorted(self._sup, c, side="right")) if self._sup.size else 0
        for i in range(start, len(self._inf)):
            pa, pb = self._inf[i], self._sup[i]
            if pa >= d:
                break
            if pb <= c:
                continue
            if pa >= c and pb <= d:
                out.append((i, 1.0))
            else:
                out.append((i, (pb - c) / (pb - pa)))
        return out

    def _query(self, interval, resolution):
        lo, hi = self.support
        inf_, sup_ = interval
        if sup_ < lo or inf_ > hi:
            return None
        a, b = max(inf_, lo), min(sup_, hi)
        if b - a == 0.0:
            return None
        res = resolution if resolution is not None else max(self.resolution,
                                                            param_to_resolution(b - a) + 2)
        return a, b, res

    def log_signature(self, interval_or_inf=None, sup=None, resolution=None, *, depth=None,
                      ctx=None) -> DenseLie:
        interval, res = LieIncrementStream._parse(self, interval_or_inf, sup, resolution)
        if depth is not None and int(depth) != self.depth:
            raise ValueError("a piecewise abelian stream answers at the depth of its pieces")
        t = _device.torch()
        zero = DenseLie(np.zeros(self._lb.size()), self._lb)
        q = self._query(interval, res)
        if q is None:
            return zero
        a, b, res = q
        groups = []
        for k, n in to_dyadic_intervals(a, b, res):
            c, d = math.ldexp(k, -n), math.ldexp(k + 1, -n)
            groups.append(self._terms(c, d))
        # dyadic results: one term is the (scaled) piece, several are CBH'd
        results = [None] * len(groups)
        multi = [g for g in groups if len(g) > 1]
        if multi:
            T = max(len(g) for g in multi)
            X = t.zeros((T, len(multi), self._lb.size()), dtype=t.float64, device="cuda")
            for m, g in enumerate(multi):
                for j, (i, w) in enumerate(g):
                    X[j, m] = self._lies[i] * w
            cbh = _cbh_rows(X, self._tb, self._lb)
        mi = 0
        for gi, g in enumerate(groups):
            if len(g) == 1:
                i, w = g[0]
                results[gi] = self._lies[i] * w
            elif len(g) > 1:
                results[gi] = cbh[mi]
                mi += 1
        kept = [r for r in results if r is not None and bool(t.any(r != 0))]
        if not kept:
            return zero
        if len(kept) == 1:
            return DenseLie(_device.to_host(kept[0]), self._lb)
        X = t.stack(kept).unsqueeze(1)
        return DenseLie(_device.to_host(_cbh_rows(X, self._tb, self._lb)[0]), self._lb)

    def signature(self, interval_or_inf=None, sup=None, resolution=None, *, depth=None,
                  ctx=None) -> DenseFreeTensor:
        """exp of the log-signature (Stream::signature, stream.cpp:156-171)."""
        from . import compute as rc
        ls = self.log_signature(interval_or_inf, sup, resolution, depth=depth, ctx=ctx)
        t = _device.torch()
        L = t.from_numpy(np.asarray(ls.data, dtype=np.float64).reshape(1, -1)).cuda()
        X = rc.lie_to_tensor(rc.Lie(L, self._lb), self._tb).data
        E = rc.ft_exp(rc.FreeTensor(X, self._tb)).data
        return DenseFreeTensor(_device.to_host(E)[0], self._tb)

    def simplify(self, partition: Partition, resolution=None, *, depth=None,
                 ctx=None) -> "PiecewiseAbelianStream":
        """Stream.simplify of a piecewise abelian stream: its log-signature
        on every sub-interval becomes a piece."""
        res = self.resolution if resolution is None else _resolution_arg(resolution)
        n = len(partition)
        bounds = [(partition[i].inf, partition[i].sup) for i in range(n)]
        lies = np.stack([np.asarray(self.log_signature(RealInterval(*bd), res).data)
                         for bd in bounds]) if n else np.zeros((0, self._lb.size()))
        return PiecewiseAbelianStream(np.array([bd[0] for bd in bounds]),
                                      np.array([bd[1] for bd in bounds]), lies, self.width,
                                      self.depth, res, self.support)
