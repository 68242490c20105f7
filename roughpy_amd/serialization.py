"""Binary format for streams and dyadic signature caches.

The reference serialises streams with cereal binary archives: pickling a
stream stores its archive bytes (roughpy/src/streams/stream.cpp:498-519),
PiecewiseAbelianStream archives its (interval, Lie) pieces
(streams/include/roughpy/streams/piecewise_abelian_stream.h:88-92), and the
dyadic caching layer dumps its node map to a file named by the cache id
(streams/src/dyadic_caching_layer.cpp:168-217).  Cereal's archive layout is
an implementation detail of that library, so this module defines its own
layout carrying the same content, laid out for bulk device transfers: one
fixed header, then flat little-endian arrays that load straight into
device memory.

Layout (all integers and floats little-endian):

    magic    8 bytes  b"RPYAMD\\x00\\x01"
    version  u32      2 (version 1 files, whose caches cover [0, 1), load too)
    kind     u32      1 LieIncrementStream, 2 PiecewiseAbelianStream,
                      3 DyadicSignatureCache
    length   u64      bytes of the body that follows
    body     kind-specific fields:

    kind 1: i32 width, depth, resolution, dtype (4 float32, 8 float64);
            f64 support_inf, support_sup; i64 n;
            f64 params[n]; dtype rows[n * width]
    kind 2: i32 width, depth, resolution, 0; f64 support_inf, support_sup;
            i64 n_pieces, lie_dim; f64 infs[n], sups[n], lies[n * lie_dim]
    kind 3: i32 width, depth, resolution, levels, dtype, root_power;
            i64 n_streams, n_rows, n_nodes_values; f64 origin (version 2);
            i64 offsets[n_streams + 1]; f64 params[n_rows] (their own units);
            dtype increments[n_rows * width]; dtype nodes[n_nodes_values]
            (the cache root is [origin, origin + 2^root_power); version 1
            caches have root_power 0, no origin field and origin 0)

Loading checks the magic, version, kind and body length and raises
ValueError on a mismatch or a truncated buffer.
"""
from __future__ import annotations

import io
import struct

import numpy as np

MAGIC = b"RPYAMD\x00\x01"
VERSION = 2
KIND_LIE_INCREMENT = 1
KIND_PIECEWISE_ABELIAN = 2
KIND_DYADIC_CACHE = 3

_HEADER = struct.Struct("<8sIIQ")

__all__ = ["dumps", "loads", "dump", "load", "MAGIC", "VERSION"]


def _dtype_tag(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return 4
    if dt == np.float64:
        return 8
    raise TypeError(f"unsupported dtype {dt}")


def _tag_dtype(tag: int):
    if tag == 4:
        return np.dtype("<f4")
    if tag == 8:
        return np.dtype("<f8")
    raise ValueError(f"unknown dtype tag {tag}")


class _Reader:
    def __init__(self, buf: memoryview):
        self.buf = buf
        self.pos = 0

    def unpack(self, fmt: str):
        s = struct.Struct(fmt)
        if self.pos + s.size > len(self.buf):
            raise ValueError("truncated stream data")
        out = s.unpack_from(self.buf, self.pos)
        self.pos += s.size
        return out

    def array(self, dtype, count: int) -> np.ndarray:
        dtype = np.dtype(dtype)
        nbytes = int(count) * dtype.itemsize
        if count < 0 or self.pos + nbytes > len(self.buf):
            raise ValueError("truncated stream data")
        out = np.frombuffer(self.buf, dtype=dtype, count=int(count), offset=self.pos).copy()
        self.pos += nbytes
        return out


def _le(a: np.ndarray, dtype) -> bytes:
    return np.ascontiguousarray(a, dtype=np.dtype(dtype).newbyteorder("<")).tobytes()


def _frame(kind: int, body: bytes) -> bytes:
    return _HEADER.pack(MAGIC, VERSION, kind, len(body)) + body


def _body_lie_increment(s) -> bytes:
    out = io.BytesIO()
    out.write(struct.pack("<iiii", s.width, s.depth, s.resolution, _dtype_tag(s.dtype)))
    out.write(struct.pack("<dd", *s.support))
    n = int(s._params.shape[0])
    out.write(struct.pack("<q", n))
    out.write(_le(s._params, np.float64))
    out.write(_le(np.asarray(s._rows).reshape(n, s.width), s.dtype))
    return out.getvalue()


def _body_piecewise(p) -> bytes:
    from . import _device
    lies = _device.to_host(p._lies)
    out = io.BytesIO()
    out.write(struct.pack("<iiii", p.width, p.depth, p.resolution, 0))
    out.write(struct.pack("<dd", *p.support))
    out.write(struct.pack("<qq", lies.shape[0], lies.shape[1]))
    out.write(_le(p._inf, np.float64))
    out.write(_le(p._sup, np.float64))
    out.write(_le(lies, np.float64))
    return out.getvalue()


def _cache_parts(c):
    """Header bytes and the (array, dtype) parts of a cache, nodes last (left
    on the device)."""
    from . import _device
    offs = _device.to_host(c._offs).astype(np.int64)
    params = _device.to_host(c.params()).astype(np.float64)
    inc = _device.to_host(c._inc)
    n_vals = int(c.cache.shape[0])
    head = struct.pack("<iiiiii", c.width, c.depth, c.resolution, c.levels,
                       _dtype_tag(c.dtype), c.root_power)
    head += struct.pack("<qqqd", c.n_streams, params.shape[0], n_vals, c.origin)
    es = np.dtype(c.dtype).itemsize
    length = len(head) + offs.nbytes + params.nbytes + inc.size * es + n_vals * es
    return _HEADER.pack(MAGIC, VERSION, KIND_DYADIC_CACHE, length) + head, \
        [(offs, np.int64), (params, np.float64), (inc, c.dtype)], n_vals * es


def _cache_bytes(c) -> bytearray:
    """The cache in one preallocated buffer; the nodes are copied from the
    device straight into it."""
    from . import _device
    head, parts, node_bytes = _cache_parts(c)
    total = len(head) + sum(a.size * np.dtype(dt).itemsize for a, dt in parts) + node_bytes
    buf = bytearray(total)
    mv = memoryview(buf)
    mv[:len(head)] = head
    pos = len(head)
    for a, dt in parts:
        n = a.size * np.dtype(dt).itemsize
        np.frombuffer(mv[pos:pos + n], dtype=np.dtype(dt)).reshape(a.shape)[...] = a
        pos += n
    if node_bytes:
        t = _device.torch()
        dst = t.frombuffer(mv[pos:pos + node_bytes], dtype=c.cache.dtype)
        dst.copy_(c.cache)
    return buf


def _write_cache(c, f) -> None:
    """Streams the cache to a file, the nodes in 256 MiB device-to-host chunks."""
    from . import _device
    head, parts, node_bytes = _cache_parts(c)
    f.write(head)
    for a, dt in parts:
        f.write(_le(a, dt))
    chunk = max(1, (256 << 20) // np.dtype(c.dtype).itemsize)
    n = int(c.cache.shape[0])
    for i in range(0, n, chunk):
        f.write(_device.to_host(c.cache[i:i + chunk]).tobytes())


def dumps(obj):
    """Serialises a LieIncrementStream, PiecewiseAbelianStream or
    DyadicSignatureCache (bytes; a large cache comes back as the bytearray
    it was assembled in, to avoid one more copy)."""
    from .streams import DyadicSignatureCache, LieIncrementStream, PiecewiseAbelianStream
    if isinstance(obj, LieIncrementStream):
        return _frame(KIND_LIE_INCREMENT, _body_lie_increment(obj))
    if isinstance(obj, PiecewiseAbelianStream):
        return _frame(KIND_PIECEWISE_ABELIAN, _body_piecewise(obj))
    if isinstance(obj, DyadicSignatureCache):
        return bytes(_cache_bytes(obj)) if len(obj.cache) < (64 << 20) else _cache_bytes(obj)
    raise TypeError(f"cannot serialise {type(obj).__name__}")


def _read_header(buf: memoryview):
    if len(buf) < _HEADER.size:
        raise ValueError("truncated stream data")
    magic, version, kind, length = _HEADER.unpack_from(buf, 0)
    if magic != MAGIC:
        raise ValueError("not roughpy_amd stream data (bad magic)")
    if version not in (1, VERSION):
        raise ValueError(f"unsupported format version {version}")
    if len(buf) - _HEADER.size != length:
        raise ValueError("stream data length does not match its header")
    return kind, version, _Reader(buf[_HEADER.size:])


def loads(data):
    """Inverse of dumps."""
    from .streams import DyadicSignatureCache, LieIncrementStream, PiecewiseAbelianStream
    buf = memoryview(bytes(data) if not isinstance(data, (bytes, bytearray, memoryview))
                     else data).cast("B")
    kind, version, r = _read_header(buf)
    if kind == KIND_LIE_INCREMENT:
        width, depth, resolution, tag = r.unpack("<iiii")
        sinf, ssup = r.unpack("<dd")
        (n,) = r.unpack("<q")
        dt = _tag_dtype(tag)
        params = r.array("<f8", n).astype(np.float64)
        rows = r.array(dt, n * width).reshape(n, width).astype(dt.newbyteorder("="))
        return LieIncrementStream(params, rows, width, depth, resolution, (sinf, ssup),
                                  dt.newbyteorder("="))
    if kind == KIND_PIECEWISE_ABELIAN:
        width, depth, resolution, _ = r.unpack("<iiii")
        sinf, ssup = r.unpack("<dd")
        n, dim = r.unpack("<qq")
        infs = r.array("<f8", n)
        sups = r.array("<f8", n)
        lies = r.array("<f8", n * dim).reshape(n, dim)
        return PiecewiseAbelianStream(infs, sups, lies, width, depth, resolution, (sinf, ssup))
    if kind == KIND_DYADIC_CACHE:
        width, depth, resolution, levels, tag, root_power = r.unpack("<iiiiii")
        n_streams, n_rows, n_vals = r.unpack("<qqq")
        (origin,) = r.unpack("<d") if version >= 2 else (0.0,)
        if version < 2:
            root_power = 0
        dt = _tag_dtype(tag)
        offs = r.array("<i8", n_streams + 1)
        params = r.array("<f8", n_rows)
        inc = r.array(dt, n_rows * width).reshape(n_rows, width)
        nodes = r.array(dt, n_vals)
        return DyadicSignatureCache._from_state(inc, params, offs, nodes, width, depth,
                                                resolution, levels, dt.newbyteorder("="),
                                                origin, root_power)
    raise ValueError(f"unknown stream data kind {kind}")


def dump(obj, path) -> None:
    """Writes obj to a file (the dyadic cache is streamed array by array)."""
    from .streams import DyadicSignatureCache
    with open(path, "wb") as f:
        if isinstance(obj, DyadicSignatureCache):
            _write_cache(obj, f)
        else:
            f.write(dumps(obj))


def load(path):
    """Reads what dump wrote."""
    with open(path, "rb") as f:
        return loads(f.read())
