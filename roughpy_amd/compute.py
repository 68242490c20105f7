"""Dense free-tensor compute functions on the GPU -- a drop-in for roughpy.compute.

Same names, argument meaning and exceptions as roughpy/compute/__init__.py:
91-375.  Data may be NumPy arrays (copied to HBM and back, results written in
place where the reference writes in place) or torch tensors already on the
GPU (used in place, no copies).  Leading dimensions are batch dimensions,
as in roughpy/compute/_src/py_ternary_array_fn.hpp:12-67.  Only float32 and
float64 are supported, as in the reference.
"""
from __future__ import annotations

import typing
from dataclasses import dataclass

import numpy as np

from . import _capi, _device
from .bases import LieBasis, SparseMatrix, TensorBasis, check_lie_basis_data

__all__ = [
    "TensorBasis", "LieBasis", "SparseMatrix", "DenseFreeTensor", "DenseShuffleTensor",
    "DenseLie", "FreeTensor", "ShuffleTensor", "Lie", "check_lie_basis_data", "ft_fma",
    "ft_mul", "ft_inplace_mul", "antipode", "ft_exp", "ft_log", "ft_fmexp",
    "ft_adjoint_left_mul", "lie_to_tensor", "tensor_to_lie", "st_fma", "st_mul",
    "st_adj_mul",
]


@dataclass()
class DenseFreeTensor:
    data: typing.Any
    basis: TensorBasis


@dataclass()
class DenseShuffleTensor:
    data: typing.Any
    basis: TensorBasis


@dataclass()
class DenseLie:
    data: typing.Any
    basis: LieBasis


FreeTensor = DenseFreeTensor
ShuffleTensor = DenseShuffleTensor
Lie = DenseLie


def _check_basis_compat(out_basis, *other_bases):
    out_width = out_basis.width
    for i, basis in enumerate(other_bases):
        if basis.width != out_width:
            raise ValueError(f"Incompatible width for basis {i}")


def _dtype_code(dt: np.dtype, exc=TypeError) -> int:
    if dt == np.float64:
        return _capi.RPY_F64
    if dt == np.float32:
        return _capi.RPY_F32
    raise exc(f"Unsupported dtype {dt}")


def _shape(a):
    return tuple(a.shape)


def _check_leading(*arrays):
    lead = _shape(arrays[0])[:-1]
    for a in arrays[1:]:
        if _shape(a)[:-1] != lead:
            raise ValueError("mismatched batch dimensions")
    return lead


class _Operand:
    """A device view of one argument; writes back to NumPy on `finish`."""

    def __init__(self, a, dtype, size: int, name: str):
        if not (isinstance(a, np.ndarray) or _device.is_device_array(a)):
            if hasattr(a, "shape") and hasattr(a, "dtype"):
                a = np.asarray(a)
            else:
                raise TypeError(f"{name} must be a numpy array or a GPU tensor")
        if _device.np_dtype(a) != dtype:
            raise TypeError(f"{name} has dtype {_device.np_dtype(a)}, expected {dtype}")
        if len(_shape(a)) < 1:
            raise ValueError(f"{name} must have at least one dimension")
        if _shape(a)[-1] < size:
            raise ValueError(f"{name} has {_shape(a)[-1]} coefficients, the basis needs {size}")
        self.src = a
        self.size = size
        dev = _device.to_device(a)
        lead = _shape(a)[:-1]
        n = int(np.prod(lead)) if lead else 1
        self.dev = dev.reshape(n, _shape(a)[-1])
        self.batch = n

    @property
    def ptr(self):
        return self.dev.data_ptr()

    @property
    def stride(self):
        return int(self.dev.shape[-1])

    def finish(self):
        """Propagate device results back into the caller's array."""
        if isinstance(self.src, np.ndarray):
            np.copyto(self.src, _device.to_host(self.dev).reshape(self.src.shape))
        elif self.dev.data_ptr() != self.src.data_ptr():
            self.src.copy_(self.dev.reshape(self.src.shape))


def _new_like(shape, dtype, zero=True):
    return _device.zeros(shape, dtype) if zero else _device.empty(shape, dtype)


def _wrap_result(dev, like):
    """Return a result of the same kind (NumPy or device) as `like`."""
    if isinstance(like, np.ndarray):
        return _device.to_host(dev)
    return dev


def _bounds(lo=0, hi=-1):
    return _capi.DegreeBoundsC(int(lo), int(hi))


def _padded(a, basis_size: int, dtype, name: str):
    """Device copy of `a` zero-extended to basis_size coefficients."""
    op = _Operand(a, dtype, min(basis_size, _shape(a)[-1]), name)
    if op.stride >= basis_size:
        return op
    t = _device.torch()
    pad = _device.zeros((op.batch, basis_size), dtype)
    pad[:, : op.stride] = op.dev
    op.dev = pad
    return op


# ---------------------------------------------------------------------------
# Basic functions
# ---------------------------------------------------------------------------

def ft_fma(a: FreeTensor, b: FreeTensor, c: FreeTensor, **kwargs):
    """Free tensor fused multiply-add: a = b * c + a (in place on a.data).

    As roughpy/compute/__init__.py:145 calls it: the product is taken in
    b's basis and truncated at c's depth (out_depth = c.basis.depth).
    ``compensated=True`` with float32 data accumulates every output
    coefficient in double and rounds it to float once."""
    _check_basis_compat(a.basis, b.basis, c.basis)
    dt = _device.np_dtype(a.data)
    code = _dtype_code(dt)
    if kwargs.get("compensated") and dt == np.float32:
        code = _capi.RPY_F32_COMPENSATED
    _check_leading(a.data, b.data, c.data)
    basis = b.basis
    size = basis.size()
    out = _Operand(a.data, dt, size, "out")
    lhs = _padded(b.data, size, dt, "lhs")
    rhs = _padded(c.data, size, dt, "rhs")
    L = _capi.lib()
    _capi.check(L.rpy_dense_ft_fma(
        out.ptr, lhs.ptr, rhs.ptr, code, basis.c_struct(), _bounds(0, c.basis.depth),
        _bounds(), _bounds(), out.batch, out.stride, lhs.stride, rhs.stride, 1, 1, 1,
        _device.stream_handle()))
    out.finish()


def ft_mul(a: FreeTensor, b: FreeTensor, **kwargs) -> FreeTensor:
    """Free tensor product a * b in a's basis, with b read up to its own depth
    (rhs_depth = b.basis.depth, roughpy/compute/__init__.py:161-163)."""
    _check_basis_compat(a.basis, b.basis)
    dt = _device.np_dtype(a.data)
    code = _dtype_code(dt)
    if kwargs.get("compensated") and dt == np.float32:
        code = _capi.RPY_F32_COMPENSATED
    _check_leading(a.data, b.data)
    basis = a.basis
    size = basis.size()
    lhs = _padded(a.data, size, dt, "lhs")
    rhs = _padded(b.data, size, dt, "rhs")
    res = _new_like((lhs.batch, size), dt)
    _capi.check(_capi.lib().rpy_dense_ft_fma(
        res.data_ptr(), lhs.ptr, rhs.ptr, code, basis.c_struct(), _bounds(), _bounds(),
        _bounds(0, b.basis.depth), lhs.batch, size, lhs.stride, rhs.stride, 1, 1, 1,
        _device.stream_handle()))
    res = res.reshape(_shape(a.data)[:-1] + (size,))
    return DenseFreeTensor(_wrap_result(res, a.data), a.basis)


def ft_inplace_mul(a: FreeTensor, b: FreeTensor, **kwargs):
    """a = a * b in place, truncated at b's depth (out_depth = b.basis.depth,
    roughpy/compute/__init__.py:181)."""
    _check_basis_compat(a.basis, b.basis)
    dt = _device.np_dtype(a.data)
    code = _dtype_code(dt)
    _check_leading(a.data, b.data)
    basis = a.basis
    size = basis.size()
    lhs = _Operand(a.data, dt, size, "lhs")
    rhs = _padded(b.data, size, dt, "rhs")
    _capi.check(_capi.lib().rpy_dense_ft_inplace_mul(
        lhs.ptr, rhs.ptr, code, basis.c_struct(), _bounds(0, b.basis.depth), _bounds(),
        lhs.batch, lhs.stride, rhs.stride, 1, 1, _device.stream_handle()))
    lhs.finish()


def antipode(a: FreeTensor, **kwargs) -> FreeTensor:
    """Antipode of a free tensor (new tensor)."""
    dt = _device.np_dtype(a.data)
    code = _dtype_code(dt)
    size = a.basis.size()
    arg = _Operand(a.data, dt, size, "arg")
    res = _new_like((arg.batch, size), dt, zero=False)  # the kernel writes every coefficient
    _capi.check(_capi.lib().rpy_dense_ft_antipode(
        res.data_ptr(), arg.ptr, code, a.basis.c_struct(), _bounds(), _bounds(), arg.batch,
        size, arg.stride, 1, 1, _device.stream_handle()))
    res = res.reshape(_shape(a.data)[:-1] + (size,))
    return FreeTensor(_wrap_result(res, a.data), a.basis)


def _unary_out(x, out_basis, entry, name):
    out_basis = out_basis or x.basis
    _check_basis_compat(out_basis, x.basis)
    dt = _device.np_dtype(x.data)
    if dt not in (np.float32, np.float64):
        raise ValueError(f"Unsupported dtype {dt}")
    code = _dtype_code(dt, ValueError)
    size = out_basis.size()
    arg = _padded(x.data, size, dt, name)
    res = _new_like((arg.batch, size), dt, zero=False)  # exp and log write every coefficient
    _capi.check(entry(res.data_ptr(), arg.ptr, code, out_basis.c_struct(), _bounds(), _bounds(),
                      arg.batch, size, arg.stride, 1, 1, _device.stream_handle()))
    res = res.reshape(_shape(x.data)[:-1] + (size,))
    return FreeTensor(_wrap_result(res, x.data), out_basis)


def ft_exp(x: FreeTensor, out_basis: TensorBasis | None = None) -> FreeTensor:
    """Exponential of a free tensor (x[0] is ignored)."""
    return _unary_out(x, out_basis, _capi.lib().rpy_dense_ft_exp, "x")


def ft_log(x: FreeTensor, out_basis: TensorBasis | None = None) -> FreeTensor:
    """Logarithm of a free tensor whose unit coefficient is 1."""
    return _unary_out(x, out_basis, _capi.lib().rpy_dense_ft_log, "x")


def ft_fmexp(multiplier: FreeTensor, exponent: FreeTensor,
             out_basis: TensorBasis | None = None) -> FreeTensor:
    """Fused multiply-exponential multiplier * exp(exponent)."""
    out_basis = out_basis or multiplier.basis
    _check_basis_compat(out_basis, multiplier.basis, exponent.basis)
    dt = _device.np_dtype(multiplier.data)
    if dt not in (np.float32, np.float64):
        raise ValueError(f"Unsupported dtype {dt}")
    code = _dtype_code(dt, ValueError)
    _check_leading(multiplier.data, exponent.data)
    size = out_basis.size()
    mul = _padded(multiplier.data, size, dt, "multiplier")
    exp = _padded(exponent.data, size, dt, "exponent")
    res = _new_like((mul.batch, size), dt, zero=False)  # the kernel writes every coefficient
    _capi.check(_capi.lib().rpy_dense_ft_fmexp(
        res.data_ptr(), mul.ptr, exp.ptr, code, out_basis.c_struct(), _bounds(), _bounds(),
        _bounds(), mul.batch, size, mul.stride, exp.stride, 1, 1, 1, _device.stream_handle()))
    res = res.reshape(_shape(multiplier.data)[:-1] + (size,))
    return FreeTensor(_wrap_result(res, multiplier.data), out_basis)


def ft_adjoint_left_mul(op: FreeTensor, arg: ShuffleTensor) -> ShuffleTensor:
    """Adjoint of left multiplication by `op`, applied to the shuffle tensor `arg`.

    Taken in op's basis with out_depth = arg depth, op depth, arg depth
    (roughpy/compute/__init__.py:333-334); the result has arg's shape."""
    _check_basis_compat(op.basis, arg.basis)
    dt = _device.np_dtype(arg.data)
    code = _dtype_code(dt)
    _check_leading(op.data, arg.data)
    basis = op.basis
    size = basis.size()
    a = _padded(arg.data, size, dt, "arg")
    o = _padded(op.data, size, dt, "op")
    res = _new_like((a.batch, size), dt)
    _capi.check(_capi.lib().rpy_dense_ft_adjoint_left_mul(
        res.data_ptr(), o.ptr, a.ptr, code, basis.c_struct(), _bounds(0, arg.basis.depth),
        _bounds(0, op.basis.depth), _bounds(0, arg.basis.depth), a.batch, size, o.stride,
        a.stride, 1, 1, 1, _device.stream_handle()))
    asize = arg.basis.size()
    if size != asize:
        full = _new_like((a.batch, asize), dt)
        n = min(size, asize)
        full[:, :n] = res[:, :n]
        res = full
    res = res.reshape(_shape(arg.data)[:-1] + (asize,))
    return ShuffleTensor(_wrap_result(res, arg.data), arg.basis)


def st_fma(a: ShuffleTensor, b: ShuffleTensor, c: ShuffleTensor) -> ShuffleTensor:
    """Shuffle tensor fused multiply-add: a += b (shuffle) c, in place on a.data;
    returns a (roughpy/compute/__init__.py:203-216)."""
    _check_basis_compat(a.basis, b.basis, c.basis)
    dt = _device.np_dtype(a.data)
    code = _dtype_code(dt)
    _check_leading(a.data, b.data, c.data)
    size = a.basis.size()
    out = _Operand(a.data, dt, size, "out")
    lhs = _padded(b.data, size, dt, "lhs")
    rhs = _padded(c.data, size, dt, "rhs")
    _capi.check(_capi.lib().rpy_dense_st_fma(
        out.ptr, lhs.ptr, rhs.ptr, code, a.basis.c_struct(), out.batch, out.stride, lhs.stride,
        rhs.stride, 1, 1, 1, _device.stream_handle()))
    out.finish()
    return a


def st_mul(lhs: ShuffleTensor, rhs: ShuffleTensor) -> ShuffleTensor:
    """Shuffle product (roughpy/compute/__init__.py:219-232)."""
    _check_basis_compat(lhs.basis, rhs.basis)
    dt = _device.np_dtype(lhs.data)
    _dtype_code(dt)
    res = _new_like(_shape(lhs.data), dt)
    st_fma(ShuffleTensor(res, lhs.basis), lhs, rhs)
    return ShuffleTensor(_wrap_result(res, lhs.data), lhs.basis)


def st_adj_mul(op: ShuffleTensor, arg: FreeTensor) -> FreeTensor:
    """Adjoint of shuffle multiplication by `op`, applied to `arg`:
    <st_adj_mul(op, arg), x> = <arg, op (shuffle) x>
    (roughpy_compute/dense/basic/shuffle_tensor_adjoint_mul.hpp:13-70)."""
    _check_basis_compat(op.basis, arg.basis)
    dt = _device.np_dtype(arg.data)
    code = _dtype_code(dt)
    _check_leading(op.data, arg.data)
    size = arg.basis.size()
    a = _Operand(arg.data, dt, size, "arg")
    o = _padded(op.data, size, dt, "op")
    res = _new_like((a.batch, size), dt)
    _capi.check(_capi.lib().rpy_dense_st_adj_mul(
        res.data_ptr(), o.ptr, a.ptr, code, arg.basis.c_struct(), a.batch, size, o.stride,
        a.stride, 1, 1, 1, _device.stream_handle()))
    res = res.reshape(_shape(arg.data)[:-1] + (size,))
    return FreeTensor(_wrap_result(res, arg.data), arg.basis)


def _apply_map(mat: SparseMatrix, x, dt, out_cols_basis_size, scale):
    code = _dtype_code(dt)
    a = _Operand(x, dt, mat.cols, "arg")
    res = _new_like((a.batch, mat.rows), dt)
    _capi.check(_capi.lib().rpy_dense_sparse_map(
        res.data_ptr(), a.ptr, mat.device_struct(), code, float(scale), a.batch, mat.rows,
        a.stride, 1, 1, _device.stream_handle()))
    return res.reshape(_shape(x)[:-1] + (mat.rows,))


def lie_to_tensor(arg: Lie, tensor_basis: TensorBasis | None = None,
                  scale_factor=None) -> FreeTensor:
    """Embedding of a Lie element in the free tensor algebra."""
    dt = _device.np_dtype(arg.data)
    l2t = arg.basis.get_l2t_matrix(dt)
    tensor_basis = tensor_basis or TensorBasis(arg.basis.width, arg.basis.depth)
    if tensor_basis.width != arg.basis.width:
        raise ValueError("tensor basis width does not match the Lie basis")
    res = _apply_map(l2t, arg.data, dt, tensor_basis.size(),
                     1.0 if scale_factor is None else scale_factor)
    size = tensor_basis.size()
    if res.shape[-1] != size:
        t = _device.torch()
        full = _device.zeros(tuple(res.shape[:-1]) + (size,), dt)
        n = min(size, res.shape[-1])
        full[..., :n] = res[..., :n]
        res = full
    return FreeTensor(_wrap_result(res, arg.data), tensor_basis)


def tensor_to_lie(arg: FreeTensor, lie_basis: LieBasis | None = None,
                  scale_factor=None) -> Lie:
    """Projection of a free tensor onto the Lie algebra (Dynkin map / degree)."""
    lie_basis = lie_basis or LieBasis(arg.basis.width, arg.basis.depth)
    dt = _device.np_dtype(arg.data)
    t2l = lie_basis.get_t2l_matrix(dt)
    x = arg.data
    need = t2l.cols
    if _shape(x)[-1] != need:
        op = _padded(x, need, dt, "arg")
        x = op.dev[:, :need].reshape(_shape(arg.data)[:-1] + (need,))
    res = _apply_map(t2l, x, dt, need, 1.0 if scale_factor is None else scale_factor)
    return Lie(_wrap_result(res, arg.data), lie_basis)
