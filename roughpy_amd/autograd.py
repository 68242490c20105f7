"""Differentiable signatures and log-signatures on the GPU (torch.autograd).

* ``signature(increments, offsets, width, depth)`` -- batched signatures of
  uneven streams whose backward is the hand-written reverse-Chen
  vector-Jacobian product (rpy_stream_signature_vjp): prefixes are rebuilt
  by multiplying with exp(-x), nothing is stored per step.
* ``ft_mul(a, b, depth)`` -- batched truncated tensor product with backward
  grad_a = R*_b(g), grad_b = L*_a(g) (adjoints of right / left
  multiplication; the left one is free_tensor_adjoint_left_mul.hpp:10-77).
* ``log_signature(...)`` -- t2l(log(S)) (lite_context.h:899-902), the log
  taken in float64 by rpy_dense_ft_log (free_tensor_log.hpp:12-40) and the
  t2l projection a sparse map.  Its backward is hand-written: t2l^T (the
  transposed sparse map), the reverse Horner chain of the log
  (rpy_dense_ft_log_vjp), then reverse Chen (rpy_stream_signature_vjp).

All arithmetic runs in the library's HIP kernels; torch only carries device
memory, streams and the autograd graph.
"""
from __future__ import annotations

import numpy as np

from . import _capi, _device
from .bases import LieBasis, SparseMatrix, TensorBasis

__all__ = ["signature", "log_signature", "ft_mul", "sparse_map"]


def _torch():
    return _device.require_gpu()


def _code(t) -> int:
    torch = _torch()
    if t.dtype == torch.float64:
        return _capi.RPY_F64
    if t.dtype == torch.float32:
        return _capi.RPY_F32
    raise TypeError(f"unsupported dtype {t.dtype}")


def _make_signature_fn():
    torch = _torch()

    class _Signature(torch.autograd.Function):
        @staticmethod
        def forward(ctx, increments, offsets, width, depth):
            inc = increments.contiguous()
            offs = offsets.to(torch.int64).contiguous()
            n = int(offs.shape[0]) - 1
            tsize = int(_capi.lib().rpy_tensor_size(width, depth))
            out = torch.empty((n, tsize), dtype=inc.dtype, device=inc.device)
            _capi.check(_capi.lib().rpy_stream_signature_batch(
                out.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, width, depth, _code(inc),
                _device.stream_handle()))
            ctx.save_for_backward(inc, offs, out)
            ctx.wd = (width, depth)
            return out

        @staticmethod
        def backward(ctx, grad_out):
            inc, offs, out = ctx.saved_tensors
            width, depth = ctx.wd
            g = grad_out.to(inc.dtype).contiguous()
            grad = torch.zeros_like(inc)
            n = int(offs.shape[0]) - 1
            _capi.check(_capi.lib().rpy_stream_signature_vjp(
                grad.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, width, depth,
                out.data_ptr(), g.data_ptr(), _code(inc), _device.stream_handle()))
            return grad, None, None, None

    return _Signature


def _make_mul_fn():
    torch = _torch()

    class _FtMul(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a, b, basis):
            a = a.contiguous()
            b = b.contiguous()
            n = a.shape[0]
            size = basis.size()
            out = torch.zeros_like(a)
            B = _capi.DegreeBoundsC
            _capi.check(_capi.lib().rpy_dense_ft_fma(
                out.data_ptr(), a.data_ptr(), b.data_ptr(), _code(a), basis.c_struct(),
                B(0, -1), B(0, -1), B(0, -1), n, size, size, size, 1, 1, 1,
                _device.stream_handle()))
            ctx.save_for_backward(a, b)
            ctx.basis = basis
            return out

        @staticmethod
        def backward(ctx, g):
            a, b = ctx.saved_tensors
            basis = ctx.basis
            g = g.contiguous()
            n, size = a.shape[0], basis.size()
            ga = gb = None
            L = _capi.lib()
            if ctx.needs_input_grad[0]:
                ga = torch.zeros_like(a)
                _capi.check(L.rpy_dense_ft_adjoint_right_mul(
                    ga.data_ptr(), b.data_ptr(), g.data_ptr(), _code(a), basis.c_struct(), n,
                    size, size, size, 1, 1, 1, _device.stream_handle()))
            if ctx.needs_input_grad[1]:
                gb = torch.zeros_like(b)
                B = _capi.DegreeBoundsC
                _capi.check(L.rpy_dense_ft_adjoint_left_mul(
                    gb.data_ptr(), a.data_ptr(), g.data_ptr(), _code(b), basis.c_struct(),
                    B(0, -1), B(0, -1), B(0, -1), n, size, size, size, 1, 1, 1,
                    _device.stream_handle()))
            return ga, gb, None

    return _FtMul


def _apply_map(mat: SparseMatrix, x):
    """y = M x per row of the device tensor x (float64 or float32)."""
    torch = _torch()
    out = torch.zeros((x.shape[0], mat.rows), dtype=x.dtype, device=x.device)
    _capi.check(_capi.lib().rpy_dense_sparse_map(
        out.data_ptr(), x.data_ptr(), mat.device_struct(), _code(x), 1.0, x.shape[0],
        mat.rows, x.shape[1], 1, 1, _device.stream_handle()))
    return out


def _make_map_fn():
    torch = _torch()

    class _Map(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, mat, mat_t):
            ctx.mat_t = mat_t
            return _apply_map(mat, x.contiguous())

        @staticmethod
        def backward(ctx, g):
            return _apply_map(ctx.mat_t, g.contiguous()), None, None

    return _Map


def _make_logsig_fn():
    torch = _torch()

    class _LogSignature(torch.autograd.Function):
        """t2l(log(S(x))) with the hand-written backward: the cotangent g of
        the Lie coefficients goes back through t2l^T (sparse map), then the
        reverse Horner chain of the log (rpy_dense_ft_log_vjp), then reverse
        Chen (rpy_stream_signature_vjp); float64 for the map and the log."""

        @staticmethod
        def forward(ctx, increments, offsets, width, depth, lie_basis):
            inc = increments.contiguous()
            offs = offsets.to(torch.int64).contiguous()
            n = int(offs.shape[0]) - 1
            tb = TensorBasis(width, depth)
            size = tb.size()
            L = _capi.lib()
            S = torch.empty((n, size), dtype=inc.dtype, device=inc.device)
            _capi.check(L.rpy_stream_signature_batch(
                S.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, width, depth, _code(inc),
                _device.stream_handle()))
            # the float64 log and the Hall projection (one kernel for the
            # fixed shapes, the log kept in LDS; float32 signatures are
            # widened as they load)
            t2l = lie_basis.get_t2l_matrix(np.float64)
            out = torch.empty((n, lie_basis.size()), dtype=torch.float64, device=inc.device)
            _capi.check(L.rpy_dense_log_to_lie(
                out.data_ptr(), S.data_ptr(), _code(inc), width, depth, t2l.device_struct(), n,
                lie_basis.size(), size, _device.stream_handle()))
            ctx.save_for_backward(inc, offs, S)
            ctx.args = (width, depth, lie_basis, tb)
            return out

        @staticmethod
        def backward(ctx, g):
            inc, offs, S = ctx.saved_tensors
            width, depth, lie_basis, tb = ctx.args
            n, size = S.shape
            L = _capi.lib()
            t2l = lie_basis.get_t2l_matrix(np.float64)
            if getattr(t2l, "_transposed", None) is None:
                t2l._transposed = _transpose(t2l)
            H = _apply_map(t2l._transposed, g.to(torch.float64).contiguous())
            S64 = S if S.dtype == torch.float64 else S.double()
            hS = torch.empty((n, size), dtype=torch.float64, device=S.device)
            _capi.check(L.rpy_dense_ft_log_vjp(
                hS.data_ptr(), S64.data_ptr(), H.data_ptr(), _capi.RPY_F64, tb.c_struct(), n,
                size, size, size, _device.stream_handle()))
            hS = hS.to(inc.dtype).contiguous()
            grad = torch.zeros_like(inc)
            _capi.check(L.rpy_stream_signature_vjp(
                grad.data_ptr(), inc.data_ptr(), offs.data_ptr(), n, width, depth, S.data_ptr(),
                hS.data_ptr(), _code(inc), _device.stream_handle()))
            return grad, None, None, None, None

    return _LogSignature


_FNS = {}


def _fn(name):
    if name not in _FNS:
        _FNS[name] = {"sig": _make_signature_fn, "mul": _make_mul_fn, "map": _make_map_fn,
                      "logsig": _make_logsig_fn}[name]()
    return _FNS[name]


def signature(increments, offsets, width: int, depth: int):
    """(n, tensor_size) signatures, differentiable w.r.t. increments."""
    torch = _torch()
    inc = _device.to_device(increments) if not isinstance(increments, torch.Tensor) else increments
    offs = _device.to_device(np.asarray(offsets, dtype=np.int64)) if not isinstance(
        offsets, torch.Tensor) else offsets.to(inc.device)
    return _fn("sig").apply(inc, offs, int(width), int(depth))


def ft_mul(a, b, depth: int = None, basis: TensorBasis = None):
    """Batched truncated tensor product a (x) b of (n, size) device tensors."""
    if basis is None:
        raise ValueError("ft_mul needs the tensor basis")
    return _fn("mul").apply(a, b, basis)


def _transpose(m: SparseMatrix) -> SparseMatrix:
    rows = np.repeat(np.arange(m.rows), np.diff(m.offsets))
    order = np.lexsort((rows, m.indices))
    cols_sorted = m.indices[order]
    offs = np.zeros(m.cols + 1, dtype=np.int64)
    np.add.at(offs, cols_sorted + 1, 1)
    offs = np.cumsum(offs)
    return SparseMatrix(m.cols, m.rows, offs, rows[order], m.data[order])


def sparse_map(mat: SparseMatrix, x, mat_t: SparseMatrix = None):
    """y = M x per row of x (device), differentiable (backward applies M^T)."""
    return _fn("map").apply(x, mat, mat_t if mat_t is not None else _transpose(mat))


def log_signature(increments, offsets, width: int, depth: int, lie_basis: LieBasis = None):
    """(n, lie_size) float64 log-signatures in the Hall basis, differentiable
    w.r.t. the increments through the hand-written backward."""
    torch = _torch()
    lie_basis = lie_basis or LieBasis(width, depth)
    inc = _device.to_device(increments) if not isinstance(increments, torch.Tensor) else increments
    offs = _device.to_device(np.asarray(offsets, dtype=np.int64)) if not isinstance(
        offsets, torch.Tensor) else offsets.to(inc.device)
    return _fn("logsig").apply(inc, offs, int(width), int(depth), lie_basis)
