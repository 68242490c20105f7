"""roughpy_amd -- RoughPy's dense free-tensor hot path on AMD MI355X (gfx950).

Modules:
  compute  -- drop-in for roughpy.compute (ft_fma, ft_exp, ft_log, ...)
  streams  -- LieIncrementStream and batched signatures / log-signatures
  bases    -- TensorBasis, LieBasis (Hall basis), sparse l2t / t2l maps
  algebra  -- the main library's FreeTensor, ShuffleTensor, Lie, Context and
              get_context objects over the same kernels

All algebra (products, exp, log, maps, signatures) runs in hand-written HIP
kernels in _lib/libroughpy_amd.so; only sums and scalings of single host
elements are done on the host.
"""
from . import _capi
from .bases import LieBasis, SparseMatrix, TensorBasis, check_lie_basis_data

__version__ = "0.1.0"

__all__ = ["TensorBasis", "LieBasis", "SparseMatrix", "check_lie_basis_data", "native_library"]


def native_library():
    """The loaded native library (raises ImportError if it was not built)."""
    return _capi.lib()
