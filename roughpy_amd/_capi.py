"""ctypes binding of libroughpy_amd.so (the C-ABI declared in include/roughpy_amd.h).

Loading the library needs no GPU; every compute entry point does.  Errors
come back as status codes and are raised here with the exception class the
reference bindings use for the same condition.
"""
from __future__ import annotations

import ctypes
import os
from functools import lru_cache

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libroughpy_amd.so")

RPY_OK = 0
RPY_ERR_INVALID_ARGUMENT = 1
RPY_ERR_SHAPE = 2
RPY_ERR_DTYPE = 3
RPY_ERR_BASIS = 4
RPY_ERR_HIP = 5
RPY_ERR_OUT_OF_MEMORY = 6
RPY_ERR_UNSUPPORTED = 7

RPY_F32 = 0
RPY_F64 = 1
RPY_F32_COMPENSATED = 2

# rpy_kernel_op / rpy_kernel_kind
RPY_OP_SIGNATURE = 0
RPY_OP_SIGNATURE_VJP = 1
RPY_KIND_GENERIC = 0
RPY_KIND_LANE_CHUNK = 1
RPY_KIND_STREAM_GROUP = 2
RPY_KIND_WAVE_VJP = 3
RPY_KIND_BLOCK_MERGE = 4
RPY_KIND_STREAM_TABLE = 5

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dbl = ctypes.c_double
_p_i64 = ctypes.POINTER(ctypes.c_int64)
_p_dbl = ctypes.POINTER(ctypes.c_double)


class TensorBasisC(ctypes.Structure):
    _fields_ = [("width", _i32), ("depth", _i32), ("degree_begin", _p_i64)]


class DegreeBoundsC(ctypes.Structure):
    _fields_ = [("min_degree", _i32), ("max_degree", _i32)]


class KernelVariantC(ctypes.Structure):
    """rpy_kernel_variant: kind, prefix, block, waves_per_eu."""
    _fields_ = [("kind", _i32), ("prefix", _i32), ("block", _i32), ("waves_per_eu", _i32)]

    def astuple(self):
        return (self.kind, self.prefix, self.block, self.waves_per_eu)


class LieBasisC(ctypes.Structure):
    _fields_ = [("width", _i32), ("depth", _i32), ("degree_begin", _p_i64), ("data", _p_i64)]


class SparseMatrixC(ctypes.Structure):
    _fields_ = [("rows", _i64), ("cols", _i64), ("nnz", _i64), ("offsets", _vp),
                ("indices", _vp), ("data", _vp), ("compressed_rows", _i32)]


_TB = ctypes.POINTER(TensorBasisC)
_LB = ctypes.POINTER(LieBasisC)
_SM = ctypes.POINTER(SparseMatrixC)
_DB = DegreeBoundsC

# name -> (restype, argtypes); the complete exported surface.
SIGNATURES = {
    "rpy_last_error": (ctypes.c_char_p, []),
    "rpy_version": (_i32, []),
    "rpy_device_count": (_i32, []),
    "rpy_release_workspace": (ctypes.c_int, []),
    "rpy_tensor_size": (_i64, [_i32, _i32]),
    "rpy_lie_basis_size": (_i64, [_i32, _i32]),
    "rpy_lie_basis_build": (ctypes.c_int, [_i32, _i32, _p_i64, _p_i64]),
    "rpy_check_lie_basis_data": (ctypes.c_int, [_LB]),
    "rpy_l2t_matrix": (ctypes.c_int, [_LB, _p_i64, _p_i64, _p_i64, _p_dbl]),
    "rpy_t2l_matrix": (ctypes.c_int, [_LB, _p_i64, _p_i64, _p_i64, _p_dbl]),
    "rpy_dense_ft_fma": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _DB, _DB, _DB,
                                        _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp]),
    "rpy_dense_ft_inplace_mul": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _TB, _DB, _DB,
                                                _i64, _i64, _i64, _i64, _i64, _vp]),
    "rpy_dense_ft_antipode": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _TB, _DB, _DB, _i64, _i64,
                                             _i64, _i64, _i64, _vp]),
    "rpy_dense_ft_adjoint_left_mul": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _DB, _DB,
                                                     _DB, _i64, _i64, _i64, _i64, _i64, _i64,
                                                     _i64, _vp]),
    "rpy_dense_ft_exp": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _TB, _DB, _DB, _i64, _i64, _i64,
                                        _i64, _i64, _vp]),
    "rpy_dense_ft_fmexp": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _DB, _DB, _DB, _i64,
                                          _i64, _i64, _i64, _i64, _i64, _i64, _vp]),
    "rpy_dense_ft_log": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _TB, _DB, _DB, _i64, _i64, _i64,
                                        _i64, _i64, _vp]),
    "rpy_dense_ft_log_vjp": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int,
                                            ctypes.POINTER(TensorBasisC), _i64, _i64, _i64, _i64,
                                            _vp]),
    "rpy_dense_sparse_map": (ctypes.c_int, [_vp, _vp, _SM, ctypes.c_int, _dbl, _i64, _i64,
                                            _i64, _i64, _i64, _vp]),
    "rpy_stream_signature_batch": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32,
                                                  ctypes.c_int, _vp]),
    "rpy_stream_signature_batch_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32,
                                                          ctypes.c_int, _vp]),
    "rpy_stream_log_signature_batch": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32,
                                                      ctypes.c_int, _SM, _vp]),
    "rpy_dense_log_to_lie": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i32, _i32, _SM, _i64, _i64,
                                            _i64, _vp]),
    "rpy_param_to_resolution": (_i32, [_dbl]),
    "rpy_to_dyadic_intervals": (ctypes.c_int, [_dbl, _dbl, _i32, _p_i64,
                                               ctypes.POINTER(_i32), _i32,
                                               ctypes.POINTER(_i32)]),
    "rpy_stream_select_rows": (ctypes.c_int, [_p_dbl, _i64, _i32, _dbl, _dbl, _i32, _p_i64,
                                              _i64, _p_i64]),
    "rpy_stream_select_rows_batch": (ctypes.c_int, [_p_dbl, _i64, _i32, _dbl, _dbl, _p_dbl,
                                                    _p_dbl, _i64, _i32, _p_i64, _p_i64, _i64]),
    "rpy_debug_signature_host_emulation": (ctypes.c_int, [_p_dbl, _p_dbl, _i64, _i32, _i32,
                                                          _i32]),
    "rpy_debug_signature_prefix": (_i32, [_i32, _i32, _i32]),
    "rpy_tune_variants": (_i32, [_i32, _i32, _i32, ctypes.c_int, ctypes.POINTER(KernelVariantC),
                                 _i32]),
    "rpy_tune_select": (ctypes.c_int, [_i32, _i32, _i32, ctypes.c_int,
                                       ctypes.POINTER(KernelVariantC)]),
    "rpy_tune_config": (ctypes.c_int, [_i32, _i32, _i32, ctypes.c_int,
                                       ctypes.POINTER(KernelVariantC), ctypes.POINTER(_i32)]),
    "rpy_dyadic_cache_elements": (_i64, [_i64, _i32, _i32, _i32]),
    "rpy_dyadic_cache_build": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32,
                                              _i32, ctypes.c_int, _vp]),
    "rpy_dyadic_cache_query": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32,
                                              _i32, _vp, _vp, _vp, _i64, ctypes.c_int, _vp]),
    "rpy_debug_dyadic_cover": (_i32, [_i64, _i64, _i32, _p_i64, ctypes.POINTER(_i32), _i32]),
    "rpy_signature_gram": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, _vp]),
    "rpy_stream_signature_vjp": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _vp,
                                                ctypes.c_int, _vp]),
    "rpy_debug_vjp_host_emulation": (ctypes.c_int, [_p_dbl, _p_dbl, _i64, _i32, _i32, _i32,
                                                    _p_dbl, _p_dbl]),
    "rpy_dense_st_fma": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _i64, _i64, _i64,
                                        _i64, _i64, _i64, _i64, _vp]),
    "rpy_dense_st_adj_mul": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _i64, _i64, _i64,
                                            _i64, _i64, _i64, _i64, _vp]),
    "rpy_dense_ft_adjoint_right_mul": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _TB, _i64,
                                                      _i64, _i64, _i64, _i64, _i64, _i64, _vp]),
}


@lru_cache(maxsize=None)
def lib() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"roughpy_amd native library not built ({LIB_PATH}); run "
            "`python -m roughpy_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


_EXC = {
    RPY_ERR_INVALID_ARGUMENT: ValueError,
    RPY_ERR_SHAPE: ValueError,
    RPY_ERR_DTYPE: TypeError,
    RPY_ERR_BASIS: ValueError,
    RPY_ERR_HIP: RuntimeError,
    RPY_ERR_OUT_OF_MEMORY: MemoryError,
    RPY_ERR_UNSUPPORTED: NotImplementedError,
}


def check(rc: int) -> None:
    if rc != RPY_OK:
        msg = lib().rpy_last_error()
        msg = msg.decode() if msg else "unknown error"
        raise _EXC.get(rc, RuntimeError)(msg)


def i64p(arr):
    return arr.ctypes.data_as(_p_i64)


def dblp(arr):
    return arr.ctypes.data_as(_p_dbl)
