"""ctypes binding of libmtl4g.so (include/mtl4g.h).

The library is the only compute path: if it is missing, importing this module
raises -- there is no CPU fallback.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtl4g.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mtl4g.h")

OK, MAX_ITER = 0, 1
ERR_INVALID, ERR_NOT_SQUARE, ERR_SIZE, ERR_MISSING_DIAG = -1, -2, -3, -4
ERR_BREAKDOWN, ERR_NONFINITE, ERR_OOM, ERR_HIP, ERR_UNSUPPORTED, ERR_TIMEOUT = -5, -6, -7, -8, -9, -10
ASSIGN, PLUS, MINUS = 0, 1, 2
PC_IDENTITY, PC_DIAGONAL, PC_ILU0, PC_IC0 = 0, 1, 2, 3
F64, F32, BF16 = 0, 1, 2
CG, BICGSTAB, CGS, TFQMR, BICGSTAB_2, BICGSTAB_ELL, GMRES = 0, 1, 2, 3, 4, 5, 6


class Iteration(C.Structure):
    _fields_ = [("norm_r0", C.c_double), ("max_iter", C.c_int32), ("rtol", C.c_double),
                ("atol", C.c_double), ("i", C.c_int32), ("error", C.c_int32),
                ("resid", C.c_double)]


IR_RELIABLE, IR_RESTART = 0, 1


class IRParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("value_format", C.c_int32), ("mode", C.c_int32),
                ("inner_rtol", C.c_double), ("inner_max_iter", C.c_int32), ("max_outer", C.c_int32)]


class IRStats(C.Structure):
    _fields_ = [("outer", C.c_int32), ("inner_total", C.c_int32), ("relres", C.c_double)]


SPMV_TILE, SPMV_VECTOR, SPMV_STREAM, SPMV_SCALAR, SPMV_STREAM2, SPMV_SLICED, SPMV_DIA, SPMV_DIA_SYM = 0, 1, 2, 3, 4, 5, 6, 7
SPMV_PANEL = 8
ORDER_TREE, ORDER_MTL4 = 0, 1
SWEEP_AUTO, SWEEP_LEVELS = 0, 1


class BatchItem(C.Structure):
    _fields_ = [("sys", C.c_int32), ("max_iter", C.c_int32), ("b_off", C.c_int64),
                ("x_off", C.c_int64), ("ws_off", C.c_int64), ("rtol", C.c_double),
                ("atol", C.c_double)]


class BatchResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("error", C.c_int32), ("status", C.c_int32),
                ("refreshes", C.c_int32), ("resid", C.c_double), ("norm_r0", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("streams", C.c_int32), ("max_items", C.c_int32), ("precision", C.c_int32),
                ("pc_kind", C.c_int32), ("max_iter", C.c_int32), ("big_rows", C.c_int32),
                ("rtol", C.c_double), ("cap_bytes", C.c_double), ("per_request", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("solves", C.c_int64), ("launches", C.c_int32),
                ("max_inflight", C.c_int32), ("peak_bytes", C.c_double), ("errors", C.c_int32),
                ("big_solves", C.c_int32)]


def declared_functions():
    """Names of every function include/mtl4g.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(mtl4g_\w+)\s*\(", text, re.M)))


_P, _I32, _I64, _D, _INT = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_int
_PP = C.POINTER(C.c_void_p)

_SIGS = {
    "mtl4g_status_string": (C.c_char_p, [_INT]),
    "mtl4g_last_error": (C.c_char_p, []),
    "mtl4g_version": (_INT, []),
    "mtl4g_set_device": (_INT, [_INT]),
    "mtl4g_set_reduction_order": (_INT, [_INT]),
    "mtl4g_get_reduction_order": (_INT, []),
    "mtl4g_set_product_timing": (_INT, [_INT]),
    "mtl4g_product_timing": (_INT, [C.POINTER(_D), C.POINTER(C.c_int64)]),
    "mtl4g_csr_create": (_INT, [_PP, _I32, _I32, _I64, _P, _P, _P, _INT]),
    "mtl4g_csr_destroy": (_INT, [_P]),
    "mtl4g_csr_info": (_INT, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64)]),
    "mtl4g_csr_set_values": (_INT, [_P, _P, _INT]),
    "mtl4g_csr_tune": (_INT, [_P, _I32, _I32]),
    "mtl4g_spmv": (_INT, [_P, _P, _P, _INT, _P]),
    "mtl4g_spmv_f32": (_INT, [_P, _INT, _P, _P, _INT, _P]),
    "mtl4g_spmv_dot": (_INT, [_P, _P, _P, C.POINTER(_D), _P]),
    "mtl4g_spmv_dot_noalloc": (_INT, [_P, _P, _P, C.POINTER(_D), _P]),
    "mtl4g_dot": (_INT, [_I64, _P, _P, C.POINTER(_D), _P]),
    "mtl4g_unary_dot": (_INT, [_I64, _P, C.POINTER(_D), _P]),
    "mtl4g_two_norm": (_INT, [_I64, _P, C.POINTER(_D), _P]),
    "mtl4g_axpy": (_INT, [_I64, _D, _P, _P, _P]),
    "mtl4g_axpy_unary_dot": (_INT, [_I64, _D, _P, _P, C.POINTER(_D), _P]),
    "mtl4g_pc_create": (_INT, [_PP, _P, _INT, _INT]),
    "mtl4g_pc_destroy": (_INT, [_P]),
    "mtl4g_pc_solve": (_INT, [_P, _P, _P, _P]),
    "mtl4g_pc_adjoint_solve": (_INT, [_P, _P, _P, _P]),
    "mtl4g_pc_factor": (_INT, [_P, _INT, C.POINTER(_I32), C.POINTER(_I64), _P, _P, _P]),
    "mtl4g_pc_levels": (_INT, [_P, C.POINTER(_I32), C.POINTER(_I32)]),
    "mtl4g_set_sweep_schedule": (_INT, [_INT]),
    "mtl4g_pc_sweeps": (_INT, [_P, C.POINTER(_INT), C.POINTER(_INT)]),
    "mtl4g_set_chain_tile": (_INT, [_I32, _I32]),
    "mtl4g_pc_chain_tile": (_INT, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)]),
    "mtl4g_pc_refactor": (_INT, [_P, _P]),
    "mtl4g_cg": (_INT, [_P, _P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_bicgstab": (_INT, [_P, _P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_cgs": (_INT, [_P, _P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_tfqmr": (_INT, [_P, _P, _P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_bicgstab_2": (_INT, [_P, _P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_bicgstab_ell": (_INT, [_P, _P, _P, _P, _P, _I32, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_gmres": (_INT, [_P, _P, _P, _P, _P, _I32, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_solver_create": (_INT, [_PP, _P, _P, _INT]),
    "mtl4g_solver_destroy": (_INT, [_P]),
    "mtl4g_solver_solve": (_INT, [_P, _P, _P, C.POINTER(Iteration), _P, _I32, _P]),
    "mtl4g_ir_solve": (_INT, [_P, _P, _P, _P, C.POINTER(Iteration), C.POINTER(IRParams),
                              C.POINTER(IRStats), _P]),
    "mtl4g_malloc": (_INT, [_PP, C.c_size_t]),
    "mtl4g_free": (_INT, [_P]),
    "mtl4g_memcpy": (_INT, [_P, _P, C.c_size_t, _INT, _P]),
    "mtl4g_memset": (_INT, [_P, _INT, C.c_size_t, _P]),
    "mtl4g_stream_synchronize": (_INT, [_P]),
    "mtl4g_csr_set_kernel": (_INT, [_P, _INT, _I32, _I32]),
    "mtl4g_csr_get_kernel": (_INT, [_P, C.POINTER(_INT), C.POINTER(_I32), C.POINTER(_I32)]),
    "mtl4g_csr_grid": (_INT, [_P, _INT, C.POINTER(_I32)]),
    "mtl4g_csr_autotune": (_INT, [_P, C.POINTER(_D), _P]),
    "mtl4g_csr_prepare": (_INT, [_P, _INT, _P]),
    "mtl4g_batch_create": (_INT, [_PP, _I32, _P, _P, _P, _P, _INT]),
    "mtl4g_batch_destroy": (_INT, [_P]),
    "mtl4g_batch_info": (_INT, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I64)]),
    "mtl4g_batch_graph_info": (_INT, [_P, C.POINTER(_I64), C.POINTER(_I64)]),
    "mtl4g_batch_workspace_elems": (_I64, [_I32, _INT]),
    "mtl4g_batch_set_shape": (_INT, [_P, _INT, _I32, _I32]),
    "mtl4g_batch_get_shape": (_INT, [_P, _INT, C.POINTER(_I32), C.POINTER(_I32)]),
    "mtl4g_batch_solve": (_INT, [_P, _INT, _INT, _P, _P, _D, _D, _I32, C.POINTER(BatchResult), _P]),
    "mtl4g_batch_tune": (_INT, [_P, _INT, _INT, _P, _P, _P, _D, _I32, C.POINTER(_D), _P]),
    "mtl4g_batch_solve_items": (_INT, [_P, _INT, _INT, _P, _I32, _P, _P, _P, _P, _P, _P]),
    "mtl4g_server_create": (_INT, [_PP, _P, C.POINTER(ServerParams)]),
    "mtl4g_server_destroy": (_INT, [_P]),
    "mtl4g_server_run": (_INT, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(ServerStats)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmtl4g.so is not built (run __graft_entry__.build() or "
                              "python -m mtl4_amd.build); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


class MtlError(RuntimeError):
    """Base of the errors the C ABI reports (MTL4's exception types)."""
    status = None


class logic_error(MtlError):
    status = ERR_INVALID


class matrix_not_square(MtlError):
    status = ERR_NOT_SQUARE


class incompatible_size(MtlError):
    status = ERR_SIZE


class missing_diagonal(MtlError):
    status = ERR_MISSING_DIAG


class unexpected_orthogonality(MtlError):
    status = ERR_BREAKDOWN


class nonfinite_error(MtlError):
    status = ERR_NONFINITE


class device_error(MtlError):
    status = ERR_HIP


_ERRORS = {c.status: c for c in (logic_error, matrix_not_square, incompatible_size,
                                 missing_diagonal, unexpected_orthogonality, nonfinite_error,
                                 device_error)}


def check(status, what=""):
    """Raise the MTL4-named exception for a negative status; return status >= 0."""
    if status < 0:
        L = lib()
        msg = "%s: %s (%s)" % (what, L.mtl4g_status_string(status).decode(),
                               L.mtl4g_last_error().decode())
        raise _ERRORS.get(status, MtlError)(msg)
    return status
