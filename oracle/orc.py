"""ctypes binding of the CPU oracle (oracle/mtl4_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product
(mtl4_amd/) never imports it.

Each wrapper names the MTL4 source the C routine restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmtl4_oracle.so")

ASSIGN, PLUS, MINUS = 0, 1, 2
PC_IDENTITY, PC_DIAGONAL, PC_ILU0, PC_IC0 = 0, 1, 2, 3
PC_KINDS = {"identity": 0, "diagonal": 1, "ilu_0": 2, "ic_0": 3}

OK, NOT_SQUARE, SIZE_MISMATCH, MISSING_DIAG, BREAKDOWN, OOM, INVALID = 0, -1, -2, -3, -4, -5, -6


class CSR(C.Structure):
    _fields_ = [("n", C.c_int32), ("ncols", C.c_int32),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class Iter(C.Structure):
    _fields_ = [("norm_r0", C.c_double), ("max_iter", C.c_int32),
                ("rtol", C.c_double), ("atol", C.c_double),
                ("i", C.c_int32), ("error", C.c_int32), ("resid", C.c_double)]


def build():
    """Compile the oracle (gcc, oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P, D, I32, I64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
        L.orc_laplacian_setup.restype = I64
        L.orc_laplacian_setup.argtypes = [I32, I32, P, P, P]
        L.orc_spmv.restype = None
        L.orc_spmv.argtypes = [P, P, P, C.c_int]
        L.orc_dot.restype = D
        L.orc_dot.argtypes = [I64, P, P, C.c_int]
        L.orc_unary_dot.restype = D
        L.orc_unary_dot.argtypes = [I64, P, C.c_int]
        L.orc_two_norm.restype = D
        L.orc_two_norm.argtypes = [I64, P]
        L.orc_fused_spmv_dot.restype = D
        L.orc_fused_spmv_dot.argtypes = [P, P, P]
        L.orc_fused_axpy_unary_dot.restype = D
        L.orc_fused_axpy_unary_dot.argtypes = [I64, D, P, P]
        L.orc_pc_create.restype = P
        L.orc_pc_create.argtypes = [C.c_int, P, C.c_int, C.POINTER(C.c_int)]
        L.orc_pc_free.restype = None
        L.orc_pc_free.argtypes = [P]
        L.orc_pc_solve.restype = C.c_int
        L.orc_pc_solve.argtypes = [P, P, P]
        L.orc_pc_adjoint_solve.restype = C.c_int
        L.orc_pc_adjoint_solve.argtypes = [P, P, P]
        L.orc_pc_factor.restype = I64
        L.orc_pc_factor.argtypes = [P, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p)]
        L.orc_lower_unit.restype = None
        L.orc_lower_unit.argtypes = [P, P, P]
        L.orc_lower_invdiag_last.restype = C.c_int
        L.orc_lower_invdiag_last.argtypes = [P, P, P]
        L.orc_upper_invdiag_first.restype = None
        L.orc_upper_invdiag_first.argtypes = [P, P, P]
        L.orc_cg.restype = C.c_int
        L.orc_cg.argtypes = [P, P, P, P, C.POINTER(Iter), P, I32]
        L.orc_cg_diag_omp.restype = C.c_int
        L.orc_cg_diag_omp.argtypes = [P, P, P, P, C.POINTER(Iter), C.c_int]
        L.orc_omp_threads.restype = C.c_int
        L.orc_omp_threads.argtypes = []
        L.orc_bicgstab.restype = C.c_int
        L.orc_bicgstab.argtypes = [P, P, P, P, C.POINTER(Iter), P, I32]
        L.orc_cgs.restype = C.c_int
        L.orc_cgs.argtypes = [P, P, P, P, C.POINTER(Iter), P, I32]
        L.orc_tfqmr.restype = C.c_int
        L.orc_tfqmr.argtypes = [P, P, P, P, P, C.POINTER(Iter), P, I32]
        L.orc_bicgstab_2.restype = C.c_int
        L.orc_bicgstab_2.argtypes = [P, P, P, P, C.POINTER(Iter), P, I32]
        L.orc_bicgstab_ell.restype = C.c_int
        L.orc_bicgstab_ell.argtypes = [P, P, P, P, P, I32, C.POINTER(Iter), P, I32]
        L.orc_gmres.restype = C.c_int
        L.orc_gmres.argtypes = [P, P, P, P, P, I32, C.POINTER(Iter), P, I32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Matrix:
    """CSR operand held in numpy arrays (int32 indices, float64 values)."""

    def __init__(self, rowptr, col, val, ncols=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = _f64(val)
        self.n = len(self.rowptr) - 1
        self.ncols = self.n if ncols is None else int(ncols)
        self.c = CSR(self.n, self.ncols, _p(self.rowptr).value, _p(self.col).value,
                     _p(self.val).value)
        self.c._keep = (self.rowptr, self.col, self.val)   # the struct holds raw pointers

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def ref(self):
        return C.byref(self.c)


def laplacian_setup(m, n):
    """mtl::mat::laplacian_setup(A, m, n), matrix/laplacian_setup.hpp:25-44."""
    L = lib()
    nnz = L.orc_laplacian_setup(m, n, None, None, None)
    rp = np.zeros(m * n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    va = np.zeros(nnz, np.float64)
    L.orc_laplacian_setup(m, n, _p(rp), _p(ci), _p(va))
    return Matrix(rp, ci, va)


def spmv(A, x, y=None, mode=ASSIGN):
    x = _f64(x)
    y = np.zeros(A.n) if y is None else y
    assert y.dtype == np.float64 and y.flags.c_contiguous
    lib().orc_spmv(A.ref(), _p(x), _p(y), mode)
    return y


def dot(a, b, unroll=4):
    a, b = _f64(a), _f64(b)
    return lib().orc_dot(len(a), _p(a), _p(b), unroll)


def unary_dot(a, unroll=8):
    a = _f64(a)
    return lib().orc_unary_dot(len(a), _p(a), unroll)


def two_norm(a):
    a = _f64(a)
    return lib().orc_two_norm(len(a), _p(a))


def fused_spmv_dot(A, p):
    p = _f64(p)
    q = np.zeros(A.n)
    d = lib().orc_fused_spmv_dot(A.ref(), _p(p), _p(q))
    return q, d


def fused_axpy_unary_dot(alpha, q, r):
    q = _f64(q)
    assert r.dtype == np.float64 and r.flags.c_contiguous
    return lib().orc_fused_axpy_unary_dot(len(r), alpha, _p(q), _p(r))


class Preconditioner:
    """itl::pc::{identity,diagonal,ilu_0,ic_0}<Matrix, Value>(A)."""

    def __init__(self, kind, A, value_bytes=8):
        self.kind = PC_KINDS[kind] if isinstance(kind, str) else kind
        self.A = A
        st = C.c_int(0)
        self.h = lib().orc_pc_create(self.kind, A.ref(), value_bytes, C.byref(st))
        self.status = st.value
        if not self.h:
            raise RuntimeError("orc_pc_create failed with status %d" % st.value)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pc_free(self.h)
            self.h = None

    def solve(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        st = lib().orc_pc_solve(self.h, _p(x), _p(y))
        if st:
            raise RuntimeError("orc_pc_solve status %d" % st)
        return y

    def adjoint_solve(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        st = lib().orc_pc_adjoint_solve(self.h, _p(x), _p(y))
        if st:
            raise RuntimeError("orc_pc_adjoint_solve status %d" % st)
        return y

    def factor(self, which):
        """which: 'inv_diag', 'L' or 'U' -> numpy arrays (copies)."""
        w = {"inv_diag": 0, "L": 1, "U": 2}[which]
        rp, ci, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nnz = lib().orc_pc_factor(self.h, w, C.byref(rp), C.byref(ci), C.byref(va))
        if nnz < 0:
            raise ValueError("factor %s not present" % which)
        val = np.ctypeslib.as_array(C.cast(va, C.POINTER(C.c_double)), (nnz,)).copy()
        if w == 0:
            return val
        n = self.A.n
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (n + 1,)).copy()
        col = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
        return Matrix(rowptr, col, val)


def iteration(b=None, max_iter=1000, rtol=1e-10, atol=0.0, norm_r0=None, i0=0):
    """basic_iteration(b, max_iter, rtol, atol), basic_iteration.hpp:31-39."""
    if norm_r0 is None:
        norm_r0 = abs(two_norm(b))
    return Iter(norm_r0, max_iter, rtol, atol, i0, 0, 0.0)


def _solve(fn, A, b, P, x0, it, hist_cap):
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    hist = np.full(max(hist_cap, 1), np.nan)
    st = fn(A.ref(), P.h if P is not None else None, _p(x), _p(b), C.byref(it), _p(hist), hist_cap)
    nh = int(np.argmax(np.isnan(hist))) if np.isnan(hist).any() else len(hist)
    return x, st, hist[:min(nh, hist_cap)]


def cg(A, b, P=None, x0=None, it=None, hist_cap=0, **kw):
    """itl::cg(A, x, b, P, iter), itl/krylov/cg.hpp:38-111.

    Returns (x, status, iter, history)."""
    it = iteration(b, **kw) if it is None else it
    x, st, hist = _solve(lib().orc_cg, A, b, P, x0, it, hist_cap)
    return x, st, it, hist


def cg_diag_omp(A, b, threads=0, x0=None, it=None, **kw):
    """itl::cg + pc::diagonal as MTL4 runs it built with MTL_WITH_OPENMP
    (mtl4_oracle_omp.c); threads > 0 sets the OpenMP team size.
    Returns (x, status, iter)."""
    it = iteration(b, **kw) if it is None else it
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    inv = Preconditioner("diagonal", A).factor("inv_diag")
    st = lib().orc_cg_diag_omp(A.ref(), _p(inv), _p(x), _p(b), C.byref(it), threads)
    return x, st, it


def omp_threads():
    return lib().orc_omp_threads()


def bicgstab(A, b, P=None, x0=None, it=None, hist_cap=0, **kw):
    """itl::bicgstab(A, x, b, P, iter), itl/krylov/bicgstab.hpp:27-78."""
    it = iteration(b, **kw) if it is None else it
    x, st, hist = _solve(lib().orc_bicgstab, A, b, P, x0, it, hist_cap)
    return x, st, it, hist


def cgs(A, b, P=None, x0=None, it=None, hist_cap=0, **kw):
    """itl::cgs(A, x, b, P, iter), itl/krylov/cgs.hpp:24-62."""
    it = iteration(b, **kw) if it is None else it
    x, st, hist = _solve(lib().orc_cgs, A, b, P, x0, it, hist_cap)
    return x, st, it, hist


def tfqmr(A, b, L=None, R=None, x0=None, it=None, hist_cap=0, **kw):
    """itl::tfqmr(A, x, b, L, R, iter), itl/krylov/tfqmr.hpp:32-96."""
    it = iteration(b, **kw) if it is None else it
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    hist = np.full(max(hist_cap, 1), np.nan)
    st = lib().orc_tfqmr(A.ref(), L.h if L is not None else None, R.h if R is not None else None,
                         _p(x), _p(b), C.byref(it), _p(hist), hist_cap)
    nh = int(np.argmax(np.isnan(hist))) if np.isnan(hist).any() else len(hist)
    return x, st, it, hist[:min(nh, hist_cap)]


def bicgstab_2(A, b, L=None, x0=None, it=None, hist_cap=0, **kw):
    """itl::bicgstab_2(A, x, b, L, iter), itl/krylov/bicgstab_2.hpp:37-130."""
    it = iteration(b, **kw) if it is None else it
    x, st, hist = _solve(lib().orc_bicgstab_2, A, b, L, x0, it, hist_cap)
    return x, st, it, hist


def bicgstab_ell(A, b, L=None, R=None, l=2, x0=None, it=None, hist_cap=0, **kw):
    """itl::bicgstab_ell(A, x, b, L, R, iter, l), itl/krylov/bicgstab_ell.hpp:37-134."""
    it = iteration(b, **kw) if it is None else it
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    hist = np.full(max(hist_cap, 1), np.nan)
    st = lib().orc_bicgstab_ell(A.ref(), L.h if L is not None else None, R.h if R is not None else None,
                                _p(x), _p(b), int(l), C.byref(it), _p(hist), hist_cap)
    nh = int(np.argmax(np.isnan(hist))) if np.isnan(hist).any() else len(hist)
    return x, st, it, hist[:min(nh, hist_cap)]


def gmres(A, b, L=None, R=None, restart=30, x0=None, it=None, hist_cap=0, **kw):
    """itl::gmres(A, x, b, L, R, iter, restart) (restart > 0) or itl::gmres_full
    (restart == 0), itl/krylov/gmres.hpp:37-137."""
    it = iteration(b, **kw) if it is None else it
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    hist = np.full(max(hist_cap, 1), np.nan)
    st = lib().orc_gmres(A.ref(), L.h if L is not None else None, R.h if R is not None else None,
                         _p(x), _p(b), int(restart), C.byref(it), _p(hist), hist_cap)
    nh = int(np.argmax(np.isnan(hist))) if np.isnan(hist).any() else len(hist)
    return x, st, it, hist[:min(nh, hist_cap)]
