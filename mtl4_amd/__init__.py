"""mtl4_amd -- MTL4's sparse Krylov path on MI355X (gfx950).

Python mirror of MTL4's interface for this path, over the C ABI of
include/mtl4g.h (libmtl4g.so, hand-written HIP kernels):

    A = compressed2D(n, n, rowptr, col, val)      # mtl::compressed2D<double>
    x = dense_vector(n); b = A * xt               # mtl::dense_vector<double>, w = A*v
    P = pc.diagonal(A)                            # itl::pc::{identity,diagonal,ilu_0,ic_0}
    it = basic_iteration(b, 2000, 1e-10)          # itl::basic_iteration<double>
    err = cg(A, x, b, P, it)                      # itl::cg / itl::bicgstab

Device memory and streams come from torch (plumbing only); every operation
runs in libmtl4g.so and fails loudly when it is missing.
"""
import ctypes as C
import sys

import numpy as np

from . import _lib
from ._lib import (ASSIGN, BF16, BICGSTAB, CG, F32, F64, MINUS, ORDER_MTL4, ORDER_TREE, PLUS,
                   MtlError, check, incompatible_size, logic_error, matrix_not_square,
                   missing_diagonal, nonfinite_error, unexpected_orthogonality)

__all__ = ["compressed2D", "dense_vector", "pc", "solve", "adjoint_solve", "basic_iteration",
           "cyclic_iteration", "noisy_iteration", "cg", "bicgstab", "cgs", "tfqmr", "bicgstab_2",
           "bicgstab_ell", "gmres", "gmres_full", "dot",
           "unary_dot",
           "two_norm", "num_rows", "num_cols", "size", "laplacian_setup", "ir_solve",
           "MtlError", "matrix_not_square", "incompatible_size", "missing_diagonal",
           "unexpected_orthogonality", "nonfinite_error", "logic_error"]


def _torch():
    import torch
    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _L():
    return _lib.lib()


def set_reduction_order(order):
    """ORDER_TREE (default: parallel, deterministic) or ORDER_MTL4 (MTL4's
    sequential unrolled order: FP64 reductions and solves then reproduce
    MTL4's CPU bits; meant for validation on small systems)."""
    check(_L().mtl4g_set_reduction_order(order), "set_reduction_order")


def get_reduction_order():
    return _L().mtl4g_get_reduction_order()


class reduction_order:
    """with reduction_order(ORDER_MTL4): ... -- sets the order, restores it."""

    def __init__(self, order):
        self.order = order

    def __enter__(self):
        self.saved = get_reduction_order()
        set_reduction_order(self.order)
        return self

    def __exit__(self, *exc):
        set_reduction_order(self.saved)
        return False


# ---------------------------------------------------------------- containers
class dense_vector:
    """mtl::dense_vector<double> (vector/dense_vector.hpp:50-269), device resident.

    dense_vector(n), dense_vector(n, value), dense_vector(array_or_tensor)."""

    def __init__(self, n_or_data, value=0.0, dtype=None):
        torch = _torch()
        dt = dtype or torch.float64
        if isinstance(n_or_data, (int, np.integer)):
            self.t = torch.full((int(n_or_data),), float(value), dtype=dt, device="cuda")
        elif isinstance(n_or_data, torch.Tensor):
            self.t = n_or_data.to(device="cuda", dtype=dt).contiguous()
        else:
            self.t = torch.as_tensor(np.ascontiguousarray(n_or_data), dtype=dt).to("cuda")

    @property
    def ptr(self):
        return C.c_void_p(self.t.data_ptr())

    def size(self):
        return self.t.numel()

    def __len__(self):
        return self.t.numel()

    def numpy(self):
        return self.t.cpu().numpy()

    def __getitem__(self, i):
        return float(self.t[i])

    def __setitem__(self, i, v):
        self.t[i] = v

    def assign(self, other):
        """v = w (copy)."""
        self.t.copy_(other.t if isinstance(other, dense_vector) else _torch().as_tensor(other))
        return self


def size(v):
    return v.size()


class compressed2D:
    """mtl::compressed2D<double> with int32 indices (matrix/compressed2D.hpp:258-653).

    compressed2D(m, n, rowptr, col, val) copies the CSR arrays to the device,
    as the raw-pointer constructor (compressed2D.hpp:355-364) copies them."""

    def __init__(self, nrows, ncols=None, rowptr=None, col=None, val=None):
        ncols = nrows if ncols is None else ncols
        if rowptr is None:
            rowptr = np.zeros(nrows + 1, np.int32)
            col = np.zeros(0, np.int32)
            val = np.zeros(0)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        if len(self.rowptr) != nrows + 1:
            raise incompatible_size("rowptr must have nrows+1 entries")
        h = C.c_void_p()
        check(_L().mtl4g_csr_create(C.byref(h), nrows, ncols, len(self.col),
                                    self.rowptr.ctypes.data_as(C.c_void_p),
                                    self.col.ctypes.data_as(C.c_void_p),
                                    self.val.ctypes.data_as(C.c_void_p), 0), "compressed2D")
        self.h = h
        self._n, self._m = nrows, ncols

    @classmethod
    def from_csr(cls, rowptr, col, val, ncols=None):
        n = len(rowptr) - 1
        return cls(n, n if ncols is None else ncols, rowptr, col, val)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(_lib, "_lib", None) is not None:
            _lib._lib.mtl4g_csr_destroy(h)
            self.h = None

    def num_rows(self):
        return self._n

    def num_cols(self):
        return self._m

    def nnz(self):
        return len(self.col)

    def set_values(self, val):
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        check(_L().mtl4g_csr_set_values(self.h, self.val.ctypes.data_as(C.c_void_p), 0))

    def tune(self, rows_per_block=0, nnz_per_block=0):
        """LDS-tile CSR kernel with this row-block partition."""
        check(_L().mtl4g_csr_tune(self.h, rows_per_block, nnz_per_block), "tune")

    def set_kernel(self, kernel, group=0, rows_per_block=0):
        check(_L().mtl4g_csr_set_kernel(self.h, kernel, group, rows_per_block), "set_kernel")

    def get_kernel(self):
        k, g, r = C.c_int(), C.c_int32(), C.c_int32()
        check(_L().mtl4g_csr_get_kernel(self.h, C.byref(k), C.byref(g), C.byref(r)))
        return k.value, g.value, r.value

    def grid(self, value_format=F64):
        """Workgroups of the persistent SpMV grid (0: not a persistent kernel)."""
        g = C.c_int32()
        check(_L().mtl4g_csr_grid(self.h, value_format, C.byref(g)))
        return g.value

    def autotune(self):
        """Time the CSR kernel candidates on this matrix, keep the fastest (us)."""
        us = C.c_double()
        check(_L().mtl4g_csr_autotune(self.h, C.byref(us), _stream()), "autotune")
        return us.value

    def prepare(self, value_format):
        """Build the F32 or BF16 value copy and tune its product now rather
        than in the first mixed-precision solve (F64: nothing to do)."""
        check(_L().mtl4g_csr_prepare(self.h, value_format, _stream()), "prepare")

    def mult(self, v, w, assign=ASSIGN):
        """w (=,+=,-=) A*v  (operation/mat_vec_mult.hpp:613-683)."""
        if v.size() != self._m or w.size() != self._n:
            raise incompatible_size("A*v: sizes %d x %d, v %d, w %d" % (self._n, self._m, v.size(), w.size()))
        check(_L().mtl4g_spmv(self.h, v.ptr, w.ptr, assign, _stream()), "A*v")
        return w

    def __mul__(self, v):
        w = dense_vector(self._n)
        return self.mult(v, w)


def num_rows(A):
    return A.num_rows()


def num_cols(A):
    return A.num_cols()


def laplacian_setup(m, n):
    """mtl::mat::laplacian_setup(A, m, n) (matrix/laplacian_setup.hpp:25-44)."""
    from .generators import laplacian_setup as gen
    rp, ci, va = gen(m, n)
    return compressed2D(m * n, m * n, rp, ci, va)


# ---------------------------------------------------------------- reductions
def dot(v, w):
    """dot(v, w) (operation/dot.hpp:226-232)."""
    if v.size() != w.size():
        raise incompatible_size("dot")
    r = C.c_double()
    check(_L().mtl4g_dot(v.size(), v.ptr, w.ptr, C.byref(r), _stream()), "dot")
    return r.value


def unary_dot(v):
    r = C.c_double()
    check(_L().mtl4g_unary_dot(v.size(), v.ptr, C.byref(r), _stream()), "unary_dot")
    return r.value


def two_norm(v):
    r = C.c_double()
    check(_L().mtl4g_two_norm(v.size(), v.ptr, C.byref(r), _stream()), "two_norm")
    return r.value


# ---------------------------------------------------------------- preconditioners
class _PC:
    kind = None

    def __init__(self, A, value=F64):
        if value in (np.float32, "float", "float32"):
            value = F32
        if value in (np.float64, "double", "float64"):
            value = F64
        h = C.c_void_p()
        check(_L().mtl4g_pc_create(C.byref(h), A.h, self.kind, value), type(self).__name__)
        self.h, self.A, self.value = h, A, value

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(_lib, "_lib", None) is not None:
            _lib._lib.mtl4g_pc_destroy(h)
            self.h = None

    def solve(self, x, y=None):
        y = dense_vector(x.size()) if y is None else y
        check(_L().mtl4g_pc_solve(self.h, x.ptr, y.ptr, _stream()), "solve")
        return y

    def adjoint_solve(self, x, y=None):
        y = dense_vector(x.size()) if y is None else y
        check(_L().mtl4g_pc_adjoint_solve(self.h, x.ptr, y.ptr, _stream()), "adjoint_solve")
        return y

    def _factor(self, which):
        n, nnz = C.c_int32(), C.c_int64()
        check(_L().mtl4g_pc_factor(self.h, which, C.byref(n), C.byref(nnz), None, None, None))
        if nnz.value < 0:
            raise logic_error("factor not present")
        if which == 0:
            val = np.zeros(nnz.value)
            check(_L().mtl4g_pc_factor(self.h, 0, None, None, None, None,
                                       val.ctypes.data_as(C.c_void_p)))
            return val
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        va = np.zeros(nnz.value)
        check(_L().mtl4g_pc_factor(self.h, which, None, None, rp.ctypes.data_as(C.c_void_p),
                                   ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p)))
        return rp, ci, va

    def get_L(self):
        return self._factor(1)

    def get_U(self):
        return self._factor(2)

    def inv_diag(self):
        return self._factor(0)

    def levels(self):
        lo, up = C.c_int32(), C.c_int32()
        check(_L().mtl4g_pc_levels(self.h, C.byref(lo), C.byref(up)))
        return lo.value, up.value

    def refactor(self):
        """Factor again from the matrix's current values (same pattern), on
        the device with the plan made at construction."""
        check(_L().mtl4g_pc_refactor(self.h, _stream()), "refactor")
        return self


class _PCNamespace:
    class identity(_PC):
        """itl::pc::identity (itl/pc/identity.hpp:24-71)."""
        kind = _lib.PC_IDENTITY

    class diagonal(_PC):
        """itl::pc::diagonal (itl/pc/diagonal.hpp:28-103)."""
        kind = _lib.PC_DIAGONAL

    class ilu_0(_PC):
        """itl::pc::ilu_0<Matrix, Value> (itl/pc/ilu_0.hpp:34-92)."""
        kind = _lib.PC_ILU0

    class ic_0(_PC):
        """itl::pc::ic_0<Matrix, Value> (itl/pc/ic_0.hpp:39-194)."""
        kind = _lib.PC_IC0


pc = _PCNamespace


def solve(P, x):
    """solve(P, x) (itl/pc/solver.hpp:23-52)."""
    return P.solve(x)


def adjoint_solve(P, x):
    return P.adjoint_solve(x)


# ---------------------------------------------------------------- iteration control
class basic_iteration:
    """itl::basic_iteration<double> (itl/iteration/basic_iteration.hpp:23-164).

    basic_iteration(r0_vector_or_norm, max_iter, rtol, atol=0)."""

    def __init__(self, r0, max_iter, rtol, atol=0.0):
        self.norm_r0_ = abs(two_norm(r0)) if isinstance(r0, dense_vector) else float(r0)
        self.max_iter = int(max_iter)
        self.rtol_, self.atol_ = float(rtol), float(atol)
        self.i, self.error, self.resid_ = 0, 0, 0.0
        self.is_finished = False
        self.quite, self.suppress = False, False
        self.history = []

    # accessors of basic_iteration
    def iterations(self): return self.i
    def resid(self): return self.resid_
    def relresid(self): return self.resid_ / self.norm_r0_ if self.norm_r0_ else float("nan")
    def norm_r0(self): return self.norm_r0_
    def set_norm_r0(self, v): self.norm_r0_ = float(v)
    def tol(self): return self.rtol_
    def atol(self): return self.atol_
    def max_iterations(self): return self.max_iter
    def set_max_iterations(self, m): self.max_iter = int(m)
    def restart(self): self.i = 0
    def first(self): return self.i <= 1
    def is_converged(self): return self.is_finished and self.error == 0
    def set_quite(self, q): self.quite = q
    def suppress_resume(self, s): self.suppress = s
    def error_code(self): return self.error
    def __int__(self): return self.error_code()
    def fail(self, code, msg=""):
        self.error = code
        return self.error_code()

    def _state(self):
        return _lib.Iteration(self.norm_r0_, self.max_iter, self.rtol_, self.atol_, self.i, 0, 0.0)

    def _update(self, st, hist, method, ell=2):
        self.i, self.error, self.resid_ = st.i, st.error, st.resid
        self.is_finished = True
        self.history = list(hist) if hist is not None else []

    def _wants_history(self, per_iteration=3):
        return 0


class cyclic_iteration(basic_iteration):
    """itl::cyclic_iteration (itl/iteration/cyclic_iteration.hpp:23-100): prints
    'iteration i: resid r' every `cycle` iterations and a final resume."""

    def __init__(self, r0, max_iter, rtol, atol=0.0, cycle=100, out=None):
        super().__init__(r0, max_iter, rtol, atol)
        self.cycle = int(cycle)
        self.out = out or sys.stdout
        self._i0 = 0
        self.multi_print = False

    def set_multi_print(self, m):
        """Print every finished() call of a printing iteration, also several
        with one iteration number (cyclic_iteration.hpp:44-50)."""
        self.multi_print = bool(m)

    def is_multi_print(self):
        return self.multi_print

    def _wants_history(self, per_iteration=3):
        return per_iteration * max(self.max_iter - self.i, 0) + 4

    def _state(self):
        self._i0 = self.i
        return super()._state()

    def _update(self, st, hist, method, ell=2):
        super()._update(st, hist, method)
        if self.quite:
            return
        # finished() calls: CG one per iteration; BiCGSTAB r at the loop head,
        # then s (after ++iter) and r again per iteration; BiCGSTAB(2) the
        # loop head, then r_hat[0] and r_hat[1] after ++iter; BiCGSTAB(l) the
        # loop head, then r_hat[0..l-1]
        counts, i = [], self._i0
        for k in range(len(self.history)):
            if method == _lib.BICGSTAB and k > 0 and k % 2 == 1:
                i += 1
            elif method == _lib.BICGSTAB_2 and k % 3 == 1:
                i += 1
            elif method == _lib.BICGSTAB_ELL and k % (ell + 1) == 1:
                i += 1
            elif method == _lib.GMRES and k % 2 == 1:   # start, end of each cycle of `ell` steps
                i = min(i + (ell if ell > 0 else self.max_iter), self.max_iter)
            elif method in (_lib.CG, _lib.CGS) and k > 0:
                i += 1
            elif method == _lib.TFQMR and k > 1:   # finished(rt) before the loop, then tau
                i += 1
            counts.append(i)
        last = -1
        for i, r in zip(counts, self.history):
            if i % self.cycle == 0 and (self.multi_print or i != last):
                print("iteration %d: resid %g" % (i, r), file=self.out)
                last = i

    def error_code(self):
        if not self.suppress:
            it = max(self.i, 1)
            rel = self.relresid()
            conv = rel ** (1.0 / it) if rel == rel and rel > 0 else float("nan")
            print("finished! error code = %d\n%d iterations\n%g is actual final residual. \n"
                  "%g is actual relative tolerance achieved. \nRelative tol: %g  Absolute tol: %g\n"
                  "Convergence:  %g" % (self.error, self.i, self.resid_, rel, self.rtol_,
                                        self.atol_, conv), file=self.out)
        return self.error


class noisy_iteration(cyclic_iteration):
    """itl::noisy_iteration (itl/iteration/noisy_iteration.hpp:21-30): cycle 1."""

    def __init__(self, r0, max_iter, rtol, atol=0.0, out=None):
        super().__init__(r0, max_iter, rtol, atol, 1, out)


# ---------------------------------------------------------------- solvers
def _krylov(fn, method, A, x, b, P, it, R=None, ell=2):
    if A.num_rows() != A.num_cols():
        raise matrix_not_square("Krylov solve")
    if x.size() != A.num_cols() or b.size() != A.num_rows():
        raise incompatible_size("Krylov solve")
    st = it._state()
    cap = it._wants_history(ell + 1 if method == _lib.BICGSTAB_ELL else 3)
    hist = np.full(max(cap, 1), np.nan)
    ph = P.h if P is not None else None
    hp = hist.ctypes.data_as(C.c_void_p) if cap else None
    if method == _lib.TFQMR:
        status = fn(A.h, ph, R.h if R is not None else None, x.ptr, b.ptr, C.byref(st), hp, cap, _stream())
    elif method in (_lib.BICGSTAB_ELL, _lib.GMRES):
        status = fn(A.h, ph, R.h if R is not None else None, x.ptr, b.ptr, int(ell), C.byref(st), hp, cap,
                    _stream())
    else:
        status = fn(A.h, ph, x.ptr, b.ptr, C.byref(st), hp, cap, _stream())
    h = hist[:cap]
    h = h[~np.isnan(h)] if cap else None
    it._update(st, h, method, ell)
    check(status, {CG: "cg", BICGSTAB: "bicgstab", _lib.CGS: "cgs", _lib.BICGSTAB_2: "bicgstab_2",
                   _lib.BICGSTAB_ELL: "bicgstab_ell", _lib.GMRES: "gmres"}.get(method, "tfqmr"))
    return it.error_code()


def cg(A, x, b, P=None, it=None):
    """itl::cg(A, x, b, P, iter) or cg(A, x, b, iter) (itl/krylov/cg.hpp:38-111)."""
    if isinstance(P, basic_iteration):
        P, it = None, P
    return _krylov(_L().mtl4g_cg, CG, A, x, b, P, it)


def bicgstab(A, x, b, P=None, it=None):
    """itl::bicgstab(A, x, b, P, iter) (itl/krylov/bicgstab.hpp:27-78)."""
    if isinstance(P, basic_iteration):
        P, it = None, P
    return _krylov(_L().mtl4g_bicgstab, BICGSTAB, A, x, b, P, it)


def cgs(A, x, b, P=None, it=None):
    """itl::cgs(A, x, b, P, iter) (itl/krylov/cgs.hpp:24-62)."""
    if isinstance(P, basic_iteration):
        P, it = None, P
    return _krylov(_L().mtl4g_cgs, _lib.CGS, A, x, b, P, it)


def tfqmr(A, x, b, L=None, R=None, it=None):
    """itl::tfqmr(A, x, b, L, R, iter) (itl/krylov/tfqmr.hpp:32-96): left
    preconditioner L, right preconditioner R (None: identity)."""
    return _krylov(_L().mtl4g_tfqmr, _lib.TFQMR, A, x, b, L, it, R)


def bicgstab_2(A, x, b, L=None, it=None):
    """itl::bicgstab_2(A, x, b, L, iter) (itl/krylov/bicgstab_2.hpp:37-130):
    BiCGSTAB(l) with l = 2, left preconditioner L (None: identity).  Returns
    3 (iter.fail(3)) on omega == 0."""
    if isinstance(L, basic_iteration):
        L, it = None, L
    return _krylov(_L().mtl4g_bicgstab_2, _lib.BICGSTAB_2, A, x, b, L, it)


def bicgstab_ell(A, x, b, L, R, it, l):
    """itl::bicgstab_ell(A, x, b, L, R, iter, l) (itl/krylov/bicgstab_ell.hpp:
    37-134): BiCGSTAB(l), 1 <= l <= 8, left preconditioner L and right
    preconditioner R (None: identity).  Returns 3 (iter.fail(3)) on omega == 0."""
    return _krylov(_L().mtl4g_bicgstab_ell, _lib.BICGSTAB_ELL, A, x, b, L, it, R, ell=l)


def gmres(A, x, b, L, R, it, restart):
    """itl::gmres(A, x, b, L, R, iter, restart) (itl/krylov/gmres.hpp:122-137):
    restarted GMRES, left preconditioner L and right preconditioner R (None:
    identity), at most 256 Krylov vectors per cycle.  Returns 2
    (iter.fail(2)) when no direction was found."""
    if int(restart) < 1:
        raise logic_error("gmres: restart must be positive")
    return _krylov(_L().mtl4g_gmres, _lib.GMRES, A, x, b, L, it, R, ell=int(restart))


def gmres_full(A, x, b, L, R, it):
    """itl::gmres_full(A, x, b, L, R, iter) (itl/krylov/gmres.hpp:37-119): one
    cycle of min(n, iterations left) Arnoldi steps."""
    return _krylov(_L().mtl4g_gmres, _lib.GMRES, A, x, b, L, it, R, ell=0)


class Solver:
    """Reusable device workspace for repeated solves with one (A, P)."""

    def __init__(self, A, P=None, method=CG):
        h = C.c_void_p()
        check(_L().mtl4g_solver_create(C.byref(h), A.h, P.h if P is not None else None, method))
        self.h, self.A, self.P, self.method = h, A, P, method

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(_lib, "_lib", None) is not None:
            _lib._lib.mtl4g_solver_destroy(h)
            self.h = None

    def solve(self, x, b, it):
        st = it._state()
        status = _L().mtl4g_solver_solve(self.h, x.ptr, b.ptr, C.byref(st), None, 0, _stream())
        it._update(st, None, self.method)
        check(status, "solve")
        return it.error


def ir_solve(A, x, b, it, P32=None, method=CG, value_format=F32, mode="reliable",
             inner_rtol=None, inner_max_iter=10000, max_outer=50):
    """Mixed-precision refinement: FP64 residual, correction and stopping norm
    on the device; Krylov recurrence in FP32 arithmetic (FP32 or BF16 values).
    mode "reliable": one FP32 recurrence with FP64 residual refreshes when the
    FP32 residual fell by inner_rtol (default 0.1); mode "restart": classic
    refinement with inner solves to inner_rtol (default 1e-5).
    Returns (error_code, stats dict)."""
    md = _lib.IR_RELIABLE if mode == "reliable" else _lib.IR_RESTART
    if inner_rtol is None:
        inner_rtol = 0.1 if md == _lib.IR_RELIABLE else 1e-5
    prm = _lib.IRParams(method, value_format, md, inner_rtol, inner_max_iter, max_outer)
    stats = _lib.IRStats()
    st = it._state()
    status = _L().mtl4g_ir_solve(A.h, P32.h if P32 is not None else None, x.ptr, b.ptr,
                                 C.byref(st), C.byref(prm), C.byref(stats), _stream())
    it._update(st, None, method)
    check(status, "ir_solve")
    return it.error, {"outer": stats.outer, "inner_total": stats.inner_total,
                      "relres": stats.relres}


class Batch:
    """Many independent systems, solved by one persistent launch.

    Batch(systems) with systems a list of (rowptr, col, val) CSR triples (local
    column indices); kept in HBM as one block-diagonal matrix."""

    def __init__(self, systems):
        sys_row = [0]
        rps, cols, vals = [], [], []
        nnz = 0
        for (rp, ci, va) in systems:
            rp = np.asarray(rp, np.int64)
            off = sys_row[-1]
            rps.append(rp[:-1] + nnz)
            cols.append(np.asarray(ci, np.int64) + off)
            vals.append(np.asarray(va, np.float64))
            nnz += int(rp[-1])
            sys_row.append(off + len(rp) - 1)
        self.sys_row = np.asarray(sys_row, np.int32)
        self.rowptr = np.concatenate(rps + [np.array([nnz])]).astype(np.int32)
        self.col = np.concatenate(cols).astype(np.int32)
        self.val = np.concatenate(vals)
        self.nsys = len(systems)
        self.nrows = int(self.sys_row[-1])
        h = C.c_void_p()
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        check(_L().mtl4g_batch_create(C.byref(h), self.nsys, p(self.sys_row), p(self.rowptr),
                                      p(self.col), p(self.val), 0), "batch_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(_lib, "_lib", None) is not None:
            _lib._lib.mtl4g_batch_destroy(h)
            self.h = None

    def solve(self, x, b, pc_kind=_lib.PC_DIAGONAL, precision=F64, rtol=1e-10, atol=0.0,
              max_iter=100000):
        """Every system once; x, b device vectors laid out like the rows.
        Returns a list of result dicts, one per system."""
        res = (_lib.BatchResult * self.nsys)()
        check(_L().mtl4g_batch_solve(self.h, pc_kind, precision, x.ptr, b.ptr, rtol, atol,
                                     max_iter, res, _stream()), "batch_solve")
        return [{"iterations": r.iterations, "error": r.error, "status": r.status,
                 "refreshes": r.refreshes, "resid": r.resid, "norm_r0": r.norm_r0} for r in res]

    def tune(self, x, x0, b, pc_kind=_lib.PC_DIAGONAL, precision=F64, rtol=1e-10, max_iter=100000):
        ms = C.c_double()
        check(_L().mtl4g_batch_tune(self.h, pc_kind, precision, x.ptr, x0.ptr, b.ptr, rtol,
                                    max_iter, C.byref(ms), _stream()), "batch_tune")
        return ms.value, self.shape(precision)

    def shape(self, precision=F64):
        t, pcu = C.c_int32(), C.c_int32()
        check(_L().mtl4g_batch_get_shape(self.h, precision, C.byref(t), C.byref(pcu)))
        return t.value, pcu.value

    def graph_info(self):
        """(captures, replays) of the solve-all graph."""
        c, r = C.c_int64(), C.c_int64()
        check(_L().mtl4g_batch_graph_info(self.h, C.byref(c), C.byref(r)), "graph_info")
        return c.value, r.value

    def set_shape(self, precision, threads, per_cu):
        check(_L().mtl4g_batch_set_shape(self.h, precision, threads, per_cu), "set_shape")


class Server:
    """Serving loop over the matrices of a Batch (continuous batching)."""

    def __init__(self, batch, precision=F32, pc_kind=_lib.PC_DIAGONAL, rtol=1e-10,
                 max_iter=20000, streams=4, max_items=4096, big_rows=65536, cap_bytes=200e9,
                 per_request=False):
        self.batch = batch
        self.prm = _lib.ServerParams(streams, max_items, precision, pc_kind, max_iter, big_rows,
                                     rtol, cap_bytes, int(per_request))
        h = C.c_void_p()
        check(_L().mtl4g_server_create(C.byref(h), batch.h, C.byref(self.prm)), "server_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and getattr(_lib, "_lib", None) is not None:
            _lib._lib.mtl4g_server_destroy(h)
            self.h = None

    @staticmethod
    def solver(method, pc_kind):
        """Solver code of a request: MTL4G_SERVER_SOLVER(method, pc_kind)."""
        return (int(method) << 8) | int(pc_kind)

    def run(self, mat, nrhs, arrival_s, seed, want_x=False, solver=None):
        """Replay a trace; returns (latency_s, iterations per rhs, x list or None, stats).
        solver: per-request codes (Server.solver(method, pc_kind)); None: cg
        with the server's pc_kind."""
        mat = np.ascontiguousarray(mat, np.int32)
        nrhs = np.ascontiguousarray(nrhs, np.int32)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        sd = np.ascontiguousarray(seed, np.uint64)
        sv = None if solver is None else np.ascontiguousarray(solver, np.int32)
        n = len(mat)
        lat = np.zeros(n)
        its = np.zeros(int(nrhs.sum()), np.int32)
        sizes = np.diff(self.batch.sys_row)[mat].astype(np.int64)
        xoff = np.zeros(n + 1, np.int64)
        np.cumsum(sizes * nrhs, out=xoff[1:])
        xo = np.zeros(int(xoff[-1])) if want_x else None
        st = _lib.ServerStats()
        p = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        check(_L().mtl4g_server_run(self.h, n, p(mat), p(nrhs), p(sv), p(arr), p(sd), p(lat), p(its),
                                    p(xo), p(xoff) if want_x else None, C.byref(st)), "server_run")
        xs = None
        if want_x:
            xs = [xo[xoff[r]:xoff[r + 1]].reshape(nrhs[r], sizes[r]) for r in range(n)]
        stats = {f: getattr(st, f) for f, _ in _lib.ServerStats._fields_}
        return lat, its, xs, stats
