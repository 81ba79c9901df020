"""Pins the CPU oracle (oracle/mtl4_oracle.c) before anything is checked
against it: the known-answer values MTL4's own tests assert
(tests/golden/mtl4_kats.json), closed forms, textbook factorizations and
hand-derived cases where the MTL4 summation order gives a distinct value.
CPU only."""
import ctypes
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

import orc

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mtl4_kats.json")))


def dense(A):
    D = np.zeros((A.n, A.ncols))
    for r in range(A.n):
        for j in range(A.rowptr[r], A.rowptr[r + 1]):
            D[r, A.col[j]] += A.val[j]
    return D


def from_dense(D):
    rp, ci, va = [0], [], []
    for r in range(D.shape[0]):
        for c in range(D.shape[1]):
            if D[r, c] != 0:
                ci.append(c)
                va.append(D[r, c])
        rp.append(len(ci))
    return orc.Matrix(rp, ci, va, ncols=D.shape[1])


def from_entries(n, entries):
    D = np.zeros((n, n))
    for r, c, v in entries:
        D[r, c] += v
    return from_dense(D)


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("m,n", KATS["laplacian_setup"]["grids"])
def test_laplacian_setup_kat(m, n):
    A = orc.laplacian_setup(m, n)
    D = dense(A)
    assert A.nnz == 5 * m * n - 2 * m - 2 * n
    if m > 1 and n > 1:      # laplacian_setup_test.cpp:36-42
        assert D[0, 0] == 4 and D[0, 1] == -1 and D[0, n] == -1 and D[1, 0] == -1 and D[n, 0] == -1
        if n > 2:
            assert D[0, 2] == 0 and D[2, 0] == 0
    # closed form: kron(I_m, T_n) + kron(T_m, I_n) with T = tridiag(-1, 2, -1)
    T = lambda k: 2 * np.eye(k) - np.eye(k, k=1) - np.eye(k, k=-1)
    assert np.array_equal(D, np.kron(np.eye(m), T(n)) + np.kron(T(m), np.eye(n)))
    for r in range(A.n):
        cols = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0)


def test_config1_size():
    A = orc.laplacian_setup(64, 64)
    assert (A.n, A.nnz) == (4096, 20224)


# ---------------------------------------------------------------- SpMV
def test_matrix_vector_product_kat():
    k = KATS["matrix_vector_product"]
    A = orc.laplacian_setup(*k["grid"])
    v = dense(A)[k["row"]].copy()
    w = orc.spmv(A, v)
    for i, e in k["assign"].items():
        assert w[int(i)] == e
    orc.spmv(A, v, w, orc.PLUS)
    for i, e in k["plus"].items():
        assert w[int(i)] == e
    orc.spmv(A, v, w, orc.MINUS)
    for i, e in k["minus"].items():
        assert w[int(i)] == e


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_vs_dense(seed):
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (37, 37)) * (rng.random((37, 37)) < 0.2)
    A = from_dense(D)
    x = rng.uniform(-1, 1, 37)
    assert np.allclose(orc.spmv(A, x), D @ x, rtol=1e-14, atol=1e-14)
    y0 = rng.uniform(-1, 1, 37)
    y = y0.copy()
    orc.spmv(A, x, y, orc.MINUS)
    assert np.allclose(y, y0 - D @ x, rtol=1e-14, atol=1e-14)


def test_spmv_row_order():
    # one row holding 1e16, 1, -1e16, 1: left-to-right sum in column order is
    # ((1e16+1)-1e16)+1 = 1 (the first +1 is absorbed)
    A = orc.Matrix([0, 4], [0, 1, 2, 3], [1e16, 1.0, -1e16, 1.0], ncols=4)
    assert orc.spmv(A, np.ones(4))[0] == 1.0


# ---------------------------------------------------------------- reductions
@pytest.mark.parametrize("unroll", KATS["dot"]["unrolls"])
def test_dot_kat(unroll):
    u = np.arange(1, KATS["dot"]["size"] + 1, dtype=np.float64)
    assert orc.dot(u, u, unroll) == KATS["dot"]["expected"]


@pytest.mark.parametrize("unroll", KATS["unary_dot"]["unrolls"])
def test_unary_dot_kat(unroll):
    u = np.arange(1, KATS["unary_dot"]["size"] + 1, dtype=np.float64)
    assert orc.unary_dot(u, unroll) == KATS["unary_dot"]["expected"]
    assert abs(orc.two_norm(u) ** 2 - 285.0) < 1e-3


def test_dot4_grouping():
    # dot<4>: ((t0+t1)+(t2+t3)): (1e16+1) + (-1e16+1) = 1e16 - 1e16 = 0,
    # whereas a left-to-right sum gives 1 and dot<1> (dot_simple) gives 1.
    a = np.array([1e16, 1.0, -1e16, 1.0])
    assert orc.dot(a, np.ones(4), 4) == 0.0
    assert orc.dot(a, np.ones(4), 1) == 1.0
    # remainder goes into t0: n=5, a4 joins t0 = 1e16 -> absorbed
    a5 = np.array([1e16, 0.0, -1e16, 0.0, 1.0])
    assert orc.dot(a5, np.ones(5), 4) == 0.0


def test_unary_dot_right_nested():
    # reduction<8>::finish nests to the right: 1e16 + (1+(1+...)) = 1e16 + 7
    # rounds to 1e16+8; any left-to-right order absorbs each +1 (ties to even).
    x = np.array([1e8, 1, 1, 1, 1, 1, 1, 1.0])
    assert orc.unary_dot(x, 8) == 1e16 + 8
    assert orc.unary_dot(x, 1) == 1e16
    # two_norm<4>: t0 + (t1 + (t2 + t3)) on [1e16, 1, 1, 1] (squares) = 1e16+4 (exact: 1e16+3 ties->+4)
    y = np.array([1e8, 1, 1, 1.0])
    assert orc.two_norm(y) == math.sqrt(1e16 + 4)


def test_fused_kats():
    k = KATS["fuse"]
    n = k["n"]
    B = from_dense(2.0 * np.eye(n))
    v = np.ones(n)
    w, d = orc.fused_spmv_dot(B, v)
    assert abs(d - k["spmv_dot"]) < 1e-12 and np.all(w == 2.0)
    r = np.full(n, 6.0)
    q = np.full(n, 2.0)
    rho = orc.fused_axpy_unary_dot(k["alpha"], q, r)
    assert abs(rho - k["axpy_unary_dot"]) < 1e-9
    x7 = np.full(n, 7.0)
    assert orc.unary_dot(x7, 4) == k["unary_dot_7"]
    assert abs(orc.two_norm(x7) - k["two_norm_7"]) < 1e-4
    assert abs(orc.dot(r, np.full(n, 2.0)) - k["dot_r_x2"]) < 1e-9
    r2 = k["alpha"] * q
    assert abs(orc.dot(r2, q) - k["dot_alpha_q"]) < 1e-9


def test_fused_dot_order():
    # fused lazy_dot sums tmp0+tmp1+tmp2+tmp3 from the left:
    # ((1e16+1)+(-1e16))+1 = 1, unlike dot<4>'s grouping (which gives 0)
    A = from_dense(np.eye(4))
    p = np.array([1e8, 1.0, 1.0, 1.0])
    # p*q = p^2: 1e16, 1, 1, 1 -> ((1e16+1)+1)+1 = 1e16 (ties to even each time)
    q, d = orc.fused_spmv_dot(A, p)
    assert d == 1e16
    # fused unary_dot: (t0+t1)+(t2+t3) = (1e16+1)+2 = 1e16+2
    r = np.array([1e8, 1.0, 1.0, 1.0])
    assert orc.fused_axpy_unary_dot(0.0, np.zeros(4), r) == 1e16 + 2


def test_reductions_vs_fsum():
    rng = np.random.default_rng(5)
    for n in (0, 1, 3, 4, 7, 1000, 1001):
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        exact = math.fsum(a * b)
        assert abs(orc.dot(a, b) - exact) <= 1e-13 * max(1.0, np.abs(a * b).sum())
        assert abs(orc.unary_dot(a) - math.fsum(a * a)) <= 1e-13 * max(1.0, (a * a).sum())


# ---------------------------------------------------------------- trisolves / PCs
def test_inverse_trisolve_kat():
    k = KATS["inverse_trisolve"]
    n = k["n"]
    U = from_entries(n, k["upper_entries"])
    x = np.array(k["x"])
    b = dense(U) @ x
    Ud = dense(U)
    Uinv = Ud.copy()
    np.fill_diagonal(Uinv, 1.0 / np.diag(Ud))      # invert_diagonal(A)
    y = np.zeros(n)
    orc.lib().orc_upper_invdiag_first(from_dense(Uinv).ref(), orc._p(b), orc._p(y))
    assert abs(y[k["check_index"]] - k["check_value"]) < k["tol"]
    assert np.allclose(y, x, rtol=1e-14)
    Lt = from_dense(Uinv.T.copy())                   # B = trans(A), diagonal last
    bt = Ud.T @ x
    z = np.zeros(n)
    assert orc.lib().orc_lower_invdiag_last(Lt.ref(), orc._p(bt), orc._p(z)) == 0
    assert abs(z[k["check_index"]] - k["check_value"]) < k["tol"]
    assert np.allclose(z, x, rtol=1e-14)


def test_trisolve_vs_scipy():
    rng = np.random.default_rng(3)
    n = 40
    Lw = np.tril(rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.3), -1)
    v = rng.standard_normal(n)
    w = np.zeros(n)
    orc.lib().orc_lower_unit(from_dense(Lw).ref(), orc._p(v), orc._p(w))
    assert np.allclose(w, sla.solve_triangular(Lw + np.eye(n), v, lower=True, unit_diagonal=True))
    dia = rng.uniform(1, 2, n)
    Uw = np.triu(rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.3), 1)
    Us = Uw + np.diag(1.0 / dia)                     # stored inverse diagonal
    w2 = np.zeros(n)
    orc.lib().orc_upper_invdiag_first(from_dense(Us).ref(), orc._p(v), orc._p(w2))
    assert np.allclose(w2, sla.solve_triangular(Uw + np.diag(dia), v, lower=False))


def test_ic0_golden_vector():
    k = KATS["ic_0_solve"]
    A = orc.laplacian_setup(*k["grid"])
    P = orc.Preconditioner("ic_0", A, k["value_bytes"])
    y = P.solve(np.array(k["x"], dtype=np.float64))
    assert np.linalg.norm(y - np.array(k["y"])) <= k["tol"]
    assert np.linalg.norm(y - np.array(k["y"])) <= 2e-5   # printed to 6 significant digits


def dense_ic0(A):
    """Textbook IC(0) of ic_0_cg_test.cpp:18-40 (U with the true diagonal)."""
    U = np.triu(A).copy()
    n = len(U)
    for k in range(n):
        dia = U[k, k] = math.sqrt(U[k, k])
        for i in range(k + 1, n):
            U[k, i] /= dia
            d = U[k, i]
            for j in range(k + 1, i + 1):
                if U[j, i] != 0.0:
                    U[j, i] -= d * U[k, j]
    return U


@pytest.mark.parametrize("grid", [(3, 3), (4, 5), (6, 6)])
def test_ic0_vs_dense(grid):
    A = orc.laplacian_setup(*grid)
    P = orc.Preconditioner("ic_0", A, 8)
    Ud = dense_ic0(dense(A))
    Us = dense(P.factor("U"))
    # off-diagonals equal the textbook factor; stored diagonal is inverted
    assert np.allclose(np.triu(Us, 1), np.triu(Ud, 1), atol=1e-14)
    assert np.allclose(np.diag(Us), 1.0 / np.diag(Ud), rtol=1e-14)
    k = KATS["ic_0_cg"]
    if grid == (3, 3):
        for (i, j) in k["u_checks"]:
            ref = Ud[i, j] if i != j else 1.0 / Ud[i, j]
            assert abs(Us[i, j] - ref) < k["tol"]
    # L = trans(U)
    assert np.array_equal(dense(P.factor("L")), Us.T)


def test_ic0_tridiagonal_is_exact():
    # no fill-in is dropped for a tridiagonal SPD matrix: IC(0) is Cholesky
    n = 30
    D = 3 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    A = from_dense(D)
    P = orc.Preconditioner("ic_0", A)
    x = np.random.default_rng(1).standard_normal(n)
    assert np.allclose(P.solve(D @ x), x, rtol=1e-12)


def glibc_rand_matrix(n):
    """ilu_0_test.cpp:16-28 with glibc rand() from its default seed."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    D = np.eye(n)
    pattern = np.eye(n, dtype=bool)
    for _ in range(2 * n):
        r = libc.rand() % n
        c = libc.rand() % n
        D[r, c] += -1
        D[r, r] += 1
        pattern[r, c] = True
    rp, ci, va = [0], [], []
    for r in range(n):
        for c in range(n):
            if pattern[r, c]:
                ci.append(c)
                va.append(D[r, c])
        rp.append(len(ci))
    return orc.Matrix(rp, ci, va), D


def dense_ilu0(D):
    """Textbook ILU(0) of ilu_0_test.cpp:31-50."""
    LU = D.copy()
    n = len(LU)
    for i in range(1, n):
        for k in range(i):
            LU[i, k] /= LU[k, k]
            for j in range(k + 1, n):
                if LU[i, j] != 0:
                    LU[i, j] -= LU[i, k] * LU[k, j]
    return LU


def test_ilu0_rand_kat():
    k = KATS["ilu_0_rand"]
    A, D = glibc_rand_matrix(k["n"])
    P = orc.Preconditioner("ilu_0", A)
    LU = dense_ilu0(D)
    Ls, Us = dense(P.factor("L")), dense(P.factor("U"))
    assert abs(LU[2, 1] - Ls[2, 1]) <= k["dense_tol"]
    assert abs(LU[2, 2] - 1.0 / Us[2, 2]) <= k["dense_tol"]
    n = k["n"]
    x = np.full(n, 3.0)
    Ut = np.triu(Us, 1) + np.diag(1.0 / np.diag(Us))    # U with its true diagonal
    Px = (Ls + np.eye(n)) @ (Ut @ x)
    assert np.linalg.norm(P.solve(Px) - x) <= k["tol"]
    PTx = Ut.T @ ((Ls + np.eye(n)).T @ x)
    assert np.linalg.norm(P.adjoint_solve(PTx) - x) <= k["tol"]


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_ilu0_vs_dense_random(seed):
    rng = np.random.default_rng(seed)
    n = 25
    D = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.2)
    np.fill_diagonal(D, np.abs(D).sum(1) + 1.0)
    A = from_dense(D)
    P = orc.Preconditioner("ilu_0", A)
    LU = dense_ilu0(D)
    Ls, Us = dense(P.factor("L")), dense(P.factor("U"))
    assert np.allclose(np.tril(LU, -1), Ls, atol=1e-12)
    assert np.allclose(np.triu(LU, 1), np.triu(Us, 1), atol=1e-12)
    assert np.allclose(np.diag(LU), 1.0 / np.diag(Us), rtol=1e-12)
    # solve is (L+I)^-1 then U^-1 of those factors
    v = rng.standard_normal(n)
    Ut = np.triu(Us, 1) + np.diag(1.0 / np.diag(Us))
    assert np.allclose(P.solve((Ls + np.eye(n)) @ (Ut @ v)), v, rtol=1e-10)
    assert np.allclose(P.adjoint_solve(Ut.T @ ((Ls + np.eye(n)).T @ v)), v, rtol=1e-10)


def test_ilu0_banded_is_exact():
    n = 20
    D = 4 * np.eye(n) - np.eye(n, k=1) - 2 * np.eye(n, k=-1)
    P = orc.Preconditioner("ilu_0", from_dense(D))
    x = np.random.default_rng(2).standard_normal(n)
    assert np.allclose(P.solve(D @ x), x, rtol=1e-12)


def test_diagonal_pc():
    A = orc.laplacian_setup(4, 5)
    P = orc.Preconditioner("diagonal", A)
    x = np.arange(20.0)
    assert np.array_equal(P.solve(x), (1.0 / 4.0) * x)


def test_missing_diagonal():
    A = orc.Matrix([0, 1, 2], [1, 0], [1.0, 1.0])
    with pytest.raises(RuntimeError):
        orc.Preconditioner("ic_0", A)
    with pytest.raises(RuntimeError):
        orc.Preconditioner("ilu_0", A)


# ---------------------------------------------------------------- solvers
def test_config1_cg_diagonal():
    A = orc.laplacian_setup(64, 64)
    xt = np.random.default_rng(0).uniform(-1, 1, A.n)
    b = orc.spmv(A, xt)
    P = orc.Preconditioner("diagonal", A)
    x, st, it, _ = orc.cg(A, b, P, max_iter=2000, rtol=1e-10)
    assert st == 0 and it.error == 0
    # independent restatement of the same recurrences: 201 iterations
    assert abs(it.i - 201) <= 2
    assert it.resid <= 1e-10 * np.linalg.norm(b)
    true_res = np.linalg.norm(b - dense_mul(A, x)) / np.linalg.norm(b)
    assert true_res < 1e-9
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-7


def dense_mul(A, x):
    return orc.spmv(A, x)


def test_cg_identity_smoke_and_exact():
    k = KATS["cg_identity"]
    A = orc.laplacian_setup(*k["grid"])
    b = orc.spmv(A, np.ones(A.n))
    x, st, it, _ = orc.cg(A, b, None, max_iter=k["max_iter"], rtol=k["rtol"])
    assert st == 0 and np.allclose(x, 1.0, rtol=1e-9)
    assert np.allclose(x, np.linalg.solve(dense(A), b), rtol=1e-9)


def test_ilu0_cg_kat():
    k = KATS["ilu_0_cg"]
    A = orc.laplacian_setup(*k["grid"])
    b = orc.spmv(A, np.ones(A.n))
    P = orc.Preconditioner("ilu_0", A)
    x, st, it, _ = orc.cg(A, b, P, max_iter=k["max_iter"], rtol=k["rtol"])
    assert st == 0 and it.i <= k["max_iterations_allowed"]


def test_ic0_cg_kat():
    k = KATS["ic_0_cg"]
    A = orc.laplacian_setup(*k["grid"])
    b = orc.spmv(A, np.ones(A.n))
    P = orc.Preconditioner("ic_0", A, 4)
    x, st, it, _ = orc.cg(A, b, P, max_iter=k["max_iter"], rtol=k["rtol"])
    assert st == 0 and np.allclose(x, 1.0, atol=1e-4)


def test_absolute_termination_kat():
    k = KATS["absolute_termination"]
    A = orc.laplacian_setup(*k["grid"])
    b = orc.spmv(A, np.ones(A.n))
    P = orc.Preconditioner("diagonal", A)
    x, st, it, _ = orc.cg(A, b, P, max_iter=k["max_iter"], rtol=0.0, atol=k["atol"])
    assert st == 0
    assert np.linalg.norm(b - orc.spmv(A, x)) <= k["atol"]


def test_bicgstab_diagonal_kat():
    k = KATS["bicgstab_diagonal"]
    A = orc.laplacian_setup(*k["grid"])
    b = orc.spmv(A, np.ones(A.n))
    P = orc.Preconditioner("diagonal", A)
    x, st, it, _ = orc.bicgstab(A, b, P, max_iter=k["max_iter"], rtol=k["rtol"])
    assert st == 0 and it.resid <= k["rtol"] * np.linalg.norm(b)


@pytest.mark.parametrize("pc", ["identity", "diagonal", "ilu_0"])
def test_bicgstab_nonsymmetric(pc):
    rng = np.random.default_rng(11)
    n = 60
    D = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.1)
    np.fill_diagonal(D, np.abs(D).sum(1) + 1.0)
    A = from_dense(D)
    xt = rng.standard_normal(n)
    b = D @ xt
    P = None if pc == "identity" else orc.Preconditioner(pc, A)
    x, st, it, _ = orc.bicgstab(A, b, P, max_iter=200, rtol=1e-12)
    assert st == 0 and np.allclose(x, xt, rtol=1e-8, atol=1e-9)


def test_iteration_semantics():
    A = orc.laplacian_setup(8, 8)
    P = orc.Preconditioner("diagonal", A)
    b = orc.spmv(A, np.ones(A.n))
    # max_iter reached -> error 1 ("Too many iterations", basic_iteration.hpp:43-48)
    x, st, it, _ = orc.cg(A, b, P, max_iter=3, rtol=1e-12)
    assert st == 1 and it.error == 1 and it.i == 3
    # b = 0: ||r0|| = 0 and resid 0 <= atol 0 -> converged at 0 iterations (:90-91)
    x, st, it, _ = orc.cg(A, np.zeros(A.n), P, max_iter=10, rtol=1e-10)
    assert st == 0 and it.i == 0
    # history holds one residual per finished() call
    x, st, it, h = orc.cg(A, b, P, max_iter=100, rtol=1e-8, hist_cap=200)
    assert len(h) == it.i + 1 and h[0] == pytest.approx(np.linalg.norm(b))


def test_bicgstab_breakdown():
    # A = 0 off a zero diagonal gives gamma = dot(rtilde, A*phat) = 0
    A = orc.Matrix([0, 1, 2], [0, 1], [0.0, 0.0])
    x, st, it, _ = orc.bicgstab(A, np.ones(2), None, max_iter=10, rtol=1e-8)
    assert st == orc.BREAKDOWN


# ---------------------------------------------------------------- BiCGSTAB pinned
# Three checks fix orc_bicgstab (bicgstab.hpp:27-78) independently of its
# recurrences, each run on the oracle and on the NumPy restatement in
# tests/golden/bicgstab_numpy.py; that restatement's four plausible mistakes
# (beta = 0, p = r + beta*p, omega = t.s/s.s, no x += alpha*phat at the
# half-step exit) must each fail at least one of them.  A solver here is
# f(rowptr, col, val, b, pc, x0, max_iter, rtol) -> (x, status, iterations,
# residuals handed to finished()).
def _orc_bicgstab(rp, ci, va, b, pc, x0, max_iter, rtol):
    A = orc.Matrix(rp, ci, va)
    P = None if pc == "identity" else orc.Preconditioner(pc, A)
    x, st, it, h = orc.bicgstab(A, b, P, x0=x0, max_iter=max_iter, rtol=rtol, hist_cap=4 * max_iter + 8)
    return x, st, it.i, h


def _numpy_bicgstab(mutant=None):
    import bicgstab_numpy as bn

    def solve(rp, ci, va, b, pc, x0, max_iter, rtol):
        A = bn.Csr(rp, ci, np.asarray(va, dtype=float))
        inv = None if pc == "identity" else 1.0 / A.diag
        x, err, it = bn.bicgstab(A, np.asarray(b, dtype=float), inv, x0, max_iter, rtol, mutant=mutant)
        return x, err, it.i, np.array(it.history)
    return solve


def _csr(D):
    rp, ci, va = [0], [], []
    for r in range(D.shape[0]):
        nz = np.nonzero(D[r])[0]
        ci.extend(int(c) for c in nz)
        va.extend(float(D[r, c]) for c in nz)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va)


def _check_bicgstab_first_residuals(solve, pc):
    """The six residuals of the first 2 1/2 iterations, from Krylov
    characterizations rather than the recurrences.  bicgstab preconditions
    from the right: phat = M^-1 p, v = A phat, so with K = A M^-1, r0 = b
    and the shadow residual rtilde = r0:
      |r0|; |s1| with s1 the Petrov-Galerkin BiCG residual of step 1
      (s1 in r0 + span{K r0}, orthogonal to rtilde); |r1| = min over w of
      |s1 - w K s1|; |s2| = |(I - w1 K) phi2| with phi2 the BiCG residual of
      step 2 (orthogonal to rtilde and K^T rtilde); |r2| likewise; |s3| =
      |(I - w2 K)(I - w1 K) phi3|.  max_iter = 3 stops at the half-step
      check of iteration 3 with x += alpha*phat, so b - A x = s3."""
    n = 40
    D, rng = _nonsym(n, 21)
    b = rng.standard_normal(n)
    Minv = np.diag(1.0 / np.diag(D)) if pc == "diagonal" else np.eye(n)
    K = D @ Minv
    x, st, its, h = solve(*_csr(D), b, pc, None, 3, 1e-14)
    assert st == 1 and its == 3 and len(h) == 6, (st, its, len(h))
    r0 = b.copy()
    want, Q = [np.linalg.norm(r0)], np.eye(n)
    for k in (1, 2, 3):
        s = Q @ _pg_residual(K, r0, r0, k)
        want.append(np.linalg.norm(s))
        if k < 3:
            t = K @ s
            w = (t @ s) / (t @ t)
            want.append(np.linalg.norm(s - w * t))
            Q = (np.eye(n) - w * K) @ Q
    assert np.allclose(h, want, rtol=1e-9, atol=0), (h, want)
    assert np.linalg.norm(b - D @ x) == pytest.approx(want[5], rel=1e-8)


def _check_bicgstab_half_step_exit(solve, pc):
    """A circulant tridiagonal (-1, 4, -1) has the constant vector as an
    eigenvector (eigenvalue 2, and 1/2 after the diagonal preconditioner):
    for b and x0 constant, s1 = r0 - alpha*A*phat is exactly zero, the
    half-step check stops the first iteration, and x = x0 + alpha*phat =
    1/2 exactly (every intermediate is a power of two times an integer)."""
    n = 16
    D = 4.0 * np.eye(n)
    for i in range(n):
        D[i, (i + 1) % n] = D[i, (i - 1) % n] = -1.0
    for x0 in (None, np.full(n, 0.25)):
        x, st, its, h = solve(*_csr(D), np.ones(n), pc, x0, 10, 1e-12)
        assert st == 0 and its == 1 and len(h) == 2 and h[1] == 0.0, (st, its, h)
        assert np.array_equal(x, np.full(n, 0.5)), x


def _check_bicgstab_finite_termination(solve, k):
    """With k distinct eigenvalues BiCG's residual polynomial of degree k
    annihilates r0: s_k = 0 in exact arithmetic, so the half-step check of
    iteration k stops the solve, after 2k checks, and x then solves A x = b
    (only if the half-step exit adds alpha*phat)."""
    vals = np.repeat([1.0, 2.0, 5.0, 3.0, 7.0][:k], 6)
    n = len(vals)
    b = np.random.default_rng(k).uniform(0.5, 1.5, n)
    x, st, its, h = solve(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), vals, b,
                          "identity", None, 20, 1e-11)
    assert st == 0 and its == k and len(h) == 2 * k, (st, its, len(h))
    assert min(h[:-1]) > 1e-6 and h[-1] <= 1e-11 * np.linalg.norm(b)
    assert np.allclose(x, b / vals, rtol=1e-10, atol=0)


_BICGSTAB_CHECKS = [lambda s: _check_bicgstab_first_residuals(s, "identity"),
                    lambda s: _check_bicgstab_first_residuals(s, "diagonal"),
                    lambda s: _check_bicgstab_half_step_exit(s, "identity"),
                    lambda s: _check_bicgstab_half_step_exit(s, "diagonal"),
                    lambda s: _check_bicgstab_finite_termination(s, 1),
                    lambda s: _check_bicgstab_finite_termination(s, 3)]


@pytest.mark.parametrize("pc", ["identity", "diagonal"])
def test_bicgstab_checked_residuals(pc):
    _check_bicgstab_first_residuals(_orc_bicgstab, pc)


@pytest.mark.parametrize("pc", ["identity", "diagonal"])
def test_bicgstab_half_step_exit(pc):
    _check_bicgstab_half_step_exit(_orc_bicgstab, pc)


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_bicgstab_finite_termination(k):
    _check_bicgstab_finite_termination(_orc_bicgstab, k)


def test_bicgstab_checks_pass_numpy_restatement():
    for check in _BICGSTAB_CHECKS:
        check(_numpy_bicgstab())


@pytest.mark.parametrize("mutant", ["beta0", "no_omega_v", "omega_ss", "no_half_step_x"])
def test_bicgstab_checks_reject_mutants(mutant):
    failed = 0
    for check in _BICGSTAB_CHECKS:
        try:
            check(_numpy_bicgstab(mutant))
        except AssertionError:
            failed += 1
    assert failed >= 1, mutant


_BICGSTAB_FIXTURE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bicgstab_fixture.json")))


@pytest.mark.parametrize("case", _BICGSTAB_FIXTURE["cases"], ids=lambda c: c["name"])
def test_bicgstab_fixture(case):
    """The oracle against tests/golden/bicgstab_fixture.json (written by the
    NumPy restatement with MTL4's reduction orders): status, iteration
    count, every checked residual and x, bit for bit."""
    def f(hs):
        return np.array([float.fromhex(v) for v in hs])
    sy = _BICGSTAB_FIXTURE["systems"][case["system"]]
    x0 = None if sy["x0"] is None else f(sy["x0"])
    x, st, its, h = _orc_bicgstab(np.array(sy["rowptr"], np.int32), np.array(sy["col"], np.int32), f(sy["val"]),
                                  f(sy["b"]), case["pc"], x0, case["max_iter"], case["rtol"])
    assert (st, its) == (case["error"], case["iterations"])
    assert np.array_equal(h, f(case["history"]))
    assert np.array_equal(x, f(case["x"]))


# ---------------------------------------------------------------- CGS
def test_cgs_ic0_reference_case():
    """The reference's ic_0_cgs_test.cpp case: 10x10 laplacian_setup, ic_0,
    x = 1, b = A*x, cyclic_iteration(b, N, 1e-6): converges, and the
    recurrence residual it stops on is the true one within rounding."""
    A = orc.laplacian_setup(10, 10)
    b = orc.spmv(A, np.ones(A.n))
    x, st, it, hist = orc.cgs(A, b, orc.Preconditioner("ic_0", A), max_iter=A.n, rtol=1e-6, hist_cap=200)
    assert st == 0 and it.i < A.n
    true = np.linalg.norm(b - orc.spmv(A, x))
    assert it.resid <= 1e-6 * np.linalg.norm(b)
    assert abs(true - it.resid) <= 1e-9 * np.linalg.norm(b)
    assert np.allclose(x, np.ones(A.n), atol=1e-4)


def test_cgs_finite_termination():
    """In exact arithmetic CGS, like BiCG, ends after as many steps as A has
    distinct eigenvalues: a diagonal matrix with three distinct values needs
    three iterations (a wrong sign or a dropped term in any update breaks
    this)."""
    vals = np.repeat([1.0, 2.0, 5.0], 7)
    n = len(vals)
    A = orc.Matrix(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), vals)
    b = np.random.default_rng(3).uniform(0.5, 1.5, n)
    x, st, it, _ = orc.cgs(A, b, None, max_iter=10, rtol=1e-12)
    assert st == 0 and it.i == 3
    assert np.allclose(x, b / vals, rtol=1e-12)


@pytest.mark.parametrize("pc", ["identity", "diagonal", "ilu_0"])
def test_cgs_nonsymmetric(pc):
    rng = np.random.default_rng(12)
    n = 60
    D = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.1)
    np.fill_diagonal(D, np.abs(D).sum(1) + 1.0)
    A = from_dense(D)
    xt = rng.standard_normal(n)
    b = D @ xt
    P = None if pc == "identity" else orc.Preconditioner(pc, A)
    x, st, it, _ = orc.cgs(A, b, P, max_iter=200, rtol=1e-12)
    assert st == 0 and np.allclose(x, xt, rtol=1e-8, atol=1e-9)


# ---------------------------------------------------------------- TFQMR
@pytest.mark.parametrize("pc", ["identity", "ilu_0", "ic_0", "diagonal"])
def test_tfqmr_reference_case(pc):
    """The reference's tfqmr_test.cpp: 10x10 laplacian_setup, b = 1, x = 0.5
    (10.5 for the diagonal case), the same preconditioner left and right,
    cyclic_iteration(b, 1000, 1e-8); it requires |A x - b| <= 1e-6."""
    A = orc.laplacian_setup(10, 10)
    b = np.ones(A.n)
    P = None if pc == "identity" else orc.Preconditioner(pc, A)
    x0 = np.full(A.n, 10.5 if pc == "diagonal" else 0.5)
    x, st, it, _ = orc.tfqmr(A, b, P, P, x0=x0, max_iter=1000, rtol=1e-8)
    assert st == 0
    assert np.linalg.norm(orc.spmv(A, x) - b) <= 1e-6


def test_tfqmr_finite_termination():
    """TFQMR's iterates span the same Krylov space as CGS's: on a diagonal
    matrix with two distinct eigenvalues the quasi-residual tau reaches zero
    (to rounding) within two outer iterations and x is exact."""
    vals = np.repeat([1.0, 3.0], 9)
    n = len(vals)
    A = orc.Matrix(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), vals)
    b = np.random.default_rng(4).uniform(0.5, 1.5, n)
    x, st, it, _ = orc.tfqmr(A, b, None, None, max_iter=10, rtol=1e-12)
    assert st == 0 and it.i <= 2
    assert np.allclose(x, b / vals, rtol=1e-10)


@pytest.mark.parametrize("lr", [("ilu_0", None), (None, "ilu_0"), ("diagonal", "ilu_0")])
def test_tfqmr_nonsymmetric_left_right(lr):
    rng = np.random.default_rng(13)
    n = 60
    D = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.1)
    np.fill_diagonal(D, np.abs(D).sum(1) + 1.0)
    A = from_dense(D)
    xt = rng.standard_normal(n)
    b = D @ xt
    L = orc.Preconditioner(lr[0], A) if lr[0] else None
    R = orc.Preconditioner(lr[1], A) if lr[1] else None
    x, st, it, _ = orc.tfqmr(A, b, L, R, max_iter=400, rtol=1e-13)
    assert st == 0
    # with R the loop iterates on R x; x = solve(R, x) at the end recovers x
    assert np.allclose(x, xt, rtol=1e-7, atol=1e-8)


def test_tfqmr_empty_is_logic_error():
    A = orc.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    x, st, it, _ = orc.tfqmr(A, np.zeros(0), None, None, max_iter=3, rtol=1e-8)
    assert st == orc.INVALID


# ---------------------------------------------------------------- BiCGSTAB(2)
def _pg_residual(M, r0, rt, k):
    """The BiCG residual after k steps from its characterization: r in r0 +
    span{M r0, ..., M^k r0} with r orthogonal to rt, M^T rt, ..., (M^T)^(k-1) rt
    (Petrov-Galerkin), solved densely."""
    V = np.column_stack([np.linalg.matrix_power(M, p) @ r0 for p in range(1, k + 1)])
    W = np.column_stack([np.linalg.matrix_power(M.T, p) @ rt for p in range(k)])
    return r0 - V @ np.linalg.solve(W.T @ V, W.T @ r0)


def _b2_pc_residual(A, P, b, x):
    r = b - orc.spmv(A, x)
    return np.linalg.norm(P.solve(r) if P is not None else r)


def _nonsym(n, seed, density=0.2):
    rng = np.random.default_rng(seed)
    D = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < density)
    np.fill_diagonal(D, np.abs(D).sum(1) + 1.0)
    return D, rng


@pytest.mark.parametrize("pc", ["identity", "diagonal"])
def test_bicgstab_2_checked_residuals(pc):
    """Every residual bicgstab_2 hands to finished() in its first one and a
    half iterations, from the Krylov characterizations instead of the
    recurrences: with M = L^-1 A, r0 = L^-1 b and the shadow residual
    rtilde = b/|b| (taken before solve(L, .), bicgstab_2.hpp:61-63),
      head:        |r0|
      j=0 check:   |r_1|,   r_1 the Petrov-Galerkin BiCG residual of step 1
      j=1 check:   |M r_2|  (the check is on r_hat[1] = M r_hat[0])
      next head:   min over g of |r_2 - g1 M r_2 - g2 M^2 r_2| (the MR part)
      next j=0:    |Q(M) r_3| with Q(t) = 1 - g1 t - g2 t^2 from that minimum.
    A wrong sign in rho_0 = -omega*rho_0, a dropped term of the MR
    coefficients or a copy instead of the in-place orthogonalization each
    moves one of these."""
    n = 40
    D, rng = _nonsym(n, 21)
    A = from_dense(D)
    b = rng.standard_normal(n)
    P = orc.Preconditioner(pc, A) if pc != "identity" else None
    Minv = np.diag(1.0 / np.diag(D)) if P is not None else np.eye(n)
    M = Minv @ D
    x, st, it, h = orc.bicgstab_2(A, b, P, max_iter=2, rtol=1e-14, hist_cap=20)
    assert st == 1 and it.i == 2 and len(h) == 5   # stops at the j=0 check of iteration 2
    r0 = Minv @ b
    rt = b / np.linalg.norm(b)
    r2 = _pg_residual(M, r0, rt, 2)
    G = np.column_stack([M @ r2, M @ M @ r2])
    g = np.linalg.lstsq(G, r2, rcond=None)[0]
    r3 = _pg_residual(M, r0, rt, 3)
    q3 = r3 - g[0] * (M @ r3) - g[1] * (M @ (M @ r3))
    want = [np.linalg.norm(r0), np.linalg.norm(_pg_residual(M, r0, rt, 1)), np.linalg.norm(M @ r2),
            np.linalg.norm(r2 - G @ g), np.linalg.norm(q3)]
    assert np.allclose(h, want, rtol=1e-10, atol=0)
    # the exit at an inner check returns before that step's x update: x has
    # the residual of the previous check, the loop head
    assert _b2_pc_residual(A, P, b, x) == pytest.approx(h[3], rel=1e-10)


@pytest.mark.parametrize("k,iters,calls", [(2, 1, 3), (3, 2, 5), (4, 2, 6), (5, 3, 8)])
def test_bicgstab_2_finite_termination(k, iters, calls):
    """With k distinct eigenvalues the BiCG residual of step k is zero, and
    the first check after that step stops the solve: step s is the inner
    step j = (s-1) % 2 of iteration (s+1)//2 (head, j=0, j=1 per iteration)."""
    vals = np.repeat([1.0, 2.0, 5.0, 3.0, 7.0][:k], 6)
    n = len(vals)
    A = orc.Matrix(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), vals)
    b = np.random.default_rng(k).uniform(0.5, 1.5, n)
    x, st, it, h = orc.bicgstab_2(A, b, None, max_iter=10, rtol=1e-11, hist_cap=40)
    assert st == 0 and it.i == iters and len(h) == calls
    assert h[-1] <= 1e-11 * np.linalg.norm(b) and min(h[:-1]) > 1e-6
    assert _b2_pc_residual(A, None, b, x) == pytest.approx(h[-2], rel=1e-8)


@pytest.mark.parametrize("pc", ["identity", "diagonal", "ilu_0"])
@pytest.mark.parametrize("start", ["zero", "nonzero"])
def test_bicgstab_2_solution_tracks_residual(pc, start):
    """On a nonsymmetric system solved to 1e-12: the returned x (after the
    shift back by x0) has the preconditioned residual of the loop head when
    the head stopped, of the check before otherwise; and it solves A x = b."""
    n = 80
    D, rng = _nonsym(n, 31, 0.1)
    A = from_dense(D)
    xt = rng.standard_normal(n)
    b = D @ xt
    x0 = np.zeros(n) if start == "zero" else rng.standard_normal(n)
    P = orc.Preconditioner(pc, A) if pc != "identity" else None
    x, st, it, h = orc.bicgstab_2(A, b, P, x0=x0, max_iter=200, rtol=1e-12, hist_cap=700)
    assert st == 0
    want = h[-1] if len(h) % 3 == 1 else h[-2]
    assert _b2_pc_residual(A, P, b, x) == pytest.approx(want, rel=1e-6, abs=1e-14 * np.linalg.norm(b))
    assert np.allclose(x, xt, rtol=1e-6, atol=1e-7)


def test_bicgstab_2_reference_case():
    """The reference's ic_0_bicgstab_2_test.cpp case: 10x10 laplacian_setup,
    ic_0, x = 1, b = A*x, x = 0, cyclic_iteration(b, 200, 1e-6)."""
    A = orc.laplacian_setup(10, 10)
    b = orc.spmv(A, np.ones(A.n))
    P = orc.Preconditioner("ic_0", A)
    x, st, it, h = orc.bicgstab_2(A, b, P, max_iter=200, rtol=1e-6, hist_cap=700)
    assert st == 0 and 2 <= it.i < 20
    assert h[-1] <= 1e-6 * np.linalg.norm(b)
    assert np.allclose(x, 1.0, atol=1e-4)


def test_bicgstab_2_omega_breakdown():
    """A 3x3 integer system where the minimal-residual step's last
    coefficient is zero: omega == 0 is iter.fail(3), returned before the
    shift back.  Pinned by least squares: min |r_2 - g1 A r_2 - g2 A^2 r_2|
    has g2 = 0, and the returned x is the shifted problem's BiCG iterate
    (A x = r0 - r_2) without x0 added."""
    D = np.array([[1.0, 0.0, -2.0], [0.0, -2.0, 1.0], [2.0, 2.0, -1.0]])
    x_init = np.array([1.0, 0.0, 0.0])
    r0 = np.array([-2.0, 0.0, 0.0])
    b = r0 + D @ x_init
    x, st, it, h = orc.bicgstab_2(from_dense(D), b, None, x0=x_init, max_iter=5, rtol=1e-14, hist_cap=10)
    assert st == 3 and it.i == 1 and len(h) == 3
    r2 = _pg_residual(D, r0, r0 / np.linalg.norm(r0), 2)
    G = np.column_stack([D @ r2, D @ D @ r2])
    g = np.linalg.lstsq(G, r2, rcond=None)[0]
    assert abs(g[1]) <= 1e-12 * abs(g[0])
    assert np.allclose(D @ x, r0 - r2, atol=1e-12)   # x0 = x_init not added
    assert np.linalg.norm(D @ x_init) > 1.0


def test_bicgstab_2_empty_is_logic_error():
    A = orc.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    x, st, it, _ = orc.bicgstab_2(A, np.zeros(0), None, max_iter=3, rtol=1e-8)
    assert st == orc.INVALID


# ---------------------------------------------------------------- BiCGSTAB(l)
@pytest.mark.parametrize("pc", ["identity", "diagonal", "ilu_0"])
def test_bicgstab_ell_2_is_bicgstab_2(pc):
    """bicgstab_ell with l = 2 and no right preconditioner runs
    bicgstab_2's arithmetic step for step: same bits, same history."""
    D, rng = _nonsym(80, 31, 0.1)
    A = from_dense(D)
    b = D @ rng.standard_normal(80)
    P = orc.Preconditioner(pc, A) if pc != "identity" else None
    x2, s2, i2, h2 = orc.bicgstab_2(A, b, P, max_iter=100, rtol=1e-12, hist_cap=400)
    xe, se, ie, he = orc.bicgstab_ell(A, b, P, None, 2, max_iter=100, rtol=1e-12, hist_cap=400)
    assert s2 == se == 0 and i2.i == ie.i
    assert np.array_equal(x2, xe) and np.array_equal(h2, he)


@pytest.mark.parametrize("l", [1, 3, 4])
@pytest.mark.parametrize("side", ["left", "right"])
def test_bicgstab_ell_checked_residuals(l, side):
    """The residuals bicgstab_ell hands to finished() in its first iteration
    and the next head and inner check, from Krylov characterizations: with
    M = L^-1 A and r0 = L^-1 b (left) or M = A R^-1 and r0 = b (right, the
    iteration runs on R x), rtilde = b/|b|:
      head |r0|; inner check j: |M^j r_(j+1)| (r_hat[j] = M^j r_hat[0]);
      next head: min over g of |r_l - sum_i g_i M^i r_l|;
      next j=0 check: |Q(M) r_(l+1)|, Q(t) = 1 - sum_i g_i t^i,
    r_k the Petrov-Galerkin BiCG residual of step k."""
    n = 40
    D, rng = _nonsym(n, 21)
    A = from_dense(D)
    b = rng.standard_normal(n)
    P = orc.Preconditioner("diagonal", A)
    Dinv = np.diag(1.0 / np.diag(D))
    if side == "left":
        M, r0 = Dinv @ D, Dinv @ b
        x, st, it, h = orc.bicgstab_ell(A, b, P, None, l, max_iter=2, rtol=1e-14, hist_cap=40)
    else:
        M, r0 = D @ Dinv, b.copy()
        x, st, it, h = orc.bicgstab_ell(A, b, None, P, l, max_iter=2, rtol=1e-14, hist_cap=40)
    assert st == 1 and it.i == 2 and len(h) == l + 3
    rt = b / np.linalg.norm(b)
    want = [np.linalg.norm(r0)]
    for j in range(l):
        want.append(np.linalg.norm(np.linalg.matrix_power(M, j) @ _pg_residual(M, r0, rt, j + 1)))
    rl = _pg_residual(M, r0, rt, l)
    G = np.column_stack([np.linalg.matrix_power(M, i) @ rl for i in range(1, l + 1)])
    g = np.linalg.lstsq(G, rl, rcond=None)[0]
    want.append(np.linalg.norm(rl - G @ g))
    rn = _pg_residual(M, r0, rt, l + 1)
    want.append(np.linalg.norm(rn - sum(g[i - 1] * (np.linalg.matrix_power(M, i) @ rn) for i in range(1, l + 1))))
    assert np.allclose(h, want, rtol=1e-9, atol=0)


@pytest.mark.parametrize("l", [1, 3, 4, 8])
@pytest.mark.parametrize("side", ["left", "right"])
def test_bicgstab_ell_solution(l, side):
    """Solved to 1e-12 on a nonsymmetric system with ILU(0) on either side:
    x (after x = solve(R, x) + x0) solves A x = b; when the loop head or the
    first inner check stopped it, x carries exactly the preconditioned
    residual that check (or the loop head before it) measured."""
    n = 80
    D, rng = _nonsym(n, 33, 0.1)
    A = from_dense(D)
    xt = rng.standard_normal(n)
    b = D @ xt
    x0 = rng.standard_normal(n)
    P = orc.Preconditioner("ilu_0", A)
    L, R = (P, None) if side == "left" else (None, P)
    x, st, it, h = orc.bicgstab_ell(A, b, L, R, l, x0=x0, max_iter=300, rtol=1e-12, hist_cap=3000)
    assert st == 0
    assert np.allclose(x, xt, rtol=1e-6, atol=1e-7)
    k = (len(h) - 1) % (l + 1)   # 0: loop head, else inner check k-1
    if k <= 1:
        # the recurrence drifts from the true residual by rounding, more so
        # the more vectors the minimal-residual step combines
        r = b - orc.spmv(A, x)
        got = np.linalg.norm(P.solve(r) if side == "left" else r)
        assert got == pytest.approx(h[-1] if k == 0 else h[-2], rel=1e-6, abs=1e-11 * l * np.linalg.norm(b))


def test_bicgstab_ell_reference_case():
    """bicgstab_ell_output_test.cpp: 10x10 laplacian_setup, b = 1, x = 0.5,
    cyclic_iteration(b, 100, 0, 1e-8), l = 8: |A x - b| <= 1e-6."""
    A = orc.laplacian_setup(10, 10)
    b = np.ones(A.n)
    x, st, it, _ = orc.bicgstab_ell(A, b, None, None, 8, x0=np.full(A.n, 0.5), max_iter=100, rtol=0.0,
                                    atol=1e-8)
    assert st == 0
    assert np.linalg.norm(orc.spmv(A, x) - b) <= 1e-6


def test_bicgstab_ell_breakdown_and_bad_l():
    D = np.array([[1.0, 0.0, -2.0], [0.0, -2.0, 1.0], [2.0, 2.0, -1.0]])
    x_init = np.array([1.0, 0.0, 0.0])
    b = np.array([-2.0, 0.0, 0.0]) + D @ x_init
    x2, s2, _, _ = orc.bicgstab_2(from_dense(D), b, None, x0=x_init, max_iter=5, rtol=1e-14)
    xe, se, it, _ = orc.bicgstab_ell(from_dense(D), b, None, None, 2, x0=x_init, max_iter=5, rtol=1e-14)
    assert s2 == se == 3 and it.i == 1 and np.array_equal(x2, xe)
    A = orc.laplacian_setup(3, 3)
    _, st, _, _ = orc.bicgstab_ell(A, np.ones(9), None, None, 0, max_iter=5, rtol=1e-8)
    assert st == orc.INVALID
    E = orc.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    _, st, _, _ = orc.bicgstab_ell(E, np.zeros(0), None, None, 2, max_iter=3, rtol=1e-8)
    assert st == orc.INVALID


# ---------------------------------------------------------------- GMRES
def test_gmres_reference_cases():
    """gmres_test.cpp, the cases it asserts: laplacian_setup(2, 2), b = 1;
    gmres(1) from x = 5 within one iteration, gmres(8), (16), (32) from x =
    2.5: |A x - b| <= 5e-5.  And gmres_preconditioned_test.cpp's GMRES(4)
    with left ILU(0) from x = (2, 3, 4, 8), which converges."""
    A = orc.laplacian_setup(2, 2)
    b = np.ones(4)
    x, st, it, _ = orc.gmres(A, b, None, None, 1, x0=np.full(4, 5.0), max_iter=1, rtol=1e-8)
    assert np.linalg.norm(orc.spmv(A, x) - b) <= 5e-5
    for restart, iters in [(8, 8), (16, 32), (32, 32)]:
        x, st, it, _ = orc.gmres(A, b, None, None, restart, x0=np.full(4, 2.5), max_iter=iters, rtol=1e-8)
        assert np.linalg.norm(orc.spmv(A, x) - b) <= 5e-5
    P = orc.Preconditioner("ilu_0", A)
    x, st, it, _ = orc.gmres(A, b, P, None, 4, x0=np.array([2.0, 3.0, 4.0, 8.0]), max_iter=100, rtol=1e-8)
    assert st == 0 and np.linalg.norm(orc.spmv(A, x) - b) <= 1e-5


@pytest.mark.parametrize("side", ["identity", "left", "right"])
def test_gmres_1_cycles_are_minimal_residual(side):
    """GMRES(1): one Krylov vector per cycle, no earlier rotation to apply,
    so every cycle is textbook minimal residual: x_new minimizes the
    (preconditioned) residual over x + R^-1 span{r} -- computed here by
    dense least squares.  The residuals handed to finished() are, per cycle,
    the preconditioned |r| at its start and |b - A x_new| at its end."""
    n = 40
    D, rng = _nonsym(n, 41)
    A = from_dense(D)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    Dinv = np.diag(1.0 / np.diag(D))
    P = orc.Preconditioner("diagonal", A)
    L, R = {"identity": (None, None), "left": (P, None), "right": (None, P)}[side]
    Lm = Dinv if side == "left" else np.eye(n)
    Rm = Dinv if side == "right" else np.eye(n)
    cycles = 4
    x, st, it, h = orc.gmres(A, b, L, R, 1, x0=x0, max_iter=cycles, rtol=1e-14, hist_cap=20)
    assert st == 1 and it.i == cycles and len(h) == 2 * cycles
    xs = x0.copy()
    for cyc in range(cycles):
        r = Lm @ (b - D @ xs)
        Mr = Lm @ D @ Rm @ r
        xs = xs + Rm @ r * (Mr @ r) / (Mr @ Mr)
        assert h[2 * cyc] == pytest.approx(np.linalg.norm(r), rel=1e-9)
        assert h[2 * cyc + 1] == pytest.approx(np.linalg.norm(b - D @ xs), rel=1e-9)
    assert np.allclose(x, xs, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("restart", [2, 5])
@pytest.mark.parametrize("side", ["left", "right"])
def test_gmres_update_in_krylov_space(restart, side):
    """Whatever coefficients a cycle's small solve yields, its update of x is
    R^-1 V y with V the Arnoldi basis of K_m(M, r): x_new - x_old must lie in
    R^-1 K_m (projection residual at rounding level), and the end-of-cycle
    residual is |b - A x_new|.  (With m >= 2 the coefficients follow
    gmres.hpp:84-94, which rotates with column k-1 of H; no textbook identity
    fixes them.)"""
    n = 40
    D, rng = _nonsym(n, 43)
    A = from_dense(D)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    Dinv = np.diag(1.0 / np.diag(D))
    P = orc.Preconditioner("diagonal", A)
    L, R = (P, None) if side == "left" else (None, P)
    Lm = Dinv if side == "left" else np.eye(n)
    Rm = Dinv if side == "right" else np.eye(n)
    x, st, it, h = orc.gmres(A, b, L, R, restart, x0=x0, max_iter=restart, rtol=1e-14, hist_cap=10)
    assert it.i == restart and len(h) == 2
    r = Lm @ (b - D @ x0)
    assert h[0] == pytest.approx(np.linalg.norm(r), rel=1e-12)
    assert h[1] == pytest.approx(np.linalg.norm(b - D @ x), rel=1e-9)
    M = Lm @ D @ Rm
    K = np.column_stack([np.linalg.matrix_power(M, p) @ r for p in range(restart)])
    Q = np.linalg.qr(K)[0]
    u = np.linalg.solve(Rm, x - x0)
    assert np.linalg.norm(u - Q @ (Q.T @ u)) <= 1e-9 * np.linalg.norm(u)
    assert np.linalg.norm(u) > 1e-3


def test_gmres_no_direction():
    """GMRES(1) on a rotation: A r is orthogonal to r, the one-dimensional
    least-squares correction is zero, so g[0] = 0 after the rotation and
    the cycle finds no direction: iter.fail(2), x unchanged (gmres.hpp:
    99, 113-114)."""
    D = np.array([[0.0, 1.0], [-1.0, 0.0]])
    b = np.array([1.0, 0.0])
    x, st, it, h = orc.gmres(from_dense(D), b, None, None, 1, max_iter=10, rtol=1e-12, hist_cap=10)
    assert st == 2 and it.i == 1 and np.array_equal(x, np.zeros(2)) and len(h) == 1


def test_gmres_bad_arguments():
    E = orc.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    _, st, _, _ = orc.gmres(E, np.zeros(0), None, None, 4, max_iter=3, rtol=1e-8)
    assert st == orc.INVALID
    _, st, _, _ = orc.gmres(orc.laplacian_setup(2, 2), np.ones(4), None, None, -1, max_iter=3, rtol=1e-8)
    assert st == orc.INVALID


# ---------------------------------------------------------------- GMRES pinned
# From two Krylov vectors per cycle on, the coefficients follow gmres.hpp's
# own rotation order (rotation i rebuilt from column k-1 of H and applied to
# the whole of H), a pivoted LU that scatters its permutation and works in
# place on the block it factors: no textbook identity fixes them.  They are
# pinned by a second restatement, tests/golden/gmres_numpy.py, written with
# NumPy from the reference files: the oracle and it must agree bit for bit
# (status, iterations, every residual handed to finished(), x), and each of
# the restatement's four plausible mistakes must change the result on at
# least one of the systems below, so the systems exercise the branches those
# mistakes touch:
#   * diagonally dominant sparse systems, restart 2, 3, 5 and full, with
#     pc::diagonal on no side, the left, the right or both;
#   * dense Gaussian systems, whose Hessenberg LUs pivot in cycles of three
#     rows or more (scatter and gather permutations differ there);
#   * a singular 3x3 system whose first LU attempts hit zero pivots, so the
#     range shrinks and each attempt starts from the block the last one left.
def _gauss_system(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 9))
    D = rng.standard_normal((n, n))
    return D, rng.standard_normal(n)


def _gmres_cases():
    cases = []
    for seed in (1, 2):
        D, rng = _nonsym(30, seed, 0.3)
        b, x0 = rng.standard_normal(30), rng.standard_normal(30)
        for restart in (2, 3, 5, 0):
            for side in ("identity", "left", "right", "both"):
                cases.append((f"dominant{seed}-restart{restart}-{side}", D, b, x0, restart, side, 40))
    for seed in (3, 11, 12, 17):
        D, b = _gauss_system(seed)
        for restart in (3, 4, 0):
            cases.append((f"gauss{seed}-restart{restart}", D, b, None, restart, "identity", 3 * len(b)))
    D = np.array([[0.0, -1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    cases.append(("singular3-full", D, np.array([0.0, -1.0, 1.0]), None, 0, "identity", 6))
    return cases


_GMRES_CASES = _gmres_cases()


def _orc_gmres_case(D, b, x0, restart, side, max_iter):
    A = orc.Matrix(*_csr(D))
    P = orc.Preconditioner("diagonal", A) if side != "identity" else None
    L, R = {"identity": (None, None), "left": (P, None), "right": (None, P), "both": (P, P)}[side]
    x, st, it, h = orc.gmres(A, b, L, R, restart, x0=x0, max_iter=max_iter, rtol=1e-12, hist_cap=4 * max_iter + 8)
    return x, st, it.i, h


def _numpy_gmres_case(D, b, x0, restart, side, max_iter, mutant=None):
    import bicgstab_numpy as bn
    import gmres_numpy as gn
    A = bn.Csr(*_csr(D))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / A.diag
        il, ir = {"identity": (None, None), "left": (inv, None), "right": (None, inv), "both": (inv, inv)}[side]
        return gn.gmres(A, b, restart, il, ir, x0, max_iter, 1e-12, 0.0, mutant)


def _same_run(u, v):
    return (u[1], u[2]) == (v[1], v[2]) and np.array_equal(u[3], v[3], equal_nan=True) \
        and np.array_equal(u[0], v[0], equal_nan=True)


@pytest.mark.parametrize("case", _GMRES_CASES, ids=lambda c: c[0])
def test_gmres_matches_numpy_restatement(case):
    _, D, b, x0, restart, side, max_iter = case
    o = _orc_gmres_case(D, b, x0, restart, side, max_iter)
    p = _numpy_gmres_case(D, b, x0, restart, side, max_iter)
    assert _same_run(o, p), (o[1:3], p[1:3], o[3], p[3])


def test_gmres_singular_block_case():
    """The singular 3x3 case: the oracle returns a finite x after the range
    shrank (the restatement also agrees bit for bit above)."""
    _, D, b, x0, restart, side, max_iter = _GMRES_CASES[-1]
    x, st, its, h = _orc_gmres_case(D, b, x0, restart, side, max_iter)
    assert st == 0 and its == 3 and np.isfinite(x).all() and len(h) == 2


@pytest.mark.parametrize("mutant", ["textbook_rotations", "gather_permute", "no_reorth", "lu_on_copy"])
def test_gmres_cases_reject_restatement_mutants(mutant):
    changed = [c[0] for c in _GMRES_CASES
               if not _same_run(_numpy_gmres_case(*c[1:]), _numpy_gmres_case(*c[1:], mutant=mutant))]
    assert changed, mutant


def test_cg_diag_omp_matches_sequential_cg():
    """The OpenMP restatement of MTL4's cg + diagonal (the all-core CPU
    baseline) solves config 1 like the sequential oracle: iteration count
    within 2, solution within 1e-12 relative, with 1 and with 4 threads."""
    O = orc.laplacian_setup(64, 64)
    xt = np.random.default_rng(0).uniform(-1, 1, O.n)
    b = orc.spmv(O, xt)
    xs, sts, its, _ = orc.cg(O, b, orc.Preconditioner("diagonal", O), max_iter=2000, rtol=1e-10)
    assert sts == 0
    for th in (1, 4):
        xo, sto, ito = orc.cg_diag_omp(O, b, threads=th, max_iter=2000, rtol=1e-10)
        assert sto == 0 and abs(ito.i - its.i) <= 2, (th, ito.i, its.i)
        assert np.linalg.norm(xo - xs) <= 1e-12 * np.linalg.norm(xs)
    # the iteration limit is error 1, as in basic_iteration
    _, st, it = orc.cg_diag_omp(O, b, threads=2, max_iter=5, rtol=1e-10)
    assert st == 1 and it.i == 5
