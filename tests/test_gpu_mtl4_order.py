"""MTL4-order reductions: with mtl4g_set_reduction_order(MTL4G_ORDER_MTL4)
every FP64 reduction sums in MTL4's unrolled sequential order, so the
reductions and whole cg / bicgstab solves must reproduce the CPU oracle
(MTL4's algorithm at -ffp-contract=off) bit for bit: same iteration count,
same final residual, same solution bits.

The data are chosen order-sensitive: magnitudes spread over 16 decades,
and in test_reductions_order_sensitive 1e16 beside ones, where the oracle
itself shows that another accumulator layout gives other bits."""
import ctypes as C

import numpy as np
import pytest

import orc
from helpers import family, pair

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 33, 1000, 4099, 100003]


@pytest.fixture(scope="module")
def m():
    import mtl4_amd
    return mtl4_amd


@pytest.fixture
def mtl4_order(m):
    with m.reduction_order(m.ORDER_MTL4):
        yield
    assert m.get_reduction_order() == m.ORDER_TREE


def dv(m, a):
    return m.dense_vector(np.asarray(a, dtype=np.float64))


def wide(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 8, n)


def test_order_switch(m):
    assert m.get_reduction_order() == m.ORDER_TREE
    with pytest.raises(m.logic_error):
        m.set_reduction_order(7)
    with m.reduction_order(m.ORDER_MTL4):
        assert m.get_reduction_order() == m.ORDER_MTL4
    assert m.get_reduction_order() == m.ORDER_TREE


@pytest.mark.parametrize("n", SIZES)
def test_reductions_bitwise(m, mtl4_order, n):
    a, b = wide(n, n), wide(n, n + 1)
    va, vb = dv(m, a), dv(m, b)
    assert m.dot(va, vb) == orc.dot(a, b, 4)
    assert m.unary_dot(va) == orc.unary_dot(a, 8)      # unary_dot<8> is MTL4's default
    assert m.two_norm(va) == orc.two_norm(a)           # two_norm<4>


@pytest.mark.parametrize("n", [9, 4099, 100003])
def test_reductions_order_sensitive(m, mtl4_order, n):
    """Vectors whose sums change with the accumulator layout (1e16 next to
    ones): the device must give MTL4's answer, not another order's."""
    a = np.ones(n)
    a[0] = 1e8
    assert orc.unary_dot(a, 8) != orc.unary_dot(a, 4)
    assert m.unary_dot(dv(m, a)) == orc.unary_dot(a, 8)
    assert m.two_norm(dv(m, a)) == orc.two_norm(a)
    x, y = np.ones(n), np.ones(n)
    x[0], x[2] = 1e16, -1e16
    assert orc.dot(x, y, 4) != orc.dot(x, y, 1)
    assert m.dot(dv(m, x), dv(m, y)) == orc.dot(x, y, 4)


@pytest.mark.parametrize("n", SIZES)
def test_fused_axpy_unary_dot_bitwise(m, mtl4_order, n):
    q, r = wide(n, 3 * n), wide(n, 3 * n + 1)
    alpha = 0.7310585786300049
    rg = dv(m, r)
    rho = C.c_double()
    m._lib.check(m._lib.lib().mtl4g_axpy_unary_dot(n, alpha, dv(m, q).ptr, rg.ptr, C.byref(rho),
                                                   m._stream()))
    ro = r.copy()
    rho_o = orc.fused_axpy_unary_dot(alpha, q, ro)
    assert np.array_equal(rg.numpy(), ro)
    assert rho.value == rho_o


@pytest.mark.parametrize("fam", ["lap2d_64", "random_3000", "long_rows", "diff3d_16"])
def test_fused_spmv_dot_bitwise(m, mtl4_order, fam):
    rp, ci, va = family(fam, seed=5)
    O, A = pair(rp, ci, va)
    p = wide(A.num_rows(), 17)
    q = m.dense_vector(A.num_rows())
    d = C.c_double()
    m._lib.check(m._lib.lib().mtl4g_spmv_dot(A.h, dv(m, p).ptr, q.ptr, C.byref(d), m._stream()))
    qo, do = orc.fused_spmv_dot(O, p)
    assert np.array_equal(q.numpy(), qo)
    assert d.value == do


def _solve_both(m, method, fam, kind, seed, max_iter=3000):
    rp, ci, va = family(fam, seed=seed)
    O, A = pair(rp, ci, va)
    n = A.num_rows()
    xt = np.random.default_rng(seed + 100).uniform(-1, 1, n)
    b = orc.spmv(O, xt)
    OP = None if kind == "identity" else orc.Preconditioner(kind, O)
    xo, st, ito, _ = getattr(orc, method)(O, b, OP, max_iter=max_iter, rtol=1e-10)
    P = getattr(m.pc, kind)(A)
    x = m.dense_vector(n)
    it = m.basic_iteration(dv(m, b), max_iter, 1e-10)
    err = getattr(m, method)(A, x, dv(m, b), P, it)
    return (err, it, x.numpy()), (st, ito, xo)


@pytest.mark.parametrize("fam", ["lap2d_64", "diff3d_16", "spd_2000", "lap3d_12"])
@pytest.mark.parametrize("kind", ["identity", "diagonal", "ic_0", "ilu_0"])
def test_cg_bitwise(m, mtl4_order, fam, kind):
    (err, it, x), (st, ito, xo) = _solve_both(m, "cg", fam, kind, seed=2)
    assert err == st == 0
    assert it.norm_r0() == ito.norm_r0
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("kind", ["identity", "diagonal", "ilu_0"])
def test_bicgstab_bitwise(m, mtl4_order, fam, kind):
    (err, it, x), (st, ito, xo) = _solve_both(m, "bicgstab", fam, kind, seed=4, max_iter=2000)
    assert err == st == 0
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("kind", ["identity", "diagonal", "ilu_0", "ic_0"])
def test_cgs_bitwise(m, mtl4_order, fam, kind):
    """itl::cgs in MTL4's reduction order: iteration count, final residual
    and solution bit for bit with the oracle's restatement of cgs.hpp."""
    if kind == "ic_0" and fam != "lap2d_64":
        pytest.skip("ic_0 needs a symmetric matrix")
    (err, it, x), (st, ito, xo) = _solve_both(m, "cgs", fam, kind, seed=5, max_iter=2000)
    assert err == st == 0
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("lr", [("identity", "identity"), ("ilu_0", "identity"), ("identity", "ilu_0"),
                                ("diagonal", "diagonal")])
def test_tfqmr_bitwise(m, mtl4_order, fam, lr):
    """itl::tfqmr with left and right preconditioners in MTL4's reduction
    order: iteration count, final quasi-residual and solution (after the
    shift back by solve(R, x)) bit for bit with the oracle."""
    rp, ci, va = family(fam, seed=7)
    O, A = pair(rp, ci, va)
    n = A.num_rows()
    xt = np.random.default_rng(107).uniform(-1, 1, n)
    b = orc.spmv(O, xt)
    OL, OR = [None if k == "identity" else orc.Preconditioner(k, O) for k in lr]
    xo, st, ito, _ = orc.tfqmr(O, b, OL, OR, max_iter=2000, rtol=1e-10)
    L, R = [getattr(m.pc, k)(A) for k in lr]
    x = m.dense_vector(n)
    it = m.basic_iteration(dv(m, b), 2000, 1e-10)
    err = m.tfqmr(A, x, dv(m, b), L, R, it)
    assert err == st == 0
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(x.numpy(), xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("kind", ["identity", "diagonal", "ilu_0", "ic_0"])
@pytest.mark.parametrize("start", ["zero", "nonzero"])
def test_bicgstab_2_bitwise(m, mtl4_order, fam, kind, start):
    """itl::bicgstab_2 in MTL4's reduction order: iteration count, every
    residual handed to finished() (loop head and both inner steps), and the
    solution after the shift back bit for bit with the oracle; a nonzero
    start takes the shifted-problem branch (bicgstab_2.hpp:55-59)."""
    if kind == "ic_0" and fam != "lap2d_64":
        pytest.skip("ic_0 needs a symmetric matrix")
    rp, ci, va = family(fam, seed=9)
    O, A = pair(rp, ci, va)
    n = A.num_rows()
    rng = np.random.default_rng(109)
    xt = rng.uniform(-1, 1, n)
    b = orc.spmv(O, xt)
    x0 = np.zeros(n) if start == "zero" else rng.uniform(-1, 1, n)
    OP = None if kind == "identity" else orc.Preconditioner(kind, O)
    xo, st, ito, ho = orc.bicgstab_2(O, b, OP, x0=x0, max_iter=2000, rtol=1e-10, hist_cap=7000)
    x = dv(m, x0)
    it = m.cyclic_iteration(dv(m, b), 2000, 1e-10, 0.0, 10**9)
    it.set_quite(True)
    it.suppress_resume(True)
    err = m.bicgstab_2(A, x, dv(m, b), getattr(m.pc, kind)(A), it)
    assert err == st == 0
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(np.asarray(it.history), ho)
    assert np.array_equal(x.numpy(), xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("l", [1, 3, 8])
@pytest.mark.parametrize("lr", [("identity", "identity"), ("ilu_0", "identity"), ("identity", "ilu_0"),
                                ("diagonal", "ilu_0")])
def test_bicgstab_ell_bitwise(m, mtl4_order, fam, l, lr):
    """itl::bicgstab_ell in MTL4's reduction order, l = 1, 3, 8, left and
    right preconditioners, nonzero start: iteration count, every residual
    handed to finished() and the solution (after x = solve(R, x) + x0) bit
    for bit with the oracle."""
    rp, ci, va = family(fam, seed=11)
    O, A = pair(rp, ci, va)
    n = A.num_rows()
    rng = np.random.default_rng(111)
    b = orc.spmv(O, rng.uniform(-1, 1, n))
    x0 = rng.uniform(-1, 1, n)
    OL, OR = [None if k == "identity" else orc.Preconditioner(k, O) for k in lr]
    xo, st, ito, ho = orc.bicgstab_ell(O, b, OL, OR, l, x0=x0, max_iter=1000, rtol=1e-10, hist_cap=10000)
    L, R = [getattr(m.pc, k)(A) for k in lr]
    x = dv(m, x0)
    it = m.cyclic_iteration(dv(m, b), 1000, 1e-10, 0.0, 10**9)
    it.set_quite(True)
    it.suppress_resume(True)
    err = m.bicgstab_ell(A, x, dv(m, b), L, R, it, l)
    assert err == st
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(np.asarray(it.history), ho)
    assert np.array_equal(x.numpy(), xo)


@pytest.mark.parametrize("fam", ["convdiff_40", "lap2d_64", "random_3000"])
@pytest.mark.parametrize("restart", [1, 2, 5, 16, 0])
@pytest.mark.parametrize("lr", [("identity", "identity"), ("ilu_0", "identity"), ("identity", "ilu_0"),
                                ("diagonal", "ilu_0")])
def test_gmres_bitwise(m, mtl4_order, fam, restart, lr):
    """itl::gmres (restart > 0) and itl::gmres_full (restart 0: one cycle of
    min(n, max_iter) steps) in MTL4's reduction order, nonzero start:
    status, iteration count, the residuals at each cycle's start and end,
    and the solution bit for bit with the oracle -- converged or not."""
    rp, ci, va = family(fam, seed=13)
    O, A = pair(rp, ci, va)
    n = A.num_rows()
    rng = np.random.default_rng(113)
    b = orc.spmv(O, rng.uniform(-1, 1, n))
    x0 = rng.uniform(-1, 1, n)
    OL, OR = [None if k == "identity" else orc.Preconditioner(k, O) for k in lr]
    max_iter = 24 if restart == 0 else 60
    xo, st, ito, ho = orc.gmres(O, b, OL, OR, restart, x0=x0, max_iter=max_iter, rtol=1e-10, hist_cap=400)
    L, R = [getattr(m.pc, k)(A) for k in lr]
    x = dv(m, x0)
    it = m.cyclic_iteration(dv(m, b), max_iter, 1e-10, 0.0, 10**9)
    it.set_quite(True)
    it.suppress_resume(True)
    if restart == 0:
        err = m.gmres_full(A, x, dv(m, b), L, R, it)
    else:
        err = m.gmres(A, x, dv(m, b), L, R, it, restart)
    assert err == st
    assert it.iterations() == ito.i
    assert it.resid() == ito.resid
    assert np.array_equal(np.asarray(it.history), ho)
    assert np.array_equal(x.numpy(), xo)


def _small_gmres_system(name):
    if name == "singular3":
        return np.array([[0.0, -1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]), np.array([0.0, -1.0, 1.0])
    rng = np.random.default_rng(int(name[5:]))
    n = int(rng.integers(4, 9))
    D = rng.standard_normal((n, n))
    return D, rng.standard_normal(n)


@pytest.mark.parametrize("name,restart", [("gauss3", 0), ("gauss3", 4), ("gauss11", 4), ("gauss17", 3),
                                          ("singular3", 0)])
def test_gmres_pivoting_and_singular_blocks_bitwise(m, mtl4_order, name, restart):
    """Dense Gaussian systems, whose Hessenberg least-squares solves pivot in
    cycles of three rows or more (MTL4's permute scatters, which differs
    from a gather there), and a singular 3x3 whose first LU attempts meet
    zero pivots, so the range shrinks and each attempt starts from the
    block the previous one left: the device against the oracle bit for
    bit (the oracle is pinned on the same systems by a NumPy restatement)."""
    D, b = _small_gmres_system(name)
    n = len(b)
    rows, cols = np.nonzero(D)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    O = orc.Matrix(rp, cols.astype(np.int32), D[rows, cols].copy())
    A = m.compressed2D(n, n, rp, cols.astype(np.int32), D[rows, cols].copy())
    max_iter = 2 * n if name == "singular3" else 3 * n
    xo, st, ito, ho = orc.gmres(O, b, None, None, restart, max_iter=max_iter, rtol=1e-12, hist_cap=400)
    x = m.dense_vector(n)
    it = m.cyclic_iteration(dv(m, b), max_iter, 1e-12, 0.0, 10**9)
    it.set_quite(True)
    it.suppress_resume(True)
    if restart == 0:
        err = m.gmres_full(A, x, dv(m, b), None, None, it)
    else:
        err = m.gmres(A, x, dv(m, b), None, None, it, restart)
    assert err == st
    assert it.iterations() == ito.i
    assert np.array_equal(np.asarray(it.history), ho)
    assert np.array_equal(x.numpy(), xo)


def test_bicgstab_2_omega_breakdown(m, mtl4_order):
    """omega == 0 is iter.fail(3) (bicgstab_2.hpp:107): status 3 after one
    iteration and x returned without the shift back x += x0, bit for bit
    with the oracle (whose own breakdown case is pinned in the CPU tests)."""
    D = np.array([[1.0, 0.0, -2.0], [0.0, -2.0, 1.0], [2.0, 2.0, -1.0]])
    x_init = np.array([1.0, 0.0, 0.0])
    b = np.array([-2.0, 0.0, 0.0]) + D @ x_init
    rows, cols = np.nonzero(D)
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=3))]).astype(np.int32)
    O = orc.Matrix(rp, cols.astype(np.int32), D[rows, cols].copy())
    A = m.compressed2D(3, 3, rp, cols.astype(np.int32), D[rows, cols].copy())
    xo, st, ito, _ = orc.bicgstab_2(O, b, None, x0=x_init, max_iter=5, rtol=1e-14)
    assert st == 3 and ito.i == 1
    x = dv(m, x_init)
    it = m.basic_iteration(dv(m, b), 5, 1e-14)
    assert m.bicgstab_2(A, x, dv(m, b), None, it) == 3
    assert it.iterations() == 1 and it.error_code() == 3
    assert np.array_equal(x.numpy(), xo)


def test_cg_config1_bitwise(m, mtl4_order):
    """CG + diagonal on the 64x64 Laplacian, the reference's CPU test case."""
    rp, ci, va = family("lap2d_64")
    O, A = pair(rp, ci, va)
    xt = np.random.default_rng(0).uniform(-1, 1, 4096)
    b = orc.spmv(O, xt)
    xo, st, ito, ho = orc.cg(O, b, orc.Preconditioner("diagonal", O), max_iter=2000, rtol=1e-10,
                             hist_cap=4000)
    x = m.dense_vector(4096)
    it = m.basic_iteration(dv(m, b), 2000, 1e-10)
    assert m.cg(A, x, dv(m, b), m.pc.diagonal(A), it) == st == 0
    assert it.iterations() == ito.i and it.resid() == ito.resid
    assert np.array_equal(x.numpy(), xo)


def test_solver_handle_follows_order(m):
    """A solver made in tree order and run in MTL4 order gets its terms
    buffer on first use and then matches the oracle bit for bit."""
    rp, ci, va = family("lap2d_64")
    O, A = pair(rp, ci, va)
    b = orc.spmv(O, np.ones(4096))
    S = m.Solver(A, m.pc.diagonal(A), method=m.CG)
    x = m.dense_vector(4096)
    it = m.basic_iteration(dv(m, b), 2000, 1e-10)
    S.solve(x, dv(m, b), it)
    with m.reduction_order(m.ORDER_MTL4):
        x2 = m.dense_vector(4096)
        it2 = m.basic_iteration(dv(m, b), 2000, 1e-10)
        S.solve(x2, dv(m, b), it2)
    xo, st, ito, _ = orc.cg(O, b, orc.Preconditioner("diagonal", O), max_iter=2000, rtol=1e-10)
    assert it2.iterations() == ito.i and np.array_equal(x2.numpy(), xo)
    assert abs(it.iterations() - ito.i) <= 2
