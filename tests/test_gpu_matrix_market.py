"""The reference's Matrix Market test matrices on the device: SpMV bit for
bit against the oracle, and CG / BiCGSTAB on the SPD stiffness matrix
bcsstk01 (48x48, condition number about 1e6) bit for bit in MTL4's
reduction order and within the CG bar (+-2 iterations) in tree order."""
import os

import numpy as np
import pytest

import orc
from mtl4_amd import io as mio

pytestmark = pytest.mark.gpu
MM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "matrix_market")


@pytest.fixture(scope="module")
def m():
    import mtl4_amd
    return mtl4_amd


def load(m, name):
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, name + ".mtx"))
    return orc.Matrix(rp, ci, va, ncols=nc), m.compressed2D(nr, nc, rp, ci, va)


@pytest.mark.parametrize("name", ["bcsstk01", "plskz362", "jgl009", "bcspwr02", "dense_read"])
def test_spmv_bitwise(m, name):
    O, A = load(m, name)
    x = np.random.default_rng(1).uniform(-1, 1, A.num_cols())
    y = A * m.dense_vector(x)
    assert np.array_equal(y.numpy(), orc.spmv(O, x))


@pytest.mark.parametrize("kind", ["identity", "diagonal", "ic_0", "ilu_0"])
def test_cg_bcsstk01(m, kind):
    O, A = load(m, "bcsstk01")
    n = A.num_rows()
    b = orc.spmv(O, np.random.default_rng(2).uniform(-1, 1, n))
    OP = None if kind == "identity" else orc.Preconditioner(kind, O)
    xo, st, ito, _ = orc.cg(O, b, OP, max_iter=5000, rtol=1e-10)
    assert st == 0
    for order in (m.ORDER_MTL4, m.ORDER_TREE):
        with m.reduction_order(order):
            x = m.dense_vector(n)
            it = m.basic_iteration(m.dense_vector(b), 5000, 1e-10)
            assert m.cg(A, x, m.dense_vector(b), getattr(m.pc, kind)(A), it) == 0
        if order == m.ORDER_MTL4:
            assert it.iterations() == ito.i and np.array_equal(x.numpy(), xo)
        else:
            assert abs(it.iterations() - ito.i) <= 2 or kind == "identity"
            assert np.linalg.norm(x.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def test_bicgstab_ilu0_bcsstk01(m):
    O, A = load(m, "bcsstk01")
    n = A.num_rows()
    b = orc.spmv(O, np.random.default_rng(3).uniform(-1, 1, n))
    xo, st, ito, _ = orc.bicgstab(O, b, orc.Preconditioner("ilu_0", O), max_iter=2000, rtol=1e-10)
    with m.reduction_order(m.ORDER_MTL4):
        x = m.dense_vector(n)
        it = m.basic_iteration(m.dense_vector(b), 2000, 1e-10)
        err = m.bicgstab(A, x, m.dense_vector(b), m.pc.ilu_0(A), it)
    assert err == st == 0 and it.iterations() == ito.i
    assert np.array_equal(x.numpy(), xo)
