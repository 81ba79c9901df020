"""mtl4g_spmv_dot_noalloc (fused SpMV + dot whose reduction scratch is a fixed
per-device array) against mtl4g_spmv_dot, which allocates its scratch per
call: q bit for bit and the same dot, on consecutive calls that share the
scratch (the second on a smaller matrix), in the default and in MTL4's
reduction order, plus the size limit of the MTL4 order."""
import ctypes as C

import numpy as np
import pytest

from helpers import family, pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mtl4_amd
    return mtl4_amd


def _both(m, A, p, seed):
    L = m._lib.lib()
    n = A.num_rows()
    pd = m.dense_vector(np.asarray(p, dtype=np.float64))
    q_ref, q_na = m.dense_vector(n), m.dense_vector(n)
    # stale values in the outputs must not matter
    q_na.assign(np.random.default_rng(seed).uniform(-9, 9, n))
    d_ref, d_na = C.c_double(), C.c_double(np.nan)
    m._lib.check(L.mtl4g_spmv_dot(A.h, pd.ptr, q_ref.ptr, C.byref(d_ref), m._stream()), "spmv_dot")
    m._lib.check(L.mtl4g_spmv_dot_noalloc(A.h, pd.ptr, q_na.ptr, C.byref(d_na), m._stream()),
                 "spmv_dot_noalloc")
    return q_ref.numpy(), d_ref.value, q_na.numpy(), d_na.value


@pytest.mark.parametrize("order", ["tree", "mtl4"])
def test_noalloc_matches_allocating_entry_point(m, order):
    big = family("diff3d_16", seed=2)        # 4096 rows
    small = family("random_3000", seed=4)    # 3000 rows, other pattern
    ctx = m.reduction_order(m.ORDER_MTL4) if order == "mtl4" else None
    if ctx:
        ctx.__enter__()
    try:
        for k, (rp, ci, va) in enumerate((big, small)):
            _, A = pair(rp, ci, va)
            p = np.random.default_rng(10 + k).uniform(-1, 1, A.num_rows())
            q_ref, d_ref, q_na, d_na = _both(m, A, p, k)
            print("spmv_dot_noalloc order=%s call=%d n=%d" % (order, k, A.num_rows()), flush=True)
            assert np.array_equal(q_na, q_ref)
            assert d_na == d_ref
    finally:
        if ctx:
            ctx.__exit__(None, None, None)


def test_noalloc_every_kernel_shape(m):
    # the tuned layouts (sliced, diagonal) are read through the same scratch
    rp, ci, va = family("lap3d_12", seed=0)
    _, A = pair(rp, ci, va)
    p = np.random.default_rng(3).uniform(-1, 1, A.num_rows())
    L = m._lib
    for kernel, group, rows in ((L.SPMV_TILE, 0, 0), (L.SPMV_SCALAR, 0, 0), (L.SPMV_VECTOR, 8, 0),
                                (L.SPMV_STREAM, 4, 0), (L.SPMV_SLICED, 4, 4), (L.SPMV_DIA, 4, 8),
                                (L.SPMV_DIA_SYM, 4, 4)):
        A.set_kernel(kernel, group, rows)
        q_ref, d_ref, q_na, d_na = _both(m, A, p, kernel)
        assert np.array_equal(q_na, q_ref), kernel
        assert d_na == d_ref, kernel


def test_noalloc_mtl4_order_size_limit(m):
    n = (1 << 20) + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    va = np.full(n, 2.0)
    _, A = pair(rp, ci, va)
    p = m.dense_vector(np.ones(n))
    q = m.dense_vector(n)
    d = C.c_double()
    L = m._lib.lib()
    with m.reduction_order(m.ORDER_MTL4):
        st = L.mtl4g_spmv_dot_noalloc(A.h, p.ptr, q.ptr, C.byref(d), m._stream())
        assert st == m._lib.ERR_UNSUPPORTED
        assert b"fixed scratch" in L.mtl4g_last_error()
    # the default order has no such limit
    m._lib.check(L.mtl4g_spmv_dot_noalloc(A.h, p.ptr, q.ptr, C.byref(d), m._stream()), "noalloc")
    assert d.value == 2.0 * n
