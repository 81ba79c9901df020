"""Batched solves of many uneven systems in one persistent launch, checked
system by system against the oracle's itl::cg, and for independence of the
batch a system runs in (same bits alone and in the mix)."""
import numpy as np
import pytest

import orc
from mtl4_amd import generators as gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mtl4_amd
    return mtl4_amd


def make_systems(count, seed=0, nmin=64, nmax=3000):
    """The three families of the batched workload at test sizes."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        fam = k % 3
        if fam == 0:
            a = max(2, int(np.sqrt(n)))
            out.append(gen.laplacian_setup(a, max(2, n // a)))
        elif fam == 1:
            a = max(2, int(round(n ** (1 / 3))))
            out.append(gen.laplacian3d(a, a, max(2, n // (a * a))))
        else:
            out.append(gen.random_spd(n, seed=seed * 1000 + k))
    return out


def rhs(systems, seed=1):
    rng = np.random.default_rng(seed)
    bs = []
    for rp, ci, va in systems:
        xt = rng.uniform(-1, 1, len(rp) - 1)
        bs.append(gen.csr_matvec(rp, ci, va, xt))
    return bs


@pytest.mark.parametrize("pc", ["diagonal", "identity"])
def test_batch_fp64_matches_oracle(m, pc):
    systems = make_systems(60, seed=3)
    bs = rhs(systems)
    B = m.Batch(systems)
    b = m.dense_vector(np.concatenate(bs))
    x = m.dense_vector(B.nrows)
    kind = m._lib.PC_DIAGONAL if pc == "diagonal" else m._lib.PC_IDENTITY
    res = B.solve(x, b, pc_kind=kind, precision=m.F64, rtol=1e-10, max_iter=5000)
    xs = x.numpy()
    for k, ((rp, ci, va), bk) in enumerate(zip(systems, bs)):
        O = orc.Matrix(rp, ci, va)
        P = orc.Preconditioner(pc, O) if pc == "diagonal" else None
        xo, st, it, _ = orc.cg(O, bk, P, max_iter=5000, rtol=1e-10)
        r = res[k]
        assert r["error"] == 0 and r["status"] == 0 and st == 0
        assert abs(r["iterations"] - it.i) <= 2, (k, r, it.i)
        xk = xs[B.sys_row[k]:B.sys_row[k + 1]]
        assert np.linalg.norm(xk - xo) <= 1e-9 * np.linalg.norm(xo), k
        assert abs(r["norm_r0"] - np.linalg.norm(bk)) <= 1e-12 * np.linalg.norm(bk)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_batch_mixed_reaches_fp64_tolerance(m, prec):
    systems = make_systems(45, seed=5)
    if prec == "bf16":   # BF16 values: only the well-conditioned family converges
        systems = [s for k, s in enumerate(systems) if k % 3 == 2]
    bs = rhs(systems, seed=2)
    B = m.Batch(systems)
    b = m.dense_vector(np.concatenate(bs))
    x = m.dense_vector(B.nrows)
    res = B.solve(x, b, precision=m.F32 if prec == "f32" else m.BF16, rtol=1e-10, max_iter=20000)
    xs = x.numpy()
    for k, ((rp, ci, va), bk) in enumerate(zip(systems, bs)):
        assert res[k]["error"] == 0 and res[k]["status"] == 0, (k, res[k])
        xk = xs[B.sys_row[k]:B.sys_row[k + 1]]
        relres = np.linalg.norm(bk - gen.csr_matvec(rp, ci, va, xk)) / np.linalg.norm(bk)
        assert relres <= 1.0e-10 * 1.01, (k, relres)
        assert res[k]["refreshes"] >= 1


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_alone_equals_in_mix(m, prec):
    systems = make_systems(30, seed=7)
    bs = rhs(systems, seed=4)
    p = m.F64 if prec == "f64" else m.F32
    B = m.Batch(systems)
    x = m.dense_vector(B.nrows)
    res = B.solve(x, m.dense_vector(np.concatenate(bs)), precision=p, rtol=1e-10)
    xs = x.numpy()
    for k in (0, 7, 29):
        Bk = m.Batch([systems[k]])
        xk = m.dense_vector(Bk.nrows)
        rk = Bk.solve(xk, m.dense_vector(bs[k]), precision=p, rtol=1e-10)[0]
        assert rk["iterations"] == res[k]["iterations"]
        assert np.array_equal(xk.numpy(), xs[B.sys_row[k]:B.sys_row[k + 1]])


def test_batch_tune_and_shapes(m):
    systems = make_systems(40, seed=9)
    bs = rhs(systems, seed=6)
    B = m.Batch(systems)
    b = m.dense_vector(np.concatenate(bs))
    x = m.dense_vector(B.nrows)
    x0 = m.dense_vector(B.nrows)
    ms, shape = B.tune(x, x0, b, precision=m.F64, rtol=1e-10)
    assert ms > 0 and shape[0] in (256, 512, 1024)
    ref = None
    for threads, per_cu in ((256, 1), (512, 2), (1024, 1)):
        B.set_shape(m.F64, threads, per_cu)
        x.t.zero_()
        B.solve(x, b, precision=m.F64, rtol=1e-10)
        if ref is None:
            ref = x.numpy()
        elif threads == 256:
            assert np.array_equal(x.numpy(), ref)


def test_batch_rejects_coupled_systems(m):
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 1, 1], np.int32)   # row 0 of system 0 reaches into system 1
    va = np.ones(3)
    import ctypes as C
    h = C.c_void_p()
    sys_row = np.array([0, 1, 2], np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    st = m._lib.lib().mtl4g_batch_create(C.byref(h), 2, p(sys_row), p(rp), p(ci), p(va), 0)
    assert st == m._lib.ERR_INVALID


def test_batch_graph_replays_match_direct_launch(m):
    """mtl4g_batch_solve records the launch once as a graph and replays it:
    three replays with eager work and synchronizations in between must each
    give the bits of a direct (uncaptured) launch of the same items."""
    import ctypes as C
    import torch
    systems = make_systems(24, seed=13)
    bs = rhs(systems, seed=8)
    B = m.Batch(systems)
    b = m.dense_vector(np.concatenate(bs))
    x = m.dense_vector(B.nrows)
    outs = []
    for rep in range(3):
        x.t.zero_()
        res = B.solve(x, b, precision=m.F64, rtol=1e-10)
        outs.append((x.numpy(), [r["iterations"] for r in res]))
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
        torch.cuda.synchronize()
        assert float(junk.abs().sum()) > 0          # eager work between replays
    assert B.graph_info() == (1, 3)
    # direct launch of the same items (device item list from the last solve)
    L = m._lib.lib()
    items = (m._lib.BatchItem * B.nsys)()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    order = sorted(range(B.nsys), key=lambda k: -(len(systems[k][1]) * np.sqrt(len(systems[k][0]) - 1)))
    ws = np.concatenate([[0], np.cumsum([L.mtl4g_batch_workspace_elems(len(s[0]) - 1, m.F64) for s in systems])])
    for j, k in enumerate(order):
        items[j] = m._lib.BatchItem(k, 100000, int(B.sys_row[k]), int(B.sys_row[k]), int(ws[k]), 1e-10, 0.0)
    ditems = torch.empty(C.sizeof(items), dtype=torch.uint8, device="cuda")
    ditems.copy_(torch.frombuffer(bytearray(items), dtype=torch.uint8))
    work = torch.zeros(int(ws[-1]), dtype=torch.float64, device="cuda")
    dres = torch.zeros(B.nsys * C.sizeof(m._lib.BatchResult), dtype=torch.uint8, device="cuda")
    queue = torch.zeros(1, dtype=torch.int32, device="cuda")
    xd = m.dense_vector(B.nrows)
    m._lib.check(L.mtl4g_batch_solve_items(B.h, m._lib.PC_DIAGONAL, m.F64, ditems.data_ptr(), B.nsys,
                                           b.ptr, xd.ptr, work.data_ptr(), dres.data_ptr(),
                                           queue.data_ptr(), m._stream()))
    torch.cuda.synchronize()
    for xs, its in outs:
        assert np.array_equal(xs, xd.numpy())
        assert its == outs[0][1]
