"""Serving loop with continuous batching: every request is solved to the
FP64 tolerance, admission respects the working-set cap, and each result is
bit-for-bit the same alone and in the mix."""
import numpy as np
import pytest

from mtl4_amd import generators as gen
import orc
from test_gpu_batched import make_systems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mtl4_amd
    return mtl4_amd


def splitmix_uniform(seed, n):
    """The documented right-hand-side generator of the serving loop:
    splitmix64 of (seed, i), mapped to [-1, 1)."""
    M = (1 << 64) - 1
    out = np.empty(n)
    for i in range(n):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = (z >> 11) * (2.0 / 9007199254740992.0) - 1.0
    return out


@pytest.fixture(scope="module")
def setup(m):
    systems = make_systems(36, seed=11, nmin=64, nmax=2500)
    systems.append(gen.laplacian3d(20, 20, 18))     # "big": whole-GPU engine
    B = m.Batch(systems)
    return systems, B


def trace(nsys, nreq, rate, seed):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    mat = rng.integers(0, nsys, nreq)
    nrhs = rng.integers(1, 5, nreq)
    seeds = rng.integers(1, 2 ** 62, nreq, dtype=np.uint64)
    return mat, nrhs, arrivals, seeds


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_serving_solves_every_request(m, setup, prec):
    systems, B = setup
    S = m.Server(B, precision=m.F32 if prec == "f32" else m.F64, rtol=1e-10, big_rows=5000,
                 streams=3, max_items=256, cap_bytes=2e9)
    mat, nrhs, arr, seeds = trace(B.nsys, 80, 4000.0, 1)
    lat, its, xs, st = S.run(mat, nrhs, arr, seeds, want_x=True)
    assert st["errors"] == 0 and st["solves"] == nrhs.sum()
    assert st["big_solves"] == nrhs[mat == B.nsys - 1].sum()
    assert np.all(lat > 0)
    for r in range(0, 80, 7):
        rp, ci, va = systems[mat[r]]
        for k in range(nrhs[r]):
            b = splitmix_uniform(int(seeds[r]) * 131 + k, len(rp) - 1)
            res = np.linalg.norm(b - gen.csr_matvec(rp, ci, va, xs[r][k])) / np.linalg.norm(b)
            assert res <= 1.0e-10 * 1.01, (r, k, res)


def test_serving_result_alone_equals_in_mix(m, setup):
    systems, B = setup
    S = m.Server(B, precision=m.F32, rtol=1e-10, big_rows=5000, streams=4, max_items=128, cap_bytes=2e9)
    mat, nrhs, arr, seeds = trace(B.nsys, 60, 8000.0, 2)
    mat[5] = B.nsys - 1                      # one request on the whole-GPU engine
    lat, its, xs, st = S.run(mat, nrhs, arr, seeds, want_x=True)
    assert st["launches"] >= 2
    for r in (0, 5, 17, 42):
        l1, i1, x1, _ = S.run(mat[r:r + 1], nrhs[r:r + 1], np.zeros(1), seeds[r:r + 1], want_x=True)
        assert np.array_equal(x1[0], xs[r]), r


def test_serving_admission_under_small_cap(m, setup):
    systems, B = setup
    # room for three requests of 2 right-hand sides (n <= 2600: b, x and 3
    # workspace vectors of 2624 doubles each): later ones wait for memory
    sizes = np.diff(B.sys_row)
    cap = float(3 * 2 * (5 * 2624) * 8)
    S = m.Server(B, precision=m.F64, rtol=1e-10, big_rows=10 ** 9, streams=4, max_items=64,
                 cap_bytes=cap)
    rng = np.random.default_rng(3)
    small = np.where(sizes < 2600)[0]
    mat = rng.choice(small, 30)
    nrhs = np.ones(30, np.int32) * 2
    arr = np.zeros(30)                        # all at once
    seeds = rng.integers(1, 2 ** 62, 30, dtype=np.uint64)
    lat, its, xs, st = S.run(mat, nrhs, arr, seeds)
    assert st["errors"] == 0 and st["solves"] == 60
    assert st["peak_bytes"] <= cap
    # every right-hand side holds b, x and the workspace (5 vectors, each
    # padded to 64 elements): the requests cannot all be in flight at once
    r64 = lambda k: (k + 63) // 64 * 64
    total = sum(int(nrhs[i]) * 5 * r64(int(sizes[mat[i]])) * 8 for i in range(30))
    assert total > cap
    assert st["launches"] >= int(np.ceil(total / cap))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_serving_concurrent_big_workers(m, prec):
    """Several whole-GPU workers solving systems of different sizes at once
    (each worker's workspace is re-sized and cleared on its own stream):
    every request converges, with the same bits as when it runs alone."""
    systems = [gen.laplacian3d(a, a, a - 1) for a in (14, 16, 18, 20, 22)]
    B = m.Batch(systems)
    rng = np.random.default_rng(9)
    n = 40
    mat = rng.integers(0, len(systems), n)
    nrhs = rng.integers(1, 4, n)
    arr = np.zeros(n)                         # all at once
    seeds = rng.integers(1, 2 ** 62, n, dtype=np.uint64)
    S = m.Server(B, precision=m.F32 if prec == "f32" else m.F64, rtol=1e-10, big_rows=1000,
                 streams=4, max_items=64, cap_bytes=2e9)
    lat, its, xs, st = S.run(mat, nrhs, arr, seeds, want_x=True)
    assert st["errors"] == 0 and st["big_solves"] == nrhs.sum(), st
    for r in range(n):
        rp, ci, va = systems[mat[r]]
        for k in range(nrhs[r]):
            b = splitmix_uniform(int(seeds[r]) * 131 + k, len(rp) - 1)
            res = np.linalg.norm(b - gen.csr_matvec(rp, ci, va, xs[r][k])) / np.linalg.norm(b)
            assert res <= 1.0e-10 * 1.01, (r, k, res)
    for r in (0, 13, 27, 39):
        _, _, x1, _ = S.run(mat[r:r + 1], nrhs[r:r + 1], np.zeros(1), seeds[r:r + 1], want_x=True)
        assert np.array_equal(x1[0], xs[r]), r


SOLVERS = [("cg", "identity"), ("cg", "diagonal"), ("cg", "ilu_0"),
           ("bicgstab", "identity"), ("bicgstab", "diagonal"), ("bicgstab", "ilu_0")]


def _codes(m, names):
    meth = {"cg": m.CG, "bicgstab": m.BICGSTAB}
    pcs = {"identity": m._lib.PC_IDENTITY, "diagonal": m._lib.PC_DIAGONAL, "ilu_0": m._lib.PC_ILU0}
    return np.array([m.Server.solver(meth[a], pcs[b]) for a, b in names], np.int32)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_serving_solver_mix(m, setup, prec):
    """Requests of every solver (cg / bicgstab with identity, diagonal or
    ilu_0) arrive mixed: each is solved to the FP64 tolerance, on the
    batched kernel or (the big system) the whole-GPU engine, and a request's
    result is the same bits alone and in the mix."""
    systems, B = setup
    S = m.Server(B, precision=m.F32 if prec == "f32" else m.F64, rtol=1e-10, big_rows=5000,
                 streams=4, max_items=128, cap_bytes=2e9)
    mat, nrhs, arr, seeds = trace(B.nsys, 90, 6000.0, 4)
    rng = np.random.default_rng(5)
    names = [SOLVERS[i] for i in rng.integers(0, len(SOLVERS), len(mat))]
    mat[3] = B.nsys - 1                      # big system with bicgstab + ilu_0
    names[3] = ("bicgstab", "ilu_0")
    mat[8] = B.nsys - 1                      # and with cg + ilu_0
    names[8] = ("cg", "ilu_0")
    codes = _codes(m, names)
    print("serving solver mix %s: %d requests" % (prec, len(mat)), flush=True)
    lat, its, xs, st = S.run(mat, nrhs, arr, seeds, want_x=True, solver=codes)
    assert st["errors"] == 0 and st["solves"] == nrhs.sum(), st
    for r in range(len(mat)):
        rp, ci, va = systems[mat[r]]
        for k in range(nrhs[r]):
            b = splitmix_uniform(int(seeds[r]) * 131 + k, len(rp) - 1)
            res = np.linalg.norm(b - gen.csr_matvec(rp, ci, va, xs[r][k])) / np.linalg.norm(b)
            assert res <= 1.0e-10 * 1.01, (r, k, names[r], res)
    picks = [3, 8] + [r for r in range(len(mat)) if names[r] == ("bicgstab", "ilu_0")][:3] + \
            [r for r in range(len(mat)) if names[r] == ("cg", "ilu_0")][:2] + [0]
    for r in picks:
        _, _, x1, _ = S.run(mat[r:r + 1], nrhs[r:r + 1], np.zeros(1), seeds[r:r + 1], want_x=True,
                            solver=codes[r:r + 1])
        assert np.array_equal(x1[0], xs[r]), (r, names[r])


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_serving_fp64_ilu0_matches_oracle(m, setup, method):
    """FP64 requests with ilu_0 on the batched kernel (one workgroup per
    system: MTL4's ILU(0) factors from the device factorization, level-
    scheduled sweeps inside the workgroup) agree with the oracle's solve of
    the same system and right-hand side: iteration count within 2, solution
    within what the residual tolerance allows."""
    systems, B = setup
    S = m.Server(B, precision=m.F64, rtol=1e-10, big_rows=10 ** 9, streams=2, max_items=64,
                 cap_bytes=2e9)
    rng = np.random.default_rng(6)
    idx = rng.choice(B.nsys - 1, 8, replace=False)
    n = len(idx)
    seeds = rng.integers(1, 2 ** 62, n, dtype=np.uint64)
    code = _codes(m, [(method, "ilu_0")])[0]
    lat, its, xs, st = S.run(idx, np.ones(n, np.int32), np.zeros(n), seeds, want_x=True,
                             solver=np.full(n, code, np.int32))
    assert st["errors"] == 0
    for k, r in enumerate(idx):
        rp, ci, va = systems[r]
        O = orc.Matrix(rp, ci, va)
        b = splitmix_uniform(int(seeds[k]) * 131, len(rp) - 1)
        fn = orc.cg if method == "cg" else orc.bicgstab
        xo, sto, ito, _ = fn(O, b, orc.Preconditioner("ilu_0", O), max_iter=20000, rtol=1e-10)
        assert sto == 0
        assert abs(int(its[k]) - ito.i) <= 2, (r, its[k], ito.i)
        assert np.linalg.norm(xs[k][0] - xo) <= 1e-6 * np.linalg.norm(xo), r


@pytest.mark.parametrize("big", [False, True])
def test_serving_mixed_bicgstab_laplacian(m, big):
    """Mixed-precision BiCGSTAB + diagonal on a 233 x 234 Laplacian (54,522
    rows, values exact in FP32): a refresh that kept the old directions made
    this solve diverge (20,000 iterations, residual 1e21); with the
    recurrence restarted at every refresh it converges, on the batched
    kernel and on the whole-GPU engine."""
    systems = [gen.laplacian_setup(233, 234), gen.laplacian_setup(40, 41)]
    B = m.Batch(systems)
    S = m.Server(B, precision=m.F32, rtol=1e-10, big_rows=1000 if big else 10 ** 9, streams=4,
                 max_items=64, cap_bytes=2e9)
    n = 6
    mat = np.zeros(n, np.int32)
    seeds = (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(7919))
    code = _codes(m, [("bicgstab", "diagonal")])[0]
    lat, its, xs, st = S.run(mat, np.ones(n, np.int32), np.zeros(n), seeds, want_x=True,
                             solver=np.full(n, code, np.int32))
    assert st["errors"] == 0, (st, its)
    rp, ci, va = systems[0]
    for k in range(n):
        b = splitmix_uniform(int(seeds[k]) * 131, len(rp) - 1)
        res = np.linalg.norm(b - gen.csr_matvec(rp, ci, va, xs[k][0])) / np.linalg.norm(b)
        assert res <= 1.0e-10 * 1.01, (k, res, its[k])
        assert its[k] < 3000, its[k]
