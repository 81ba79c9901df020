"""Serving loop with continuous batching: every request is solved to the
FP64 tolerance, admission respects the working-set cap, and each result is
bit-for-bit the same alone and in the mix."""
import numpy as np
import pytest

from mtl4_amd import generators as gen
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
