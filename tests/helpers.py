"""Shared test helpers: one CSR input handed to both the oracle and the HIP
library, plus the matrix families the tests sweep over."""
import numpy as np

import orc
from mtl4_amd import generators as gen


def pair(rowptr, col, val, ncols=None):
    """(oracle matrix, device compressed2D) built from the same arrays."""
    import mtl4_amd as m
    n = len(rowptr) - 1
    nc = n if ncols is None else ncols
    return orc.Matrix(rowptr, col, val, ncols=nc), m.compressed2D(n, nc, rowptr, col, val)


def family(name, seed=0):
    """Small instances of every family the workloads use."""
    if name == "lap2d_64":
        return gen.laplacian_setup(64, 64)
    if name == "lap2d_5":
        return gen.laplacian_setup(5, 5)
    if name == "lap3d_12":
        return gen.laplacian3d(12, 11, 10)
    if name == "diff3d_16":
        return gen.diffusion3d(16, seed)
    if name == "random_3000":
        return gen.random_sparse(3000, seed)
    if name == "convdiff_40":
        return gen.convection_diffusion2d(40, 10.0, seed)
    if name == "spd_2000":
        return gen.random_spd(2000, seed)
    if name == "long_rows":
        # rows longer than the LDS tile, empty rows and a dense row
        rng = np.random.default_rng(seed)
        n = 700
        lens = rng.integers(0, 9, n)
        lens[5] = 0
        lens[100] = 5000 // 7
        lens[101] = n
        lens[n - 1] = 0
        rp = np.zeros(n + 1, np.int64)
        cols, vals = [], []
        for r in range(n):
            c = np.sort(rng.choice(n, size=min(lens[r], n), replace=False))
            cols.append(c)
            vals.append(rng.uniform(-1, 1, c.size))
            rp[r + 1] = rp[r] + c.size
        return rp.astype(np.int32), np.concatenate(cols).astype(np.int32), np.concatenate(vals)
    raise KeyError(name)


FAMILIES = ["lap2d_5", "lap2d_64", "lap3d_12", "diff3d_16", "random_3000", "convdiff_40",
            "spd_2000", "long_rows"]
