"""Matrix generators for the workloads (host, numpy; int32 CSR, rows sorted).

laplacian_setup restates mtl::mat::laplacian_setup (matrix/laplacian_setup.hpp:
25-44).  The other families have no MTL4 counterpart; they are defined here:

* diffusion3d(n, seed): 7-point finite-volume diffusion on an n^3 grid, natural
  ordering idx = i + n*(j + n*k).  Every face (interior and Dirichlet boundary)
  carries kappa = 10**U[-2, 2]; an interior face couples its two cells with
  -kappa, the diagonal of a cell is the sum of its six face kappas.  Symmetric,
  strictly diagonally dominant on boundary cells, SPD.  nnz = 7 n^3 - 6 n^2.
  With all kappa = 1 it is the constant-coefficient 7-point Laplacian.
* random_sparse(n, seed, mean=16): row i draws l_i ~ Poisson(mean) clipped to
  [1, 256] columns uniformly (duplicates merged), values U[-1, 1]; the
  diagonal is set to sum|offdiag| + 1 (inserted when not drawn).
* convection_diffusion2d(m, peclet, seed): -Laplace(u) + w.grad(u) on the unit
  square, m^2 interior points, h = 1/(m+1), |w| = peclet, wind angle
  theta ~ U[0, 2 pi) from the seed, first-order upwind convection; rows are
  scaled by h^2, row = i*m + j as in laplacian_setup.  Nonsymmetric M-matrix.
* random_spd(n, seed, nnz_per_row=12): symmetric pattern with about
  nnz_per_row entries per row, off-diagonal values U[-1, 0), diagonal
  sum|offdiag| + 1.
"""
import numpy as np


def _compress(cols, vals, valid):
    """Row-major compression of (n, k) candidate arrays (k sorted offsets)."""
    counts = valid.sum(axis=1)
    rowptr = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    col = cols[valid].astype(np.int32)
    val = vals[valid].astype(np.float64)
    return rowptr.astype(np.int32), col, val


def laplacian_setup(m, n):
    """5-point Laplacian on an m x n grid, Dirichlet (laplacian_setup.hpp:25-44)."""
    N = m * n
    r = np.arange(N, dtype=np.int64)
    i, j = r // n, r % n
    offs = [-n, -1, 0, 1, n]
    cols = np.stack([r + o for o in offs], axis=1)
    valid = np.stack([i > 0, j > 0, np.ones(N, bool), j < n - 1, i < m - 1], axis=1)
    vals = np.stack([np.full(N, -1.0), np.full(N, -1.0), np.full(N, 4.0),
                     np.full(N, -1.0), np.full(N, -1.0)], axis=1)
    return _compress(cols, vals, valid)


def laplacian3d(nx, ny, nz):
    """Constant-coefficient 7-point Laplacian, idx = i + nx*(j + ny*k)."""
    N = nx * ny * nz
    r = np.arange(N, dtype=np.int64)
    i, j, k = r % nx, (r // nx) % ny, r // (nx * ny)
    sxy = nx * ny
    offs = [-sxy, -nx, -1, 0, 1, nx, sxy]
    cols = np.stack([r + o for o in offs], axis=1)
    valid = np.stack([k > 0, j > 0, i > 0, np.ones(N, bool), i < nx - 1, j < ny - 1,
                      k < nz - 1], axis=1)
    vals = np.where(np.arange(7) == 3, 6.0, -1.0)[None, :].repeat(N, 0)
    return _compress(cols, vals, valid)


def diffusion3d(n, seed=0, lo=-2.0, hi=2.0):
    """Variable-coefficient 7-point diffusion (see module docstring)."""
    rng = np.random.default_rng(seed)
    N = n ** 3
    # face kappas: kx[k, j, i] is the face at x-index i (between cells i-1 and i)
    kx = 10.0 ** rng.uniform(lo, hi, (n, n, n + 1))
    ky = 10.0 ** rng.uniform(lo, hi, (n, n + 1, n))
    kz = 10.0 ** rng.uniform(lo, hi, (n + 1, n, n))
    west, east = kx[:, :, :-1], kx[:, :, 1:]
    south, north = ky[:, :-1, :], ky[:, 1:, :]
    down, up = kz[:-1, :, :], kz[1:, :, :]
    diag = (west + east + south + north + down + up).reshape(N)
    r = np.arange(N, dtype=np.int64)
    i, j, k = r % n, (r // n) % n, r // (n * n)
    s2 = n * n
    offs = [-s2, -n, -1, 0, 1, n, s2]
    cols = np.stack([r + o for o in offs], axis=1)
    valid = np.stack([k > 0, j > 0, i > 0, np.ones(N, bool), i < n - 1, j < n - 1, k < n - 1], axis=1)
    vals = np.stack([-down.reshape(N), -south.reshape(N), -west.reshape(N), diag,
                     -east.reshape(N), -north.reshape(N), -up.reshape(N)], axis=1)
    return _compress(cols, vals, valid)


def random_sparse(n, seed=0, mean=16, max_len=256):
    """Random rows, Poisson(mean) lengths clipped to [1, max_len], dominant diagonal."""
    rng = np.random.default_rng(seed)
    lens = np.clip(rng.poisson(mean, n), 1, max_len).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    cols = rng.integers(0, n, rows.size, dtype=np.int64)
    keys = np.unique(np.concatenate([rows * n + cols, np.arange(n, dtype=np.int64) * (n + 1)]))
    r = keys // n
    c = keys % n
    vals = rng.uniform(-1.0, 1.0, keys.size)
    isdiag = r == c
    vals[isdiag] = 0.0
    absum = np.bincount(r, weights=np.abs(vals), minlength=n)
    vals[isdiag] = absum + 1.0
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rowptr[1:])
    return rowptr.astype(np.int32), c.astype(np.int32), vals


def convection_diffusion2d(m, peclet=10.0, seed=0):
    """Upwind convection-diffusion on an m x m grid (see module docstring)."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * np.pi)
    wx, wy = peclet * np.cos(theta), peclet * np.sin(theta)
    h = 1.0 / (m + 1)
    N = m * m
    r = np.arange(N, dtype=np.int64)
    i, j = r // m, r % m            # i: slow (y direction), j: fast (x direction)
    # upwind: flux from the side the wind comes from
    cw = -1.0 - h * max(wx, 0.0)    # west  (j-1)
    ce = -1.0 - h * max(-wx, 0.0)   # east  (j+1)
    cs = -1.0 - h * max(wy, 0.0)    # south (i-1)
    cn = -1.0 - h * max(-wy, 0.0)   # north (i+1)
    cd = 4.0 + h * (abs(wx) + abs(wy))
    offs = [-m, -1, 0, 1, m]
    cols = np.stack([r + o for o in offs], axis=1)
    valid = np.stack([i > 0, j > 0, np.ones(N, bool), j < m - 1, i < m - 1], axis=1)
    vals = np.stack([np.full(N, cs), np.full(N, cw), np.full(N, cd), np.full(N, ce),
                     np.full(N, cn)], axis=1)
    return _compress(cols, vals, valid)


def random_spd(n, seed=0, nnz_per_row=12):
    """Symmetric, diagonally dominant (hence SPD) random sparse matrix."""
    rng = np.random.default_rng(seed)
    k = max(1, (nnz_per_row - 1) // 2)
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = rng.integers(0, n, rows.size, dtype=np.int64)
    keep = rows != cols
    rows, cols = rows[keep], cols[keep]
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    ukeys = np.unique(lo * n + hi)
    u = ukeys // n
    c = ukeys % n
    v = rng.uniform(-1.0, 0.0, ukeys.size)
    allr = np.concatenate([u, c, np.arange(n)])
    allc = np.concatenate([c, u, np.arange(n)])
    allv = np.concatenate([v, v, np.zeros(n)])
    order = np.argsort(allr * n + allc, kind="stable")
    allr, allc, allv = allr[order], allc[order], allv[order]
    isdiag = allr == allc
    absum = np.bincount(allr, weights=np.abs(allv), minlength=n)
    allv[isdiag] = absum + 1.0
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(allr, minlength=n), out=rowptr[1:])
    return rowptr.astype(np.int32), allc.astype(np.int32), allv


def csr_matvec(rowptr, col, val, x):
    """Reference-free host product for generating right-hand sides."""
    import scipy.sparse as sp
    n = len(rowptr) - 1
    A = sp.csr_matrix((val, col, rowptr), shape=(n, len(x)))
    return A @ x
