"""Writes tests/golden/bicgstab_fixture.json: inputs and outputs of
itl::bicgstab (itl/krylov/bicgstab.hpp:27-78) on small systems, computed by
the NumPy restatement in tests/golden/bicgstab_numpy.py (not by the C
oracle, which tests/test_oracle.py then checks against this file bit for
bit: iteration count, status, every residual handed to finished(), x).

Every value is stored as float.hex, so the file round-trips exactly.
Run:  python tests/golden/make_bicgstab_fixture.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bicgstab_numpy as bn  # noqa: E402


def hexes(a):
    return [float(v).hex() for v in np.asarray(a, dtype=float).ravel()]


def to_csr(D):
    rp, ci, va = [0], [], []
    for r in range(D.shape[0]):
        for c in np.nonzero(D[r])[0]:
            ci.append(int(c))
            va.append(float(D[r, c]))
        rp.append(len(ci))
    return rp, ci, va


def convdiff(m, peclet, theta):
    """Upwind 5-point -Laplace(u) + w.grad(u) on an m x m interior grid of
    the unit square (h = 1/(m+1)), rows in natural order, each row's
    entries in column order."""
    h = 1.0 / (m + 1)
    wx, wy = peclet * np.cos(theta), peclet * np.sin(theta)
    n = m * m
    D = np.zeros((n, n))
    for i in range(m):
        for j in range(m):
            r = i * m + j
            cW = -1.0 / h ** 2 - max(wx, 0.0) / h
            cE = -1.0 / h ** 2 + min(wx, 0.0) / h
            cS = -1.0 / h ** 2 - max(wy, 0.0) / h
            cN = -1.0 / h ** 2 + min(wy, 0.0) / h
            D[r, r] = 4.0 / h ** 2 + abs(wx) / h + abs(wy) / h
            if j > 0:
                D[r, r - 1] = cW
            if j < m - 1:
                D[r, r + 1] = cE
            if i > 0:
                D[r, r - m] = cS
            if i < m - 1:
                D[r, r + m] = cN
    return D


def cases():
    rng = np.random.default_rng(2024)
    out = []
    # nonsymmetric convection-diffusion, 10 x 10 grid, Peclet 10 and 200,
    # rows scaled by factors in [0.5, 2] so the diagonal varies
    for pe, theta in ((10.0, 0.7), (200.0, 2.3)):
        D = convdiff(10, pe, theta)
        D = np.diag(rng.uniform(0.5, 2.0, D.shape[0])) @ D
        xt = rng.uniform(-1.0, 1.0, D.shape[0])
        for pc in ("identity", "diagonal"):
            out.append(("convdiff10_pe%d_%s" % (pe, pc), D, D @ xt, None, pc, 1e-10, 1000))
    # random nonsymmetric, diagonally dominant, n = 70, nonzero start
    n = 70
    D = rng.uniform(-1.0, 1.0, (n, n)) * (rng.random((n, n)) < 0.08)
    np.fill_diagonal(D, np.abs(D).sum(1) * rng.uniform(0.6, 1.4, n) + 0.1)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    for pc in ("identity", "diagonal"):
        out.append(("random70_x0_%s" % pc, D, b, x0, pc, 1e-12, 1000))
    # max_iter reached: error 1 after 7 iterations
    out.append(("random70_x0_maxiter", D, b, x0, "diagonal", 1e-14, 7))
    # 5 distinct eigenvalues on a non-diagonal matrix (similarity by a
    # well-conditioned triangular T): BiCG terminates at step 5, so the
    # half-step check of iteration 5 stops the solve
    n = 30
    lam = np.repeat([1.0, 2.5, 4.0, 7.0, 9.5], 6)
    T = np.eye(n) + np.triu(rng.uniform(-0.3, 0.3, (n, n)), 1) * (rng.random((n, n)) < 0.2)
    D = T @ np.diag(lam) @ np.linalg.inv(T)
    D[np.abs(D) < 1e-13] = 0.0
    out.append(("five_eigenvalues_identity", D, rng.uniform(0.5, 1.5, n), None, "identity", 1e-9, 50))
    return out


def main():
    data = {"source": "tests/golden/bicgstab_numpy.py (NumPy restatement of itl/krylov/bicgstab.hpp:27-78)",
            "systems": {}, "cases": []}
    for name, D, b, x0, pc, rtol, max_iter in cases():
        rp, ci, va = to_csr(D)
        A = bn.Csr(rp, ci, np.array(va))
        inv = None if pc == "identity" else 1.0 / A.diag
        x, err, it = bn.bicgstab(A, np.asarray(b, float), inv, x0, max_iter, rtol)
        system = name.rsplit("_", 1)[0]   # the cases of one system differ in pc / max_iter only
        if system not in data["systems"]:
            data["systems"][system] = {"rowptr": rp, "col": ci, "val": hexes(va), "b": hexes(b),
                                       "x0": None if x0 is None else hexes(x0)}
        data["cases"].append({
            "name": name, "system": system, "pc": pc, "rtol": rtol, "max_iter": max_iter,
            "iterations": it.i, "error": err, "history": hexes(it.history), "x": hexes(x)})
        print("%-28s n=%3d iterations %3d error %d checks %d last %.3e"
              % (name, A.n, it.i, err, len(it.history), it.history[-1]))
    with open(os.path.join(HERE, "bicgstab_fixture.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
