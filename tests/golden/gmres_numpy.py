"""A second restatement of itl::gmres_full and itl::gmres
(itl/krylov/gmres.hpp:37-137), written with NumPy and independent of
oracle/mtl4_oracle.c.

It keeps MTL4's arithmetic so that its results can be compared with the
oracle's bit for bit (row sums, dot and two_norm as in bicgstab_numpy):
  * the Arnoldi step: modified Gram-Schmidt, one reorthogonalisation pass,
    then V[k+1] *= 1/H[k+1][k] (gmres.hpp:64-81);
  * rotation i < k of step k is rebuilt from column k-1 of H, as the
    current H holds it, and applied to the whole of H: rows i, i+1 become
    trans(G) * rows, then columns i, i+1 become columns * G, with
    G = [c s; -s c] from mtl::mat::givens::set_rotation
    (gmres.hpp:84-85, operation/givens.hpp:38-72);
  * the new rotation (c[k], s[k]) zeroes H[k+1][k] and turns g[k], g[k+1]
    as mtl::vec::givens::trafo does (gmres.hpp:87-94, givens.hpp:92-101);
  * k shrinks while |g[k-1]| <= atol, then lu_solve on the leading
    range x range block of H, shrinking the range while the factorisation
    finds a zero pivot (gmres.hpp:99-109).  lu_solve works in place on that
    block (the copy of a sub-matrix view shares its memory,
    detail/contiguous_memory_block.hpp:295-302): partial pivoting on the
    first entry of largest magnitude, whole block rows swapped, every pivot
    but the last checked against 0 (operation/lu.hpp:66-89,
    operation/max_abs_pos.hpp:31-56); permute(P, b) scatters,
    b'[P[i]] = b[i] (operation/permute.hpp:24-35); then the unit lower and
    the upper row sweeps, the upper one dividing by the diagonal
    (lower_trisolve.hpp:96-109, upper_trisolve.hpp:101,110-126);
  * x += solve(R, sum_j V[j]*y[j]), accumulated from zero in j order
    (operation/mat_vec_mult.hpp:352-360), then terminate(b - A*x);
  * gmres(restart): cycles of gmres_full on a copy of the iteration limited
    to i + restart, each folded back by update_progress
    (basic_iteration.hpp:145-155).

`mutant` selects one plausible mistake so that tests can show their
checks reject it: "textbook_rotations" (rotate only the new column with the
stored (c[i], s[i]), the usual GMRES), "gather_permute" (b'[i] = b[P[i]]),
"no_reorth" (skip the second Gram-Schmidt pass), "lu_on_copy" (each lu_solve
attempt starts from the unmodified block).

Test infrastructure: used by tests/ only.
"""
import math

import numpy as np

from bicgstab_numpy import Iteration, dot, two_norm


class _Iteration(Iteration):
    def finished_value(self, resid):
        self.history.append(resid)
        self.resid = resid
        return self._decide(resid)

    def finished(self, r):
        return self.finished_value(two_norm(r))

    def _decide(self, resid):
        if self.norm_r0 == 0:
            conv = resid <= self.atol
        else:
            conv = resid <= self.rtol * self.norm_r0 or resid <= self.atol
        if conv:
            return True
        if self.i >= self.max_iter:
            self.error = 1
            return True
        return False


class Singular(Exception):
    pass


def _rotation(a, b):
    """mtl::mat::givens::set_rotation(a, b) -> G (2x2)."""
    if b == 0.0:
        c, s = 1.0, 0.0
    elif abs(b) > abs(a):
        t = -a / b
        s = 1.0 / math.sqrt(1.0 + t * t)
        c = s * t
    else:
        t = -b / a
        c = 1.0 / math.sqrt(1.0 + t * t)
        s = c * t
    return np.array([[c, s], [-s, c]])


def _trafo(H, G, i):
    """H[i:i+2, :] = trans(G) * H[i:i+2, :]; H[:, i:i+2] = H[:, i:i+2] * G,
    each entry a two-term product sum (elementwise, no fused multiply-add)."""
    rows = H[i:i + 2, :].copy()
    H[i:i + 2, :] = np.stack([G[0, a] * rows[0] + G[1, a] * rows[1] for a in range(2)])
    cols = H[:, i:i + 2].copy()
    H[:, i:i + 2] = np.stack([cols[:, 0] * G[0, b] + cols[:, 1] * G[1, b] for b in range(2)], axis=1)


def _lu_solve(B, rhs, mutant=None):
    """lu_solve(B, rhs) with B factored in place (B is a view)."""
    m = B.shape[0]
    P = np.arange(m)
    for i in range(m - 1):
        rmax = i + int(np.argmax(np.abs(B[i:, i])))
        B[[i, rmax], :] = B[[rmax, i], :]
        P[[i, rmax]] = P[[rmax, i]]
        if abs(B[i, i]) <= 0.0:
            raise Singular
        B[i + 1:, i] = B[i + 1:, i] / B[i, i]
        B[i + 1:, i + 1:] = B[i + 1:, i + 1:] - np.outer(B[i + 1:, i], B[i, i + 1:])
    pb = np.empty(m)
    if mutant == "gather_permute":
        pb[:] = rhs[P]
    else:
        pb[P] = rhs
    z = np.empty(m)
    for r in range(m):
        acc = pb[r]
        for c in range(r):
            acc = acc - B[r, c] * z[c]
        z[r] = acc
    y = np.empty(m)
    for r in range(m - 1, -1, -1):
        acc = z[r]
        for c in range(r + 1, m):
            acc = acc - B[r, c] * y[c]
        y[r] = acc / B[r, r]
    return y


def gmres_full(A, b, x, it, inv_l=None, inv_r=None, mutant=None):
    """itl::gmres_full(A, x, b, L, R, iter); L, R are pc::diagonal given as
    inverse diagonals, or identity (None).  Updates x and it in place."""
    solve_l = (lambda v: v.copy()) if inv_l is None else (lambda v: inv_l * v)
    solve_r = (lambda v: v.copy()) if inv_r is None else (lambda v: inv_r * v)
    n = A.n
    kmax = min(n, it.max_iter - it.i)
    V = np.zeros((kmax + 1, n))
    H = np.zeros((kmax + 1, kmax))
    g = np.zeros(kmax + 1)
    cs = np.zeros(kmax + 1)
    sn = np.zeros(kmax + 1)
    r = solve_l(b - A.mult(x))
    g[0] = two_norm(r)
    if it.finished_value(g[0]):
        return
    V[0] = r / g[0]
    k = 0
    while k < kmax:
        V[k + 1] = solve_l(A.mult(solve_r(V[k])))
        for j in range(k + 1):
            H[j, k] = dot(V[j], V[k + 1])
            V[k + 1] = V[k + 1] - H[j, k] * V[j]
        H[k + 1, k] = two_norm(V[k + 1])
        if mutant != "no_reorth":
            for j in range(k + 1):
                hr = dot(V[k + 1], V[j])
                H[j, k] = H[j, k] + hr
                V[k + 1] = V[k + 1] - hr * V[j]
            H[k + 1, k] = two_norm(V[k + 1])
        if H[k + 1, k] != 0.0:
            V[k + 1] = V[k + 1] * (1.0 / H[k + 1, k])
        if mutant == "textbook_rotations":
            for i in range(k):
                h0, h1 = H[i, k], H[i + 1, k]
                H[i, k] = cs[i] * h0 - sn[i] * h1
                H[i + 1, k] = sn[i] * h0 + cs[i] * h1
        else:
            for i in range(k):
                _trafo(H, _rotation(H[i, k - 1], H[i + 1, k - 1]), i)
        nu = math.sqrt(H[k, k] * H[k, k] + H[k + 1, k] * H[k + 1, k])
        if nu != 0.0:
            cs[k] = H[k, k] / nu
            sn[k] = -H[k + 1, k] / nu
            H[k, k] = cs[k] * H[k, k] - sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k], g[k + 1] = cs[k] * g[k] - sn[k] * g[k + 1], sn[k] * g[k] + cs[k] * g[k + 1]
        k += 1
        it.i += 1
    while k > 0 and abs(g[k - 1]) <= it.atol:
        k -= 1
    m = k
    y = None
    while m > 0:
        block = H[:m, :m].copy() if mutant == "lu_on_copy" else H[:m, :m]
        try:
            y = _lu_solve(block, g[:m].copy(), mutant)
            break
        except Singular:
            m -= 1
    if m == 0:
        it.error = 2
        return
    u = np.zeros(n)
    for j in range(m):
        u = u + V[j] * y[j]
    x += solve_r(u)
    it.finished(b - A.mult(x))


def gmres(A, b, restart, inv_l=None, inv_r=None, x0=None, max_iter=100, rtol=1e-10, atol=0.0, mutant=None):
    """itl::gmres(A, x, b, L, R, iter, restart) for restart > 0, gmres_full
    for restart == 0, with iter = basic_iteration(b, max_iter, rtol, atol).
    Returns (x, error, iterations, residuals handed to finished())."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=float)
    it = _Iteration(two_norm(b), max_iter, rtol, atol)
    if restart == 0:
        gmres_full(A, b, x, it, inv_l, inv_r, mutant)
        return x, it.error, it.i, np.array(it.history)
    history = []
    while True:
        inner = _Iteration(it.norm_r0, min(it.i + restart, it.max_iter), rtol, atol)
        inner.i = it.i
        gmres_full(A, b, x, inner, inv_l, inv_r, mutant)
        history.extend(inner.history)
        it.i = inner.i
        if inner.error > 1:
            it.error = inner.error
            break
        if it._decide(inner.resid):
            break
    return x, it.error, it.i, np.array(history)
