"""A second restatement of itl::bicgstab (itl/krylov/bicgstab.hpp:27-78),
written with NumPy array operations and independent of oracle/mtl4_oracle.c.

It keeps MTL4's arithmetic so that its results can be compared with the
oracle's bit for bit:
  * y = A*x sums each row in column order from 0.0
    (operation/mat_vec_mult.hpp:656-682);
  * dot(v, w) = sfunctor::dot<4>: accumulator k takes i = k (mod 4) of the
    unrolled block, the remainder goes into accumulator 0, the result is
    ((t0+t1)+(t2+t3)) (operation/dot.hpp:136-147, 226-232);
  * two_norm(v) = reduction<4, two_norm_functor>: the same accumulator
    layout, finished to the right, t0+(t1+(t2+t3)), then sqrt
    (vector/reduction.hpp:28-82,154-173, operation/two_norm.hpp:33-57);
  * pc::diagonal stores reciprocal(A[i][i]) and applies
    y[i] = inv_diag[i]*x[i] (itl/pc/diagonal.hpp:37-45,57-64);
  * basic_iteration::finished (itl/iteration/basic_iteration.hpp:52-93):
    converged first, then i >= max_iter sets error 1.

`mutant` selects one plausible mistake, so that tests can show that their
checks reject it: "beta0" (beta = 0), "no_omega_v" (p = r + beta*p),
"omega_ss" (omega = (t.s)/(s.s)) and "no_half_step_x" (the half-step exit
forgets x += alpha*phat).  The fixture script uses mutant=None only.

Test infrastructure: used by tests/ and tests/golden/make_bicgstab_fixture.py.
"""
import numpy as np


class Csr:
    """Row-major CSR with every row's entries kept in column order, padded
    to the longest row for vectorised row sums (padding adds 0.0*x[0] at a
    row's end, which leaves a sum that started from +0.0 unchanged)."""

    def __init__(self, rowptr, col, val):
        rowptr = np.asarray(rowptr, dtype=np.int64)
        self.n = len(rowptr) - 1
        lens = np.diff(rowptr)
        w = int(lens.max()) if self.n else 0
        self.C = np.zeros((self.n, w), dtype=np.int64)
        self.V = np.zeros((self.n, w))
        for r in range(self.n):
            a, e = rowptr[r], rowptr[r + 1]
            self.C[r, :e - a] = col[a:e]
            self.V[r, :e - a] = val[a:e]
        self.diag = np.array([self.V[r][self.C[r] == r].sum() if (self.C[r, :lens[r]] == r).any() else 0.0
                              for r in range(self.n)])

    def mult(self, x):
        y = np.zeros(self.n)
        for k in range(self.C.shape[1]):
            y = y + self.V[:, k] * x[self.C[:, k]]
        return y


def _acc4(prod):
    """The four accumulators of an unroll-4 loop, each summed left to right
    (np.add.accumulate is sequential), remainder into accumulator 0."""
    n = len(prod)
    nb = 4 * (n // 4)
    t = [float(np.add.accumulate(prod[k:nb:4])[-1]) if nb else 0.0 for k in range(4)]
    for i in range(nb, n):
        t[0] = t[0] + float(prod[i])
    return t


def dot(a, b):
    t = _acc4(a * b)
    return (t[0] + t[1]) + (t[2] + t[3])


def two_norm(a):
    t = _acc4(a * a)
    return float(np.sqrt(t[0] + (t[1] + (t[2] + t[3]))))


class Iteration:
    def __init__(self, norm_r0, max_iter, rtol, atol=0.0):
        self.norm_r0, self.max_iter, self.rtol, self.atol = norm_r0, max_iter, rtol, atol
        self.i, self.error, self.history = 0, 0, []

    def finished(self, r):
        resid = two_norm(r)
        self.history.append(resid)
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


class Breakdown(Exception):
    """unexpected_orthogonality (bicgstab.hpp:46,51,59)."""


def bicgstab(A, b, inv_diag=None, x0=None, max_iter=1000, rtol=1e-10, atol=0.0, mutant=None):
    """itl::bicgstab(A, x, b, M, iter) with M = identity (inv_diag None) or
    pc::diagonal (inv_diag = 1/diag).  Returns (x, error, iter)."""
    n = A.n
    it = Iteration(two_norm(b), max_iter, rtol, atol)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=float)
    solve = (lambda v: v.copy()) if inv_diag is None else (lambda v: inv_diag * v)
    rho_1 = rho_2 = alpha = omega = 0.0
    p = np.zeros(n)
    v = np.zeros(n)
    r = b - A.mult(x)
    rtilde = r.copy()
    while not it.finished(r):
        it.i += 1
        rho_1 = dot(rtilde, r)
        if rho_1 == 0.0:
            raise Breakdown("rho_1")
        if it.i <= 1:
            p = r.copy()
        else:
            if omega == 0.0:
                raise Breakdown("omega")
            beta = (rho_1 / rho_2) * (alpha / omega)
            if mutant == "beta0":
                beta = 0.0
            if mutant == "no_omega_v":
                p = r + beta * p
            else:
                p = r + beta * (p - omega * v)
        phat = solve(p)
        v = A.mult(phat)
        gamma = dot(rtilde, v)
        if gamma == 0.0:
            raise Breakdown("gamma")
        alpha = rho_1 / gamma
        s = r - alpha * v
        if it.finished(s):
            if mutant != "no_half_step_x":
                x = x + alpha * phat
            break
        shat = solve(s)
        t = A.mult(shat)
        omega = dot(t, s) / (dot(s, s) if mutant == "omega_ss" else dot(t, t))
        x = x + (omega * shat + alpha * phat)
        r = s - omega * t
        rho_2 = rho_1
    return x, it.error, it
