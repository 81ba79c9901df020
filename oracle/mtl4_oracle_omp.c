/*
 * mtl4_oracle_omp.c -- itl::cg + pc::diagonal as MTL4 runs it when built
 * with MTL_WITH_OPENMP, for the all-core CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg (and its
 * CPU test), never by the product.
 *
 * With MTL_WITH_OPENMP the lazy/fused evaluation is off
 * (utility/index_evaluatable.hpp:35-60), so every statement of cg.hpp:74-111
 * is its own pass; the passes MTL4 parallelises are
 *   - the CSR product, four-row blocks shared by an omp for
 *     (operation/mat_vec_mult.hpp:648-674), the remainder rows after it;
 *   - dot<4>: per thread four accumulators over an omp for, combined in an
 *     omp critical section, the remainder added last (operation/dot.hpp:
 *     112-134);
 *   - unary_dot: the same with reduction<8> (vector/reduction.hpp:115-150);
 *   - element-wise vector assignments: one omp for (vector/
 *     vec_vec_aop_expr.hpp:89-96).
 * pc::diagonal's solve is a plain loop (itl/pc/diagonal.hpp:57-64).  The
 * order in which threads enter the critical sections is not fixed, so the
 * last bits of the reductions (and the iteration count, by a step or two)
 * can vary between runs, as they do in MTL4.
 */
#include "mtl4_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void spmv_omp(const orc_csr* A, const double* v, double* w)
{
    const int32_t nr = A->n, nrb = nr / 4 * 4;
    const int32_t* rp = A->rowptr;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < nrb; i += 4) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
        for (int32_t j = rp[i]; j != rp[i + 1]; ++j) t0 += A->val[j] * v[A->col[j]];
        for (int32_t j = rp[i + 1]; j != rp[i + 2]; ++j) t1 += A->val[j] * v[A->col[j]];
        for (int32_t j = rp[i + 2]; j != rp[i + 3]; ++j) t2 += A->val[j] * v[A->col[j]];
        for (int32_t j = rp[i + 3]; j != rp[i + 4]; ++j) t3 += A->val[j] * v[A->col[j]];
        w[i] = t0;
        w[i + 1] = t1;
        w[i + 2] = t2;
        w[i + 3] = t3;
    }
    for (int32_t i = nrb; i < nr; ++i) {
        double t = 0.0;
        for (int32_t j = rp[i]; j != rp[i + 1]; ++j) t += A->val[j] * v[A->col[j]];
        w[i] = t;
    }
}

/* dot<4>: tmp00..tmp03 take i = k (mod 4) inside each thread's share */
static double dot_omp(int64_t n, const double* a, const double* b)
{
    const int64_t nb = n / 4 * 4;
    double result = 0.0;
#pragma omp parallel
    {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nb; i += 4) {
            t0 += a[i] * b[i];
            t1 += a[i + 1] * b[i + 1];
            t2 += a[i + 2] * b[i + 2];
            t3 += a[i + 3] * b[i + 3];
        }
#pragma omp critical
        result += ((t0 + t1) + (t2 + t3)) + ((0.0 + 0.0) + (0.0 + 0.0));
    }
    for (int64_t i = nb; i < n; i++) result += a[i] * b[i];
    return result;
}

/* unary_dot's reduction<8> of squares: t0 + (t1 + (... + (t6 + t7))) per
 * thread (vector/reduction.hpp:115-150, unary_dot.hpp:56-60) */
static double unary_dot_omp(int64_t n, const double* a)
{
    const int64_t nb = n / 8 * 8;
    double result = 0.0;
#pragma omp parallel
    {
        double t[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nb; i += 8)
            for (int k = 0; k < 8; k++) t[k] += a[i + k] * a[i + k];
#pragma omp critical
        result += t[0] + (t[1] + (t[2] + (t[3] + (t[4] + (t[5] + (t[6] + t[7]))))));
    }
    for (int64_t i = nb; i < n; i++) result += a[i] * a[i];
    return result;
}

int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* itl::cg with pc::diagonal (cg.hpp:74-111), statement by statement.
 * threads > 0 sets the OpenMP team size.  basic_iteration semantics as
 * orc_cg: finished(r) on ||r||; i >= max_iter is error 1.              */
int orc_cg_diag_omp(const orc_csr* A, const double* inv_diag, double* x, const double* b,
                    orc_iter* it, int threads)
{
    const int32_t n = A->n;
    if (A->n != A->ncols) return -1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    double* w = (double*)malloc((size_t)(n ? n : 1) * 4 * sizeof(double));
    if (!w) return -5;
    double *p = w, *q = w + n, *r = w + 2 * (size_t)n, *z = w + 3 * (size_t)n;
    double rho = 0.0, rho_1 = 0.0;
    it->error = 0;
    spmv_omp(A, x, q);                                   /* r = b - A*x */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) r[i] = b[i] - q[i];
    double resid = sqrt(fabs(dot_omp(n, r, r)));         /* rr = dot(r, r) */
    for (;;) {
        it->resid = resid;
        const int conv = it->norm_r0 == 0.0 ? resid <= it->atol
                                            : (resid <= it->rtol * it->norm_r0 || resid <= it->atol);
        if (conv) break;
        if (it->i >= it->max_iter) { it->error = 1; break; }
        ++it->i;
        for (int32_t i = 0; i < n; i++) z[i] = inv_diag[i] * r[i];   /* z = solve(P, r) */
        rho = dot_omp(n, r, z);
        if (it->i <= 1) {
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; i++) p[i] = z[i];
        } else {
            const double beta = rho / rho_1;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        }
        spmv_omp(A, p, q);
        const double alpha = rho / dot_omp(n, p, q);
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; i++) x[i] += alpha * p[i];
        rho_1 = rho;
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; i++) r[i] -= alpha * q[i];
        resid = sqrt(fabs(unary_dot_omp(n, r)));
    }
    free(w);
    return it->error;
}
