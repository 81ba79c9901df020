/*
 * mtl4_oracle.c -- CPU restatement of MTL4's sparse Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Build: gcc -std=c99 -O3 -fPIC -shared -ffp-contract=off (oracle/Makefile).
 * -ffp-contract=off keeps every a*b+c as a rounded product followed by a
 * rounded sum, as MTL4's default (non -march=native) x86-64 build computes it.
 *
 * Citations are to files of the MTL4 tree (boost/numeric/...).
 */
#include "mtl4_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_NOT_SQUARE (-1)
#define ORC_SIZE_MISMATCH (-2)
#define ORC_MISSING_DIAG (-3)
#define ORC_BREAKDOWN (-4)
#define ORC_OOM (-5)
#define ORC_INVALID (-6)   /* MTL4's logic_error */

/* ------------------------------------------------------------------ */
/* laplacian_setup: mtl/matrix/laplacian_setup.hpp:25-44.  The inserter
 * (compressed2D.hpp:1072-1151) leaves every row sorted by column, so the
 * row of grid point (i,j) reads  row-n, row-1, row, row+1, row+n.       */
int64_t orc_laplacian_setup(int32_t m, int32_t n, int32_t* rowptr, int32_t* col, double* val)
{
    int64_t nnz = 0;
    int32_t row = 0;
    if (rowptr) rowptr[0] = 0;
    for (int32_t i = 0; i < m; i++)
        for (int32_t j = 0; j < n; j++, row++) {
            int32_t c[5]; double v[5]; int k = 0;
            if (i > 0)     { c[k] = row - n; v[k++] = -1.0; }
            if (j > 0)     { c[k] = row - 1; v[k++] = -1.0; }
            c[k] = row; v[k++] = 4.0;
            if (j < n - 1) { c[k] = row + 1; v[k++] = -1.0; }
            if (i < m - 1) { c[k] = row + n; v[k++] = -1.0; }
            for (int t = 0; t < k; t++, nnz++) {
                if (col) col[nnz] = c[t];
                if (val) val[nnz] = v[t];
            }
            if (rowptr) rowptr[row + 1] = (int32_t)nnz;
        }
    return nnz;
}

/* ------------------------------------------------------------------ */
/* One row of smat_cvec_mult, operation/mat_vec_mult.hpp:656-682:
 * tmp starts at zero and accumulates A.data[j]*v[minor[j]] in column order. */
static double row_sum(const orc_csr* A, int32_t i, const double* x)
{
    double tmp = 0.0;
    for (int32_t j = A->rowptr[i]; j != A->rowptr[i + 1]; ++j)
        tmp += A->val[j] * x[A->col[j]];
    return tmp;
}

/* operation/mat_vec_mult.hpp:613-683.  Every variant (4-row unrolled main
 * loop, very-sparse :416-443, empty-row adaptive :545-611) forms the row sum
 * first and then applies Assign::first_update (operation/assign_mode.hpp):
 * =, += or -=, so the result per row is the same in all of them.        */
void orc_spmv(const orc_csr* A, const double* x, double* y, int mode)
{
    for (int32_t i = 0; i < A->n; i++) {
        double t = row_sum(A, i, x);
        if (mode == ORC_ASSIGN) y[i] = t;
        else if (mode == ORC_PLUS) y[i] += t;
        else y[i] -= t;
    }
}

/* ------------------------------------------------------------------ */
/* sfunctor::dot<Unroll>::apply, operation/dot.hpp:136-147 (non-OpenMP):
 * partial tmp0k takes indices i = k (mod Unroll) of the unrolled block,
 * the remainder goes into tmp00, result ((t0+t1)+(t2+t3))+((t4+t5)+(t6+t7)). */
double orc_dot(int64_t n, const double* a, const double* b, int unroll)
{
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (unroll < 1) unroll = 1;
    if (unroll > 8) unroll = 8;
    if (unroll == 1) { /* dot_class<1> -> dot_simple, dot.hpp:153-163 */
        double s = 0.0;
        for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
        return s;
    }
    int64_t ib = unroll * (n / unroll);
    for (int64_t i = 0; i < ib; i += unroll)
        for (int k = 0; k < unroll; k++) t[k] += a[i + k] * b[i + k];
    for (int64_t i = ib; i < n; i++) t[0] += a[i] * b[i];
    return ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
}

/* reduction<Unroll, two_norm_functor>::apply, vector/reduction.hpp:154-173
 * with impl::reduction (:28-82): accumulator k takes i = k (mod Unroll),
 * the remainder goes into tmp00, and finish() nests to the right:
 * t0 + (t1 + (t2 + ... + (t[U-2] + t[U-1]))).                           */
double orc_unary_dot(int64_t n, const double* a, int unroll)
{
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (unroll < 1) unroll = 1;
    if (unroll > 8) unroll = 8;
    int64_t ib = unroll * (n / unroll);
    for (int64_t i = 0; i < ib; i += unroll)
        for (int k = 0; k < unroll; k++) t[k] += a[i + k] * a[i + k];
    for (int64_t i = ib; i < n; i++) t[0] += a[i] * a[i];
    for (int k = unroll - 2; k >= 0; k--) t[k] += t[k + 1];
    return t[0];
}

/* two_norm(v) = two_norm<4>(v), operation/two_norm.hpp:33-57 */
double orc_two_norm(int64_t n, const double* a)
{
    return sqrt(orc_unary_dot(n, a, 4));
}

/* (lazy(q)= A*p) || (lazy(alpha)= lazy_dot(p,q)): fused_expr.hpp:57-77 runs
 * both evaluators per index with 4-way unrolling; row_mat_cvec_index_evaluator
 * (:36-44) writes q[i], dot_index_evaluator (:22-39) adds p[i]*q[i] into
 * tmp[i mod 4] (remainder rows into tmp[0]) and sums tmp0+tmp1+tmp2+tmp3
 * from the left in its destructor.                                      */
double orc_fused_spmv_dot(const orc_csr* A, const double* p, double* q)
{
    double t[4] = {0, 0, 0, 0};
    int64_t n = A->n, sb = n >> 2 << 2;
    for (int64_t i = 0; i < sb; i += 4)
        for (int k = 0; k < 4; k++) {
            q[i + k] = row_sum(A, (int32_t)(i + k), p);
            t[k] += p[i + k] * q[i + k];
        }
    for (int64_t i = sb; i < n; i++) {
        q[i] = row_sum(A, (int32_t)i, p);
        t[0] += p[i] * q[i];
    }
    return ((t[0] + t[1]) + t[2]) + t[3];
}

/* (lazy(r) -= alpha*q) || (lazy(rr)= lazy_unary_dot(r)): the reduction
 * index evaluator (vector/reduction_index_evaluator.hpp:22-41) finishes
 * (t0+t1)+(t2+t3).                                                      */
double orc_fused_axpy_unary_dot(int64_t n, double alpha, const double* q, double* r)
{
    double t[4] = {0, 0, 0, 0};
    int64_t sb = n >> 2 << 2;
    for (int64_t i = 0; i < sb; i += 4)
        for (int k = 0; k < 4; k++) {
            r[i + k] -= alpha * q[i + k];
            t[k] += r[i + k] * r[i + k];
        }
    for (int64_t i = sb; i < n; i++) {
        r[i] -= alpha * q[i];
        t[0] += r[i] * r[i];
    }
    return (t[0] + t[1]) + (t[2] + t[3]);
}

/* ------------------------------------------------------------------ */
/* Owned CSR used for factors. */
typedef struct {
    int32_t n;
    int32_t* rowptr;
    int32_t* col;
    double* val;
} own_csr;

static void own_free(own_csr* M)
{
    free(M->rowptr); free(M->col); free(M->val);
    M->rowptr = NULL; M->col = NULL; M->val = NULL;
}

static int own_alloc(own_csr* M, int32_t n, int64_t nnz)
{
    M->n = n;
    M->rowptr = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    M->col = (int32_t*)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    M->val = (double*)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    return (M->rowptr && M->col && M->val) ? ORC_OK : ORC_OOM;
}

static orc_csr as_view(const own_csr* M)
{
    orc_csr v;
    v.n = M->n; v.ncols = M->n; v.rowptr = M->rowptr; v.col = M->col; v.val = M->val;
    return v;
}

/* compressed2D_indexer::offset (compressed2D.hpp:181-204): search row r for
 * column c; -1 if absent (operator() then reads zero).                  */
static int64_t find_entry(const int32_t* rowptr, const int32_t* col, int32_t r, int32_t c)
{
    int32_t lo = rowptr[r], hi = rowptr[r + 1];
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (col[mid] < c) lo = mid + 1; else hi = mid;
    }
    return (lo < rowptr[r + 1] && col[lo] == c) ? lo : -1;
}

/* Copy the entries of A with lo_off <= col-row <= hi_off (bands.hpp:34-40;
 * strict_lower = bands(-inf,0), upper = bands(0,inf)), rounding values to
 * float when value_bytes == 4 (conversion into compressed2D<float>).    */
static int copy_band(const orc_csr* A, int strict_lower, own_csr* out, int value_bytes)
{
    int64_t nnz = 0;
    for (int32_t r = 0; r < A->n; r++)
        for (int32_t j = A->rowptr[r]; j < A->rowptr[r + 1]; j++)
            if (strict_lower ? (A->col[j] < r) : (A->col[j] >= r)) nnz++;
    if (own_alloc(out, A->n, nnz) != ORC_OK) return ORC_OOM;
    int64_t k = 0;
    for (int32_t r = 0; r < A->n; r++) {
        for (int32_t j = A->rowptr[r]; j < A->rowptr[r + 1]; j++)
            if (strict_lower ? (A->col[j] < r) : (A->col[j] >= r)) {
                out->col[k] = A->col[j];
                out->val[k] = value_bytes == 4 ? (double)(float)A->val[j] : A->val[j];
                k++;
            }
        out->rowptr[r + 1] = (int32_t)k;
    }
    return ORC_OK;
}

/* Explicit transpose (trans(U) copied into compressed2D, ic_0.hpp:57). */
static int transpose(const own_csr* U, own_csr* L)
{
    int32_t n = U->n;
    int64_t nnz = U->rowptr[n];
    if (own_alloc(L, n, nnz) != ORC_OK) return ORC_OOM;
    for (int64_t j = 0; j < nnz; j++) L->rowptr[U->col[j] + 1]++;
    for (int32_t r = 0; r < n; r++) L->rowptr[r + 1] += L->rowptr[r];
    int32_t* fill = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
    if (!fill) return ORC_OOM;
    memcpy(fill, L->rowptr, ((size_t)n + 1) * sizeof(int32_t));
    for (int32_t r = 0; r < n; r++)              /* rows ascending -> columns */
        for (int32_t j = U->rowptr[r]; j < U->rowptr[r + 1]; j++) { /* sorted */
            int32_t d = fill[U->col[j]]++;
            L->col[d] = r;
            L->val[d] = U->val[j];
        }
    free(fill);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Triangular sweeps. */

/* lower_trisolve_t version 5 (unit diagonal, strictly lower compact CSR),
 * operation/lower_trisolve.hpp:174-189.                                 */
void orc_lower_unit(const orc_csr* L, const double* v, double* w)
{
    for (int32_t r = 0; r < L->n; r++) {
        double rr = v[r];
        for (int32_t j = L->rowptr[r]; j != L->rowptr[r + 1]; ++j)
            rr -= L->val[j] * w[L->col[j]];
        w[r] = rr;
    }
}

/* lower_trisolve_t version 6 (inverse diagonal stored last in the row),
 * operation/lower_trisolve.hpp:193-209.                                 */
int orc_lower_invdiag_last(const orc_csr* L, const double* v, double* w)
{
    for (int32_t r = 0; r < L->n; r++) {
        int32_t j0 = L->rowptr[r], j1 = L->rowptr[r + 1];
        if (j0 == j1) return ORC_MISSING_DIAG;
        --j1;
        if (L->col[j1] != r) return ORC_MISSING_DIAG;
        double dia = L->val[j1];
        double rr = v[r];
        for (; j0 != j1; ++j0) rr -= L->val[j0] * w[L->col[j0]];
        w[r] = rr * dia;
    }
    return ORC_OK;
}

/* upper_trisolve_t version 3 (inverse diagonal stored first in the row),
 * operation/upper_trisolve.hpp:161-176 with row_update :101-103.         */
void orc_upper_invdiag_first(const orc_csr* U, const double* v, double* w)
{
    for (int32_t r = U->n; r-- > 0;) {
        int32_t j0 = U->rowptr[r];
        const int32_t cj1 = U->rowptr[r + 1];
        double rr = v[r], dia = U->val[j0++];
        for (; j0 != cj1; ++j0) rr -= U->val[j0] * w[U->col[j0]];
        w[r] = rr * dia;
    }
}

/* Column-oriented sweeps on a transposed view (adjoint solves): generic
 * lower version 4 (lower_trisolve.hpp:154-170) applied to trans(U) with the
 * inverse diagonal, and generic upper version 2 (upper_trisolve.hpp:128-147)
 * applied to trans(L) with a unit diagonal.                              */
static void adj_lower_on_upper(const own_csr* U, const double* v, double* w)
{
    int32_t n = U->n;
    memmove(w, v, (size_t)n * sizeof(double));
    for (int32_t r = 0; r < n; r++) {
        int32_t j = U->rowptr[r];
        double rr = w[r] *= U->val[j];        /* diagonal first in row r of U */
        for (++j; j < U->rowptr[r + 1]; j++) w[U->col[j]] -= U->val[j] * rr;
    }
}

static void adj_upper_on_lower(const own_csr* L, const double* v, double* w)
{
    int32_t n = L->n;
    memmove(w, v, (size_t)n * sizeof(double));
    for (int32_t r = n; r-- > 0;) {
        double rr = w[r];                       /* unit diagonal */
        for (int32_t j = L->rowptr[r]; j < L->rowptr[r + 1]; j++) w[L->col[j]] -= L->val[j] * rr;
    }
}

/* ------------------------------------------------------------------ */
/* Preconditioners. */
struct orc_pc {
    int kind;
    int32_t n;
    double* inv_diag;     /* diagonal */
    own_csr L, U;         /* ilu_0: L strict lower, U upper; ic_0: L = trans(U) */
    double* y0;           /* workspace (MTL keeps a function-local static) */
};

/* diagonal(A): inv_diag[i]= reciprocal(A[i][i]), itl/pc/diagonal.hpp:37-45;
 * with Value = float the reciprocal is stored in a dense_vector<float>.  */
static int setup_diagonal(orc_pc* P, const orc_csr* A, int value_bytes)
{
    P->inv_diag = (double*)malloc((size_t)A->n * sizeof(double));
    if (!P->inv_diag) return ORC_OOM;
    for (int32_t i = 0; i < A->n; i++) {
        int64_t k = find_entry(A->rowptr, A->col, i, i);
        double a = k >= 0 ? A->val[k] : 0.0;
        P->inv_diag[i] = 1.0 / a;
        if (value_bytes == 4) P->inv_diag[i] = (double)(float)P->inv_diag[i];
    }
    return ORC_OK;
}

/* ilu_0_factorizer::factorize, itl/pc/ilu_0.hpp:45-82, then
 * invert_diagonal (operation/invert_diagonal.hpp:47-54), L= strict_lower(LU),
 * U= upper(LU) (ilu_0.hpp:79-81).                                       */
static int setup_ilu0(orc_pc* P, const orc_csr* A, int value_bytes)
{
    int32_t n = A->n;
    int64_t nnz = A->rowptr[n];
    own_csr LU;
    if (own_alloc(&LU, n, nnz) != ORC_OK) return ORC_OOM;
    memcpy(LU.rowptr, A->rowptr, ((size_t)n + 1) * sizeof(int32_t));
    memcpy(LU.col, A->col, (size_t)nnz * sizeof(int32_t));
    memcpy(LU.val, A->val, (size_t)nnz * sizeof(double));
    double* inv_dia = (double*)calloc((size_t)n, sizeof(double));
    if (!inv_dia) { own_free(&LU); return ORC_OOM; }
    for (int32_t i = 0; i < n; i++) {
        int64_t dpos = find_entry(LU.rowptr, LU.col, i, i);
        if (dpos < 0) { free(inv_dia); own_free(&LU); return ORC_MISSING_DIAG; }
        for (int32_t kc = LU.rowptr[i]; kc < LU.rowptr[i + 1]; kc++) {
            int32_t k = LU.col[kc];
            if (k == i) break;
            double aik = LU.val[kc] * inv_dia[k];
            LU.val[kc] = aik;
            for (int32_t jc = kc + 1; jc < LU.rowptr[i + 1]; jc++) {
                int64_t kj = find_entry(LU.rowptr, LU.col, k, LU.col[jc]);
                double ukj = kj >= 0 ? LU.val[kj] : 0.0;
                LU.val[jc] = LU.val[jc] - aik * ukj;
            }
        }
        inv_dia[i] = 1.0 / LU.val[dpos];
    }
    free(inv_dia);
    for (int32_t i = 0; i < n; i++) {            /* invert_diagonal(LU) */
        int64_t d = find_entry(LU.rowptr, LU.col, i, i);
        LU.val[d] = 1.0 / LU.val[d];
    }
    orc_csr v = as_view(&LU);
    int st = copy_band(&v, 1, &P->L, value_bytes);
    if (st == ORC_OK) st = copy_band(&v, 0, &P->U, value_bytes);
    own_free(&LU);
    return st;
}

/* ic_0::factorizer::factorize, itl/pc/ic_0.hpp:135-186.  For row k the
 * diagonal becomes 1/sqrt(U[k][k]); each later entry U[k][i] is scaled by it
 * (d) and the entries of row i of U with column in (k, i] are updated by
 * U(j,i) -= d*U[k][j].  Row i of an upper factor holds only columns >= i,
 * so the single match is the diagonal (i,i).  Factorization runs in double
 * and is rounded afterwards for float storage (ic_0.hpp:122-130).        */
static int setup_ic0(orc_pc* P, const orc_csr* A, int value_bytes)
{
    int32_t n = A->n;
    own_csr U;
    int st = copy_band(A, 0, &U, 8);
    if (st != ORC_OK) return st;
    for (int32_t k = 0; k < n; k++) {
        int32_t ic = U.rowptr[k], iend = U.rowptr[k + 1];
        if (ic == iend || U.col[ic] != k) { own_free(&U); return ORC_MISSING_DIAG; }
        double inv_dia = 1.0 / sqrt(U.val[ic]);
        U.val[ic] = inv_dia;
        for (++ic; ic < iend; ++ic) {
            double d = U.val[ic] * inv_dia;
            U.val[ic] = d;
            int32_t i = U.col[ic];
            int32_t jc = U.rowptr[i], jend = U.rowptr[i + 1];
            while (jc < jend && U.col[jc] <= k) ++jc;
            while (jend > jc && U.col[jend - 1] > i) --jend;
            for (; jc < jend; ++jc) {
                int32_t j = U.col[jc];
                int64_t ji = find_entry(U.rowptr, U.col, j, i);
                int64_t kj = find_entry(U.rowptr, U.col, k, j);
                if (ji < 0) continue;
                U.val[ji] -= d * (kj >= 0 ? U.val[kj] : 0.0);
            }
        }
    }
    if (value_bytes == 4)
        for (int64_t j = 0; j < U.rowptr[n]; j++) U.val[j] = (double)(float)U.val[j];
    P->U = U;
    return transpose(&P->U, &P->L);
}

orc_pc* orc_pc_create(int kind, const orc_csr* A, int value_bytes, int* status)
{
    int st = ORC_OK;
    orc_pc* P = (orc_pc*)calloc(1, sizeof(orc_pc));
    if (!P) { if (status) *status = ORC_OOM; return NULL; }
    P->kind = kind;
    P->n = A->n;
    if (A->n != A->ncols) st = ORC_NOT_SQUARE;
    else if (kind == ORC_PC_DIAGONAL) st = setup_diagonal(P, A, value_bytes);
    else if (kind == ORC_PC_ILU0) st = setup_ilu0(P, A, value_bytes);
    else if (kind == ORC_PC_IC0) st = setup_ic0(P, A, value_bytes);
    if (st == ORC_OK && (kind == ORC_PC_ILU0 || kind == ORC_PC_IC0)) {
        P->y0 = (double*)malloc((size_t)(A->n ? A->n : 1) * sizeof(double));
        if (!P->y0) st = ORC_OOM;
    }
    if (status) *status = st;
    if (st != ORC_OK) { orc_pc_free(P); return NULL; }
    return P;
}

void orc_pc_free(orc_pc* P)
{
    if (!P) return;
    free(P->inv_diag); free(P->y0);
    own_free(&P->L); own_free(&P->U);
    free(P);
}

int64_t orc_pc_factor(const orc_pc* P, int which, const int32_t** rowptr,
                      const int32_t** col, const double** val)
{
    if (which == 0) { *rowptr = NULL; *col = NULL; *val = P->inv_diag; return P->inv_diag ? P->n : 0; }
    const own_csr* M = which == 1 ? &P->L : &P->U;
    if (!M->rowptr) return -1;
    *rowptr = M->rowptr; *col = M->col; *val = M->val;
    return M->rowptr[M->n];
}

/* solve(P, x): identity copies (identity.hpp:40-45); diagonal scales
 * (diagonal.hpp:57-64); ilu solves L then U (ilu.hpp:73-91); ic_0 solves
 * trans(U) then U (ic_0.hpp:69-87).                                      */
int orc_pc_solve(const orc_pc* P, const double* x, double* y)
{
    int32_t n = P->n;
    switch (P->kind) {
    case ORC_PC_IDENTITY:
        memmove(y, x, (size_t)n * sizeof(double));
        return ORC_OK;
    case ORC_PC_DIAGONAL:
        for (int32_t i = 0; i < n; i++) y[i] = P->inv_diag[i] * x[i];
        return ORC_OK;
    case ORC_PC_ILU0: {
        orc_csr L = as_view(&P->L), U = as_view(&P->U);
        orc_lower_unit(&L, x, P->y0);
        orc_upper_invdiag_first(&U, P->y0, y);
        return ORC_OK;
    }
    case ORC_PC_IC0: {
        orc_csr L = as_view(&P->L), U = as_view(&P->U);
        int st = orc_lower_invdiag_last(&L, x, P->y0);
        if (st != ORC_OK) return st;
        orc_upper_invdiag_first(&U, P->y0, y);
        return ORC_OK;
    }
    }
    return ORC_SIZE_MISMATCH;
}

/* adjoint_solve: diagonal uses conj(inv_diag) (diagonal.hpp:76-83); ilu
 * solves trans(U) then trans(L) (ilu.hpp:105-115); ic_0 is self-adjoint
 * (ic_0.hpp:97-102).                                                     */
int orc_pc_adjoint_solve(const orc_pc* P, const double* x, double* y)
{
    if (P->kind == ORC_PC_ILU0) {
        adj_lower_on_upper(&P->U, x, P->y0);
        adj_upper_on_lower(&P->L, P->y0, y);
        return ORC_OK;
    }
    return orc_pc_solve(P, x, y);
}

/* ------------------------------------------------------------------ */
/* Iteration control: basic_iteration::finished(Real), converged() and
 * check_max(), itl/iteration/basic_iteration.hpp:43-48,60-65,85-93.     */
static int finished(orc_iter* it, double r, double* hist, int32_t cap, int32_t* nh)
{
    if (hist && *nh < cap) hist[*nh] = r;
    ++*nh;
    it->resid = r;
    int conv = it->norm_r0 == 0 ? (r <= it->atol)
                                : (r <= it->rtol * it->norm_r0 || r <= it->atol);
    if (conv) return 1;
    if (it->i >= it->max_iter) { it->error = 1; return 1; }
    return 0;
}

/* r = b - A*x through compute_summand (operation/compute_summand.hpp:42-55):
 * the product goes to a temporary first, then b[i] - tmp[i].            */
static void residual(const orc_csr* A, const double* x, const double* b, double* r, double* tmp)
{
    orc_spmv(A, x, tmp, ORC_ASSIGN);
    for (int32_t i = 0; i < A->n; i++) r[i] = b[i] - tmp[i];
}

/* itl::cg, itl/krylov/cg.hpp:38-71 (identity) and :74-111 (preconditioned). */
int orc_cg(const orc_csr* A, const orc_pc* P, double* x, const double* b,
           orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    double* w = (double*)malloc((size_t)(n ? n : 1) * 4 * sizeof(double));
    if (!w) return ORC_OOM;
    double *p = w, *q = w + n, *r = w + 2 * (size_t)n, *z = w + 3 * (size_t)n;
    double rho = 0, rho_1 = 0, rr, alpha, alpha_1;
    int st = ORC_OK;
    it->error = 0;

    residual(A, x, b, r, q);
    if (!P || P->kind == ORC_PC_IDENTITY) {              /* cg.hpp:38-71 */
        rho = orc_dot(n, r, r, 4);
        while (!finished(it, sqrt(fabs(rho)), hist, hist_cap, &nh)) {
            ++it->i;
            if (it->i <= 1) memcpy(p, r, (size_t)n * sizeof(double));
            else {
                double beta = rho / rho_1;
                for (int32_t i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
            }
            alpha_1 = orc_fused_spmv_dot(A, p, q);
            alpha = rho / alpha_1;
            for (int32_t i = 0; i < n; i++) x[i] += alpha * p[i];
            rho_1 = rho;
            rho = orc_fused_axpy_unary_dot(n, alpha, q, r);
        }
    } else {                                             /* cg.hpp:74-111 */
        rr = orc_dot(n, r, r, 4);
        while (!finished(it, sqrt(fabs(rr)), hist, hist_cap, &nh)) {
            ++it->i;
            st = orc_pc_solve(P, r, z);                  /* runs first, then dot: */
            if (st != ORC_OK) break;                     /* fused_expr.hpp:139-143 */
            rho = orc_dot(n, r, z, 4);
            if (it->i <= 1) memcpy(p, z, (size_t)n * sizeof(double));
            else {
                double beta = rho / rho_1;
                for (int32_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            }
            alpha_1 = orc_fused_spmv_dot(A, p, q);
            alpha = rho / alpha_1;
            for (int32_t i = 0; i < n; i++) x[i] += alpha * p[i];
            rho_1 = rho;
            rr = orc_fused_axpy_unary_dot(n, alpha, q, r);
        }
    }
    free(w);
    return st != ORC_OK ? st : it->error;
}

/* itl::bicgstab, itl/krylov/bicgstab.hpp:27-78, including the half-step exit
 * (:64-67) and the unexpected_orthogonality throws (:46,51,59), reported
 * here as ORC_BREAKDOWN.                                                 */
int orc_bicgstab(const orc_csr* A, const orc_pc* P, double* x, const double* b,
                 orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    size_t sn = (size_t)(n ? n : 1);
    double* w = (double*)malloc(sn * 8 * sizeof(double));
    if (!w) return ORC_OOM;
    double *p = w, *phat = w + sn, *s = w + 2 * sn, *shat = w + 3 * sn, *t = w + 4 * sn,
           *v = w + 5 * sn, *r = w + 6 * sn, *rtilde = w + 7 * sn;
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, gamma, omega = 0;
    int st = ORC_OK;
    it->error = 0;
    memset(p, 0, sn * sizeof(double));
    memset(v, 0, sn * sizeof(double));

    residual(A, x, b, r, t);
    memcpy(rtilde, r, (size_t)n * sizeof(double));
    while (!finished(it, orc_two_norm(n, r), hist, hist_cap, &nh)) {
        ++it->i;
        rho_1 = orc_dot(n, rtilde, r, 4);
        if (rho_1 == 0.0) { st = ORC_BREAKDOWN; break; }
        if (it->i <= 1) memcpy(p, r, (size_t)n * sizeof(double));
        else {
            if (omega == 0.0) { st = ORC_BREAKDOWN; break; }
            beta = (rho_1 / rho_2) * (alpha / omega);
            for (int32_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        if (P) { st = orc_pc_solve(P, p, phat); if (st) break; }
        else memcpy(phat, p, (size_t)n * sizeof(double));
        orc_spmv(A, phat, v, ORC_ASSIGN);
        gamma = orc_dot(n, rtilde, v, 4);
        if (gamma == 0.0) { st = ORC_BREAKDOWN; break; }
        alpha = rho_1 / gamma;
        for (int32_t i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        if (finished(it, orc_two_norm(n, s), hist, hist_cap, &nh)) {
            for (int32_t i = 0; i < n; i++) x[i] += alpha * phat[i];
            break;
        }
        if (P) { st = orc_pc_solve(P, s, shat); if (st) break; }
        else memcpy(shat, s, (size_t)n * sizeof(double));
        orc_spmv(A, shat, t, ORC_ASSIGN);
        omega = orc_dot(n, t, s, 4) / orc_dot(n, t, t, 4);
        for (int32_t i = 0; i < n; i++) x[i] += omega * shat[i] + alpha * phat[i];
        for (int32_t i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
        rho_2 = rho_1;
    }
    free(w);
    return st != ORC_OK ? st : it->error;
}

/* itl::cgs, itl/krylov/cgs.hpp:24-62.  rho_1 == 0 is iter.fail(2, ...)
 * (basic_iteration.hpp:129-136): the error code is set and the loop goes
 * on; there is no test on dot(rtilde, vhat).                           */
int orc_cgs(const orc_csr* A, const orc_pc* P, double* x, const double* b,
            orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    size_t sn = (size_t)(n ? n : 1);
    double* w = (double*)malloc(sn * 9 * sizeof(double));
    if (!w) return ORC_OOM;
    double *p = w, *phat = w + sn, *q = w + 2 * sn, *qhat = w + 3 * sn, *vhat = w + 4 * sn,
           *u = w + 5 * sn, *uhat = w + 6 * sn, *r = w + 7 * sn, *rtilde = w + 8 * sn;
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    int st = ORC_OK;
    it->error = 0;
    memset(p, 0, sn * sizeof(double));
    memset(q, 0, sn * sizeof(double));

    residual(A, x, b, r, qhat);                          /* r(b - A * x), rtilde = r */
    memcpy(rtilde, r, (size_t)n * sizeof(double));
    while (!finished(it, orc_two_norm(n, r), hist, hist_cap, &nh)) {
        ++it->i;
        rho_1 = orc_dot(n, rtilde, r, 4);
        if (rho_1 == 0.0) it->error = 2;                 /* iter.fail(2, "cgs breakdown") */
        if (it->i <= 1) {
            memcpy(u, r, (size_t)n * sizeof(double));
            memcpy(p, u, (size_t)n * sizeof(double));
        } else {
            beta = rho_1 / rho_2;
            for (int32_t i = 0; i < n; i++) u[i] = r[i] + beta * q[i];
            for (int32_t i = 0; i < n; i++) p[i] = u[i] + beta * (q[i] + beta * p[i]);
        }
        if (P) { st = orc_pc_solve(P, p, phat); if (st) break; }
        else memcpy(phat, p, (size_t)n * sizeof(double));
        orc_spmv(A, phat, vhat, ORC_ASSIGN);
        alpha = rho_1 / orc_dot(n, rtilde, vhat, 4);
        for (int32_t i = 0; i < n; i++) q[i] = u[i] - alpha * vhat[i];
        for (int32_t i = 0; i < n; i++) u[i] += q[i];
        if (P) { st = orc_pc_solve(P, u, uhat); if (st) break; }
        else memcpy(uhat, u, (size_t)n * sizeof(double));
        for (int32_t i = 0; i < n; i++) x[i] += alpha * uhat[i];
        orc_spmv(A, uhat, qhat, ORC_ASSIGN);
        for (int32_t i = 0; i < n; i++) r[i] -= alpha * qhat[i];
        rho_2 = rho_1;
    }
    free(w);
    return st != ORC_OK ? st : it->error;
}

/* itl::tfqmr, itl/krylov/tfqmr.hpp:32-96, with left preconditioner L and
 * right preconditioner R (NULL: identity).  sigma == 0 and rho == 0 are
 * iter.fail(1, ...) returned at once; rt = A*x - b at the end of each
 * iteration (:93) is computed as there, though nothing reads it.        */
static int tf_solve(const orc_pc* P, const double* in, double* out, int32_t n)
{
    if (P) return orc_pc_solve(P, in, out);
    memcpy(out, in, (size_t)n * sizeof(double));
    return ORC_OK;
}

int orc_tfqmr(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
              orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    if (n == 0) return ORC_INVALID;                       /* "empty rhs vector" */
    size_t sn = (size_t)n;
    double* wk = (double*)calloc(sn * 11, sizeof(double));
    if (!wk) return ORC_OOM;
    double *rt = wk, *r = wk + sn, *u1 = wk + 2 * sn, *u2 = wk + 3 * sn, *y1 = wk + 4 * sn,
           *y2 = wk + 5 * sn, *w = wk + 6 * sn, *d = wk + 7 * sn, *v = wk + 8 * sn,
           *tmp = wk + 9 * sn, *tmp2 = wk + 10 * sn;
    double theta = 0, eta = 0, tau, rho, rhon, sigma, alpha, beta, c;
    int st = ORC_OK;
    it->error = 0;
#define TF_TRY(e) do { st = (e); if (st != ORC_OK) goto out; } while (0)
    TF_TRY(tf_solve(R, x, tmp, n));                      /* rt(b - A*Vector(solve(R, x))) */
    residual(A, tmp, b, rt, tmp2);
    TF_TRY(tf_solve(L, rt, r, n));                       /* r(solve(L, rt)) */
    if (finished(it, orc_two_norm(n, rt), hist, hist_cap, &nh)) goto out;
    memcpy(w, r, sn * sizeof(double));                   /* y1 = w = r */
    memcpy(y1, w, sn * sizeof(double));
    TF_TRY(tf_solve(R, y1, tmp, n));
    orc_spmv(A, tmp, rt, ORC_ASSIGN);
    TF_TRY(tf_solve(L, rt, v, n));                       /* u1 = v = solve(L, rt) */
    memcpy(u1, v, sn * sizeof(double));
    tau = orc_two_norm(n, r);
    rho = tau * tau;
    while (!finished(it, tau, hist, hist_cap, &nh)) {
        ++it->i;
        sigma = orc_dot(n, r, v, 4);
        if (sigma == 0.0) { it->error = 1; goto out; }
        alpha = rho / sigma;
        for (int j = 1; j < 3; j++) {
            if (j == 1) {
                for (int32_t i = 0; i < n; i++) w[i] -= alpha * u1[i];
                double coef = theta * theta * eta / alpha;
                for (int32_t i = 0; i < n; i++) d[i] = y1[i] + coef * d[i];
            } else {
                for (int32_t i = 0; i < n; i++) y2[i] = y1[i] - alpha * v[i];
                TF_TRY(tf_solve(R, y2, tmp, n));
                orc_spmv(A, tmp, rt, ORC_ASSIGN);
                TF_TRY(tf_solve(L, rt, u2, n));
                for (int32_t i = 0; i < n; i++) w[i] -= alpha * u2[i];
                double coef = theta * theta * eta / alpha;
                for (int32_t i = 0; i < n; i++) d[i] = y2[i] + coef * d[i];
            }
            theta = orc_two_norm(n, w) / tau;
            c = 1.0 / sqrt(1.0 + theta * theta);
            tau *= theta * c;
            eta = c * c * alpha;
            for (int32_t i = 0; i < n; i++) x[i] += eta * d[i];
        }
        if (rho == 0.0) { it->error = 1; goto out; }
        rhon = orc_dot(n, r, w, 4);
        beta = rhon / rho;
        rho = rhon;
        for (int32_t i = 0; i < n; i++) y1[i] = w[i] + beta * y2[i];
        TF_TRY(tf_solve(R, y1, tmp, n));
        orc_spmv(A, tmp, rt, ORC_ASSIGN);
        TF_TRY(tf_solve(L, rt, u1, n));
        for (int32_t i = 0; i < n; i++) v[i] = u1[i] + beta * (u2[i] + beta * v[i]);
        orc_spmv(A, x, rt, ORC_ASSIGN);                  /* rt = A*x - b */
        for (int32_t i = 0; i < n; i++) rt[i] -= b[i];
    }
    TF_TRY(tf_solve(R, x, tmp, n));                      /* shift back: x = solve(R, x) */
    memcpy(x, tmp, sn * sizeof(double));
out:
#undef TF_TRY
    free(wk);
    return st != ORC_OK ? st : it->error;
}

/* itl::bicgstab_2 (l = 2), itl/krylov/bicgstab_2.hpp:37-130, with left
 * preconditioner L (NULL: identity).  The minimal-residual part runs
 * orthogonalize_factors (operation/orth.hpp:61-96, the loop without
 * MTL_WITH_FUSED_ORTHOGONALIZATION) on r_hat[1..2] in place: r_hat_tail
 * is constructed from the sub-vector r_hat[irange(1, imax)], which holds
 * r_hat's own elements (dense_vector.hpp:326-338) and is taken over, not
 * copied.  gamma_aa[1] = strict_upper(1x1) * gamma[2] + gamma[2]: the
 * naive dense mat-vec (mat_vec_mult.hpp:259-298) sums 0 * gamma[2] from
 * zero (the view reads 0 off the band), then gamma[2] is added.
 * omega == 0 is iter.fail(3) and
 * returns before the shift back x += x0; every other exit adds x0.     */
int orc_bicgstab_2(const orc_csr* A, const orc_pc* L, double* x, const double* b,
                   orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    if (n == 0) return ORC_INVALID;                       /* "empty rhs vector" */
    size_t sn = (size_t)n;
    double* wk = (double*)calloc(sn * 9, sizeof(double));
    if (!wk) return ORC_OOM;
    double* r[3] = {wk, wk + sn, wk + 2 * sn};
    double* u[3] = {wk + 3 * sn, wk + 4 * sn, wk + 5 * sn};
    double *rt = wk + 6 * sn, *y = wk + 7 * sn, *x0 = wk + 8 * sn;   /* u[0] = x0 = 0 */
    double rho_0 = 1.0, rho_1, alpha = 0.0, Gamma, beta, omega = 1.0;
    double t11, t12, t22, ga1, ga2, g1, g2, gaa1, nr;
    int st = ORC_OK;
    it->error = 0;
#define B2_TRY(e) do { st = (e); if (st != ORC_OK) goto out; } while (0)
    memcpy(r[0], b, sn * sizeof(double));                 /* r_hat[0] = b */
    if (orc_two_norm(n, x) != 0.0) {                      /* shift problem */
        orc_spmv(A, x, r[0], ORC_MINUS);
        memcpy(x0, x, sn * sizeof(double));
        memset(x, 0, sn * sizeof(double));
    }
    nr = orc_two_norm(n, r[0]);
    for (int32_t i = 0; i < n; i++) rt[i] = r[0][i] / nr; /* r0_tilde */
    B2_TRY(tf_solve(L, r[0], y, n));
    memcpy(r[0], y, sn * sizeof(double));
    while (!finished(it, orc_two_norm(n, r[0]), hist, hist_cap, &nh)) {
        ++it->i;
        rho_0 = -omega * rho_0;
        for (int j = 0; j < 2; j++) {
            rho_1 = orc_dot(n, rt, r[j], 4);
            beta = alpha * rho_1 / rho_0;
            rho_0 = rho_1;
            for (int k = 0; k <= j; k++)
                for (int32_t i = 0; i < n; i++) u[k][i] = r[k][i] - beta * u[k][i];
            orc_spmv(A, u[j], y, ORC_ASSIGN);
            B2_TRY(tf_solve(L, y, u[j + 1], n));
            Gamma = orc_dot(n, rt, u[j + 1], 4);
            alpha = rho_0 / Gamma;
            for (int k = 0; k <= j; k++)
                for (int32_t i = 0; i < n; i++) r[k][i] -= alpha * u[k + 1][i];
            if (finished(it, orc_two_norm(n, r[j]), hist, hist_cap, &nh)) goto shift_back;
            orc_spmv(A, r[j], y, ORC_ASSIGN);
            B2_TRY(tf_solve(L, y, r[j + 1], n));
            for (int32_t i = 0; i < n; i++) x[i] += alpha * u[0][i];
        }
        /* orthogonalize_factors(r_hat[1..2]) */
        t11 = orc_dot(n, r[1], r[1], 4);
        t12 = orc_dot(n, r[1], r[2], 4) / t11;
        for (int32_t i = 0; i < n; i++) r[2][i] -= t12 * r[1][i];
        t22 = orc_dot(n, r[2], r[2], 4);
        ga1 = orc_dot(n, r[1], r[0], 4) / t11;
        ga2 = orc_dot(n, r[2], r[0], 4) / t22;
        g2 = ga2;
        omega = g2;
        if (omega == 0.0) { it->error = 3; goto out; }    /* "bicg breakdown #2" */
        g1 = ga1 - (0.0 + t12 * g2);
        gaa1 = (0.0 + 0.0 * g2) + g2;                       /* naive dense mat-vec, then + gamma[2] */
        for (int32_t i = 0; i < n; i++) x[i] += g1 * r[0][i];
        for (int32_t i = 0; i < n; i++) r[0][i] -= ga2 * r[2][i];
        for (int32_t i = 0; i < n; i++) u[0][i] -= g2 * u[2][i];
        for (int32_t i = 0; i < n; i++) u[0][i] -= g1 * u[1][i];
        for (int32_t i = 0; i < n; i++) x[i] += gaa1 * r[1][i];
        for (int32_t i = 0; i < n; i++) r[0][i] -= ga1 * r[1][i];
    }
shift_back:
    for (int32_t i = 0; i < n; i++) x[i] += x0[i];
out:
#undef B2_TRY
    free(wk);
    return st != ORC_OK ? st : it->error;
}

/* itl::bicgstab_ell, itl/krylov/bicgstab_ell.hpp:37-134, with left
 * preconditioner L and right preconditioner R (NULL: identity), 1 <= l.
 * As bicgstab_2 but with l inner steps, y = A * solve(R, u_hat[j]),
 * r_hat[j+1] = solve(L, A * solve(R, r_hat[j])), and x = solve(R, x)
 * before x += x0 on every exit but iter.fail(3).  The minimal-residual part
 * orthogonalizes r_hat[1..l] in place (orth.hpp:61-96); gamma_aa[j] is the
 * naive dense mat-vec over the strictly upper view (mat_vec_mult.hpp:
 * 259-298: every column, 0 off the band, summed from zero) plus gamma[j+1]. */
int orc_bicgstab_ell(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
                     int32_t l, orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t n = A->n, nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    if (n == 0 || l < 1) return ORC_INVALID;              /* "empty rhs vector" */
    size_t sn = (size_t)n, L1 = (size_t)l + 1;
    double* wk = (double*)calloc(sn * (2 * L1 + 4), sizeof(double));
    double* sc = (double*)calloc(L1 * L1 + 3 * L1, sizeof(double));
    double** r = (double**)malloc(2 * L1 * sizeof(double*));
    if (!wk || !sc || !r) { free(wk); free(sc); free(r); return ORC_OOM; }
    double** u = r + L1;
    for (size_t j = 0; j < L1; j++) { r[j] = wk + j * sn; u[j] = wk + (L1 + j) * sn; }
    double *rt = wk + 2 * L1 * sn, *y = rt + sn, *x0 = y + sn, *tmp = x0 + sn;
    double *tau = sc, *gam = sc + L1 * L1, *ga = gam + L1, *gaa = ga + L1;
#define TAU(i, j) tau[(size_t)(i) * L1 + (j)]
    double rho_0 = 1.0, rho_1, alpha = 0.0, Gamma, beta, omega = 1.0, nr;
    int st = ORC_OK;
    it->error = 0;
#define BL_TRY(e) do { st = (e); if (st != ORC_OK) goto out; } while (0)
    memcpy(r[0], b, sn * sizeof(double));
    if (orc_two_norm(n, x) != 0.0) {
        orc_spmv(A, x, r[0], ORC_MINUS);
        memcpy(x0, x, sn * sizeof(double));
        memset(x, 0, sn * sizeof(double));
    }
    nr = orc_two_norm(n, r[0]);
    for (int32_t i = 0; i < n; i++) rt[i] = r[0][i] / nr;
    BL_TRY(tf_solve(L, r[0], y, n));
    memcpy(r[0], y, sn * sizeof(double));
    while (!finished(it, orc_two_norm(n, r[0]), hist, hist_cap, &nh)) {
        ++it->i;
        rho_0 = -omega * rho_0;
        for (int32_t j = 0; j < l; j++) {
            rho_1 = orc_dot(n, rt, r[j], 4);
            beta = alpha * rho_1 / rho_0;
            rho_0 = rho_1;
            for (int32_t k = 0; k <= j; k++)
                for (int32_t i = 0; i < n; i++) u[k][i] = r[k][i] - beta * u[k][i];
            BL_TRY(tf_solve(R, u[j], tmp, n));
            orc_spmv(A, tmp, y, ORC_ASSIGN);
            BL_TRY(tf_solve(L, y, u[j + 1], n));
            Gamma = orc_dot(n, rt, u[j + 1], 4);
            alpha = rho_0 / Gamma;
            for (int32_t k = 0; k <= j; k++)
                for (int32_t i = 0; i < n; i++) r[k][i] -= alpha * u[k + 1][i];
            if (finished(it, orc_two_norm(n, r[j]), hist, hist_cap, &nh)) goto shift_back;
            BL_TRY(tf_solve(R, r[j], r[j + 1], n));
            orc_spmv(A, r[j + 1], y, ORC_ASSIGN);
            BL_TRY(tf_solve(L, y, r[j + 1], n));
            for (int32_t i = 0; i < n; i++) x[i] += alpha * u[0][i];
        }
        /* orthogonalize_factors(r_hat[1..l]) into tau[1..l][1..l] */
        TAU(1, 1) = orc_dot(n, r[1], r[1], 4);
        for (int32_t j = 2; j <= l; j++) {
            for (int32_t i = 1; i < j; i++) {
                double t = orc_dot(n, r[i], r[j], 4) / TAU(i, i);
                TAU(i, j) = t;
                for (int32_t e = 0; e < n; e++) r[j][e] -= t * r[i][e];
            }
            TAU(j, j) = orc_dot(n, r[j], r[j], 4);
        }
        for (int32_t j = 1; j <= l; j++) ga[j] = orc_dot(n, r[j], r[0], 4) / TAU(j, j);
        gam[l] = ga[l];
        omega = gam[l];
        if (omega == 0.0) { it->error = 3; goto out; }      /* "bicg breakdown #2" */
        for (int32_t j = l - 1; j > 0; --j) {
            double sum = 0.0;
            for (int32_t i = j + 1; i <= l; ++i) sum += TAU(j, i) * gam[i];
            gam[j] = ga[j] - sum;
        }
        for (int32_t p = 0; p < l - 1; p++) {                /* gamma_aa[p+1] */
            double t = 0.0;
            for (int32_t q = 0; q < l - 1; q++) t += (q > p ? TAU(p + 1, q + 1) : 0.0) * gam[q + 2];
            gaa[p + 1] = t + gam[p + 2];
        }
        for (int32_t i = 0; i < n; i++) x[i] += gam[1] * r[0][i];
        for (int32_t i = 0; i < n; i++) r[0][i] -= ga[l] * r[l][i];
        for (int32_t i = 0; i < n; i++) u[0][i] -= gam[l] * u[l][i];
        for (int32_t j = 1; j < l; ++j) {
            for (int32_t i = 0; i < n; i++) u[0][i] -= gam[j] * u[j][i];
            for (int32_t i = 0; i < n; i++) x[i] += gaa[j] * r[j][i];
            for (int32_t i = 0; i < n; i++) r[0][i] -= ga[j] * r[j][i];
        }
    }
shift_back:
    BL_TRY(tf_solve(R, x, tmp, n));                       /* x = solve(R, x); x += x0 */
    for (int32_t i = 0; i < n; i++) x[i] = tmp[i] + x0[i];
out:
#undef BL_TRY
#undef TAU
    free(wk);
    free(sc);
    free(r);
    return st != ORC_OK ? st : it->error;
}

/* ------------------------------------------------------------------ */
/* itl::gmres_full / itl::gmres, itl/krylov/gmres.hpp:37-137.           */

/* mtl::mat::givens(H, a, b).trafo(i), operation/givens.hpp: the rotation
 * G = [c s; -s c] of set_rotation(a, b), then rows i, i+1 of H (all
 * columns) <- trans(G) * rows and columns i, i+1 (all rows) <- cols * G.
 * Each 2-term product sums from zero (0 + p0) + p1 as the naive dense
 * product does.  H is (rows x ld), row-major.                           */
static void gm_givens_trafo(double* H, int32_t rows, int32_t ld, int32_t i, double a, double b)
{
    double c, s, t;
    if (b == 0.0) { c = 1.0; s = 0.0; }
    else if (fabs(b) > fabs(a)) { t = -a / b; s = 1.0 / sqrt(1.0 + t * t); c = s * t; }
    else { t = -b / a; c = 1.0 / sqrt(1.0 + t * t); s = c * t; }
    const double ms = -s;
    for (int32_t j = 0; j < ld; j++) {
        double b0 = H[(size_t)i * ld + j], b1 = H[(size_t)(i + 1) * ld + j];
        H[(size_t)i * ld + j] = (0.0 + c * b0) + ms * b1;
        H[(size_t)(i + 1) * ld + j] = (0.0 + s * b0) + c * b1;
    }
    for (int32_t r = 0; r < rows; r++) {
        double m0 = H[(size_t)r * ld + i], m1 = H[(size_t)r * ld + i + 1];
        H[(size_t)r * ld + i] = (0.0 + m0 * c) + m1 * ms;
        H[(size_t)r * ld + i + 1] = (0.0 + m0 * s) + m1 * c;
    }
}

/* lu_solve(H[range][range], g[range]), operation/lu.hpp:62-89,117-139, on
 * the m x m leading block of H, IN PLACE: the sub-matrix is a view and
 * the copy LU(A) of a view shares its memory (contiguous_memory_block.hpp:
 * 295-302).  Partial pivoting (first row of the largest magnitude, rows
 * swapped within the block), singular when a pivot other than the last
 * has |p| <= 0.  permute(P, b) scatters: b'[P[i]] = b[i] (permute.hpp:
 * 25-35); then the unit lower and the upper sweep (upper divides by the
 * diagonal).  Returns 0, or 1 when singular (H keeps the partial work). */
static int gm_lu_solve(double* H, int32_t ld, int32_t m, const double* g, double* y, int32_t* P, double* z)
{
#define HH(r, c) H[(size_t)(r) * ld + (c)]
    for (int32_t i = 0; i < m; i++) P[i] = i;
    for (int32_t i = 0; i + 1 < m; i++) {
        int32_t rmax = i;
        double best = fabs(HH(i, i));
        for (int32_t r = i + 1; r < m; r++)
            if (fabs(HH(r, i)) > best) { best = fabs(HH(r, i)); rmax = r; }
        if (rmax != i) {
            for (int32_t c = 0; c < m; c++) { double t = HH(i, c); HH(i, c) = HH(rmax, c); HH(rmax, c) = t; }
            int32_t t = P[i]; P[i] = P[rmax]; P[rmax] = t;
        }
        if (fabs(HH(i, i)) <= 0.0) return 1;
        for (int32_t r = i + 1; r < m; r++) HH(r, i) /= HH(i, i);
        for (int32_t r = i + 1; r < m; r++)
            for (int32_t c = i + 1; c < m; c++) HH(r, c) -= HH(r, i) * HH(i, c);
    }
    for (int32_t i = 0; i < m; i++) y[P[i]] = g[i];       /* y holds the permuted rhs */
    for (int32_t r = 0; r < m; r++) {
        double rr = y[r];
        for (int32_t c = 0; c < r; c++) rr -= HH(r, c) * z[c];
        z[r] = rr;
    }
    for (int32_t r = m; r-- > 0;) {
        double rr = z[r];
        for (int32_t c = r + 1; c < m; c++) rr -= HH(r, c) * y[c];
        y[r] = rr / HH(r, r);
    }
    return 0;
#undef HH
}

/* One gmres_full call (gmres.hpp:37-119) on the iteration state `in`. */
static int gm_full(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
                   orc_iter* in, double* hist, int32_t cap, int32_t* nh)
{
    int32_t n = A->n;
    int64_t rem = (int64_t)in->max_iter - in->i;
    int32_t kmax = (rem < 0 || rem > n) ? n : (int32_t)rem;   /* min(size(x), Size(max - i)) */
    size_t sn = (size_t)n, K1 = (size_t)kmax + 1;
    double* wk = (double*)calloc(sn * (K1 + 4), sizeof(double));
    double* sc = (double*)calloc(K1 * (size_t)(kmax > 0 ? kmax : 1) + 5 * K1, sizeof(double));
    int32_t* P = (int32_t*)malloc(K1 * sizeof(int32_t));
    if (!wk || !sc || !P) { free(wk); free(sc); free(P); return ORC_OOM; }
    double *r0 = wk, *r = wk + sn, *tmp = wk + 2 * sn, *va0 = wk + 3 * sn, *V = wk + 4 * sn;
    double *H = sc, *g = sc + K1 * (size_t)(kmax > 0 ? kmax : 1), *cc = g + K1, *ss = cc + K1, *y = ss + K1,
           *z = y + K1;
    int st = ORC_OK, k;
#define GM_TRY(e) do { st = (e); if (st != ORC_OK) goto out; } while (0)
#define VV(j) (V + (size_t)(j) * sn)
#define HK(i, j) H[(size_t)(i) * kmax + (j)]
    residual(A, x, b, r0, tmp);                           /* r0(b - A*x), r(solve(L, r0)) */
    GM_TRY(tf_solve(L, r0, r, n));
    double rho = g[0] = orc_two_norm(n, r);
    if (finished(in, rho, hist, cap, nh)) goto out;
    for (int32_t i = 0; i < n; i++) VV(0)[i] = r[i] / rho;
    for (k = 0; k < kmax; ++k, ++in->i) {
        GM_TRY(tf_solve(R, VV(k), tmp, n));
        orc_spmv(A, tmp, va0, ORC_ASSIGN);
        GM_TRY(tf_solve(L, va0, VV(k + 1), n));
        double* w = VV(k + 1);
        for (int32_t j = 0; j <= k; j++) {                /* modified Gram-Schmidt */
            HK(j, k) = orc_dot(n, VV(j), w, 4);
            for (int32_t i = 0; i < n; i++) w[i] -= HK(j, k) * VV(j)[i];
        }
        HK(k + 1, k) = orc_two_norm(n, w);
        for (int32_t j = 0; j <= k; j++) {                /* reorthogonalize */
            double hr = orc_dot(n, w, VV(j), 4);
            HK(j, k) += hr;
            for (int32_t i = 0; i < n; i++) w[i] -= hr * VV(j)[i];
        }
        HK(k + 1, k) = orc_two_norm(n, w);
        if (HK(k + 1, k) != 0.0) {
            double f = 1. / HK(k + 1, k);
            for (int32_t i = 0; i < n; i++) w[i] *= f;
        }
        for (int32_t i = 0; i < k; i++) gm_givens_trafo(H, kmax + 1, kmax, i, HK(i, k - 1), HK(i + 1, k - 1));
        double nu = sqrt(HK(k, k) * HK(k, k) + HK(k + 1, k) * HK(k + 1, k));
        if (nu != 0.0) {
            cc[k] = HK(k, k) / nu;
            ss[k] = -HK(k + 1, k) / nu;
            HK(k, k) = cc[k] * HK(k, k) - ss[k] * HK(k + 1, k);
            HK(k + 1, k) = 0;
            double w1 = cc[k] * g[k] - ss[k] * g[k + 1], w2 = ss[k] * g[k] + cc[k] * g[k + 1];
            g[k] = w1;
            g[k + 1] = w2;
        }
    }
    while (k > 0 && fabs(g[k - 1]) <= in->atol) k--;
    int32_t m = k;
    for (; m > 0; --m)
        if (gm_lu_solve(H, kmax, m, g, y, P, z) == 0) break;
    if (m == 0) { in->error = 2; goto out; }             /* "GMRES did not find any direction" */
    for (int32_t i = 0; i < n; i++) {                     /* Vector(V.vector(range) * y[range]) */
        double t = 0.0;
        for (int32_t j = 0; j < m; j++) t += VV(j)[i] * y[j];
        r0[i] = t;
    }
    GM_TRY(tf_solve(R, r0, tmp, n));
    for (int32_t i = 0; i < n; i++) x[i] += tmp[i];
    residual(A, x, b, r, va0);                            /* r = b - A*x; terminate(r) */
    finished(in, orc_two_norm(n, r), hist, cap, nh);
out:
#undef GM_TRY
#undef VV
#undef HK
    free(wk);
    free(sc);
    free(P);
    return st != ORC_OK ? st : in->error;
}

/* restart > 0: itl::gmres (gmres.hpp:122-137): cycles of gmres_full on a
 * copy of iter limited to restart more iterations, each followed by
 * iter.update_progress (basic_iteration.hpp:145-155).  restart == 0:
 * gmres_full alone.  hist: the residuals gmres_full hands to finished(). */
int orc_gmres(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
              int32_t restart, orc_iter* it, double* hist, int32_t hist_cap)
{
    int32_t nh = 0;
    if (A->n != A->ncols) return ORC_NOT_SQUARE;
    if (A->n == 0 || restart < 0) return ORC_INVALID;     /* "empty rhs vector" */
    it->error = 0;
    if (restart == 0) return gm_full(A, L, R, x, b, it, hist, hist_cap, &nh);
    for (;;) {
        orc_iter inner = *it;
        inner.max_iter = it->i + restart < it->max_iter ? it->i + restart : it->max_iter;
        int st = gm_full(A, L, R, x, b, &inner, hist, hist_cap, &nh);
        if (st < 0) return st;
        it->i = inner.i;
        it->resid = inner.resid;
        if (inner.error > 1) { it->error = inner.error; break; }
        double r = it->resid;
        int conv = it->norm_r0 == 0 ? (r <= it->atol) : (r <= it->rtol * it->norm_r0 || r <= it->atol);
        if (conv) break;
        if (it->i >= it->max_iter) { it->error = 1; break; }
    }
    return it->error;
}
