/*
 * mtl4_oracle.h -- CPU restatement of MTL4's sparse Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library in mtl4_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Every routine restates the loop order of the MTL4 source it cites
 * (paths relative to the MTL4 tree), including the order in which partial
 * sums are combined, so that results are reproducible bit for bit on any
 * IEEE-754 host compiled with -ffp-contract=off (no fused multiply-add:
 * MTL4's default x86-64 build has no FMA instructions to contract into).
 *
 * Index types: int32 row pointers and column indices (MTL4's
 * unsigned_parameters, matrix/parameter.hpp:47), rows sorted by column.
 */
#ifndef MTL4_ORACLE_H
#define MTL4_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* compressed2D, row major (matrix/compressed2D.hpp:647-652) */
typedef struct {
    int32_t n;              /* rows == cols for every square operator here */
    int32_t ncols;
    const int32_t* rowptr;  /* n+1 */
    const int32_t* col;     /* nnz, sorted within each row */
    const double* val;      /* nnz */
} orc_csr;

/* basic_iteration state (itl/iteration/basic_iteration.hpp:23-164) */
typedef struct {
    double norm_r0;   /* my_norm_r0 */
    int32_t max_iter;
    double rtol, atol;
    int32_t i;        /* iteration counter; in: start value, out: final */
    int32_t error;    /* 0 converged, 1 too many iterations, 2 breakdown */
    double resid;     /* last residual handed to finished() */
} orc_iter;

enum { ORC_ASSIGN = 0, ORC_PLUS = 1, ORC_MINUS = 2 };
enum { ORC_PC_IDENTITY = 0, ORC_PC_DIAGONAL = 1, ORC_PC_ILU0 = 2, ORC_PC_IC0 = 3 };

/* matrix/laplacian_setup.hpp:25-44; returns nnz, arrays may be NULL to query */
int64_t orc_laplacian_setup(int32_t m, int32_t n, int32_t* rowptr, int32_t* col, double* val);

/* operation/mat_vec_mult.hpp:613-683 with assign_mode.hpp first_update */
void orc_spmv(const orc_csr* A, const double* x, double* y, int mode);

/* operation/dot.hpp:136-147, sfunctor::dot<Unroll> (Unroll 1..8) */
double orc_dot(int64_t n, const double* a, const double* b, int unroll);
/* vector/reduction.hpp:154-173 with two_norm_functor (Unroll 1..8) */
double orc_unary_dot(int64_t n, const double* a, int unroll);
/* operation/two_norm.hpp:52-57: two_norm<4> */
double orc_two_norm(int64_t n, const double* a);
/* cg.hpp:103: (lazy(q)= A*p) || (lazy(alpha)= lazy_dot(p,q)) */
double orc_fused_spmv_dot(const orc_csr* A, const double* p, double* q);
/* cg.hpp:108: (lazy(r) -= alpha*q) || (lazy(rr)= lazy_unary_dot(r)) */
double orc_fused_axpy_unary_dot(int64_t n, double alpha, const double* q, double* r);

/* Preconditioners (itl/pc/ diagonal, ilu_0, ic_0).  value_bytes 8: double storage,
 * 4: factor values rounded to float (ic_0<Matrix,float>, ic_0.hpp:122-130). */
typedef struct orc_pc orc_pc;
orc_pc* orc_pc_create(int kind, const orc_csr* A, int value_bytes, int* status);
void orc_pc_free(orc_pc* P);
int orc_pc_solve(const orc_pc* P, const double* x, double* y);
int orc_pc_adjoint_solve(const orc_pc* P, const double* x, double* y);
/* factor access: which 0=inv_diag, 1=L, 2=U; returns nnz (or n for inv_diag) */
int64_t orc_pc_factor(const orc_pc* P, int which, const int32_t** rowptr,
                      const int32_t** col, const double** val);

/* Triangular sweeps (operation/lower_trisolve.hpp, upper_trisolve.hpp) */
void orc_lower_unit(const orc_csr* L, const double* v, double* w);        /* v5 */
int orc_lower_invdiag_last(const orc_csr* L, const double* v, double* w); /* v6 */
void orc_upper_invdiag_first(const orc_csr* U, const double* v, double* w); /* v3 */

/* Krylov solvers (itl/krylov/cg.hpp, bicgstab.hpp).  hist (may be NULL)
 * receives the residual passed to each finished() call, hist_cap entries. */
int orc_cg(const orc_csr* A, const orc_pc* P, double* x, const double* b,
           orc_iter* it, double* hist, int32_t hist_cap);
int orc_cgs(const orc_csr* A, const orc_pc* P, double* x, const double* b,
            orc_iter* it, double* hist, int32_t hist_cap);
int orc_tfqmr(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
              orc_iter* it, double* hist, int32_t hist_cap);
int orc_bicgstab(const orc_csr* A, const orc_pc* P, double* x, const double* b,
                 orc_iter* it, double* hist, int32_t hist_cap);
int orc_bicgstab_2(const orc_csr* A, const orc_pc* L, double* x, const double* b,
                   orc_iter* it, double* hist, int32_t hist_cap);
int orc_bicgstab_ell(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
                     int32_t l, orc_iter* it, double* hist, int32_t hist_cap);
/* itl::gmres (restart > 0) / itl::gmres_full (restart == 0), gmres.hpp:37-137 */
int orc_gmres(const orc_csr* A, const orc_pc* L, const orc_pc* R, double* x, const double* b,
              int32_t restart, orc_iter* it, double* hist, int32_t hist_cap);

/* itl::cg + pc::diagonal as MTL4 runs it with MTL_WITH_OPENMP (all-core
 * CPU baseline; mtl4_oracle_omp.c).  threads > 0 sets the team size. */
int orc_cg_diag_omp(const orc_csr* A, const double* inv_diag, double* x, const double* b,
                    orc_iter* it, int threads);
int orc_omp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
