/*
 * mtl4g.h -- C ABI of the MI355X (gfx950) sparse Krylov library.
 *
 * This is the boundary between host code written against MTL4's types
 * (mtl::compressed2D, mtl::dense_vector, itl::cg, itl::bicgstab,
 * itl::pc::{identity,diagonal,ilu_0,ic_0}, itl::basic_iteration) and the
 * hand-written HIP kernels.  Plain C: opaque handles, plain pointers and
 * sizes, int status codes, no exceptions across the ABI.
 *
 * Reference interfaces are cited as path:line in the MTL4 tree.
 *
 * Memory:  "d_" arguments are device pointers (hipMalloc'd or any device
 * allocation of the current device); "h_" arguments are host pointers.
 * Vectors are dense, contiguous, FP64 unless the name says f32.
 * Streams: every call takes a hipStream_t as void* (NULL = default stream)
 * and is stream-ordered; calls that return a scalar to the host synchronize
 * that stream.
 * Threading: every call is reentrant.  Handles keep their scratch inside
 * the handle; concurrent calls on distinct handles are safe, one handle is
 * used by one thread at a time (MTL4's ilu/ic_0 solves are not reentrant
 * at all: they keep a function-local static, itl/pc/ilu.hpp:76,111,
 * itl/pc/ic_0.hpp:72).
 */
#ifndef MTL4G_H
#define MTL4G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
/* >= 0 mirrors basic_iteration::error (itl/iteration/basic_iteration.hpp:
 * 43-48: 0 converged, 1 "Too many iterations"); < 0 are the conditions MTL4
 * throws for (mtl/utility/exception.hpp:95-236, itl/utility/exception.hpp). */
#define MTL4G_OK                 0
#define MTL4G_MAX_ITER           1   /* basic_iteration error 1 */
#define MTL4G_ERR_INVALID      (-1)  /* null handle / bad size / bad enum */
#define MTL4G_ERR_NOT_SQUARE   (-2)  /* mtl::matrix_not_square */
#define MTL4G_ERR_SIZE         (-3)  /* mtl::incompatible_size */
#define MTL4G_ERR_MISSING_DIAG (-4)  /* mtl::missing_diagonal */
#define MTL4G_ERR_BREAKDOWN    (-5)  /* itl::unexpected_orthogonality */
#define MTL4G_ERR_NONFINITE    (-6)  /* NaN/Inf reached a scalar of the recurrence */
#define MTL4G_ERR_OOM          (-7)  /* device allocation failed */
#define MTL4G_ERR_HIP          (-8)  /* other HIP runtime error */
#define MTL4G_ERR_UNSUPPORTED  (-9)  /* index range or format not supported */
#define MTL4G_ERR_TIMEOUT     (-10)  /* a bounded device-side wait expired */

/* assign modes of w (=,+=,-=) A*v: mtl/operation/assign_mode.hpp:22-131 */
#define MTL4G_ASSIGN 0
#define MTL4G_PLUS   1
#define MTL4G_MINUS  2

/* preconditioner kinds: itl/pc/{identity,diagonal,ilu_0,ic_0}.hpp */
#define MTL4G_PC_IDENTITY 0
#define MTL4G_PC_DIAGONAL 1
#define MTL4G_PC_ILU0     2
#define MTL4G_PC_IC0      3

/* value formats (the Value template argument of ilu_0/ic_0, e.g.
 * ic_0<Matrix,float>, itl/pc/ic_0.hpp:38, and inner-solve storage) */
#define MTL4G_F64  0
#define MTL4G_F32  1
#define MTL4G_BF16 2   /* values stored as bfloat16, arithmetic in FP32 */

/* Krylov methods */
#define MTL4G_CG       0   /* itl::cg,       itl/krylov/cg.hpp:38-111 */
#define MTL4G_BICGSTAB 1   /* itl::bicgstab, itl/krylov/bicgstab.hpp:27-78 */
#define MTL4G_CGS      2   /* itl::cgs,      itl/krylov/cgs.hpp:24-62 (FP64 only) */
#define MTL4G_TFQMR    3   /* itl::tfqmr,    itl/krylov/tfqmr.hpp:32-96 (FP64; mtl4g_tfqmr) */
#define MTL4G_BICGSTAB_2 4 /* itl::bicgstab_2, itl/krylov/bicgstab_2.hpp:37-130 (FP64) */
#define MTL4G_BICGSTAB_ELL 5 /* itl::bicgstab_ell, itl/krylov/bicgstab_ell.hpp:37-134 (FP64; mtl4g_bicgstab_ell) */
#define MTL4G_GMRES    6   /* itl::gmres,    itl/krylov/gmres.hpp:37-137 (FP64; mtl4g_gmres) */

typedef struct mtl4g_matrix_s* mtl4g_matrix;
typedef struct mtl4g_pc_s* mtl4g_pc;
typedef struct mtl4g_solver_s* mtl4g_solver;

/* State of itl::basic_iteration<double> (basic_iteration.hpp:23-164):
 * in: norm_r0, max_iter, rtol, atol, i (start count); out: i, error, resid.
 * A solve stops when resid <= rtol*norm_r0 or resid <= atol (when
 * norm_r0 == 0: resid <= atol only), or with error 1 when i >= max_iter. */
typedef struct {
    double norm_r0;
    int32_t max_iter;
    double rtol;
    double atol;
    int32_t i;
    int32_t error;
    double resid;
} mtl4g_iteration;

/* ---- library -------------------------------------------------------- */
const char* mtl4g_status_string(int status);
/* Message of the last failing call on this thread ("" if none). */
const char* mtl4g_last_error(void);
int mtl4g_version(void);
/* Sets the device used by later calls on this thread. */
int mtl4g_set_device(int device);

/* Summation order of the FP64 reductions (dot, unary_dot, two_norm, the
 * fused reductions and those inside cg / bicgstab / solver_solve).
 * MTL4G_ORDER_TREE (default): every workgroup sums its share, the partial
 *   sums are combined in workgroup order -- deterministic, full bandwidth,
 *   but rounded differently from MTL4.
 * MTL4G_ORDER_MTL4: MTL4's own sequential order -- four accumulators taking
 *   the indices i = k (mod 4), the remainder into the first, finished as
 *   dot<4> ((t0+t1)+(t2+t3)) (operation/dot.hpp:136-147), reduction<4>
 *   t0+(t1+(t2+t3)) (vector/reduction.hpp:28-82), the fused SpMV dot
 *   ((t0+t1)+t2)+t3 (operation/fused_expr.hpp:57-77, dot_index_evaluator)
 *   and the fused update norm (t0+t1)+(t2+t3)
 *   (vector/reduction_index_evaluator.hpp:22-41).  Products and updates are
 *   computed per element as before, so every FP64 solve then reproduces
 *   MTL4's CPU result bit for bit; the sums run on four lanes, so this mode
 *   is for validation on small systems (about n/4 dependent additions per
 *   reduction).  No MTL4 counterpart; process-wide, read at each launch.
 * Mixed-precision and batched solves always use the tree order. */
#define MTL4G_ORDER_TREE 0
#define MTL4G_ORDER_MTL4 1
int mtl4g_set_reduction_order(int order);
int mtl4g_get_reduction_order(void);
/* Diagnostic (no MTL4 counterpart): with on = 1, itl::cg solves record a
 * pair of HIP events on their stream around each product q = A*p of their
 * iterations (at most 65536 per solve); mtl4g_product_timing returns the
 * summed duration in microseconds and the number of products of the solves
 * since the last call (launches past a solve's stop are not counted), and
 * clears them.  Process-wide; for one solve at a time. */
int mtl4g_set_product_timing(int on);
int mtl4g_product_timing(double* h_total_us, int64_t* h_count);

/* Device storage for dense_vector (vector/dense_vector.hpp:50-269 keeps its
 * elements in a contiguous_memory_block; here the block lives in HBM).
 * kind: 0 host->device, 1 device->host, 2 device->device. */
int mtl4g_malloc(void** d_ptr, size_t bytes);
int mtl4g_free(void* d_ptr);
int mtl4g_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int mtl4g_memset(void* d_ptr, int value, size_t bytes, void* stream);
int mtl4g_stream_synchronize(void* stream);

/* ---- compressed2D --------------------------------------------------- */
/* mtl::compressed2D<double>(m, n, nnz, Major, Minor, Entries): copies the
 * arrays (matrix/compressed2D.hpp:355-364).  int32 row pointers and column
 * indices (mtl::mat::unsigned_parameters, matrix/parameter.hpp:47), rows
 * sorted by column, no duplicates.  where: 0 = host arrays, 1 = device. */
int mtl4g_csr_create(mtl4g_matrix* A, int32_t nrows, int32_t ncols, int64_t nnz,
                     const int32_t* rowptr, const int32_t* col, const double* val,
                     int where);
int mtl4g_csr_destroy(mtl4g_matrix A);
/* num_rows(A), num_cols(A), A.nnz() (compressed2D.hpp:610-620) */
int mtl4g_csr_info(mtl4g_matrix A, int32_t* nrows, int32_t* ncols, int64_t* nnz);
/* Overwrite the values (same pattern), e.g. after an inserter with
 * update_store on existing entries (matrix/inserter.hpp:30-48). */
int mtl4g_csr_set_values(mtl4g_matrix A, const double* val, int where);
/* CSR kernel choice.  MTL4G_SPMV_VECTOR: `group` lanes (4..64) per row,
 * `rows_per_block` rows per workgroup; MTL4G_SPMV_TILE: row blocks of at
 * most rows_per_block rows and nnz_per_block non-zeros staged through LDS;
 * MTL4G_SPMV_STREAM: the same row blocks walked by a persistent grid of
 * `group` workgroups per compute unit (1..8), each keeping its next
 * block's stream in flight while it sums the current one;
 * MTL4G_SPMV_STREAM2: the stream kernel with two row blocks requested
 * ahead instead of one;
 * MTL4G_SPMV_SCALAR: one lane per row, `group` (4 or 8) values and
 * indices requested per step, no LDS;
 * MTL4G_SPMV_SLICED: a sliced-ELL copy of the matrix (slices of 64 rows,
 * each padded to its longest row, stored column-major) built on the
 * device, one lane per row, a persistent grid of `group` workgroups per
 * compute unit (1..8), `rows_per_block` = entries per lane and request
 * chunk (4 or 8; 0 = 8); refused when the padding would exceed half the
 * non-zeros;
 * MTL4G_SPMV_DIA: a diagonal (banded) copy of the matrix -- one value per
 * row and diagonal, no column indices -- for matrices whose entries lie on
 * at most 32 diagonals filling at least 4/5 of their slots (stencils);
 * `group` and `rows_per_block` as for MTL4G_SPMV_SLICED; refused otherwise.
 * MTL4G_SPMV_DIA_SYM: the same for a square matrix equal to its transpose
 * bit for bit; only the diagonals with offset >= 0 are stored, the entries
 * below the diagonal are read from their mirror images (the value stream
 * is about half); refused for other matrices.
 * MTL4G_SPMV_PANEL: the columns split into panels of `rows_per_block`
 * columns (0 = 262144: 2 MB of FP64 x, within one XCD's L2), one CSR copy
 * of each panel's entries, the stream kernel over each panel in column
 * order with every row's partial sum carried from panel to panel (a
 * stream-ordered scratch vector per call), `group` workgroups per CU; for
 * matrices whose x gathers scatter over a vector larger than the L2;
 * refused when a row's columns are not ascending.
 * All sum every row in column order.  0 = library default. */
#define MTL4G_SPMV_TILE   0
#define MTL4G_SPMV_VECTOR 1
#define MTL4G_SPMV_STREAM 2
#define MTL4G_SPMV_SCALAR 3
#define MTL4G_SPMV_STREAM2 4
#define MTL4G_SPMV_SLICED 5
#define MTL4G_SPMV_DIA 6
#define MTL4G_SPMV_DIA_SYM 7
#define MTL4G_SPMV_PANEL 8
int mtl4g_csr_tune(mtl4g_matrix A, int32_t rows_per_block, int32_t nnz_per_block);
int mtl4g_csr_set_kernel(mtl4g_matrix A, int kernel, int32_t group, int32_t rows_per_block);
int mtl4g_csr_get_kernel(mtl4g_matrix A, int* kernel, int32_t* group, int32_t* rows_per_block);
/* Workgroups of the persistent grid that y = A*x on these values
 * (MTL4G_F64 / F32 / BF16) launches with: `group` per CU clamped to what one
 * CU holds of the chosen kernel; 0 for the non-persistent kernels.  No MTL4
 * counterpart (a launch-shape diagnostic). */
int mtl4g_csr_grid(mtl4g_matrix A, int value_format, int32_t* workgroups);
/* Autotuner: times y = A*x for the candidate kernels and shapes on this
 * matrix and keeps the fastest; h_best_us (may be NULL) gets its time. */
int mtl4g_csr_autotune(mtl4g_matrix A, double* h_best_us, void* stream);
/* Builds the MTL4G_F32 or MTL4G_BF16 copy of A's values in every device
 * layout and tunes the product on them, which the first low-precision
 * product or mixed-precision solve would otherwise do (no MTL4
 * counterpart: set-up ahead of timed or latency-bound use).  MTL4G_F64 is
 * accepted and does nothing. */
int mtl4g_csr_prepare(mtl4g_matrix A, int value_format, void* stream);

/* w (=,+=,-=) A*v: smat_cvec_mult, operation/mat_vec_mult.hpp:613-683,
 * dispatched from vector/crtp_base_vector.hpp:93-104,191-200,268-277.
 * Each row is summed in column order, as MTL4 does. */
int mtl4g_spmv(mtl4g_matrix A, const double* d_v, double* d_w, int assign, void* stream);
/* The same product on the FP32 or BF16 value copy with FP32 vectors
 * (inner solves of mixed-precision refinement; no MTL4 counterpart). */
int mtl4g_spmv_f32(mtl4g_matrix A, int value_format, const float* d_v, float* d_w,
                   int assign, void* stream);
/* (lazy(q)= A*p) || (lazy(alpha)= lazy_dot(p,q)), itl/krylov/cg.hpp:103 */
int mtl4g_spmv_dot(mtl4g_matrix A, const double* d_p, double* d_q, double* h_dot, void* stream);
/* mtl4g_spmv_dot that allocates no device memory.  No MTL4 counterpart: it
 * is for callers that cannot allocate (tight serving loops, harnesses that
 * forbid hipMalloc).  Its reduction scratch is a fixed per-device array of
 * of 9.6 MiB held until the process ends; in MTL4 reduction order it
 * takes at most 1,048,576 rows (MTL4G_ERR_UNSUPPORTED beyond), in the
 * default order any size.  Calls that overlap on one device race. */
int mtl4g_spmv_dot_noalloc(mtl4g_matrix A, const double* d_p, double* d_q, double* h_dot,
                           void* stream);

/* ---- dense_vector reductions and updates ----------------------------- */
/* dot(v, w), operation/dot.hpp:226-232 */
int mtl4g_dot(int64_t n, const double* d_v, const double* d_w, double* h_result, void* stream);
/* unary_dot(v), operation/unary_dot.hpp:56-65 */
int mtl4g_unary_dot(int64_t n, const double* d_v, double* h_result, void* stream);
/* two_norm(v), operation/two_norm.hpp:52-57 */
int mtl4g_two_norm(int64_t n, const double* d_v, double* h_result, void* stream);
/* w += alpha*v (vector/vec_vec_aop_expr.hpp:61-180) */
int mtl4g_axpy(int64_t n, double alpha, const double* d_v, double* d_w, void* stream);
/* (lazy(r) -= alpha*q) || (lazy(rr)= lazy_unary_dot(r)), cg.hpp:108 */
int mtl4g_axpy_unary_dot(int64_t n, double alpha, const double* d_q, double* d_r,
                         double* h_rr, void* stream);

/* ---- preconditioners ------------------------------------------------ */
/* itl::pc::identity(A) (identity.hpp:32), diagonal(A) (diagonal.hpp:37-45),
 * ilu_0<Matrix,Value>(A) (ilu_0.hpp:34-92 with ilu.hpp:54-56), ic_0<Matrix,
 * Value>(A) (ic_0.hpp:57,135-186).  The ILU(0)/IC(0) factorization runs on
 * the device in FP64, level-scheduled, each row in MTL4's loop order (the
 * factors equal MTL4's bit for bit); value_format selects the stored factor
 * precision.  Sweeps run level-scheduled on the device.  The handle keeps
 * A: destroy it before A. */
int mtl4g_pc_create(mtl4g_pc* P, mtl4g_matrix A, int kind, int value_format);
int mtl4g_pc_destroy(mtl4g_pc P);
/* solve(P, x) into y: P.solve(x, y) (pc/solver.hpp:23-52) */
int mtl4g_pc_solve(mtl4g_pc P, const double* d_x, double* d_y, void* stream);
/* adjoint_solve(P, x) into y (diagonal.hpp:76-83, ilu.hpp:105-115) */
int mtl4g_pc_adjoint_solve(mtl4g_pc P, const double* d_x, double* d_y, void* stream);
/* Copy a factor to the host, as stored (inverted diagonals, FP64 values):
 * which 0 = inverse diagonal (n values), 1 = L, 2 = U (CSR, natural row
 * order).  Pass NULL arrays to query *n_out and *nnz_out first. */
int mtl4g_pc_factor(mtl4g_pc P, int which, int32_t* n_out, int64_t* nnz_out,
                    int32_t* h_rowptr, int32_t* h_col, double* h_val);
/* Number of dependency levels of the lower / upper sweep. */
int mtl4g_pc_levels(mtl4g_pc P, int32_t* lower_levels, int32_t* upper_levels);
/* Schedule of the ILU(0) / IC(0) triangular sweeps of preconditioners
 * created afterwards (no MTL4 counterpart; MTL4 sweeps rows in order):
 * MTL4G_SWEEP_AUTO (default): a factor of a 5- or 7-point grid structure
 * (every row depending only on the rows 1, m and m*m before it, never
 * across a grid edge) runs as a pipeline of row chains, any other factor
 * level by level; MTL4G_SWEEP_LEVELS: every factor level by level.  Both
 * compute every row in MTL4's subtraction order: the results are the same
 * bits.  mtl4g_pc_sweeps reports the schedule of a preconditioner's two
 * sweeps. */
#define MTL4G_SWEEP_AUTO   0
#define MTL4G_SWEEP_LEVELS 1
int mtl4g_set_sweep_schedule(int schedule);
int mtl4g_pc_sweeps(mtl4g_pc P, int* lower, int* upper);
/* Tile of the chain pipeline (no MTL4 counterpart): Tx x Ty row chains per
 * workgroup (one wave: Tx*Ty = 64) for factors created afterwards; (0, 0)
 * restores the choice per grid.  mtl4g_pc_chain_tile reports a
 * preconditioner's tiles (0 for a sweep that runs level by level). */
int mtl4g_set_chain_tile(int32_t tx, int32_t ty);
int mtl4g_pc_chain_tile(mtl4g_pc P, int32_t* lower_tx, int32_t* lower_ty, int32_t* upper_tx, int32_t* upper_ty);
/* Factor again from A's current values (after mtl4g_csr_set_values), same
 * pattern: ILU(0)/IC(0) rerun the device factorization on `stream` with
 * the plan made at creation (levels, positions), nothing on the host; the
 * diagonal PC recomputes its reciprocals.  No MTL4 counterpart (MTL4
 * factors only in the constructor); the result equals a new
 * mtl4g_pc_create(A) bit for bit.  Not concurrent with solves using P. */
int mtl4g_pc_refactor(mtl4g_pc P, void* stream);

/* ---- Krylov solvers ------------------------------------------------- */
/* itl::cg(A, x, b, P, iter) (cg.hpp:74-111; P == NULL or identity takes
 * the unpreconditioned form cg.hpp:38-71 as is_identity does, :80-81) and
 * itl::bicgstab(A, x, b, P, iter) (bicgstab.hpp:27-78).  x holds the start
 * vector and receives the result.  The whole loop runs on the device;
 * h_hist (may be NULL) receives the residual of each finished() call, up
 * to hist_cap values.  Returns it->error (0 or 1) or a negative status. */
int mtl4g_cg(mtl4g_matrix A, mtl4g_pc P, double* d_x, const double* d_b,
             mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
int mtl4g_bicgstab(mtl4g_matrix A, mtl4g_pc P, double* d_x, const double* d_b,
                   mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
/* itl::cgs(A, x, b, P, iter) (cgs.hpp:24-62): same conventions; rho_1 == 0
 * sets error code 2 and the loop goes on, as MTL4's iter.fail(2) does. */
int mtl4g_cgs(mtl4g_matrix A, mtl4g_pc P, double* d_x, const double* d_b,
              mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
/* itl::tfqmr(A, x, b, L, R, iter) (tfqmr.hpp:32-96): left preconditioner L
 * and right preconditioner R (either may be NULL = identity); x is shifted
 * back by solve(R, x) after the loop as MTL4 does.  A zero sigma or rho is
 * iter.fail(1): returns 1.  An empty system is MTL4's logic_error. */
int mtl4g_tfqmr(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* d_x, const double* d_b,
                mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
/* itl::bicgstab_2(A, x, b, L, iter) (bicgstab_2.hpp:37-130): BiCGSTAB(l)
 * with l = 2 and left preconditioner L (NULL = identity).  finished() runs
 * at the loop head on r_hat[0] and after each inner step on r_hat[j]; an
 * exit there returns before that step's x += alpha*u_hat[0], as in MTL4.
 * omega == 0 is iter.fail(3): returns 3 and leaves x unshifted (no x +=
 * x0).  An empty system is MTL4's logic_error. */
int mtl4g_bicgstab_2(mtl4g_matrix A, mtl4g_pc L, double* d_x, const double* d_b,
                     mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
/* itl::bicgstab_ell(A, x, b, L, R, iter, l) (bicgstab_ell.hpp:37-134):
 * BiCGSTAB(l), 1 <= l <= 8, left preconditioner L and right preconditioner
 * R (NULL = identity).  finished() runs at the loop head on r_hat[0] and
 * after each inner step j on r_hat[j]; every exit but iter.fail(3) (omega
 * == 0, returns 3) sets x = solve(R, x) + x0 as MTL4 does.  An empty
 * system is MTL4's logic_error; l outside 1..8 is MTL4G_ERR_INVALID. */
int mtl4g_bicgstab_ell(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* d_x, const double* d_b, int32_t l,
                       mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);
/* itl::gmres(A, x, b, L, R, iter, restart) (gmres.hpp:122-137) for restart
 * > 0: cycles of gmres_full on a copy of iter limited to restart more
 * iterations, each followed by iter.update_progress; restart == 0 is
 * itl::gmres_full(A, x, b, L, R, iter) (:37-119) alone.  A cycle always runs
 * min(n, iterations left) Arnoldi steps, as MTL4's does.  h_hist receives
 * the residuals a cycle hands to finished(): two_norm(solve(L, b - A*x))
 * at its start, two_norm(b - A*x) at its end.  iter.fail(2) (no direction)
 * returns 2.  At most 256 Krylov vectors per cycle (MTL4G_ERR_INVALID
 * beyond); an empty system is MTL4's logic_error. */
int mtl4g_gmres(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* d_x, const double* d_b, int32_t restart,
                mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream);

/* Reusable solver: the workspace (MTL4 allocates it per call, cg.hpp:90)
 * stays with the handle, so repeated solves allocate nothing. */
int mtl4g_solver_create(mtl4g_solver* S, mtl4g_matrix A, mtl4g_pc P, int method);
int mtl4g_solver_destroy(mtl4g_solver S);
int mtl4g_solver_solve(mtl4g_solver S, double* d_x, const double* d_b, mtl4g_iteration* it,
                       double* h_hist, int32_t hist_cap, void* stream);

/* ---- mixed-precision iterative refinement (no MTL4 counterpart) ------ */
/* The residual r = b - A*x, the solution update and the stopping norm are
 * FP64 on the device (stop: ||r|| <= rtol*norm_r0 or atol, as in
 * basic_iteration); the Krylov recurrence runs in FP32 arithmetic on FP32
 * or BF16 stored values.
 *  MTL4G_IR_RELIABLE: one FP32 CG/BiCGSTAB recurrence; its updates are
 *    accumulated into x in FP64 and r is recomputed in FP64 ("reliable
 *    updates") whenever the FP32 residual fell by inner_rtol since the last
 *    refresh, and whenever the FP32 recurrence believes it has converged.
 *    it->max_iter bounds the FP32 iterations; it->i returns them.
 *  MTL4G_IR_RESTART: classic refinement: solve A*d = r in FP32 to
 *    inner_rtol*||r|| (at most inner_max_iter iterations), x += d, repeat
 *    (at most max_outer steps); it->i returns the refinement steps. */
#define MTL4G_IR_RELIABLE 0
#define MTL4G_IR_RESTART  1

typedef struct {
    int32_t method;          /* MTL4G_CG or MTL4G_BICGSTAB */
    int32_t value_format;    /* MTL4G_F32 or MTL4G_BF16 for the FP32 recurrence */
    int32_t mode;            /* MTL4G_IR_RELIABLE or MTL4G_IR_RESTART */
    double inner_rtol;       /* reliable: refresh trigger (e.g. 0.1); restart: inner tolerance */
    int32_t inner_max_iter;  /* restart: per inner solve */
    int32_t max_outer;       /* restart: refinement steps */
} mtl4g_ir_params;

typedef struct {
    int32_t outer;           /* FP64 residual refreshes (reliable) or refinement steps */
    int32_t inner_total;     /* FP32 iterations */
    double relres;           /* final FP64 ||b - A x|| / norm_r0 */
} mtl4g_ir_stats;

/* P32 (may be NULL = identity) must be created with value_format F32. */
int mtl4g_ir_solve(mtl4g_matrix A, mtl4g_pc P32, double* d_x, const double* d_b,
                   mtl4g_iteration* it, const mtl4g_ir_params* prm, mtl4g_ir_stats* st,
                   void* stream);

/* ---- batched solves of many independent systems (no MTL4 counterpart) -- */
/* itl::cg with pc::identity or pc::diagonal (cg.hpp:38-111) run for many
 * systems by one persistent launch: each work item is solved entirely by
 * one workgroup, so results do not depend on the batch they ran in.
 * Precision MTL4G_F64: FP64 as MTL4; MTL4G_F32 / MTL4G_BF16: FP32 recurrence
 * with FP64 reliable updates (FP64 stop, see mtl4g_ir_solve). */
typedef struct mtl4g_batch_s* mtl4g_batch;

/* One work item: system sys, right-hand side at d_b + b_off, start vector and
 * result at d_x + x_off, workspace at d_work + ws_off (doubles,
 * mtl4g_batch_workspace_elems of the system size). */
typedef struct {
    int32_t sys;
    int32_t max_iter;
    int64_t b_off, x_off, ws_off;
    double rtol, atol;
} mtl4g_batch_item;

/* basic_iteration's state after an item: iterations, error (0/1), status
 * (< 0: non-finite), FP64 refreshes (mixed precision), resid, norm_r0 = ||b||. */
typedef struct {
    int32_t iterations, error, status, refreshes;
    double resid, norm_r0;
} mtl4g_batch_result;

/* Systems in one block-diagonal CSR on the host: system k owns rows
 * [sys_rowptr[k], sys_rowptr[k+1]) and only columns in that range (absolute
 * column indices, sorted rows).  The arrays are copied to the device. */
int mtl4g_batch_create(mtl4g_batch* B, int32_t nsys, const int32_t* sys_rowptr,
                       const int32_t* rowptr, const int32_t* col, const double* val, int where);
int mtl4g_batch_destroy(mtl4g_batch B);
int mtl4g_batch_info(mtl4g_batch B, int32_t* nsys, int32_t* nrows, int64_t* nnz);
int64_t mtl4g_batch_workspace_elems(int32_t n, int precision);
/* Launch shape (threads per workgroup 256/512/1024, resident workgroups per CU). */
int mtl4g_batch_set_shape(mtl4g_batch B, int precision, int32_t threads, int32_t per_cu);
int mtl4g_batch_get_shape(mtl4g_batch B, int precision, int32_t* threads, int32_t* per_cu);
/* Every system once: x and b laid out like the rows (d_x + sys_rowptr[k]);
 * systems run largest first; h_results (may be NULL) indexed by system.
 * The launch (queue reset + persistent kernel) is a captured hipGraph,
 * recorded once per preconditioner, precision, shape and buffer set and
 * replayed after that; it runs on a stream of the batch's own, ordered
 * after the work already on `stream` and before the work enqueued next. */
int mtl4g_batch_solve(mtl4g_batch B, int pc_kind, int precision, double* d_x, const double* d_b,
                      double rtol, double atol, int32_t max_iter, mtl4g_batch_result* h_results,
                      void* stream);
/* Number of graph captures and replays made by mtl4g_batch_solve so far. */
int mtl4g_batch_graph_info(mtl4g_batch B, int64_t* captures, int64_t* replays);
/* Autotuner: times the whole batch for every launch shape (x reset from
 * d_x0 each time) and keeps the fastest for this precision. */
int mtl4g_batch_tune(mtl4g_batch B, int pc_kind, int precision, double* d_x, const double* d_x0,
                     const double* d_b, double rtol, int32_t max_iter, double* h_best_ms,
                     void* stream);
/* A list of work items already on the device (d_queue: one zero-able word);
 * asynchronous on the stream.  Used by continuous batching. */
int mtl4g_batch_solve_items(mtl4g_batch B, int pc_kind, int precision,
                            const mtl4g_batch_item* d_items, int32_t nitems, const double* d_b,
                            double* d_x, double* d_work, mtl4g_batch_result* d_results,
                            unsigned* d_queue, void* stream);

/* ---- serving loop: continuous batching (no MTL4 counterpart) --------- */
/* Matrices stay resident in HBM (one mtl4g_batch).  Requests (matrix, number
 * of right-hand sides, solver) are admitted in arrival order while their
 * device footprint (b, x and workspace of every right-hand side) fits the
 * cap; the admitted requests are packed into persistent batched launches on
 * `streams` concurrent streams as soon as a stream is free (one solver per
 * launch), so launches overlap on the device.  Systems with more than
 * big_rows rows are solved by the whole-GPU engine, `streams` of them at a
 * time on streams of their own.  A solve is itl::cg or itl::bicgstab with
 * the identity, diagonal or ilu_0 preconditioner at rtol; precision as in
 * mtl4g_batch_solve (F32 / BF16: mixed precision refined to rtol).  The
 * device arena holding the requests' vectors is allocated by
 * mtl4g_server_run: the smaller of cap_bytes and the trace's total
 * footprint. */
#define MTL4G_SERVER_SOLVER(method, pc_kind) (((method) << 8) | (pc_kind))
typedef struct mtl4g_server_s* mtl4g_server;

typedef struct {
    int32_t streams;        /* concurrent batched launches (e.g. 4) */
    int32_t max_items;      /* right-hand sides per launch (e.g. 4096) */
    int32_t precision;      /* MTL4G_F64, MTL4G_F32 or MTL4G_BF16 */
    int32_t pc_kind;        /* preconditioner of requests without a solver code:
                               MTL4G_PC_IDENTITY, _DIAGONAL or _ILU0 (cg) */
    int32_t max_iter;
    int32_t big_rows;       /* larger systems use the whole-GPU engine */
    double rtol;
    double cap_bytes;       /* admission cap on the in-flight working set */
    int32_t per_request;    /* 1: one request per launch (the plain per-request
                               baseline; with streams = 1, no overlap) */
} mtl4g_server_params;

typedef struct {
    double wall_s;          /* first arrival to last completion */
    int64_t solves;         /* right-hand sides solved */
    int32_t launches;       /* batched launches */
    int32_t max_inflight;   /* most launches on the device at once */
    double peak_bytes;      /* largest admitted working set */
    int32_t errors;         /* solves that did not converge */
    int32_t big_solves;     /* solves taken by the whole-GPU engine */
} mtl4g_server_stats;

int mtl4g_server_create(mtl4g_server* S, mtl4g_batch B, const mtl4g_server_params* prm);
int mtl4g_server_destroy(mtl4g_server S);
/* Replays a request trace in real time: request r arrives arrival_s[r]
 * seconds after the start for matrix mat[r] with nrhs[r] right-hand sides
 * generated on the device from seed[r] (values uniform in [-1,1]), solved
 * by solver[r] = MTL4G_SERVER_SOLVER(MTL4G_CG or MTL4G_BICGSTAB, pc kind)
 * (solver NULL: itl::cg with params.pc_kind for every request).
 * latency_s[r] = completion - arrival.  Optional outputs: iters (one per
 * right-hand side, requests in order), x_out + x_off[r] (the solutions of
 * request r, nrhs[r] vectors of the matrix size one after the other). */
int mtl4g_server_run(mtl4g_server S, int32_t nreq, const int32_t* mat, const int32_t* nrhs,
                     const int32_t* solver, const double* arrival_s, const uint64_t* seed,
                     double* latency_s, int32_t* iters, double* x_out, const int64_t* x_off,
                     mtl4g_server_stats* st);

#ifdef __cplusplus
}
#endif
#endif /* MTL4G_H */
