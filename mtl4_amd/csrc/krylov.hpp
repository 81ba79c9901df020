// krylov.hpp -- device-resident itl::cg / itl::bicgstab.
#pragma once

#include "internal.hpp"
#include "precond.hpp"

namespace mtl4g {

// Scalars of the recurrences and of basic_iteration, kept on the device.
struct Ctl {
    double rho, rho_1, rr, alpha, alpha_1, beta, omega, gamma, rho_2;
    double theta, eta, tau;  // tfqmr
    int32_t looped;          // tfqmr: the loop was entered (x is shifted back)
    // bicgstab_2: tau of the minimal-residual part, gamma, gamma_a, gamma_aa
    double t11, t12, t22, ga1, ga2, g1, g2, gaa1, nr;
    int32_t shift;           // bicgstab_2: x was nonzero, the problem is shifted
    double norm_r0, rtol, atol, resid;
    double rmax, delta;      // mixed precision: residual at the last FP64 refresh, trigger
    int32_t i, max_iter, error, done, first, half, status;
    int32_t hist_n, hist_cap;
    // mixed precision (reliable updates): FP32 recurrence, FP64 refreshes
    int32_t mixed, pause, want_ru, ru_count, limit, init, restart;
    int32_t stall;           // consecutive FP64 refreshes that did not halve the residual
    int32_t drop;            // the FP32 recurrence went non-finite: discard its correction
    unsigned ticket[4];
};

// Element type of the vectors (VT) and of the stored matrix values (MT).
template <typename VT>
struct Workspace {
    int32_t n = 0;
    DevBuf<VT> vecs;          // method-specific vectors
    DevBuf<double> partials;  // kPartialWords
    DevBuf<unsigned> tickets; // kTicketWords, zero between launches
    DevBuf<Ctl> ctl;
    DevBuf<double> hist;
    DevBuf<unsigned> pc_counters;
    DevBuf<double> terms;     // 3n, reductions in MTL4's order (FP64 only)
    DevBuf<double> small;     // bicgstab_ell: tau, gamma, gamma_a, gamma_aa
    DevBuf<double> pc64;      // float vectors with ilu_0/ic_0: the sweeps' FP64 chains (2n)
    Ctl* h_ctl = nullptr;     // pinned mirror for polling
    ~Workspace() {
        if (h_ctl) (void)hipHostFree(h_ctl);
    }
};

// Runs a whole Krylov solve on the device.  x and b are device vectors of
// VT; value_format selects the stored matrix values for VT = float.
template <typename VT>
int krylov_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* P, int method, int value_format,
               Workspace<VT>& W, VT* x, const VT* b, mtl4g_iteration* it, double* h_hist,
               int32_t hist_cap, hipStream_t s);

// itl::tfqmr(A, x, b, L, R, iter) on the device (L, R may be NULL = identity).
int tfqmr_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, Workspace<double>& W,
              double* x, const double* b, mtl4g_iteration* it, double* h_hist, int32_t hist_cap,
              hipStream_t s);

// itl::bicgstab_2(A, x, b, L, iter) on the device (L may be NULL = identity).
int bicgstab2_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, Workspace<double>& W, double* x,
                  const double* b, mtl4g_iteration* it, double* h_hist, int32_t hist_cap,
                  hipStream_t s);

// itl::bicgstab_ell(A, x, b, L, R, iter, l) on the device, 1 <= l <= 8
// (L, R may be NULL = identity).
int bicgstab_ell_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, int32_t l,
                     Workspace<double>& W, double* x, const double* b, mtl4g_iteration* it, double* h_hist,
                     int32_t hist_cap, hipStream_t s);

// itl::gmres(A, x, b, L, R, iter, restart) (restart > 0) or itl::gmres_full
// (restart == 0) on the device, at most 256 Krylov vectors per cycle.
int gmres_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, int32_t restart,
              Workspace<double>& W, double* x, const double* b, mtl4g_iteration* it, double* h_hist,
              int32_t hist_cap, hipStream_t s);

// Mixed precision with reliable updates: products read the matrix values in
// value_format (FP32 or BF16) and the preconditioner's values as P32 stores
// them; the true residual b - A*x is recomputed with the FP64 values
// whenever the recurrence residual fell by delta (or grew tenfold) since the
// last refresh, and the stop decision is taken on that FP64 residual.
// CG: an FP32 recurrence in W, its correction x32 folded into x at each
// refresh.  BiCGSTAB: FP64 vectors in W64, x updated in place (x32 unused).
int mixed_run(mtl4g_matrix_s* A, const mtl4g_pc_s* P32, int method, int value_format,
              double delta, int32_t max_refresh, Workspace<float>& W, Workspace<double>& W64,
              float* x32, double* x, const double* b, mtl4g_iteration* it, int32_t* refreshes,
              hipStream_t s);

}  // namespace mtl4g

struct mtl4g_solver_s {
    mtl4g_matrix_s* A = nullptr;
    mtl4g_pc_s* P = nullptr;
    int method = MTL4G_CG;
    mtl4g::Workspace<double> W;
};
