// krylov.hip -- itl::cg and itl::bicgstab with the whole loop on the device.
//
// Every scalar of the recurrence (rho, rho_1, alpha, beta, omega, ...) and
// the state of basic_iteration live in a device Ctl block.  Each reduction
// is fused into the kernel that produces its operands, and the last
// workgroup of that kernel combines the partials and takes the decisions
// MTL4 takes on the host (finished(), ++iter, first(), the breakdown
// throws).  Once Ctl::done is set every later kernel exits at entry, so the
// host only enqueues chunks of iterations and polls Ctl::done between
// chunks; no host round trip sits inside an iteration.
//
// CG + diagonal (cg.hpp:74-111) runs three kernels per iteration:
//   P : p = z + beta*p            with z = inv_diag .* r formed on the fly
//   Q : q = A*p, p.q              (SpMV + dot, cg.hpp:103)      -> alpha
//   XR: x += alpha*p, r -= alpha*q, r.r, r.(inv_diag .* r)      -> stop?, beta
// i.e. B_spmv + 11 N V bytes per iteration; z is never stored.  x (touched
// by XR only) and q (written by Q, read once by XR) move with non-temporal
// loads and stores, so the matrix, p, r and inv_diag stay in the MALL from
// one iteration to the next.
// General preconditioners add pc_apply(r -> z) and a dot for rho.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "krylov.hpp"

namespace mtl4g {

enum { PCM_NONE = 0, PCM_DIAG = 1, PCM_GENERAL = 2 };

// basic_iteration::finished(Real) (basic_iteration.hpp:60-65,85-93,43-48)
// followed by ++iter when the loop goes on.  NaN/Inf residuals stop the
// solve with MTL4G_ERR_NONFINITE instead of running to max_iter.
// finished() alone (no ++iter): sets done when the loop must stop
__device__ __forceinline__ void iter_check(Ctl* c, double resid, double* hist) {
    if (c->hist_n < c->hist_cap) hist[c->hist_n] = resid;
    c->hist_n++;
    c->resid = resid;
    const bool conv = c->norm_r0 == 0 ? (resid <= c->atol)
                                      : (resid <= c->rtol * c->norm_r0 || resid <= c->atol);
    if (conv) { c->done = 1; return; }
    if (c->i >= c->max_iter) { c->error = 1; c->done = 1; return; }
    if (!isfinite(resid)) { c->status = MTL4G_ERR_NONFINITE; c->done = 1; return; }
}

__device__ __forceinline__ void iter_finished(Ctl* c, double resid, double* hist) {
    iter_check(c, resid, hist);
    if (c->done) return;
    c->i++;
    c->first = c->i <= 1;
}

__device__ __forceinline__ bool conv_test(const Ctl* c, double resid) {
    return c->norm_r0 == 0 ? (resid <= c->atol) : (resid <= c->rtol * c->norm_r0 || resid <= c->atol);
}

// A recurrence residual this far above the last FP64 residual is refreshed
// at once instead of iterating on.
constexpr double kGrowth = 10.0;
// Mixed-precision BiCGSTAB refreshes when its residual fell by this factor.
constexpr double kBicgDelta = 1e-5;

// Mixed precision, loop head on the recurrence residual: it only estimates
// convergence, so an estimated stop (or the iteration limit) pauses the
// recurrence and asks for an FP64 refresh, which decides.  A non-finite
// residual also pauses it and discards the low-precision correction; a
// residual that has grown by kGrowth since the last refresh is refreshed
// before the recurrence goes on.
__device__ __forceinline__ void mixed_check(Ctl* c, double resid32) {
    if (!isfinite(resid32)) { c->want_ru = 1; c->pause = 1; c->drop = 1; return; }
    if (conv_test(c, resid32)) { c->want_ru = 1; c->pause = 1; return; }
    if (c->i >= c->max_iter) { c->want_ru = 1; c->pause = 1; c->limit = 1; return; }
    if (resid32 > kGrowth * c->rmax) { c->want_ru = 1; c->pause = 1; return; }
    c->i++;
    c->first = c->i <= 1;
    if (resid32 <= c->delta * c->rmax) c->want_ru = 1;
}

constexpr int32_t kMaxStall = 8;

// FP64 refresh done: stop on the true residual, or resume the recurrence
// from the replaced residual (rho and beta of CG recomputed from it).
__device__ __forceinline__ void ru_finish(Ctl* c, double rr64, double rho_new, bool cg_fused) {
    const double resid = sqrt(rr64);
    c->resid = resid;
    if (conv_test(c, resid)) { c->done = 1; return; }
    if (!isfinite(resid)) { c->status = MTL4G_ERR_NONFINITE; c->done = 1; return; }
    if (c->limit || (c->init && c->i >= c->max_iter)) { c->error = 1; c->done = 1; return; }
    // stagnation: the FP32 recurrence keeps asking for refreshes (its own
    // residual estimate has fallen) while the true residual does not fall
    // any more -- its attainable accuracy is reached; stop as MTL4 stops at
    // the iteration limit instead of cycling until max_iter
    if (!c->init) {
        c->stall = resid > 0.5 * c->rmax ? c->stall + 1 : 0;
        if (c->stall >= kMaxStall) { c->error = 1; c->done = 1; return; }
    }
    if (c->pause || c->init) {
        c->i++;
        c->first = c->i <= 1;
        c->pause = 0;
        c->init = 0;
    }
    if (c->restart || c->drop) {   // BiCGSTAB half step that FP64 did not confirm, or a
        c->first = 1;              // discarded non-finite correction: restart p from r
        c->restart = 0;
        c->drop = 0;
        c->half = 0;
    }
    c->rmax = resid;
    c->want_ru = 0;
    c->ru_count++;
    if (cg_fused) {
        c->rho = rho_new;
        c->beta = c->first ? 0.0 : (double)((float)rho_new / (float)c->rho_1);
    }
}

// ---------------------------------------------------------------- CG epilogues
// r = b - A*x; r.r (and r.(d.*r) for the diagonal PC)  (cg.hpp:92-94)
template <typename VT, typename DT, int PCM>
struct EpiCgInit {
    static constexpr int NP = PCM == PCM_DIAG ? 2 : 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    struct Pre {
        VT b, d;
    };
    const VT* b;
    VT* r;
    const DT* d;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const {
        return Pre{b[row], PCM == PCM_DIAG ? static_cast<VT>(d[row]) : VT(0)};
    }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        const VT v = pre.b - s;
        r[row] = v;
        part[0] += v * v;
        if (PCM == PCM_DIAG) {
            const VT z = pre.d * v;
            part[1] += v * z;
        }
    }
    __device__ void finish(const VT* tot) const {
        c->rr = (double)tot[0];
        iter_finished(c, (double)std::sqrt(std::fabs(tot[0])), hist);
        if (!c->done && PCM != PCM_GENERAL) {
            const VT rho = PCM == PCM_DIAG ? tot[1] : tot[0];
            c->beta = c->first ? 0.0 : (double)(rho / static_cast<VT>(c->rho_1));
            c->rho = (double)rho;
        }
    }
};

// rho = dot(r, z) after z = solve(P, r)  (cg.hpp:96)
template <typename VT>
struct EpiCgRho {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    const VT *r, *z;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT* part) const { part[0] += r[i] * z[i]; }
    __device__ void finish(const VT* tot) const {
        c->beta = c->first ? 0.0 : (double)(tot[0] / static_cast<VT>(c->rho_1));
        c->rho = (double)tot[0];
    }
};

// p = z  or  p = z + (rho/rho_1)*p  (cg.hpp:98-101)
template <typename VT, typename DT, int PCM>
struct EpiCgP {
    static constexpr int NP = 0;
    const VT* r;
    const DT* d;
    const VT* z;
    VT* p;
    Ctl* c;
    VT beta;
    int first;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {
        beta = static_cast<VT>(c->beta);
        first = c->first;
    }
    struct In {
        VT zi, p;
    };
    __device__ In load(int64_t i) const {
        In in;
        if (PCM == PCM_NONE) in.zi = r[i];
        else if (PCM == PCM_DIAG) in.zi = static_cast<VT>(d[i]) * r[i];
        else in.zi = z[i];
        in.p = first ? VT(0) : p[i];
        return in;
    }
    __device__ void store(int64_t i, const In& in, VT*) const {
        p[i] = first ? in.zi : in.zi + beta * in.p;
    }
    __device__ void operator()(int64_t i, VT* part) const { store(i, load(i), part); }
    __device__ void finish(const VT*) const {}
};

// (lazy(q)= A*p) || (lazy(alpha_1)= lazy_dot(p,q)); alpha = rho/alpha_1 (cg.hpp:103-107)
template <typename VT>
struct EpiCgQ {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_FUSED_DOT; }
    struct Pre {
        VT p;
    };
    const VT* p;
    VT* q;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{p[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        __builtin_nontemporal_store(s, q + row);
        part[0] += pre.p * s;
    }
    __device__ void finish(const VT* tot) const {
        c->alpha_1 = (double)tot[0];
        c->alpha = (double)(static_cast<VT>(c->rho) / tot[0]);
        c->rho_1 = c->rho;
    }
};

// x += alpha*p; (lazy(r) -= alpha*q) || (lazy(rr)= lazy_unary_dot(r)); then
// the loop head: finished(sqrt(|rr|)), ++iter and, for the fused PCs, the next
// rho and beta  (cg.hpp:106-108, 94-101)
template <typename VT, typename DT, int PCM>
struct EpiCgXR {
    static constexpr int NP = PCM == PCM_DIAG ? 2 : 1;
    __host__ __device__ static constexpr int order(int k) { return k == 0 ? RO_FUSED_NORM : RO_DOT4; }
    VT* x;
    VT* r;
    const VT *p, *q;
    const DT* d;
    Ctl* c;
    double* hist;
    VT alpha;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() { alpha = static_cast<VT>(c->alpha); }
    struct In {
        VT x, p, r, q, d;
    };
    __device__ In load(int64_t i) const {
        return In{__builtin_nontemporal_load(x + i), p[i], r[i], __builtin_nontemporal_load(q + i),
                  PCM == PCM_DIAG ? static_cast<VT>(d[i]) : VT(0)};
    }
    __device__ void store(int64_t i, const In& in, VT* part) const {
        __builtin_nontemporal_store(in.x + alpha * in.p, x + i);
        const VT v = in.r - alpha * in.q;
        r[i] = v;
        part[0] += v * v;
        if (PCM == PCM_DIAG) {
            const VT z = in.d * v;
            part[1] += v * z;
        }
    }
    __device__ void operator()(int64_t i, VT* part) const { store(i, load(i), part); }
    __device__ void finish(const VT* tot) const {
        c->rr = (double)tot[0];
        if (c->mixed) mixed_check(c, (double)std::sqrt(std::fabs(tot[0])));
        else iter_finished(c, (double)std::sqrt(std::fabs(tot[0])), hist);
        if (!c->done && !c->pause && PCM != PCM_GENERAL) {
            const VT rho = PCM == PCM_DIAG ? tot[1] : tot[0];
            c->beta = (double)(rho / static_cast<VT>(c->rho_1));
            c->rho = (double)rho;
        }
    }
};

// ---------------------------------------------------------------- BiCGSTAB epilogues
// r = b - A*x; rtilde = r; finished(two_norm(r))  (bicgstab.hpp:40-43)
template <typename VT>
struct EpiBiInit {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    struct Pre {
        VT b;
    };
    const VT* b;
    VT *r, *rt;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{b[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        const VT v = pre.b - s;
        r[row] = v;
        rt[row] = v;
        part[0] += v * v;
    }
    __device__ void finish(const VT* tot) const {
        iter_finished(c, (double)std::sqrt(tot[0]), hist);
    }
};

// rho_1 = dot(rtilde, r); breakdown checks; beta  (bicgstab.hpp:45-53)
template <typename VT>
struct EpiBiRho {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    const VT *rt, *r;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT* part) const { part[0] += rt[i] * r[i]; }
    __device__ void finish(const VT* tot) const {
        const VT rho_1 = tot[0];
        c->rho_1 = (double)rho_1;
        if (rho_1 == VT(0)) { c->status = MTL4G_ERR_BREAKDOWN; c->done = 1; return; }
        if (!c->first) {
            if (c->omega == 0.0) { c->status = MTL4G_ERR_BREAKDOWN; c->done = 1; return; }
            c->beta = (double)((rho_1 / static_cast<VT>(c->rho_2)) *
                               (static_cast<VT>(c->alpha) / static_cast<VT>(c->omega)));
        }
    }
};

// p = r  or  p = r + beta*(p - omega*v)  (bicgstab.hpp:48-53)
template <typename VT>
struct EpiBiP {
    static constexpr int NP = 0;
    const VT *r, *v;
    VT* p;
    Ctl* c;
    VT beta, omega;
    int first;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {
        beta = static_cast<VT>(c->beta);
        omega = static_cast<VT>(c->omega);
        first = c->first;
    }
    __device__ void operator()(int64_t i, VT*) const {
        p[i] = first ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
    }
    __device__ void finish(const VT*) const {}
};

// v = A*phat; gamma = dot(rtilde, v); alpha = rho_1/gamma  (bicgstab.hpp:56-60)
template <typename VT>
struct EpiBiV {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    struct Pre {
        VT rt;
    };
    const VT* rt;
    VT* v;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{rt[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        v[row] = s;
        part[0] += pre.rt * s;
    }
    __device__ void finish(const VT* tot) const {
        c->gamma = (double)tot[0];
        if (tot[0] == VT(0)) { c->status = MTL4G_ERR_BREAKDOWN; c->done = 1; return; }
        c->alpha = (double)(static_cast<VT>(c->rho_1) / tot[0]);
    }
};

// s = r - alpha*v; half-step exit if finished(s)  (bicgstab.hpp:61-67)
template <typename VT>
struct EpiBiS {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    const VT *r, *v;
    VT* s;
    Ctl* c;
    double* hist;
    VT alpha;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() { alpha = static_cast<VT>(c->alpha); }
    __device__ void operator()(int64_t i, VT* part) const {
        const VT t = r[i] - alpha * v[i];
        s[i] = t;
        part[0] += t * t;
    }
    __device__ void finish(const VT* tot) const {
        if (c->mixed) {   // estimated stop: x += alpha*phat, then let FP64 decide
            const double r = (double)std::sqrt(tot[0]);
            if (!isfinite(r) || conv_test(c, r)) {
                c->half = 1;
                c->want_ru = 1;
                c->pause = 1;
                c->restart = 1;
            }
            return;
        }
        const int i_before = c->i;
        iter_finished(c, (double)std::sqrt(tot[0]), hist);
        if (c->done) c->half = 1;
        else c->i = i_before;   // finished(s) does not count an iteration
    }
};

// x += alpha*phat after a half-step exit; runs once (half 1 -> 2)
template <typename VT>
struct EpiBiHalf {
    static constexpr int NP = 1;
    const VT* phat;
    VT* x;
    Ctl* c;
    VT alpha;
    __device__ bool skip() const { return c->half != 1; }
    __device__ void prepare() { alpha = static_cast<VT>(c->alpha); }
    __device__ void operator()(int64_t i, VT*) const { x[i] = x[i] + alpha * phat[i]; }
    __device__ void finish(const VT*) const { c->half = 2; }
};

// t = A*shat; omega = dot(t,s)/dot(t,t)  (bicgstab.hpp:68-70)
template <typename VT>
struct EpiBiT {
    static constexpr int NP = 2;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    struct Pre {
        VT s;
    };
    const VT* s;
    VT* t;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{s[row]}; }
    __device__ void operator()(int row, VT v, const Pre& pre, VT* part) const {
        t[row] = v;
        part[0] += v * pre.s;
        part[1] += v * v;
    }
    __device__ void finish(const VT* tot) const { c->omega = (double)(tot[0] / tot[1]); }
};

// x += omega*shat + alpha*phat; r = s - omega*t; rho_2 = rho_1; loop head
// finished(r)  (bicgstab.hpp:72-75, 43)
template <typename VT>
struct EpiBiXR {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    VT *x, *r;
    const VT *shat, *phat, *s, *t;
    Ctl* c;
    double* hist;
    VT alpha, omega;
    __device__ bool skip() const { return c->done != 0 || c->pause != 0; }
    __device__ void prepare() {
        alpha = static_cast<VT>(c->alpha);
        omega = static_cast<VT>(c->omega);
    }
    __device__ void operator()(int64_t i, VT* part) const {
        x[i] = x[i] + (omega * shat[i] + alpha * phat[i]);
        const VT v = s[i] - omega * t[i];
        r[i] = v;
        part[0] += v * v;
    }
    __device__ void finish(const VT* tot) const {
        c->rho_2 = c->rho_1;
        if (c->mixed) mixed_check(c, (double)std::sqrt(tot[0]));
        else iter_finished(c, (double)std::sqrt(tot[0]), hist);
    }
};

// ---------------------------------------------------------------- CGS epilogues
// itl::cgs, itl/krylov/cgs.hpp:26-62.  Vectors p, phat, q, vhat, u, uhat,
// r, rtilde; r = b - A*x, rtilde = r and the loop head finished(r) are
// EpiBiInit's.

// rho_1 = dot(rtilde, r); rho_1 == 0 is iter.fail(2) -- the error code
// is set and the loop goes on, as in MTL4; beta = rho_1/rho_2  (cgs.hpp:35-42)
template <typename VT>
struct EpiCgsRho {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    const VT *rt, *r;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT* part) const { part[0] += rt[i] * r[i]; }
    __device__ void finish(const VT* tot) const {
        const VT rho_1 = tot[0];
        c->rho_1 = (double)rho_1;
        if (rho_1 == VT(0)) c->error = 2;
        if (!c->first) c->beta = (double)(rho_1 / static_cast<VT>(c->rho_2));
    }
};

// p = u = r, or u = r + beta*q; p = u + beta*(q + beta*p)  (cgs.hpp:37-43)
template <typename VT>
struct EpiCgsUP {
    static constexpr int NP = 0;
    const VT *r, *q;
    VT *u, *p;
    Ctl* c;
    VT beta;
    int first;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        beta = static_cast<VT>(c->beta);
        first = c->first;
    }
    __device__ void operator()(int64_t i, VT*) const {
        if (first) {
            u[i] = r[i];
            p[i] = r[i];
        } else {
            const VT qi = q[i];
            const VT ui = r[i] + beta * qi;
            u[i] = ui;
            p[i] = ui + beta * (qi + beta * p[i]);
        }
    }
    __device__ void finish(const VT*) const {}
};

// vhat = A*phat; alpha = rho_1 / dot(rtilde, vhat)  (cgs.hpp:45-46; no
// breakdown test there: a zero dot yields a non-finite alpha)
template <typename VT>
struct EpiCgsV {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    struct Pre {
        VT rt;
    };
    const VT* rt;
    VT* v;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{rt[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        v[row] = s;
        part[0] += pre.rt * s;
    }
    __device__ void finish(const VT* tot) const {
        c->gamma = (double)tot[0];
        c->alpha = (double)(static_cast<VT>(c->rho_1) / tot[0]);
    }
};

// q = u - alpha*vhat; u += q  (cgs.hpp:47-49)
template <typename VT>
struct EpiCgsQU {
    static constexpr int NP = 0;
    const VT* v;
    VT *q, *u;
    Ctl* c;
    VT alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { alpha = static_cast<VT>(c->alpha); }
    __device__ void operator()(int64_t i, VT*) const {
        const VT ui = u[i];
        const VT qi = ui - alpha * v[i];
        q[i] = qi;
        u[i] = ui + qi;
    }
    __device__ void finish(const VT*) const {}
};

// x += alpha*uhat; qhat = A*uhat; r -= alpha*qhat; rho_2 = rho_1; loop
// head finished(r)  (cgs.hpp:52-57, 32)
template <typename VT>
struct EpiCgsXR {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    struct Pre {
        VT x, uh, r;
    };
    VT *x, *r;
    const VT* uh;
    Ctl* c;
    double* hist;
    VT alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { alpha = static_cast<VT>(c->alpha); }
    __device__ Pre preload(int row) const { return Pre{x[row], uh[row], r[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        x[row] = pre.x + alpha * pre.uh;
        const VT v = pre.r - alpha * s;
        r[row] = v;
        part[0] += v * v;
    }
    __device__ void finish(const VT* tot) const {
        c->rho_2 = c->rho_1;
        iter_finished(c, (double)std::sqrt(tot[0]), hist);
    }
};

// ---------------------------------------------------------------- TFQMR epilogues
// itl::tfqmr, itl/krylov/tfqmr.hpp:32-96.  Vectors rt, r, u1, u2, y1, y2,
// w, d, v, a temporary for solve(R, .) and the preconditioners' scratch.

// rt = b - A*solve(R, x); the check finished(rt) before the loop (:44-49)
template <typename VT>
struct EpiTfInit {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    struct Pre {
        VT b;
    };
    const VT* b;
    VT* rt;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{b[row]}; }
    __device__ void operator()(int row, VT s, const Pre& pre, VT* part) const {
        const VT v = pre.b - s;
        rt[row] = v;
        part[0] += v * v;
    }
    __device__ void finish(const VT* tot) const { iter_check(c, (double)std::sqrt(tot[0]), hist); }
};

// y = A*v, skipped once the solve is done
template <typename VT>
struct EpiTfAv {
    static constexpr int NP = 0;
    struct Pre {};
    VT* y;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int) const { return Pre{}; }
    __device__ void operator()(int row, VT s, const Pre&, VT*) const { y[row] = s; }
    __device__ void finish(const VT*) const {}
};

// a = src (and b = src when b is given)
template <typename VT>
struct EpiTfCopy {
    static constexpr int NP = 0;
    const VT* src;
    VT *a, *b;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT*) const {
        const VT v = src[i];
        a[i] = v;
        if (b) b[i] = v;
    }
    __device__ void finish(const VT*) const {}
};

// tau = two_norm(r); rho = tau*tau; loop head finished(tau)  (:51-55)
template <typename VT>
struct EpiTfTau {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    const VT* r;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT* part) const { part[0] += r[i] * r[i]; }
    __device__ void finish(const VT* tot) const {
        const VT tau = std::sqrt(tot[0]);
        c->tau = (double)tau;
        c->rho = (double)(tau * tau);
        c->looped = 1;
        iter_finished(c, (double)tau, hist);
    }
};

// sigma = dot(r, v); sigma == 0 is iter.fail(1) and returns; alpha  (:58-62)
template <typename VT>
struct EpiTfSigma {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    const VT *r, *v;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, VT* part) const { part[0] += r[i] * v[i]; }
    __device__ void finish(const VT* tot) const {
        c->gamma = (double)tot[0];
        if (tot[0] == VT(0)) { c->error = 1; c->done = 1; return; }
        c->alpha = (double)(static_cast<VT>(c->rho) / tot[0]);
    }
};

// w -= alpha*u; d = y + (theta*theta*eta/alpha)*d; theta = two_norm(w)/tau;
// c = 1/sqrt(1 + theta*theta); tau *= theta*c; eta = c*c*alpha  (:66-81)
template <typename VT>
struct EpiTfW {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_NORM4; }
    const VT *u, *y;
    VT *w, *d;
    Ctl* c;
    VT alpha, coef;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        alpha = static_cast<VT>(c->alpha);
        const VT theta = static_cast<VT>(c->theta);
        coef = theta * theta * static_cast<VT>(c->eta) / alpha;
    }
    __device__ void operator()(int64_t i, VT* part) const {
        const VT wi = w[i] - alpha * u[i];
        w[i] = wi;
        d[i] = y[i] + coef * d[i];
        part[0] += wi * wi;
    }
    __device__ void finish(const VT* tot) const {
        const VT tau = static_cast<VT>(c->tau);
        const VT theta = std::sqrt(tot[0]) / tau;
        const VT cc = VT(1) / std::sqrt(VT(1) + theta * theta);
        c->theta = (double)theta;
        c->tau = (double)(tau * (theta * cc));
        c->eta = (double)(cc * cc * static_cast<VT>(c->alpha));
    }
};

// x += eta*d (end of the first inner step); y2 = y1 - alpha*v  (:71, :68)
template <typename VT>
struct EpiTfXY2 {
    static constexpr int NP = 0;
    VT *x, *y2;
    const VT *d, *y1, *v;
    Ctl* c;
    VT eta, alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        eta = static_cast<VT>(c->eta);
        alpha = static_cast<VT>(c->alpha);
    }
    __device__ void operator()(int64_t i, VT*) const {
        x[i] = x[i] + eta * d[i];
        y2[i] = y1[i] - alpha * v[i];
    }
    __device__ void finish(const VT*) const {}
};

// x += eta*d (second inner step); rho == 0 is iter.fail(1); rhon =
// dot(r, w); beta = rhon/rho; rho = rhon; loop head finished(tau)  (:81-87, :57)
template <typename VT>
struct EpiTfRho {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int k) { return RO_DOT4; }
    VT* x;
    const VT *d, *r, *w;
    Ctl* c;
    double* hist;
    VT eta;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { eta = static_cast<VT>(c->eta); }
    __device__ void operator()(int64_t i, VT* part) const {
        x[i] = x[i] + eta * d[i];
        part[0] += r[i] * w[i];
    }
    __device__ void finish(const VT* tot) const {
        const VT rho = static_cast<VT>(c->rho);
        if (rho == VT(0)) { c->error = 1; c->done = 1; return; }
        c->beta = (double)(tot[0] / rho);
        c->rho = (double)tot[0];
        iter_finished(c, c->tau, hist);
    }
};

// y1 = w + beta*y2  (:89)
template <typename VT>
struct EpiTfY1 {
    static constexpr int NP = 0;
    VT* y1;
    const VT *w, *y2;
    Ctl* c;
    VT beta;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { beta = static_cast<VT>(c->beta); }
    __device__ void operator()(int64_t i, VT*) const { y1[i] = w[i] + beta * y2[i]; }
    __device__ void finish(const VT*) const {}
};

// v = u1 + beta*(u2 + beta*v)  (:92)
template <typename VT>
struct EpiTfV {
    static constexpr int NP = 0;
    VT* v;
    const VT *u1, *u2;
    Ctl* c;
    VT beta;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { beta = static_cast<VT>(c->beta); }
    __device__ void operator()(int64_t i, VT*) const { v[i] = u1[i] + beta * (u2[i] + beta * v[i]); }
    __device__ void finish(const VT*) const {}
};

// ---------------------------------------------------------------- BiCGSTAB(2) epilogues
// itl::bicgstab_2 with l = 2, itl/krylov/bicgstab_2.hpp:37-130.  Vectors
// r_hat[0..2], u_hat[0..2], r0_tilde, y, x0 and the preconditioner's
// scratch.  Every reduction is fused into the pass that produces (or last
// reads) its operands; with an identity preconditioner the dots on
// u_hat[j+1] = A*u_hat[j] and r_hat[j+1] = A*r_hat[j] run in the product's
// epilogue, otherwise in a pass after solve(L, y).

// two_norm(x) != 0 decides the shift of the problem  (:52-59)
struct EpiB2XNorm {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    const double* x;
    Ctl* c;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const { part[0] += x[i] * x[i]; }
    __device__ void finish(const double* tot) const { c->shift = std::sqrt(tot[0]) != 0.0; }
};

// r_hat[0] = b (-= A*x and x0 = x when shifted); two_norm(r_hat[0])  (:53-61)
struct EpiB2Init {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    struct Pre {
        double b, x;
    };
    const double *b, *x;
    double *r0, *x0;
    Ctl* c;
    int shift;
    __device__ bool skip() const { return false; }
    __device__ void prepare() { shift = c->shift; }
    __device__ Pre preload(int row) const { return Pre{b[row], x[row]}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        const double r = shift ? pre.b - s : pre.b;
        r0[row] = r;
        x0[row] = shift ? pre.x : 0.0;
        part[0] += r * r;
    }
    __device__ void finish(const double* tot) const { c->nr = std::sqrt(tot[0]); }
};

// x = 0 when shifted (after the product above has read it); r0_tilde =
// r_hat[0] / two_norm(r_hat[0]); u_hat[0] = 0  (:58, :61, :64)
struct EpiB2Start {
    static constexpr int NP = 0;
    double *x, *rt, *u0;
    const double* r0;
    Ctl* c;
    int shift;
    double nr;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {
        shift = c->shift;
        nr = c->nr;
    }
    __device__ void operator()(int64_t i, double*) const {
        if (shift) x[i] = 0.0;
        rt[i] = r0[i] / nr;
        u0[i] = 0.0;
    }
    __device__ void finish(const double*) const {}
};

// Loop head: finished(r_hat[0]) and ++iter; rho_0 = -omega*rho_0; then the
// first inner step's rho_1 = dot(r0_tilde, r_hat[0]); beta = alpha*rho_1/rho_0;
// rho_0 = rho_1  (:70-76)
__device__ __forceinline__ void b2_head(Ctl* c, const double* tot, double* hist) {
    iter_finished(c, std::sqrt(tot[0]), hist);
    if (c->done) return;
    c->rho = -c->omega * c->rho;
    c->rho_1 = tot[1];
    c->beta = c->alpha * tot[1] / c->rho;
    c->rho = tot[1];
}

// UPD: the minimal-residual updates of x, r_hat[0] and u_hat[0] (:119-126)
// in their order per element, then the loop head on the new r_hat[0]
template <bool UPD>
struct EpiB2Head {
    static constexpr int NP = 2;
    __host__ __device__ static constexpr int order(int k) { return k == 0 ? RO_NORM4 : RO_DOT4; }
    double *x, *r0, *u0;
    const double *r1, *r2, *u1, *u2, *rt;
    Ctl* c;
    double* hist;
    double g1, g2, ga1, ga2, gaa1;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        if (UPD) {
            g1 = c->g1;
            g2 = c->g2;
            ga1 = c->ga1;
            ga2 = c->ga2;
            gaa1 = c->gaa1;
        }
    }
    __device__ void operator()(int64_t i, double* part) const {
        double r = r0[i];
        if (UPD) {
            const double a = r1[i];
            x[i] = (x[i] + g1 * r) + gaa1 * a;
            r = (r - ga2 * r2[i]) - ga1 * a;
            r0[i] = r;
            u0[i] = (u0[i] - g2 * u2[i]) - g1 * u1[i];
        }
        part[0] += r * r;
        part[1] += rt[i] * r;
    }
    __device__ void finish(const double* tot) const { b2_head(c, tot, hist); }
};

// u_hat[i] = r_hat[i] - beta*u_hat[i] for i <= j  (:78-79)
struct EpiB2U {
    static constexpr int NP = 0;
    const double *r0, *r1;
    double *u0, *u1;   // u1 == nullptr: j = 0
    Ctl* c;
    double beta;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { beta = c->beta; }
    __device__ void operator()(int64_t i, double*) const {
        u0[i] = r0[i] - beta * u0[i];
        if (u1) u1[i] = r1[i] - beta * u1[i];
    }
    __device__ void finish(const double*) const {}
};

// Reductions that finish with a product (or with the pass after solve(L, .)):
//   GAMMA: Gamma = dot(r0_tilde, u_hat[j+1]); alpha = rho_0/Gamma  (:83-84)
//   RHO:   the next inner step's rho_1 = dot(r0_tilde, r_hat[1]); beta;
//          rho_0 = rho_1  (:75-76)
//   TAU:   tau[1][1] = dot(r_hat[1], r_hat[1]); tau[1][2] =
//          dot(r_hat[1], r_hat[2]) / tau[1][1]  (orth.hpp:75, :87-89)
enum { B2_NONE = 0, B2_GAMMA = 1, B2_RHO = 2, B2_TAU = 3 };
__host__ __device__ constexpr int b2_np(int mode) { return mode == B2_NONE ? 0 : mode == B2_TAU ? 2 : 1; }

__device__ __forceinline__ void b2_finish(int mode, Ctl* c, const double* tot) {
    if (mode == B2_GAMMA) {
        c->gamma = tot[0];
        c->alpha = c->rho / tot[0];
    } else if (mode == B2_RHO) {
        c->rho_1 = tot[0];
        c->beta = c->alpha * tot[0] / c->rho;
        c->rho = tot[0];
    } else if (mode == B2_TAU) {
        c->t11 = tot[0];
        c->t12 = tot[1] / tot[0];
    }
}

// out = A*in; XUP: x += alpha*u_hat[0] (:96, rows of x are not gathered);
// a = r0_tilde (GAMMA, RHO) or r_hat[1] (TAU)
template <int MODE, bool XUP>
struct EpiB2Mv {
    static constexpr int NP = b2_np(MODE);
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    struct Pre {
        double x, u, a;
    };
    const double* a;
    double* out;
    double* x;
    const double* u0;
    Ctl* c;
    double alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { alpha = c->alpha; }
    __device__ Pre preload(int row) const {
        return Pre{XUP ? x[row] : 0.0, XUP ? u0[row] : 0.0, NP ? a[row] : 0.0};
    }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        out[row] = s;
        if (XUP) x[row] = pre.x + alpha * pre.u;
        if (MODE == B2_TAU) {
            part[0] += pre.a * pre.a;
            part[1] += pre.a * s;
        } else if (NP) {
            part[0] += pre.a * s;
        }
    }
    __device__ void finish(const double* tot) const { b2_finish(MODE, c, tot); }
};

// The same reductions over v = solve(L, y)
template <int MODE>
struct EpiB2Dot {
    static constexpr int NP = b2_np(MODE);
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    const double *a, *v;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const {
        const double ai = a[i], vi = v[i];
        if (MODE == B2_TAU) {
            part[0] += ai * ai;
            part[1] += ai * vi;
        } else {
            part[0] += ai * vi;
        }
    }
    __device__ void finish(const double* tot) const { b2_finish(MODE, c, tot); }
};

// r_hat[i] -= alpha*u_hat[i+1] for i <= j; finished(r_hat[j]) without
// ++iter (:86-92); r1 == nullptr: j = 0
struct EpiB2R {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    double *r0, *r1;
    const double *u1, *u2;
    Ctl* c;
    double* hist;
    double alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { alpha = c->alpha; }
    __device__ void operator()(int64_t i, double* part) const {
        const double v0 = r0[i] - alpha * u1[i];
        r0[i] = v0;
        if (r1) {
            const double v1 = r1[i] - alpha * u2[i];
            r1[i] = v1;
            part[0] += v1 * v1;
        } else {
            part[0] += v0 * v0;
        }
    }
    __device__ void finish(const double* tot) const { iter_check(c, std::sqrt(tot[0]), hist); }
};

// Minimal-residual part: r_hat[2] -= tau[1][2]*r_hat[1]; tau[2][2] =
// dot(r_hat[2], r_hat[2]) (orth.hpp:87-93); gamma_a[j] = dot(r_hat[j],
// r_hat[0]) / tau[j][j]; omega = gamma[2] = gamma_a[2]; omega == 0 is
// iter.fail(3) (returned before the shift back); gamma[1] = gamma_a[1] -
// tau[1][2]*gamma[2]; gamma_aa[1] = (0 + 0*gamma[2]) + gamma[2]: the naive
// dense mat-vec over the strictly upper view, which reads 0 off the band
// (mat_vec_mult.hpp:259-298)  (bicgstab_2.hpp:100-117)
struct EpiB2MR {
    static constexpr int NP = 3;
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    double* r2;
    const double *r1, *r0;
    Ctl* c;
    double t12;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { t12 = c->t12; }
    __device__ void operator()(int64_t i, double* part) const {
        const double a = r1[i];
        const double v = r2[i] - t12 * a;
        r2[i] = v;
        const double z = r0[i];
        part[0] += v * v;
        part[1] += a * z;
        part[2] += v * z;
    }
    __device__ void finish(const double* tot) const {
        c->t22 = tot[0];
        c->ga1 = tot[1] / c->t11;
        c->ga2 = tot[2] / tot[0];
        c->g2 = c->ga2;
        c->omega = c->g2;
        if (c->omega == 0.0) {
            c->error = 3;
            c->done = 1;
            return;
        }
        c->g1 = c->ga1 - (0.0 + c->t12 * c->g2);
        c->gaa1 = (0.0 + 0.0 * c->g2) + c->g2;
    }
};

// x += x0: undo the shift (:90, :128)
struct EpiB2Shift {
    static constexpr int NP = 0;
    double* x;
    const double* x0;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double*) const { x[i] = x[i] + x0[i]; }
    __device__ void finish(const double*) const {}
};

// ---------------------------------------------------------------- BiCGSTAB(l) epilogues
// itl::bicgstab_ell, itl/krylov/bicgstab_ell.hpp:37-134: BiCGSTAB(2)'s
// steps for any l <= kMaxEll, with the scalars of the minimal-residual part
// (tau, gamma, gamma_a, gamma_aa) in a small device array.
constexpr int kMaxEll = 8;
struct EllVecs {
    double* r[kMaxEll + 1];
    double* u[kMaxEll + 1];
};
struct EllScalars {   // offsets into the small array: tau[(l+1)^2], gamma, gamma_a, gamma_aa
    double* sm;
    int l;
    __host__ __device__ double& tau(int i, int j) const { return sm[i * (l + 1) + j]; }
    __host__ __device__ double* gam() const { return sm + (l + 1) * (l + 1); }
    __host__ __device__ double* ga() const { return gam() + (l + 1); }
    __host__ __device__ double* gaa() const { return ga() + (l + 1); }
};

// UPD: the minimal-residual updates of x, r_hat[0], u_hat[0] (:119-131) in
// their order per element, then the loop head on the new r_hat[0]
template <bool UPD>
struct EpiEllHead {
    static constexpr int NP = 2;
    __host__ __device__ static constexpr int order(int k) { return k == 0 ? RO_NORM4 : RO_DOT4; }
    EllVecs V;
    double* x;
    const double* rt;
    EllScalars E;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const {
        double r0 = V.r[0][i];
        if (UPD) {
            const int l = E.l;
            const double *gam = E.gam(), *ga = E.ga(), *gaa = E.gaa();
            double xi = x[i] + gam[1] * r0;
            double rn = r0 - ga[l] * V.r[l][i];
            double ui = V.u[0][i] - gam[l] * V.u[l][i];
            for (int j = 1; j < l; j++) {
                ui = ui - gam[j] * V.u[j][i];
                const double rj = V.r[j][i];
                xi = xi + gaa[j] * rj;
                rn = rn - ga[j] * rj;
            }
            x[i] = xi;
            V.r[0][i] = rn;
            V.u[0][i] = ui;
            r0 = rn;
        }
        part[0] += r0 * r0;
        part[1] += rt[i] * r0;
    }
    __device__ void finish(const double* tot) const { b2_head(c, tot, hist); }
};

// u_hat[k] = r_hat[k] - beta*u_hat[k] for k <= j  (:78-79)
struct EpiEllU {
    static constexpr int NP = 0;
    EllVecs V;
    int j;
    Ctl* c;
    double beta;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { beta = c->beta; }
    __device__ void operator()(int64_t i, double*) const {
        for (int k = 0; k <= j; k++) V.u[k][i] = V.r[k][i] - beta * V.u[k][i];
    }
    __device__ void finish(const double*) const {}
};

// r_hat[k] -= alpha*u_hat[k+1] for k <= j; finished(r_hat[j]) without
// ++iter  (:86-92)
struct EpiEllR {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    EllVecs V;
    int j;
    Ctl* c;
    double* hist;
    double alpha;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { alpha = c->alpha; }
    __device__ void operator()(int64_t i, double* part) const {
        double v = 0.0;
        for (int k = 0; k <= j; k++) {
            v = V.r[k][i] - alpha * V.u[k + 1][i];
            V.r[k][i] = v;
        }
        part[0] += v * v;
    }
    __device__ void finish(const double* tot) const { iter_check(c, std::sqrt(tot[0]), hist); }
};

// One step (i, j) of orthogonalize_factors on r_hat[1..l] (orth.hpp:75-93),
// fused with the axpy of the step before: r_hat[j] -= tau[i-1][j] *
// r_hat[i-1] (i > 1), then tau[i][j] = dot(r_hat[i], r_hat[j]) / tau[i][i]
// (i < j), or tau[j][j] = dot(r_hat[j], r_hat[j]) and gamma_a[j] =
// dot(r_hat[j], r_hat[0]) / tau[j][j] (i == j; bicgstab_ell.hpp:103-104).
// After the last step (l, l): gamma, omega and gamma_aa (:106-117); omega
// == 0 is iter.fail(3).
struct EpiEllOrth {
    static constexpr int NP = 2;
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    EllVecs V;
    EllScalars E;
    int i, j;
    Ctl* c;
    double t;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { t = i > 1 ? E.tau(i - 1, j) : 0.0; }
    __device__ void operator()(int64_t e, double* part) const {
        double rj = V.r[j][e];
        if (i > 1) {
            rj = rj - t * V.r[i - 1][e];
            V.r[j][e] = rj;
        }
        if (i == j) {
            part[0] += rj * rj;
            part[1] += rj * V.r[0][e];
        } else {
            part[0] += V.r[i][e] * rj;
        }
    }
    __device__ void finish(const double* tot) const {
        if (i < j) {
            E.tau(i, j) = tot[0] / E.tau(i, i);
            return;
        }
        E.tau(j, j) = tot[0];
        E.ga()[j] = tot[1] / tot[0];
        if (j < E.l) return;
        const int l = E.l;
        double *gam = E.gam(), *ga = E.ga(), *gaa = E.gaa();
        gam[l] = ga[l];
        c->omega = gam[l];
        if (c->omega == 0.0) {
            c->error = 3;
            c->done = 1;
            return;
        }
        for (int jj = l - 1; jj > 0; --jj) {
            double sum = 0.0;
            for (int ii = jj + 1; ii <= l; ++ii) sum += E.tau(jj, ii) * gam[ii];
            gam[jj] = ga[jj] - sum;
        }
        for (int p = 0; p < l - 1; p++) {   // naive dense mat-vec over strict_upper (mat_vec_mult.hpp:259-298)
            double s = 0.0;
            for (int q = 0; q < l - 1; q++) s += (q > p ? E.tau(p + 1, q + 1) : 0.0) * gam[q + 2];
            gaa[p + 1] = s + gam[p + 2];
        }
    }
};

// x = solve(R, x) + x0 (src: the solve's output, or x itself)  (:90-91, :133)
struct EpiEllShift {
    static constexpr int NP = 0;
    double* x;
    const double *src, *x0;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double*) const { x[i] = src[i] + x0[i]; }
    __device__ void finish(const double*) const {}
};

// ---------------------------------------------------------------- GMRES epilogues
// itl::gmres_full / itl::gmres, itl/krylov/gmres.hpp:37-137.  The Krylov
// basis V (kmax + 1 vectors) and the vectors r0, r, va0 live in the
// workspace; H ((kmax+1) x kmax, row-major), g, the rotations, y and the
// reorthogonalization coefficients live in a small device array.
constexpr int kMaxGmres = 256;   // Krylov vectors per cycle
struct GmScalars {
    double* sm;
    int kmax;
    __host__ __device__ double& H(int i, int j) const { return sm[(size_t)i * kmax + j]; }
    __host__ __device__ double* g() const { return sm + (size_t)(kmax + 1) * kmax; }
    __host__ __device__ double* cs() const { return g() + (kmax + 1); }
    __host__ __device__ double* sn() const { return cs() + (kmax + 1); }
    __host__ __device__ double* y() const { return sn() + (kmax + 1); }
    __host__ __device__ double* z() const { return y() + (kmax + 1); }
    __host__ __device__ double* hr() const { return z() + (kmax + 1); }
    __host__ __device__ double* perm() const { return hr() + (kmax + 1); }   // pivots (small integers)
    __host__ __device__ double& m() const { return perm()[kmax + 1]; }      // columns solved for
    static size_t words(int kmax) { return (size_t)(kmax + 1) * kmax + 7 * (size_t)(kmax + 1) + 1; }
};

// r0 = b - A*x (compute_summand: the product first, then b - it); with
// NORM: two_norm of it handed to finished() -- the terminate(r) that ends
// a cycle (gmres.hpp:117-118); the next cycle's r0 is the same vector
template <bool NORM>
struct EpiGmRes {
    static constexpr int NP = NORM ? 1 : 0;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    struct Pre {
        double b;
    };
    const double* b;
    double* r0;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{b[row]}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        const double v = pre.b - s;
        r0[row] = v;
        if (NORM) part[0] += v * v;
    }
    __device__ void finish(const double* tot) const {
        if (NORM) iter_check(c, std::sqrt(tot[0]), hist);
    }
};

// rho = g[0] = two_norm(r); finished(rho) at the start of a cycle (:56-58)
struct EpiGmRho {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_NORM4; }
    const double* r;
    GmScalars G;
    Ctl* c;
    double* hist;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const { part[0] += r[i] * r[i]; }
    __device__ void finish(const double* tot) const {
        const double rho = std::sqrt(tot[0]);
        G.g()[0] = rho;
        c->gamma = rho;
        iter_check(c, rho, hist);
    }
};

// V.vector(0) = r / rho  (:59)
struct EpiGmV0 {
    static constexpr int NP = 0;
    const double* r;
    double* v0;
    Ctl* c;
    double rho;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { rho = c->gamma; }
    __device__ void operator()(int64_t i, double*) const { v0[i] = r[i] / rho; }
    __device__ void finish(const double*) const {}
};

// V_(k+1) = A * in (L = identity), with DOT: H[0][k] = dot(V_0, V_(k+1))
template <bool DOT>
struct EpiGmMv {
    static constexpr int NP = DOT ? 1 : 0;
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    struct Pre {
        double v0;
    };
    const double* v0;
    double* out;
    GmScalars G;
    int k;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{DOT ? v0[row] : 0.0}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        out[row] = s;
        if (DOT) part[0] += pre.v0 * s;
    }
    __device__ void finish(const double* tot) const {
        if (DOT) G.H(0, k) = tot[0];
    }
};

// One pass of the orthogonalization of w = V_(k+1) (:68-79), fused with the
// axpy of the pass before.  STEP:
//   0: H[0][k] = dot(V_0, w)                               (no axpy; L path)
//   1: w -= H[j-1][k] V_(j-1); H[j][k] = dot(V_j, w)       (MGS, 1 <= j <= k)
//   2: w -= H[k][k] V_k; H[k+1][k] = two_norm(w); hr[0] = dot(w, V_0);
//      H[0][k] += hr[0]                                    (end of MGS, reorth j = 0)
//   3: w -= hr[j-1] V_(j-1); hr[j] = dot(w, V_j); H[j][k] += hr[j]  (reorth, 1 <= j <= k)
//   4: w -= hr[k] V_k; H[k+1][k] = two_norm(w)
template <int STEP>
struct EpiGmOrth {
    static constexpr int NP = STEP == 2 ? 2 : 1;
    __host__ __device__ static constexpr int order(int q) {
        return (STEP == 4 || (STEP == 2 && q == 0)) ? RO_NORM4 : RO_DOT4;
    }
    double* w;
    const double *va, *vb;   // va: V_(j-1) or V_k for the axpy; vb: V_j for the dot
    GmScalars G;
    int j, k;
    Ctl* c;
    double t;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        t = STEP == 1 ? G.H(j - 1, k) : STEP == 2 ? G.H(k, k) : STEP == 3 ? G.hr()[j - 1] : STEP == 4 ? G.hr()[k] : 0.0;
    }
    __device__ void operator()(int64_t i, double* part) const {
        double v = w[i];
        if (STEP != 0) {
            v = v - t * va[i];
            w[i] = v;
        }
        if (STEP == 0 || STEP == 1) part[0] += vb[i] * v;
        else if (STEP == 2) {
            part[0] += v * v;
            part[1] += v * vb[i];
        } else if (STEP == 3) part[0] += v * vb[i];
        else part[0] += v * v;
    }
    __device__ void finish(const double* tot) const {
        if (STEP == 0 || STEP == 1) G.H(j, k) = tot[0];
        else if (STEP == 2) {
            G.H(k + 1, k) = std::sqrt(tot[0]);
            G.hr()[0] = tot[1];
            G.H(0, k) += tot[1];
        } else if (STEP == 3) {
            G.hr()[j] = tot[0];
            G.H(j, k) += tot[0];
        } else {
            G.H(k + 1, k) = std::sqrt(tot[0]);
        }
    }
};

// V_(k+1) *= 1. / H[k+1][k] unless H[k+1][k] == 0  (:80-81)
struct EpiGmScale {
    static constexpr int NP = 0;
    double* w;
    GmScalars G;
    int k;
    Ctl* c;
    double f;
    int on;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {
        const double h = G.H(k + 1, k);
        on = h != 0.0;
        f = on ? 1. / h : 0.0;
    }
    __device__ void operator()(int64_t i, double*) const {
        if (on) w[i] = w[i] * f;
    }
    __device__ void finish(const double*) const {}
};

// Step k's rotations (:84-94) and ++iter: for i < k the rotation of
// (H[i][k-1], H[i+1][k-1]) applied as operation/givens.hpp does -- rows
// i, i+1 of H <- trans(G) * rows, then columns i, i+1 <- columns * G, each
// 2-term product summed from zero -- then the new rotation (c_k, s_k) of
// column k and of g.  One workgroup: a rotation's row and column updates
// run across the lanes, rotations one after the other.
__global__ __launch_bounds__(256) void k_gm_givens(GmScalars G, int k, Ctl* c) {
    if (c->done) return;
    const int kmax = G.kmax;
    for (int i = 0; i < k; i++) {
        const double a = G.H(i, k - 1), b = G.H(i + 1, k - 1);
        double cc, ss, t;
        if (b == 0.0) { cc = 1.0; ss = 0.0; }
        else if (fabs(b) > fabs(a)) { t = -a / b; ss = 1.0 / sqrt(1.0 + t * t); cc = ss * t; }
        else { t = -b / a; cc = 1.0 / sqrt(1.0 + t * t); ss = cc * t; }
        const double ms = -ss;
        __syncthreads();   // every lane has read a and b
        for (int jj = threadIdx.x; jj < kmax; jj += blockDim.x) {
            const double b0 = G.H(i, jj), b1 = G.H(i + 1, jj);
            G.H(i, jj) = (0.0 + cc * b0) + ms * b1;
            G.H(i + 1, jj) = (0.0 + ss * b0) + cc * b1;
        }
        __syncthreads();
        for (int r = threadIdx.x; r <= kmax; r += blockDim.x) {
            const double m0 = G.H(r, i), m1 = G.H(r, i + 1);
            G.H(r, i) = (0.0 + m0 * cc) + m1 * ms;
            G.H(r, i + 1) = (0.0 + m0 * ss) + m1 * cc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double hkk = G.H(k, k), hk1 = G.H(k + 1, k);
        const double nu = sqrt(hkk * hkk + hk1 * hk1);
        if (nu != 0.0) {
            double* cs = G.cs();
            double* sn = G.sn();
            double* g = G.g();
            cs[k] = hkk / nu;
            sn[k] = -hk1 / nu;
            G.H(k, k) = cs[k] * hkk - sn[k] * hk1;
            G.H(k + 1, k) = 0;
            const double w1 = cs[k] * g[k] - sn[k] * g[k + 1], w2 = sn[k] * g[k] + cs[k] * g[k + 1];
            g[k] = w1;
            g[k + 1] = w2;
        }
        c->i++;
        c->first = c->i <= 1;
    }
}

// After the cycle (:98-114): drop trailing columns with |g[k-1]| <= atol,
// then lu_solve(H[range][range], g[range]) shrinking the range while a
// pivot is zero.  The LU runs in place on H, as in MTL4 (the sub-matrix is
// a view and the copy of a view shares its memory): partial pivoting with
// the first row of largest magnitude, rows swapped within the block;
// permute(P, g) scatters (g'[P[i]] = g[i]); then the unit lower and the
// upper sweep (division by the diagonal).  No direction: iter.fail(2).
__global__ __launch_bounds__(256) void k_gm_lsq(GmScalars G, int kmax_run, double atol, Ctl* c) {
    if (c->done) return;
    __shared__ int s_m, s_rmax, s_sing;
    if (threadIdx.x == 0) {
        int k = kmax_run;
        while (k > 0 && fabs(G.g()[k - 1]) <= atol) k--;
        s_m = k;
    }
    __syncthreads();
    double* P = G.perm();
    for (int m = s_m; m > 0; m--) {
        for (int i = threadIdx.x; i < m; i += blockDim.x) P[i] = i;
        __syncthreads();
        if (threadIdx.x == 0) s_sing = 0;
        for (int i = 0; i + 1 < m; i++) {
            if (threadIdx.x == 0) {
                int rmax = i;
                double best = fabs(G.H(i, i));
                for (int r = i + 1; r < m; r++)
                    if (fabs(G.H(r, i)) > best) { best = fabs(G.H(r, i)); rmax = r; }
                s_rmax = rmax;
                if (rmax != i) {
                    const double t = P[i];
                    P[i] = P[rmax];
                    P[rmax] = t;
                }
            }
            __syncthreads();
            const int rmax = s_rmax;
            if (rmax != i)
                for (int cc = threadIdx.x; cc < m; cc += blockDim.x) {
                    const double t = G.H(i, cc);
                    G.H(i, cc) = G.H(rmax, cc);
                    G.H(rmax, cc) = t;
                }
            __syncthreads();
            if (fabs(G.H(i, i)) <= 0.0) {   // every lane reads the same pivot
                if (threadIdx.x == 0) s_sing = 1;
                break;
            }
            const double piv = G.H(i, i);
            for (int r = i + 1 + threadIdx.x; r < m; r += blockDim.x) {
                const double l = G.H(r, i) / piv;
                G.H(r, i) = l;
                for (int cc = i + 1; cc < m; cc++) G.H(r, cc) -= l * G.H(i, cc);
            }
            __syncthreads();
        }
        __syncthreads();
        if (s_sing) continue;   // the next (smaller) block starts from H as left here
        if (threadIdx.x == 0) {
            double *y = G.y(), *z = G.z();
            const double* g = G.g();
            for (int i = 0; i < m; i++) y[(int)P[i]] = g[i];
            for (int r = 0; r < m; r++) {
                double rr = y[r];
                for (int cc = 0; cc < r; cc++) rr -= G.H(r, cc) * z[cc];
                z[r] = rr;
            }
            for (int r = m; r-- > 0;) {
                double rr = z[r];
                for (int cc = r + 1; cc < m; cc++) rr -= G.H(r, cc) * y[cc];
                y[r] = rr / G.H(r, r);
            }
            G.m() = m;
        }
        return;
    }
    if (threadIdx.x == 0) {   // "GMRES did not find any direction to correct x"
        G.m() = 0;
        c->error = 2;
        c->done = 1;
    }
}

// out = V.vector(range) * y[range]: zero, then += V_j * y_j column by
// column (mat_vec_mult.hpp:352-360)
struct EpiGmComb {
    static constexpr int NP = 0;
    const double* v;   // V_0; V_j = v + j*n
    int64_t n;
    double* out;
    GmScalars G;
    Ctl* c;
    int m;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() { m = (int)G.m(); }
    __device__ void operator()(int64_t i, double*) const {
        double t = 0.0;
        const double* y = G.y();
        for (int j = 0; j < m; j++) t += v[(size_t)j * n + i] * y[j];
        out[i] = t;
    }
    __device__ void finish(const double*) const {}
};

// x += solve(R, V*y)  (:115)
struct EpiGmX {
    static constexpr int NP = 0;
    double* x;
    const double* t;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double*) const { x[i] = x[i] + t[i]; }
    __device__ void finish(const double*) const {}
};

__global__ void k_gm_zero(double* p, size_t words, const Ctl* c) {
    if (c->done) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---------------------------------------------------------------- FP64 refresh
// x += double(x32); x32 = 0
struct EpiRu1 {
    static constexpr int NP = 0;
    double* x;
    float* x32;
    Ctl* c;
    int drop;
    __device__ bool skip() const { return c->done != 0 || c->want_ru == 0; }
    __device__ void prepare() { drop = c->drop; }
    __device__ void operator()(int64_t i, double*) const {
        if (!drop) x[i] = x[i] + (double)x32[i];
        x32[i] = 0.0f;
    }
    __device__ void finish(const double*) const {}
};

// r = b - A*x in FP64; r32 = float(r); ||r||^2 and, for CG with a fused
// preconditioner, rho = r32.(d .* r32) of the replaced residual
template <typename DT, int PCM>
struct EpiRu2 {
    static constexpr int NP = 2;
    struct Pre {
        double b;
        float d;
    };
    const double* b;
    float* r32;
    const DT* d;
    Ctl* c;
    int cg;
    __device__ bool skip() const { return c->done != 0 || c->want_ru == 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const {
        return Pre{b[row], PCM == PCM_DIAG ? (float)d[row] : 0.0f};
    }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        const double v = pre.b - s;
        const float f = (float)v;
        r32[row] = f;
        part[0] += v * v;
        if (PCM == PCM_DIAG) part[1] += (double)(f * (pre.d * f));
        else if (PCM == PCM_NONE) part[1] += (double)(f * f);
    }
    __device__ void finish(const double* tot) const {
        ru_finish(c, tot[0], tot[1], cg && PCM != PCM_GENERAL);
    }
};

// FP64 vectors on low-precision values (BiCGSTAB): r = b - A*x with the FP64
// values replaces the recurrence's r and the recurrence restarts from it,
// rtilde = r and p = r (bicgstab.hpp:37-38, 48-49): a replaced residual
// that the directions p, v no longer match makes BiCGSTAB erratic (a
// diverging solve was seen on a 3D Laplacian with exactly representable
// values); ||r||^2 decides in ru_finish
struct EpiRu64 {
    static constexpr int NP = 1;
    struct Pre {
        double b;
    };
    const double* b;
    double *r, *rt;
    Ctl* c;
    __device__ bool skip() const { return c->done != 0 || c->want_ru == 0; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{b[row]}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        const double v = pre.b - s;
        r[row] = v;
        rt[row] = v;
        part[0] += v * v;
    }
    __device__ void finish(const double* tot) const {
        ru_finish(c, tot[0], 0.0, false);
        if (!c->done) c->first = 1;
    }
};

__global__ void k_ctl_init(Ctl* c, Ctl init) { *c = init; }

// zero `words` 32-bit words (a kernel, not a memset: stream-ordered like
// the solve's own launches and safe inside a captured graph)
__global__ void k_zero_words(unsigned* p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

static int zero_on(void* p, size_t bytes, hipStream_t s) {
    const size_t words = bytes / 4;
    const int grid = (int)std::min<size_t>(4096, std::max<size_t>(1, (words + 255) / 256));
    hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, s, static_cast<unsigned*>(p), words);
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

// ---------------------------------------------------------------- host driver
// Buffers are (re)allocated here and cleared on the solve's stream s, so the
// clearing is ordered before the solve's kernels whatever other streams do.
template <typename VT>
static int ensure_workspace(Workspace<VT>& W, int32_t n, int nvec, const mtl4g_pc_s* P,
                            int32_t hist_cap, hipStream_t s) {
    static_assert(sizeof(VT) % 4 == 0 && sizeof(Ctl) % 4 == 0, "zeroed as 32-bit words");
    if (W.n != n || W.vecs.n < (size_t)n * nvec) {
        M4_TRY(W.vecs.alloc((size_t)std::max(n, 1) * nvec));
        M4_TRY(zero_on(W.vecs.p, W.vecs.n * sizeof(VT), s));
        W.n = n;
    }
    if (!W.partials.p) M4_TRY(W.partials.alloc(kPartialWords));
    if (!W.tickets.p) {
        M4_TRY(W.tickets.alloc(kTicketWords));
        M4_TRY(zero_on(W.tickets.p, kTicketWords * sizeof(unsigned), s));
    }
    if (!W.ctl.p) {
        M4_TRY(W.ctl.alloc(1));
        M4_TRY(zero_on(W.ctl.p, sizeof(Ctl), s));
    }
    if (!W.h_ctl) M4_HIP(hipHostMalloc((void**)&W.h_ctl, sizeof(Ctl), hipHostMallocDefault));
    if (hist_cap > 0 && W.hist.n < (size_t)hist_cap) M4_TRY(W.hist.alloc(hist_cap));
    if (!W.hist.p) M4_TRY(W.hist.alloc(1));
    if (sizeof(VT) == 8 && reduction_order() == MTL4G_ORDER_MTL4 && W.terms.n < kMaxNP * (size_t)n)
        M4_TRY(W.terms.alloc(kMaxNP * (size_t)n));
    const bool sweeps = P && (P->kind == MTL4G_PC_ILU0 || P->kind == MTL4G_PC_IC0);
    if (sweeps && sizeof(VT) == 4 && W.pc64.n < 2 * (size_t)n) M4_TRY(W.pc64.alloc(2 * (size_t)std::max(n, 1)));
    size_t words = sweeps ? pc_counter_words(P) : 1;
    if (W.pc_counters.n < words) {
        M4_TRY(W.pc_counters.alloc(words));
        M4_TRY(zero_on(W.pc_counters.p, words * sizeof(unsigned), s));
    }
    return MTL4G_OK;
}

// Events around the products of CG iterations (mtl4g_set_product_timing);
// created on first use and kept for the process.
struct ProductTimer {
    static constexpr size_t kMax = 65536;
    bool on = false;
    std::vector<hipEvent_t> ev;   // 2 per product, created on first use
    size_t used = 0;              // pairs recorded in the current solve
    double total_us = 0;
    int64_t count = 0;
};
static ProductTimer g_timer;

static void timer_mark(hipStream_t s, bool begin) {
    ProductTimer& T = g_timer;
    if (!T.on || T.used >= ProductTimer::kMax) return;
    const size_t k = 2 * T.used + (begin ? 0 : 1);
    while (T.ev.size() <= k) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        T.ev.push_back(e);
    }
    (void)hipEventRecord(T.ev[k], s);
    if (!begin) T.used++;
}

// After a solve of `iters` iterations has completed on its stream: add the
// first iters products (later launches exited at entry).
static void timer_collect(int32_t iters) {
    ProductTimer& T = g_timer;
    const size_t n = std::min(T.used, (size_t)std::max(iters, 0));
    for (size_t k = 0; k < n; k++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, T.ev[2 * k], T.ev[2 * k + 1]) == hipSuccess) {
            T.total_us += 1000.0 * ms;
            T.count++;
        }
    }
    T.used = 0;
}

// One iteration of CG.
template <typename VT, typename MT, typename DT, int PCM>
static int cg_iteration(const mtl4g_matrix_s* A, const MT* val, const mtl4g_pc_s* P, const DT* d,
                        Workspace<VT>& W, VT* x, hipStream_t s) {
    const int64_t n = A->n;
    VT* p = W.vecs.p;
    VT* q = p + n;
    VT* r = q + n;
    VT* z = r + n;
    VT* y0 = z + n;
    Ctl* c = W.ctl.p;
    VT* part = reinterpret_cast<VT*>(W.partials.p);
    VT* tm = reinterpret_cast<VT*>(W.terms.p);
    unsigned* tk = W.tickets.p;
    if (PCM == PCM_GENERAL) {
        M4_TRY(pc_apply<VT>(P, r, z, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
        M4_TRY(launch_vec<VT>(n, EpiCgRho<VT>{r, z, c}, part, tk, s, tm));
    }
    M4_TRY(launch_vec<VT>(n, EpiCgP<VT, DT, PCM>{r, d, z, p, c, VT(0), 0}, part, tk, s));
    if (sizeof(VT) == 8) timer_mark(s, true);   // the FP64 solves (mtl4g_set_product_timing)
    M4_TRY(launch_csr<VT>(A, val, p, EpiCgQ<VT>{p, q, c}, part, tk, s, tm));
    if (sizeof(VT) == 8) timer_mark(s, false);
    M4_TRY(launch_vec<VT>(n, EpiCgXR<VT, DT, PCM>{x, r, p, q, d, c, W.hist.p, VT(0)}, part, tk, s, tm));
    return MTL4G_OK;
}

template <typename VT, typename MT, typename DT, int PCM>
static int cg_init(const mtl4g_matrix_s* A, const MT* val, const DT* d, Workspace<VT>& W,
                   const VT* x, const VT* b, hipStream_t s) {
    const int64_t n = A->n;
    VT* r = W.vecs.p + 2 * n;
    Ctl* c = W.ctl.p;
    return launch_csr<VT>(A, val, x, EpiCgInit<VT, DT, PCM>{b, r, d, c, W.hist.p},
                          reinterpret_cast<VT*>(W.partials.p), W.tickets.p, s,
                          reinterpret_cast<VT*>(W.terms.p));
}

template <typename VT, typename MT>
static int bicg_init(const mtl4g_matrix_s* A, const MT* val, Workspace<VT>& W, const VT* x,
                     const VT* b, hipStream_t s) {
    const int64_t n = A->n;
    VT* r = W.vecs.p + 6 * n;
    VT* rt = W.vecs.p + 7 * n;
    Ctl* c = W.ctl.p;
    return launch_csr<VT>(A, val, x, EpiBiInit<VT>{b, r, rt, c, W.hist.p},
                          reinterpret_cast<VT*>(W.partials.p), W.tickets.p, s,
                          reinterpret_cast<VT*>(W.terms.p));
}

template <typename VT, typename MT>
static int bicg_iteration(const mtl4g_matrix_s* A, const MT* val, const mtl4g_pc_s* P,
                          Workspace<VT>& W, VT* x, hipStream_t s) {
    const int64_t n = A->n;
    VT* base = W.vecs.p;
    VT *p = base, *phat = base + n, *sv = base + 2 * n, *shat = base + 3 * n, *t = base + 4 * n,
       *v = base + 5 * n, *r = base + 6 * n, *rt = base + 7 * n, *y0 = base + 8 * n;
    const bool ident = !P || P->kind == MTL4G_PC_IDENTITY;
    if (ident) {   // solve(identity, p) is a copy: use p and s directly
        phat = p;
        shat = sv;
    }
    Ctl* c = W.ctl.p;
    VT* part = reinterpret_cast<VT*>(W.partials.p);
    VT* tm = reinterpret_cast<VT*>(W.terms.p);
    unsigned* tk = W.tickets.p;
    M4_TRY(launch_vec<VT>(n, EpiBiRho<VT>{rt, r, c}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiBiP<VT>{r, v, p, c, VT(0), VT(0), 0}, part, tk, s));
    if (!ident) M4_TRY(pc_apply<VT>(P, p, phat, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
    M4_TRY(launch_csr<VT>(A, val, phat, EpiBiV<VT>{rt, v, c}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiBiS<VT>{r, v, sv, c, W.hist.p, VT(0)}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiBiHalf<VT>{phat, x, c, VT(0)}, part, tk, s));
    if (!ident) M4_TRY(pc_apply<VT>(P, sv, shat, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
    M4_TRY(launch_csr<VT>(A, val, shat, EpiBiT<VT>{sv, t, c}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiBiXR<VT>{x, r, shat, phat, sv, t, c, W.hist.p, VT(0), VT(0)}, part, tk, s, tm));
    return MTL4G_OK;
}

template <typename VT, typename MT>
static int cgs_iteration(const mtl4g_matrix_s* A, const MT* val, const mtl4g_pc_s* P,
                         Workspace<VT>& W, VT* x, hipStream_t s) {
    const int64_t n = A->n;
    VT* base = W.vecs.p;
    VT *p = base, *phat = base + n, *q = base + 2 * n, *vhat = base + 3 * n, *u = base + 4 * n,
       *uhat = base + 5 * n, *r = base + 6 * n, *rt = base + 7 * n, *y0 = base + 8 * n;
    const bool ident = !P || P->kind == MTL4G_PC_IDENTITY;
    if (ident) {   // solve(identity, .) is a copy: use p and u directly
        phat = p;
        uhat = u;
    }
    Ctl* c = W.ctl.p;
    VT* part = reinterpret_cast<VT*>(W.partials.p);
    VT* tm = reinterpret_cast<VT*>(W.terms.p);
    unsigned* tk = W.tickets.p;
    M4_TRY(launch_vec<VT>(n, EpiCgsRho<VT>{rt, r, c}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiCgsUP<VT>{r, q, u, p, c, VT(0), 0}, part, tk, s));
    if (!ident) M4_TRY(pc_apply<VT>(P, p, phat, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
    M4_TRY(launch_csr<VT>(A, val, phat, EpiCgsV<VT>{rt, vhat, c}, part, tk, s, tm));
    M4_TRY(launch_vec<VT>(n, EpiCgsQU<VT>{vhat, q, u, c, VT(0)}, part, tk, s));
    if (!ident) M4_TRY(pc_apply<VT>(P, u, uhat, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
    M4_TRY(launch_csr<VT>(A, val, uhat, EpiCgsXR<VT>{x, r, uhat, c, W.hist.p, VT(0)}, part, tk, s, tm));
    return MTL4G_OK;
}

template <typename VT, typename MT>
static int run_typed(const mtl4g_matrix_s* A, const MT* val, const mtl4g_pc_s* P, int method,
                     Workspace<VT>& W, VT* x, const VT* b, hipStream_t s, int32_t budget) {
    // a bfloat16 inverse diagonal is applied by the preconditioner kernel
    const int pcm = (!P || P->kind == MTL4G_PC_IDENTITY)                     ? PCM_NONE
                    : P->kind == MTL4G_PC_DIAGONAL && P->fmt != MTL4G_BF16 ? PCM_DIAG
                                                                          : PCM_GENERAL;
    // inverse diagonal in the vector type of the iteration
    const VT* d = nullptr;
    if (pcm == PCM_DIAG) {
        if (sizeof(VT) == 8 && P->fmt == MTL4G_F64) d = (const VT*)P->inv_diag64.p;
        else if (sizeof(VT) == 4 && P->fmt == MTL4G_F32) d = (const VT*)P->inv_diag32.p;
    }
    auto iteration = [&]() -> int {
        if (method == MTL4G_BICGSTAB) return bicg_iteration<VT, MT>(A, val, P, W, x, s);
        if (method == MTL4G_CGS) return cgs_iteration<VT, MT>(A, val, P, W, x, s);
        if (pcm == PCM_NONE) return cg_iteration<VT, MT, VT, PCM_NONE>(A, val, P, d, W, x, s);
        if (pcm == PCM_DIAG) {
            if (d) return cg_iteration<VT, MT, VT, PCM_DIAG>(A, val, P, d, W, x, s);
            if (P->fmt == MTL4G_F32)
                return cg_iteration<VT, MT, float, PCM_DIAG>(A, val, P, P->inv_diag32.p, W, x, s);
            return cg_iteration<VT, MT, double, PCM_DIAG>(A, val, P, P->inv_diag64.p, W, x, s);
        }
        return cg_iteration<VT, MT, VT, PCM_GENERAL>(A, val, P, d, W, x, s);
    };
    if (method == MTL4G_BICGSTAB || method == MTL4G_CGS) {   // same r, rtilde slots and loop head
        M4_TRY((bicg_init<VT, MT>(A, val, W, x, b, s)));
    } else if (pcm == PCM_NONE) {
        M4_TRY((cg_init<VT, MT, VT, PCM_NONE>(A, val, d, W, x, b, s)));
    } else if (pcm == PCM_DIAG) {
        if (d) M4_TRY((cg_init<VT, MT, VT, PCM_DIAG>(A, val, d, W, x, b, s)));
        else if (P->fmt == MTL4G_F32) M4_TRY((cg_init<VT, MT, float, PCM_DIAG>(A, val, P->inv_diag32.p, W, x, b, s)));
        else M4_TRY((cg_init<VT, MT, double, PCM_DIAG>(A, val, P->inv_diag64.p, W, x, b, s)));
    } else {
        M4_TRY((cg_init<VT, MT, VT, PCM_GENERAL>(A, val, d, W, x, b, s)));
    }
    // Enqueue chunks of iterations and poll the stop flag between chunks.
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY(iteration());
        launched += chunk;
        chunk = std::min(chunk * 2, 32);
    }
    return MTL4G_OK;
}

template <typename VT>
int krylov_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* P, int method, int fmt,
               Workspace<VT>& W, VT* x, const VT* b, mtl4g_iteration* it, double* h_hist,
               int32_t hist_cap, hipStream_t s) {
    if (!A || !it || (A->n && (!x || !b))) return MTL4G_ERR_INVALID;
    if (A->n != A->m) {
        set_error("Krylov solve needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    if (P && P->n != A->n) {
        set_error("preconditioner size %d != matrix size %d", P->n, A->n);
        return MTL4G_ERR_SIZE;
    }
    if (method != MTL4G_CG && method != MTL4G_BICGSTAB && method != MTL4G_CGS) return MTL4G_ERR_INVALID;
    if (A->n == 0) {   // r is empty: finished(0) holds at once (basic_iteration.hpp:85-93)
        it->resid = 0.0;
        it->error = (it->norm_r0 == 0 || 0.0 <= it->rtol * it->norm_r0) ? 0 : 1;
        if (h_hist && hist_cap > 0) h_hist[0] = 0.0;
        return it->error;
    }
    const int nvec = method == MTL4G_CG ? 5 : 9;
    g_timer.used = 0;   // events of a solve that ended in an error are not counted
    M4_TRY(ensure_workspace(W, A->n, nvec, P, hist_cap, s));
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.hist_cap = h_hist ? hist_cap : 0;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, W.ctl.p, init);
    M4_HIP(hipGetLastError());
    const int32_t budget = std::max(0, it->max_iter - it->i) + 8;
    int st;
    if (sizeof(VT) == 8) {
        st = run_typed<VT, double>(A, A->val.p, P, method, W, x, b, s, budget);
    } else {
        M4_TRY(ensure_lowp_values(const_cast<mtl4g_matrix_s*>(A), fmt, s));
        if (fmt == MTL4G_BF16)
            st = run_typed<VT, uint16_t>(A, A->valbf.p, P, method, W, x, b, s, budget);
        else
            st = run_typed<VT, float>(A, A->val32.p, P, method, W, x, b, s, budget);
    }
    M4_TRY(st);
    if (P) M4_TRY(pc_check_error(P, s));
    M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    const Ctl& c = *W.h_ctl;
    if (method == MTL4G_CG) timer_collect(c.i - it->i);
    it->i = c.i;
    it->error = c.error;
    it->resid = c.resid;
    if (h_hist && hist_cap > 0) {
        int32_t nh = std::min(c.hist_n, hist_cap);
        if (nh > 0) M4_HIP(hipMemcpy(h_hist, W.hist.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
        for (int32_t k = nh; k < hist_cap; k++) h_hist[k] = NAN;
    }
    if (c.status != 0) {
        set_error(c.status == MTL4G_ERR_BREAKDOWN ? "unexpected orthogonality (breakdown)"
                                                   : "non-finite residual");
        return c.status;
    }
    return c.error;
}

// itl::tfqmr (tfqmr.hpp:32-96) with every scalar on the device.  The line
// rt = A*x - b at the end of each iteration (:93) computes a value nothing
// reads before it is overwritten, so it is not run.
int tfqmr_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, Workspace<double>& W,
              double* x, const double* b, mtl4g_iteration* it, double* h_hist, int32_t hist_cap,
              hipStream_t s) {
    using VT = double;
    if (!A || !it || (A->n && (!x || !b))) return MTL4G_ERR_INVALID;
    if (A->n != A->m) {
        set_error("Krylov solve needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    if ((L && L->n != A->n) || (R && R->n != A->n)) {
        set_error("preconditioner size != matrix size %d", A->n);
        return MTL4G_ERR_SIZE;
    }
    if (A->n == 0) {   // tfqmr.hpp:40
        set_error("empty rhs vector");
        return MTL4G_ERR_INVALID;
    }
    if (L && L->kind == MTL4G_PC_IDENTITY) L = nullptr;
    if (R && R->kind == MTL4G_PC_IDENTITY) R = nullptr;
    M4_TRY(ensure_workspace(W, A->n, 12, L, hist_cap, s));
    M4_TRY(ensure_workspace(W, A->n, 12, R, hist_cap, s));   // sweep counters for both
    const int64_t n = A->n;
    VT* base = W.vecs.p;
    VT *rt = base, *r = base + n, *u1 = base + 2 * n, *u2 = base + 3 * n, *y1 = base + 4 * n,
       *y2 = base + 5 * n, *w = base + 6 * n, *d = base + 7 * n, *v = base + 8 * n, *tmp = base + 9 * n,
       *y0 = base + 10 * n;
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.hist_cap = h_hist ? hist_cap : 0;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, W.ctl.p, init);
    M4_HIP(hipGetLastError());
    Ctl* c = W.ctl.p;
    VT* part = reinterpret_cast<VT*>(W.partials.p);
    VT* tm = reinterpret_cast<VT*>(W.terms.p);
    unsigned* tk = W.tickets.p;
    const double* val = A->val.p;
    // y = A * Vector(solve(R, in))
    auto a_r = [&](const VT* in, VT* y) -> int {
        const VT* src = in;
        if (R) {
            M4_TRY(pc_apply<VT>(R, in, tmp, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
            src = tmp;
        }
        return launch_csr<VT>(A, val, src, EpiTfAv<VT>{y, c}, part, tk, s);
    };
    // out = solve(L, A * Vector(solve(R, in))), through rt
    auto l_a_r = [&](const VT* in, VT* out) -> int {
        if (!L) return a_r(in, out);
        M4_TRY(a_r(in, rt));
        return pc_apply<VT>(L, rt, out, y0, &c->done, W.pc_counters.p, s);
    };
    M4_TRY(zero_on(d, (size_t)n * sizeof(VT), s));   // d(resource(x), zero)
    // rt = b - A*solve(R, x); finished(rt); r = solve(L, rt)  (:44-49)
    const VT* xs = x;
    if (R) {
        M4_TRY(pc_apply<VT>(R, x, tmp, y0, nullptr, W.pc_counters.p, s, W.pc64.p));
        xs = tmp;
    }
    M4_TRY(launch_csr<VT>(A, val, xs, EpiTfInit<VT>{b, rt, c, W.hist.p}, part, tk, s, tm));
    if (L) M4_TRY(pc_apply<VT>(L, rt, r, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
    else M4_TRY(launch_vec<VT>(n, EpiTfCopy<VT>{rt, r, nullptr, c}, part, tk, s));
    // y1 = w = r; u1 = v = solve(L, A*solve(R, y1)); tau, rho; loop head  (:50-57)
    M4_TRY(launch_vec<VT>(n, EpiTfCopy<VT>{r, y1, w, c}, part, tk, s));
    M4_TRY(l_a_r(y1, u1));
    M4_TRY(launch_vec<VT>(n, EpiTfCopy<VT>{u1, v, nullptr, c}, part, tk, s));
    M4_TRY(launch_vec<VT>(n, EpiTfTau<VT>{r, c, W.hist.p}, part, tk, s, tm));
    auto iteration = [&]() -> int {
        M4_TRY(launch_vec<VT>(n, EpiTfSigma<VT>{r, v, c}, part, tk, s, tm));
        M4_TRY(launch_vec<VT>(n, EpiTfW<VT>{u1, y1, w, d, c, VT(0), VT(0)}, part, tk, s, tm));
        M4_TRY(launch_vec<VT>(n, EpiTfXY2<VT>{x, y2, d, y1, v, c, VT(0), VT(0)}, part, tk, s));
        M4_TRY(l_a_r(y2, u2));
        M4_TRY(launch_vec<VT>(n, EpiTfW<VT>{u2, y2, w, d, c, VT(0), VT(0)}, part, tk, s, tm));
        M4_TRY(launch_vec<VT>(n, EpiTfRho<VT>{x, d, r, w, c, W.hist.p, VT(0)}, part, tk, s, tm));
        M4_TRY(launch_vec<VT>(n, EpiTfY1<VT>{y1, w, y2, c, VT(0)}, part, tk, s));
        M4_TRY(l_a_r(y1, u1));
        return launch_vec<VT>(n, EpiTfV<VT>{v, u1, u2, c, VT(0)}, part, tk, s);
    };
    const int32_t budget = std::max(0, it->max_iter - it->i) + 8;
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY(iteration());
        launched += chunk;
        chunk = std::min(chunk * 2, 32);
    }
    const Ctl& hc = *W.h_ctl;
    // shift back: x = solve(R, x)  (:95), not after the early return at :48
    if (R && hc.looped) {
        M4_TRY(pc_apply<VT>(R, x, tmp, y0, nullptr, W.pc_counters.p, s, W.pc64.p));
        M4_HIP(hipMemcpyAsync(x, tmp, (size_t)n * sizeof(VT), hipMemcpyDeviceToDevice, s));
    }
    if (L) M4_TRY(pc_check_error(L, s));
    if (R) M4_TRY(pc_check_error(R, s));
    M4_HIP(hipStreamSynchronize(s));
    it->i = hc.i;
    it->error = hc.error;
    it->resid = hc.resid;
    if (h_hist && hist_cap > 0) {
        int32_t nh = std::min(hc.hist_n, hist_cap);
        if (nh > 0) M4_HIP(hipMemcpy(h_hist, W.hist.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
        for (int32_t k = nh; k < hist_cap; k++) h_hist[k] = NAN;
    }
    if (hc.status != 0) {
        set_error("non-finite residual");
        return hc.status;
    }
    return hc.error;
}

// itl::bicgstab_2 (bicgstab_2.hpp:37-130) with every scalar on the device.
// With L = identity solve(L, y) is a copy: the products write u_hat[j+1]
// and r_hat[j+1] directly.  Ten launches per iteration with four products.
int bicgstab2_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, Workspace<double>& W, double* x,
                  const double* b, mtl4g_iteration* it, double* h_hist, int32_t hist_cap,
                  hipStream_t s) {
    if (!A || !it || (A->n && (!x || !b))) return MTL4G_ERR_INVALID;
    if (A->n != A->m) {
        set_error("Krylov solve needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    if (L && L->n != A->n) {
        set_error("preconditioner size %d != matrix size %d", L->n, A->n);
        return MTL4G_ERR_SIZE;
    }
    if (A->n == 0) {   // bicgstab_2.hpp:45
        set_error("empty rhs vector");
        return MTL4G_ERR_INVALID;
    }
    if (L && L->kind == MTL4G_PC_IDENTITY) L = nullptr;
    M4_TRY(ensure_workspace(W, A->n, 10, L, hist_cap, s));
    const int64_t n = A->n;
    double* base = W.vecs.p;
    double *r0 = base, *r1 = base + n, *r2 = base + 2 * n, *u0 = base + 3 * n, *u1 = base + 4 * n,
           *u2 = base + 5 * n, *rt = base + 6 * n, *y = base + 7 * n, *x0 = base + 8 * n,
           *y0 = base + 9 * n;
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.hist_cap = h_hist ? hist_cap : 0;
    init.rho = 1.0;     // rho_0
    init.omega = 1.0;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, W.ctl.p, init);
    M4_HIP(hipGetLastError());
    Ctl* c = W.ctl.p;
    double* part = W.partials.p;
    double* tm = W.terms.p;
    unsigned* tk = W.tickets.p;
    const double* val = A->val.p;
    // shift problem; r0_tilde; r_hat[0] = solve(L, r_hat[0]); u_hat[0] = 0  (:52-64)
    M4_TRY(launch_vec<double>(n, EpiB2XNorm{x, c}, part, tk, s, tm));
    M4_TRY(launch_csr<double>(A, val, x, EpiB2Init{b, x, r0, x0, c, 0}, part, tk, s, tm));
    M4_TRY(launch_vec<double>(n, EpiB2Start{x, rt, u0, r0, c, 0, 0.0}, part, tk, s));
    if (L) {   // y = solve(L, r_hat[0]); r_hat[0] = y: the two buffers swap roles
        M4_TRY(pc_apply<double>(L, r0, y, y0, nullptr, W.pc_counters.p, s, W.pc64.p));
        std::swap(r0, y);
    }
    M4_TRY(launch_vec<double>(n, EpiB2Head<false>{x, r0, u0, r1, r2, u1, u2, rt, c, W.hist.p}, part, tk, s, tm));
    // out = solve(L, A*in) with the reductions of MODE on out; XUP adds alpha*u_hat[0] to x
    auto product = [&](const double* in, double* out, auto mode, auto xup, const double* a) -> int {
        constexpr int MODE = decltype(mode)::value;
        constexpr bool XUP = decltype(xup)::value;
        if (!L) return launch_csr<double>(A, val, in, EpiB2Mv<MODE, XUP>{a, out, x, u0, c, 0.0}, part, tk, s, tm);
        M4_TRY((launch_csr<double>(A, val, in, EpiB2Mv<B2_NONE, XUP>{a, y, x, u0, c, 0.0}, part, tk, s)));
        M4_TRY(pc_apply<double>(L, y, out, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
        return launch_vec<double>(n, EpiB2Dot<MODE>{a, out, c}, part, tk, s, tm);
    };
    using Gamma = std::integral_constant<int, B2_GAMMA>;
    using Rho = std::integral_constant<int, B2_RHO>;
    using Tau = std::integral_constant<int, B2_TAU>;
    using NoX = std::false_type;
    using AddX = std::true_type;
    auto iteration = [&]() -> int {
        // j = 0  (:75-96; rho_1 and beta came with the loop head)
        M4_TRY(launch_vec<double>(n, EpiB2U{r0, nullptr, u0, nullptr, c, 0.0}, part, tk, s));
        M4_TRY(product(u0, u1, Gamma{}, NoX{}, rt));
        M4_TRY(launch_vec<double>(n, EpiB2R{r0, nullptr, u1, nullptr, c, W.hist.p, 0.0}, part, tk, s, tm));
        M4_TRY(product(r0, r1, Rho{}, AddX{}, rt));
        // j = 1
        M4_TRY(launch_vec<double>(n, EpiB2U{r0, r1, u0, u1, c, 0.0}, part, tk, s));
        M4_TRY(product(u1, u2, Gamma{}, NoX{}, rt));
        M4_TRY(launch_vec<double>(n, EpiB2R{r0, r1, u1, u2, c, W.hist.p, 0.0}, part, tk, s, tm));
        M4_TRY(product(r1, r2, Tau{}, AddX{}, r1));
        // minimal-residual part, then the loop head
        M4_TRY(launch_vec<double>(n, EpiB2MR{r2, r1, r0, c, 0.0}, part, tk, s, tm));
        return launch_vec<double>(n, EpiB2Head<true>{x, r0, u0, r1, r2, u1, u2, rt, c, W.hist.p},
                                  part, tk, s, tm);
    };
    const int32_t budget = std::max(0, it->max_iter - it->i) + 8;
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY(iteration());
        launched += chunk;
        chunk = std::min(chunk * 2, 32);
    }
    const Ctl& hc = *W.h_ctl;
    if (hc.error != 3) M4_TRY(launch_vec<double>(n, EpiB2Shift{x, x0}, part, tk, s));
    if (L) M4_TRY(pc_check_error(L, s));
    M4_HIP(hipStreamSynchronize(s));
    it->i = hc.i;
    it->error = hc.error;
    it->resid = hc.resid;
    if (h_hist && hist_cap > 0) {
        int32_t nh = std::min(hc.hist_n, hist_cap);
        if (nh > 0) M4_HIP(hipMemcpy(h_hist, W.hist.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
        for (int32_t k = nh; k < hist_cap; k++) h_hist[k] = NAN;
    }
    if (hc.status != 0) {
        set_error("non-finite residual");
        return hc.status;
    }
    if (hc.error == 3) set_error("bicg breakdown #2");
    return hc.error;
}

// itl::bicgstab_ell (bicgstab_ell.hpp:37-134) with every scalar on the
// device.  Per iteration: l products A*solve(R, u_hat[j]) and l products
// A*solve(R, r_hat[j]) (each followed by solve(L, .)), l(l+1)/2 passes of
// the orthogonalization, and the loop head fused with the minimal-residual
// updates.
int bicgstab_ell_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, int32_t l,
                     Workspace<double>& W, double* x, const double* b, mtl4g_iteration* it, double* h_hist,
                     int32_t hist_cap, hipStream_t s) {
    if (!A || !it || (A->n && (!x || !b))) return MTL4G_ERR_INVALID;
    if (l < 1 || l > kMaxEll) {
        set_error("bicgstab_ell: l = %d outside 1..%d", l, kMaxEll);
        return MTL4G_ERR_INVALID;
    }
    if (A->n != A->m) {
        set_error("Krylov solve needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    if ((L && L->n != A->n) || (R && R->n != A->n)) {
        set_error("preconditioner size != matrix size %d", A->n);
        return MTL4G_ERR_SIZE;
    }
    if (A->n == 0) {   // bicgstab_ell.hpp:48
        set_error("empty rhs vector");
        return MTL4G_ERR_INVALID;
    }
    if (L && L->kind == MTL4G_PC_IDENTITY) L = nullptr;
    if (R && R->kind == MTL4G_PC_IDENTITY) R = nullptr;
    const int nvec = 2 * (l + 1) + 5;
    M4_TRY(ensure_workspace(W, A->n, nvec, L, hist_cap, s));
    M4_TRY(ensure_workspace(W, A->n, nvec, R, hist_cap, s));   // sweep counters for both
    if (!W.small.p) M4_TRY(W.small.alloc(128));
    const int64_t n = A->n;
    double* base = W.vecs.p;
    EllVecs V;
    for (int j = 0; j <= kMaxEll; j++) {
        V.r[j] = j <= l ? base + (size_t)j * n : nullptr;
        V.u[j] = j <= l ? base + (size_t)(l + 1 + j) * n : nullptr;
    }
    double *rt = base + (size_t)(2 * (l + 1)) * n, *y = rt + n, *x0 = y + n, *tmp = x0 + n, *y0 = tmp + n;
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.hist_cap = h_hist ? hist_cap : 0;
    init.rho = 1.0;     // rho_0
    init.omega = 1.0;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, W.ctl.p, init);
    M4_HIP(hipGetLastError());
    Ctl* c = W.ctl.p;
    double* part = W.partials.p;
    double* tm = W.terms.p;
    unsigned* tk = W.tickets.p;
    const double* val = A->val.p;
    const EllScalars E{W.small.p, l};
    // shift problem; r0_tilde; r_hat[0] = solve(L, r_hat[0]); u_hat[0] = 0  (:52-64)
    M4_TRY(launch_vec<double>(n, EpiB2XNorm{x, c}, part, tk, s, tm));
    M4_TRY(launch_csr<double>(A, val, x, EpiB2Init{b, x, V.r[0], x0, c, 0}, part, tk, s, tm));
    M4_TRY(launch_vec<double>(n, EpiB2Start{x, rt, V.u[0], V.r[0], c, 0, 0.0}, part, tk, s));
    if (L) {   // y = solve(L, r_hat[0]); r_hat[0] = y: the two buffers swap roles
        M4_TRY(pc_apply<double>(L, V.r[0], y, y0, nullptr, W.pc_counters.p, s, W.pc64.p));
        std::swap(V.r[0], y);
    }
    M4_TRY(launch_vec<double>(n, EpiEllHead<false>{V, x, rt, E, c, W.hist.p}, part, tk, s, tm));
    // out = solve(L, A * solve(R, in)) with the reductions of MODE on out;
    // XUP adds alpha*u_hat[0] to x
    auto product = [&](const double* in, double* out, auto mode, auto xup) -> int {
        constexpr int MODE = decltype(mode)::value;
        constexpr bool XUP = decltype(xup)::value;
        if (R) {
            M4_TRY(pc_apply<double>(R, in, tmp, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
            in = tmp;
        }
        if (!L) return launch_csr<double>(A, val, in, EpiB2Mv<MODE, XUP>{rt, out, x, V.u[0], c, 0.0}, part, tk, s, tm);
        M4_TRY((launch_csr<double>(A, val, in, EpiB2Mv<B2_NONE, XUP>{rt, y, x, V.u[0], c, 0.0}, part, tk, s)));
        M4_TRY(pc_apply<double>(L, y, out, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
        if constexpr (MODE == B2_NONE) return MTL4G_OK;
        else return launch_vec<double>(n, EpiB2Dot<MODE>{rt, out, c}, part, tk, s, tm);
    };
    using Gamma = std::integral_constant<int, B2_GAMMA>;
    using Rho = std::integral_constant<int, B2_RHO>;
    using None = std::integral_constant<int, B2_NONE>;
    using NoX = std::false_type;
    using AddX = std::true_type;
    auto iteration = [&]() -> int {
        for (int j = 0; j < l; j++) {   // (:74-97; rho_1 and beta of j = 0 came with the loop head)
            M4_TRY(launch_vec<double>(n, EpiEllU{V, j, c, 0.0}, part, tk, s));
            M4_TRY(product(V.u[j], V.u[j + 1], Gamma{}, NoX{}));
            M4_TRY(launch_vec<double>(n, EpiEllR{V, j, c, W.hist.p, 0.0}, part, tk, s, tm));
            if (j + 1 < l) M4_TRY(product(V.r[j], V.r[j + 1], Rho{}, AddX{}));
            else M4_TRY(product(V.r[j], V.r[j + 1], None{}, AddX{}));
        }
        for (int j = 1; j <= l; j++)   // minimal-residual part (:100-117)
            for (int i = 1; i <= j; i++)
                M4_TRY(launch_vec<double>(n, EpiEllOrth{V, E, i, j, c, 0.0}, part, tk, s, tm));
        return launch_vec<double>(n, EpiEllHead<true>{V, x, rt, E, c, W.hist.p}, part, tk, s, tm);
    };
    const int32_t budget = std::max(0, it->max_iter - it->i) + 8;
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY(iteration());
        launched += chunk;
        chunk = std::min(chunk * 2, 32);
    }
    const Ctl& hc = *W.h_ctl;
    if (hc.error != 3) {   // x = solve(R, x); x += x0
        const double* src = x;
        if (R) {
            M4_TRY(pc_apply<double>(R, x, tmp, y0, nullptr, W.pc_counters.p, s, W.pc64.p));
            src = tmp;
        }
        M4_TRY(launch_vec<double>(n, EpiEllShift{x, src, x0}, part, tk, s));
    }
    if (L) M4_TRY(pc_check_error(L, s));
    if (R) M4_TRY(pc_check_error(R, s));
    M4_HIP(hipStreamSynchronize(s));
    it->i = hc.i;
    it->error = hc.error;
    it->resid = hc.resid;
    if (h_hist && hist_cap > 0) {
        int32_t nh = std::min(hc.hist_n, hist_cap);
        if (nh > 0) M4_HIP(hipMemcpy(h_hist, W.hist.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
        for (int32_t k = nh; k < hist_cap; k++) h_hist[k] = NAN;
    }
    if (hc.status != 0) {
        set_error("non-finite residual");
        return hc.status;
    }
    if (hc.error == 3) set_error("bicg breakdown #2");
    return hc.error;
}

// itl::gmres (restart > 0, gmres.hpp:122-137) or itl::gmres_full (restart
// == 0, :37-119).  A cycle runs kmax Arnoldi steps whatever the residual
// (as gmres_full does), then the small least-squares solve, the update of
// x and the true residual; the host reads the control block once per cycle.
int gmres_run(const mtl4g_matrix_s* A, const mtl4g_pc_s* L, const mtl4g_pc_s* R, int32_t restart,
              Workspace<double>& W, double* x, const double* b, mtl4g_iteration* it, double* h_hist,
              int32_t hist_cap, hipStream_t s) {
    if (!A || !it || (A->n && (!x || !b)) || restart < 0) return MTL4G_ERR_INVALID;
    if (A->n != A->m) {
        set_error("Krylov solve needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    if ((L && L->n != A->n) || (R && R->n != A->n)) {
        set_error("preconditioner size != matrix size %d", A->n);
        return MTL4G_ERR_SIZE;
    }
    if (A->n == 0) {   // gmres.hpp:45
        set_error("empty rhs vector");
        return MTL4G_ERR_INVALID;
    }
    if (L && L->kind == MTL4G_PC_IDENTITY) L = nullptr;
    if (R && R->kind == MTL4G_PC_IDENTITY) R = nullptr;
    const int64_t n = A->n;
    // Krylov vectors per cycle: min(size(x), iterations left), at most restart
    auto cycle_len = [&](int32_t i) -> int64_t {
        const int64_t lim = restart > 0 ? std::min<int64_t>((int64_t)i + restart, it->max_iter) : it->max_iter;
        const int64_t rem = lim - i;
        return (rem < 0 || rem > n) ? n : rem;
    };
    const int64_t kcap = restart > 0 ? std::min<int64_t>(restart, n) : cycle_len(it->i);
    if (kcap > kMaxGmres) {
        set_error("gmres: %lld Krylov vectors per cycle, at most %d here", (long long)kcap, kMaxGmres);
        return MTL4G_ERR_INVALID;
    }
    const int kvec = (int)std::max<int64_t>(kcap, 1);
    const int nvec = kvec + 1 + 4;
    M4_TRY(ensure_workspace(W, A->n, nvec, L, hist_cap, s));
    M4_TRY(ensure_workspace(W, A->n, nvec, R, hist_cap, s));
    const size_t words = GmScalars::words(kvec);
    if (W.small.n < words) M4_TRY(W.small.alloc(words));
    double* base = W.vecs.p;
    double *r0 = base, *rv = base + n, *tmp = base + 2 * n, *y0 = base + 3 * n, *V = base + 4 * n;
    auto Vj = [&](int j) { return V + (size_t)j * n; };
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.hist_cap = h_hist ? hist_cap : 0;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, W.ctl.p, init);
    M4_HIP(hipGetLastError());
    Ctl* c = W.ctl.p;
    double* part = W.partials.p;
    double* tm = W.terms.p;
    unsigned* tk = W.tickets.p;
    const double* val = A->val.p;
    // r0 = b - A*x of the first cycle; later cycles reuse terminate's r
    M4_TRY(launch_csr<double>(A, val, x, EpiGmRes<false>{b, r0, c, W.hist.p}, part, tk, s));
    int32_t i_now = it->i;
    // every cycle that does not finish the solve advances at least one
    // iteration (a cycle is shorter than restart when n < restart or when the
    // Givens estimate ends it early), so max_iter - i + 2 cycles always reach
    // the solve's own stop: convergence, or error 1 at max_iter (gmres.hpp:125-133)
    const int64_t max_cycles = (restart > 0 ? (int64_t)std::max(0, it->max_iter - it->i) + 2 : 1);
    for (int64_t cyc = 0; cyc < max_cycles; cyc++) {
        // <= kvec whenever the cycle runs (a cycle that starts at or past
        // max_iter stops at its first finished() and launches nothing that runs)
        const int kmax = (int)std::min<int64_t>(cycle_len(i_now), kvec);
        const GmScalars G{W.small.p, kmax};
        hipLaunchKernelGGL(k_gm_zero, dim3(1), dim3(256), 0, s, W.small.p, GmScalars::words(kmax), (const Ctl*)c);
        M4_HIP(hipGetLastError());
        // r = solve(L, r0); rho = g[0] = two_norm(r); finished(rho); V_0 = r / rho  (:50-59)
        const double* r = r0;
        if (L) {
            M4_TRY(pc_apply<double>(L, r0, rv, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
            r = rv;
        }
        M4_TRY(launch_vec<double>(n, EpiGmRho{r, G, c, W.hist.p}, part, tk, s, tm));
        M4_TRY(launch_vec<double>(n, EpiGmV0{r, Vj(0), c, 0.0}, part, tk, s));
        for (int k = 0; k < kmax; k++) {   // (:63-96)
            const double* in = Vj(k);
            if (R) {
                M4_TRY(pc_apply<double>(R, in, tmp, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
                in = tmp;
            }
            double* w = Vj(k + 1);
            if (!L) {
                M4_TRY((launch_csr<double>(A, val, in, EpiGmMv<true>{Vj(0), w, G, k, c}, part, tk, s, tm)));
            } else {
                M4_TRY((launch_csr<double>(A, val, in, EpiGmMv<false>{Vj(0), rv, G, k, c}, part, tk, s)));
                M4_TRY(pc_apply<double>(L, rv, w, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
                M4_TRY(launch_vec<double>(n, EpiGmOrth<0>{w, nullptr, Vj(0), G, 0, k, c, 0.0}, part, tk, s, tm));
            }
            for (int j = 1; j <= k; j++)
                M4_TRY(launch_vec<double>(n, EpiGmOrth<1>{w, Vj(j - 1), Vj(j), G, j, k, c, 0.0}, part, tk, s, tm));
            M4_TRY(launch_vec<double>(n, EpiGmOrth<2>{w, Vj(k), Vj(0), G, 0, k, c, 0.0}, part, tk, s, tm));
            for (int j = 1; j <= k; j++)
                M4_TRY(launch_vec<double>(n, EpiGmOrth<3>{w, Vj(j - 1), Vj(j), G, j, k, c, 0.0}, part, tk, s, tm));
            M4_TRY(launch_vec<double>(n, EpiGmOrth<4>{w, Vj(k), nullptr, G, 0, k, c, 0.0}, part, tk, s, tm));
            M4_TRY(launch_vec<double>(n, EpiGmScale{w, G, k, c, 0.0, 0}, part, tk, s));
            hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, s, G, k, c);
            M4_HIP(hipGetLastError());
        }
        // least squares, x += solve(R, V*y), r = b - A*x, terminate(r)  (:98-118)
        hipLaunchKernelGGL(k_gm_lsq, dim3(1), dim3(256), 0, s, G, kmax, it->atol, c);
        M4_HIP(hipGetLastError());
        M4_TRY(launch_vec<double>(n, EpiGmComb{Vj(0), n, rv, G, c, 0}, part, tk, s));
        const double* t = rv;
        if (R) {
            M4_TRY(pc_apply<double>(R, rv, tmp, y0, &c->done, W.pc_counters.p, s, W.pc64.p));
            t = tmp;
        }
        M4_TRY(launch_vec<double>(n, EpiGmX{x, t, c}, part, tk, s));
        M4_TRY(launch_csr<double>(A, val, x, EpiGmRes<true>{b, r0, c, W.hist.p}, part, tk, s, tm));
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done || restart == 0) break;
        i_now = W.h_ctl->i;
    }
    if (L) M4_TRY(pc_check_error(L, s));
    if (R) M4_TRY(pc_check_error(R, s));
    M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    const Ctl& hc = *W.h_ctl;
    if (restart > 0 && !hc.done) {
        set_error("gmres did not stop after %lld cycles", (long long)max_cycles);
        return MTL4G_ERR_INVALID;
    }
    it->i = hc.i;
    it->error = hc.error;
    it->resid = hc.resid;
    if (h_hist && hist_cap > 0) {
        int32_t nh = std::min(hc.hist_n, hist_cap);
        if (nh > 0) M4_HIP(hipMemcpy(h_hist, W.hist.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
        for (int32_t k = nh; k < hist_cap; k++) h_hist[k] = NAN;
    }
    if (hc.status != 0) {
        set_error("non-finite residual");
        return hc.status;
    }
    if (hc.error == 2) set_error("GMRES did not find any direction to correct x");
    return hc.error;
}

template int krylov_run<double>(const mtl4g_matrix_s*, const mtl4g_pc_s*, int, int, Workspace<double>&,
                                double*, const double*, mtl4g_iteration*, double*, int32_t, hipStream_t);
template int krylov_run<float>(const mtl4g_matrix_s*, const mtl4g_pc_s*, int, int, Workspace<float>&,
                               float*, const float*, mtl4g_iteration*, double*, int32_t, hipStream_t);

// ---------------------------------------------------------------- mixed precision
template <typename MT>
static int mixed_typed(mtl4g_matrix_s* A, const MT* val32, const mtl4g_pc_s* P, int method,
                       Workspace<float>& W, float* x32, double* x, const double* b, int32_t budget,
                       hipStream_t s) {
    const int64_t n = A->n;
    // a bfloat16 inverse diagonal is applied by the preconditioner kernel
    const int pcm = (!P || P->kind == MTL4G_PC_IDENTITY)                     ? PCM_NONE
                    : P->kind == MTL4G_PC_DIAGONAL && P->fmt == MTL4G_F32  ? PCM_DIAG
                                                                          : PCM_GENERAL;
    if (P && P->kind == MTL4G_PC_DIAGONAL && P->fmt == MTL4G_F64) {
        set_error("mixed precision needs a preconditioner stored in FP32 or BF16");
        return MTL4G_ERR_INVALID;
    }
    (void)method;
    const float* d = pcm == PCM_DIAG ? P->inv_diag32.p : nullptr;
    Ctl* c = W.ctl.p;
    double* part = W.partials.p;
    unsigned* tk = W.tickets.p;
    float* r32 = W.vecs.p + 2 * n;
    const int cg = 1;
    auto refresh = [&]() -> int {
        M4_TRY(launch_vec<double>(n, EpiRu1{x, x32, c, 0}, part, tk, s));
        if (pcm == PCM_DIAG)
            return launch_csr<double>(A, A->val.p, x, EpiRu2<float, PCM_DIAG>{b, r32, d, c, cg}, part, tk, s);
        if (pcm == PCM_NONE)
            return launch_csr<double>(A, A->val.p, x, EpiRu2<float, PCM_NONE>{b, r32, d, c, cg}, part, tk, s);
        return launch_csr<double>(A, A->val.p, x, EpiRu2<float, PCM_GENERAL>{b, r32, d, c, cg}, part, tk, s);
    };
    auto iteration = [&]() -> int {
        if (pcm == PCM_NONE) return cg_iteration<float, MT, float, PCM_NONE>(A, val32, P, d, W, x32, s);
        if (pcm == PCM_DIAG) return cg_iteration<float, MT, float, PCM_DIAG>(A, val32, P, d, W, x32, s);
        return cg_iteration<float, MT, float, PCM_GENERAL>(A, val32, P, d, W, x32, s);
    };
    M4_TRY(refresh());   // initial FP64 residual (Ctl::init set)
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("mixed solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY(iteration());
        M4_TRY(refresh());
        launched += chunk;
        chunk = std::min(chunk * 2, 16);
    }
    return MTL4G_OK;
}

// BiCGSTAB keeps its vectors and arithmetic in FP64 and reads the matrix
// and preconditioner values in low precision: an FP32 BiCGSTAB recurrence
// loses the shadow residual's biorthogonality and diverges once the
// residual is a few units of FP32 rounding below the start (seen on the
// 1024^2 convection-diffusion problem below a relative residual of 1e-8).
// The values the products read are still the FP32 (or BF16) copies, and the
// FP64 refreshes with the FP64 values correct for them.
template <typename MT>
static int mixed_bicg_typed(mtl4g_matrix_s* A, const MT* lowp, const mtl4g_pc_s* P, Workspace<double>& W,
                            double* x, const double* b, int32_t budget, hipStream_t s) {
    const int64_t n = A->n;
    Ctl* c = W.ctl.p;
    double* r = W.vecs.p + 6 * n;
    double* rt = W.vecs.p + 7 * n;
    auto refresh = [&]() -> int {
        return launch_csr<double>(A, A->val.p, x, EpiRu64{b, r, rt, c}, W.partials.p, W.tickets.p, s,
                                  (double*)nullptr);
    };
    M4_TRY(refresh());   // initial FP64 residual (Ctl::init set), rtilde = r
    int32_t launched = 0, chunk = 2;
    for (;;) {
        M4_HIP(hipMemcpyAsync(W.h_ctl, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        if (W.h_ctl->done) break;
        if (launched > budget) {
            set_error("mixed solver did not stop after %d launched iterations", launched);
            return MTL4G_ERR_INVALID;
        }
        for (int k = 0; k < chunk; k++) M4_TRY((bicg_iteration<double, MT>(A, lowp, P, W, x, s)));
        M4_TRY(refresh());
        launched += chunk;
        chunk = std::min(chunk * 2, 16);
    }
    return MTL4G_OK;
}

int mixed_run(mtl4g_matrix_s* A, const mtl4g_pc_s* P, int method, int fmt, double delta,
              int32_t max_refresh, Workspace<float>& W, Workspace<double>& W64, float* x32, double* x,
              const double* b, mtl4g_iteration* it, int32_t* refreshes, hipStream_t s) {
    (void)max_refresh;
    if (method != MTL4G_CG && method != MTL4G_BICGSTAB) {
        set_error("mixed precision: cg and bicgstab only");
        return MTL4G_ERR_UNSUPPORTED;
    }
    if (A->n != A->m) return MTL4G_ERR_NOT_SQUARE;
    if (P && P->n != A->n) return MTL4G_ERR_SIZE;
    if (A->n == 0) {
        it->resid = 0.0;
        it->error = 0;
        return 0;
    }
    const bool bicg = method == MTL4G_BICGSTAB;
    if (bicg) M4_TRY(ensure_workspace(W64, A->n, 9, P, 0, s));
    else M4_TRY(ensure_workspace(W, A->n, 5, P, 0, s));
    Ctl* ctl = bicg ? W64.ctl.p : W.ctl.p;
    Ctl* h_ctl = bicg ? W64.h_ctl : W.h_ctl;
    M4_TRY(ensure_lowp_values(A, fmt, s));
    if (!bicg) M4_HIP(hipMemsetAsync(x32, 0, (size_t)A->n * sizeof(float), s));
    Ctl init;
    std::memset(&init, 0, sizeof init);
    init.norm_r0 = it->norm_r0;
    init.rtol = it->rtol;
    init.atol = it->atol;
    init.max_iter = it->max_iter;
    init.i = it->i;
    init.mixed = 1;
    init.init = 1;
    init.want_ru = 1;
    // BiCGSTAB restarts at each refresh, so it refreshes only after a large
    // fall of its residual (or an estimated stop, growth, the limit)
    init.delta = bicg ? kBicgDelta : delta;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, ctl, init);
    M4_HIP(hipGetLastError());
    // a safety net against a device loop that never stops: every counted
    // iteration can leave at most one chunk (16) of launches skipped behind a
    // pause for an FP64 refresh
    const int32_t budget = (std::max(0, it->max_iter - it->i) + 8) * 16 + 64;
    int st;
    if (bicg) {
        if (fmt == MTL4G_BF16) st = mixed_bicg_typed<uint16_t>(A, A->valbf.p, P, W64, x, b, budget, s);
        else st = mixed_bicg_typed<float>(A, A->val32.p, P, W64, x, b, budget, s);
    } else {
        if (fmt == MTL4G_BF16) st = mixed_typed<uint16_t>(A, A->valbf.p, P, method, W, x32, x, b, budget, s);
        else st = mixed_typed<float>(A, A->val32.p, P, method, W, x32, x, b, budget, s);
    }
    M4_TRY(st);
    if (P) M4_TRY(pc_check_error(P, s));
    M4_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    const Ctl& c = *h_ctl;
    it->i = c.i;
    it->error = c.error;
    it->resid = c.resid;
    if (refreshes) *refreshes = c.ru_count;
    if (c.status != 0) {
        set_error(c.status == MTL4G_ERR_BREAKDOWN ? "unexpected orthogonality (breakdown)"
                                                   : "non-finite residual");
        return c.status;
    }
    return c.error;
}

}  // namespace mtl4g

using namespace mtl4g;

extern "C" {

int mtl4g_solver_create(mtl4g_solver* out, mtl4g_matrix A, mtl4g_pc P, int method) {
    if (!out || !A ||
        (method != MTL4G_CG && method != MTL4G_BICGSTAB && method != MTL4G_CGS && method != MTL4G_BICGSTAB_2))
        return MTL4G_ERR_INVALID;
    mtl4g_solver_s* S_ = new (std::nothrow) mtl4g_solver_s;
    if (!S_) return MTL4G_ERR_OOM;
    S_->A = A;
    S_->P = P;
    S_->method = method;
    *out = S_;
    return MTL4G_OK;
}

int mtl4g_solver_destroy(mtl4g_solver S_) {
    delete S_;
    return MTL4G_OK;
}

int mtl4g_solver_solve(mtl4g_solver S_, double* x, const double* b, mtl4g_iteration* it,
                       double* h_hist, int32_t hist_cap, void* stream) {
    if (!S_) return MTL4G_ERR_INVALID;
    TraceRange tr(S_->method == MTL4G_CG ? "cg" : S_->method == MTL4G_BICGSTAB ? "bicgstab"
                  : S_->method == MTL4G_CGS ? "cgs" : "bicgstab_2");
    if (S_->method == MTL4G_BICGSTAB_2)
        return bicgstab2_run(S_->A, S_->P, S_->W, x, b, it, h_hist, hist_cap, S(stream));
    return krylov_run<double>(S_->A, S_->P, S_->method, MTL4G_F64, S_->W, x, b, it, h_hist,
                              hist_cap, S(stream));
}

int mtl4g_set_product_timing(int on) {
    if (on != 0 && on != 1) {
        set_error("mtl4g_set_product_timing: on must be 0 or 1");
        return MTL4G_ERR_INVALID;
    }
    g_timer.on = on != 0;
    g_timer.used = 0;
    return MTL4G_OK;
}

int mtl4g_product_timing(double* h_total_us, int64_t* h_count) {
    if (!h_total_us || !h_count) return MTL4G_ERR_INVALID;
    *h_total_us = g_timer.total_us;
    *h_count = g_timer.count;
    g_timer.total_us = 0;
    g_timer.count = 0;
    return MTL4G_OK;
}

int mtl4g_cg(mtl4g_matrix A, mtl4g_pc P, double* x, const double* b, mtl4g_iteration* it,
             double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("cg");
    Workspace<double> W;
    return krylov_run<double>(A, P, MTL4G_CG, MTL4G_F64, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_tfqmr(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* x, const double* b, mtl4g_iteration* it,
                double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("tfqmr");
    Workspace<double> W;
    return tfqmr_run(A, L, R, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_cgs(mtl4g_matrix A, mtl4g_pc P, double* x, const double* b, mtl4g_iteration* it,
              double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("cgs");
    Workspace<double> W;
    return krylov_run<double>(A, P, MTL4G_CGS, MTL4G_F64, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_bicgstab_2(mtl4g_matrix A, mtl4g_pc L, double* x, const double* b, mtl4g_iteration* it,
                     double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("bicgstab_2");
    Workspace<double> W;
    return bicgstab2_run(A, L, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_bicgstab_ell(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* x, const double* b, int32_t l,
                       mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("bicgstab_ell");
    Workspace<double> W;
    return bicgstab_ell_run(A, L, R, l, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_gmres(mtl4g_matrix A, mtl4g_pc L, mtl4g_pc R, double* x, const double* b, int32_t restart,
                mtl4g_iteration* it, double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("gmres");
    Workspace<double> W;
    return gmres_run(A, L, R, restart, W, x, b, it, h_hist, hist_cap, S(stream));
}

int mtl4g_bicgstab(mtl4g_matrix A, mtl4g_pc P, double* x, const double* b, mtl4g_iteration* it,
                   double* h_hist, int32_t hist_cap, void* stream) {
    TraceRange tr("bicgstab");
    Workspace<double> W;
    return krylov_run<double>(A, P, MTL4G_BICGSTAB, MTL4G_F64, W, x, b, it, h_hist, hist_cap,
                              S(stream));
}

}  // extern "C"
