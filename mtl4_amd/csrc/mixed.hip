// mixed.hip -- mixed-precision iterative refinement (no MTL4 counterpart).
//
//   r = b - A*x          FP64 values, FP64 vectors, fused with ||r||^2 and
//                        the FP32 copy of r that feeds the inner solve
//   stop if ||r|| <= rtol*norm_r0 (or atol)          FP64, on the device
//   A d = r              inner CG/BiCGSTAB in FP32 arithmetic on FP32 or
//                        BF16 stored values, stopped at inner_rtol*||r||
//   x += d               FP64
// The inner solve reuses the device-resident Krylov engine with VT = float,
// so an inner iteration moves about half the bytes of an FP64 one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels.hpp"
#include "krylov.hpp"

namespace mtl4g {

struct IrCtl {
    double rr;
};

// r = b - A*x in FP64, r32 = float(r), ||r||^2
struct EpiIrResid {
    static constexpr int NP = 1;
    struct Pre {
        double b;
    };
    const double* b;
    double* r;
    float* r32;
    IrCtl* c;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{b[row]}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        const double v = pre.b - s;
        r32[row] = (float)v;
        part[0] += v * v;
    }
    __device__ void finish(const double* tot) const { c->rr = tot[0]; }
};

// x += double(d)
struct EpiIrCorrect {
    static constexpr int NP = 0;
    double* x;
    const float* d;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double*) const { x[i] = x[i] + (double)d[i]; }
    __device__ void finish(const double*) const {}
};

struct IrWorkspace {
    DevBuf<float> r32, d32;
    DevBuf<double> partials;
    DevBuf<unsigned> tickets;
    DevBuf<IrCtl> ctl;
    Workspace<float> inner;
};

}  // namespace mtl4g

using namespace mtl4g;

extern "C" int mtl4g_ir_solve(mtl4g_matrix A, mtl4g_pc P32, double* x, const double* b,
                              mtl4g_iteration* it, const mtl4g_ir_params* prm,
                              mtl4g_ir_stats* stats, void* stream) {
    if (!A || !it || !prm || (A->n && (!x || !b))) return MTL4G_ERR_INVALID;
    if (A->n != A->m) return MTL4G_ERR_NOT_SQUARE;
    if (prm->value_format != MTL4G_F32 && prm->value_format != MTL4G_BF16) {
        set_error("ir_solve: inner value format must be F32 or BF16");
        return MTL4G_ERR_INVALID;
    }
    if (P32 && P32->n != A->n) return MTL4G_ERR_SIZE;
    if (prm->method != MTL4G_CG && prm->method != MTL4G_BICGSTAB) return MTL4G_ERR_INVALID;
    TraceRange tr(prm->method == MTL4G_CG ? "cg" : "bicgstab");
    hipStream_t s = S(stream);
    const int32_t n = A->n;
    if (prm->mode == MTL4G_IR_RELIABLE) {
        Workspace<float> W32;
        Workspace<double> W64;
        DevBuf<float> x32;
        if (prm->method == MTL4G_CG) M4_TRY(x32.alloc(n));
        int32_t refreshes = 0;
        it->error = 0;
        int st = mixed_run(A, P32, prm->method, prm->value_format, prm->inner_rtol,
                           prm->max_outer, W32, W64, x32.p, x, b, it, &refreshes, s);
        if (stats) {
            stats->outer = refreshes;
            stats->inner_total = it->i;
            stats->relres = it->norm_r0 != 0 ? it->resid / it->norm_r0 : it->resid;
        }
        return st;
    }
    if (prm->mode != MTL4G_IR_RESTART) return MTL4G_ERR_INVALID;
    IrWorkspace W;
    M4_TRY(W.r32.alloc(n));
    M4_TRY(W.d32.alloc(n));
    M4_TRY(W.partials.alloc(kPartialWords));
    M4_TRY(W.tickets.alloc(kTicketWords));
    M4_HIP(hipMemsetAsync(W.tickets.p, 0, kTicketWords * sizeof(unsigned), s));
    M4_TRY(W.ctl.alloc(1));
    M4_HIP(hipMemsetAsync(W.ctl.p, 0, sizeof(IrCtl), s));
    M4_TRY(ensure_lowp_values(A, prm->value_format, s));
    int32_t outer = 0, inner_total = 0;
    double resid = 0.0;
    int status = MTL4G_OK;
    it->error = 0;
    for (;;) {
        M4_TRY(launch_csr<double>(A, A->val.p, x, EpiIrResid{b, nullptr, W.r32.p, W.ctl.p},
                                  W.partials.p, W.tickets.p, s));
        IrCtl h;
        M4_HIP(hipMemcpyAsync(&h, W.ctl.p, sizeof(IrCtl), hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        resid = std::sqrt(h.rr);
        const bool conv = it->norm_r0 == 0 ? resid <= it->atol
                                           : (resid <= it->rtol * it->norm_r0 || resid <= it->atol);
        if (conv) break;
        if (!std::isfinite(resid)) { status = MTL4G_ERR_NONFINITE; break; }
        if (outer >= prm->max_outer) { it->error = 1; break; }
        M4_HIP(hipMemsetAsync(W.d32.p, 0, (size_t)n * sizeof(float), s));
        mtl4g_iteration inner{};
        inner.norm_r0 = resid;
        inner.rtol = prm->inner_rtol;
        inner.atol = 0.0;
        inner.max_iter = prm->inner_max_iter;
        inner.i = 0;
        int st = krylov_run<float>(A, P32, prm->method, prm->value_format, W.inner, W.d32.p, W.r32.p,
                                   &inner, nullptr, 0, s);
        if (st < 0 && st != MTL4G_ERR_BREAKDOWN && st != MTL4G_ERR_NONFINITE) return st;
        inner_total += inner.i;
        M4_TRY(launch_vec<double>(n, EpiIrCorrect{x, W.d32.p}, (double*)nullptr, nullptr, s));
        outer++;
        if (st < 0 && inner.i == 0) { status = st; break; }   // no progress possible
    }
    it->i = outer;
    it->resid = resid;
    if (stats) {
        stats->outer = outer;
        stats->inner_total = inner_total;
        stats->relres = it->norm_r0 != 0 ? resid / it->norm_r0 : resid;
    }
    if (status < 0) {
        set_error("ir_solve: %s", mtl4g_status_string(status));
        return status;
    }
    return it->error;
}
