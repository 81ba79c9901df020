// batched.hip -- many independent sparse systems in one persistent launch.
//
// All systems live in one block-diagonal CSR in HBM (absolute column
// indices, int32).  A launch is a list of work items (system, right-hand
// side, start vector, workspace slice); a persistent grid of
// cus * per_cu workgroups pulls items from an atomic queue in the order
// given (largest estimated cost first) and runs each item's whole
// itl::cg + diagonal/identity inside one workgroup: row loops strided by the
// workgroup, block-level reductions in a fixed order, workgroup barriers
// between the phases of an iteration.  A system's result therefore depends
// only on the system and its right-hand side, never on which workgroup or
// which batch ran it.
//
// Precisions: FP64 (cg.hpp:74-111 as is), and FP32 recurrences on FP32 or
// BF16 values with FP64 reliable updates (x accumulated in FP64, r = b - A*x
// recomputed in FP64 when the FP32 residual fell by delta and before any
// stop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "batched.hpp"
#include "kernels.hpp"
#include "precond.hpp"

namespace mtl4g {

enum { BPC_NONE = 0, BPC_DIAG = 1, BPC_ILU = 2 };

template <int BS, typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[BS / kWave + w] = b;
    }
    __syncthreads();
    T ra = sh[0], rb = sh[BS / kWave];
#pragma unroll
    for (int i = 1; i < BS / kWave; i++) {
        ra += sh[i];
        rb += sh[BS / kWave + i];
    }
    a = ra;
    b = rb;
}

// one row of the product, column order (mat_vec_mult.hpp:656-682); the
// indices and values of eight entries are requested together, then their
// eight gathers, so a row costs two memory round trips per eight entries
template <typename AT, typename MT, typename VT>
__device__ __forceinline__ AT brow(const BatchDev& M, const MT* __restrict__ val, int grow, int r0,
                                   const VT* __restrict__ v) {
    AT s = 0;
    const int j0 = M.rowptr[grow], j1 = M.rowptr[grow + 1];
    for (int j = j0; j < j1; j += 8) {
        int c[8];
        AT m[8], g[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool in = j + u < j1;
            c[u] = in ? M.col[j + u] - r0 : 0;
            m[u] = in ? load_val<AT>(val, j + u) : AT(0);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) g[u] = j + u < j1 ? static_cast<AT>(v[c[u]]) : AT(0);
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (j + u < j1) s += m[u] * g[u];
    }
    return s;
}

__device__ __forceinline__ bool bconv(double resid, double norm_r0, double rtol, double atol) {
    return norm_r0 == 0 ? resid <= atol : (resid <= rtol * norm_r0 || resid <= atol);
}

__device__ __forceinline__ int next_item(unsigned* queue, int* s_k) {
    __syncthreads();
    if (threadIdx.x == 0) *s_k = (int)atomicAdd(queue, 1u);
    __syncthreads();
    return *s_k;
}

// One level-scheduled triangular sweep of system `sys` inside its workgroup:
// w[r] = (v[r] - sum_j T[r][c_j] * w[c_j]) [* dia], each row in MTL4's
// subtraction order (lower_trisolve.hpp:174-189 unit, upper_trisolve.hpp:
// 161-176 inverse diagonal first), a barrier between levels.  v and w are
// the system's local vectors; w is read back only after a barrier.
template <int BS, bool UNIT, typename VT>
__device__ __forceinline__ void bsweep(const BatchSweepDev& S, const VT* __restrict__ val,
                                       const VT* __restrict__ dia, int sys, int r0,
                                       const double* v, double* w) {
    const int l0 = S.lev_off[sys], l1 = S.lev_off[sys + 1] - 1;
    for (int l = l0; l < l1; l++) {
        const int a = S.lev[l], e = S.lev[l + 1];
        for (int k = a + (int)threadIdx.x; k < e; k += BS) {
            const int lr = S.perm[k] - r0;
            double acc = v[lr];
            const int j1 = S.rowptr[k + 1];
            for (int j = S.rowptr[k]; j < j1; j++) acc -= (double)val[j] * w[S.col[j] - r0];
            if (!UNIT) acc = acc * (double)dia[k];
            w[lr] = acc;
        }
        __syncthreads();
    }
}

// z = solve(ilu_0, r): unit lower sweep into y, upper sweep into z
template <int BS, bool LOW>
__device__ __forceinline__ void bilu(const BatchDev& M, int sys, int r0, const double* r, double* y, double* z) {
    if (LOW) {
        bsweep<BS, true>(M.lo, M.lo.val32, M.lo.dia32, sys, r0, r, y);
        bsweep<BS, false>(M.up, M.up.val32, M.up.dia32, sys, r0, y, z);
    } else {
        bsweep<BS, true>(M.lo, M.lo.val, M.lo.dia, sys, r0, r, y);
        bsweep<BS, false>(M.up, M.up.val, M.up.dia, sys, r0, y, z);
    }
}

// A recurrence residual this far above the last FP64 residual is refreshed.
constexpr double kBatchGrowth = 10.0;
// BiCGSTAB restarts at each refresh: it refreshes after its residual fell by
// this factor (CG: delta).
constexpr double kBatchBicgDelta = 1e-5;

// ---------------------------------------------------------------- FP64 vectors
// itl::cg (cg.hpp:74-111) with FP64 vectors.  MT = double: the FP64 solve.
// MT = float / bfloat16: the products read those values and the
// preconditioner its FP32 values, and the true residual b - A*x is
// recomputed with the FP64 values whenever the recurrence residual fell by
// delta since the last refresh (or grew tenfold, or estimates a stop), which
// then decides (reliable updates).
template <int BS, int PCM, typename MT>
__global__ __launch_bounds__(BS) void k_batch_cg64(BatchDev M, const MT* __restrict__ aval,
                                                   const mtl4g_batch_item* __restrict__ items, int nitems,
                                                   unsigned* queue, const double* __restrict__ B, double* X,
                                                   double* W, mtl4g_batch_result* out, double delta) {
    constexpr bool LOW = !std::is_same<MT, double>::value;
    __shared__ double sh[2 * (BS / kWave)];
    __shared__ int s_k;
    for (;;) {
        const int k = next_item(queue, &s_k);
        if (k >= nitems) break;
        const mtl4g_batch_item it = items[k];
        const int r0 = M.sys_row[it.sys];
        const int n = M.sys_row[it.sys + 1] - r0;
        const double* b = B + it.b_off;
        double* x = X + it.x_off;
        double* r = W + it.ws_off;
        double* p = r + n;
        double* q = p + n;
        double* z = q + n;   // ilu_0 only
        double* y = z + n;
        auto dv = [&](int i) -> double { return LOW ? (double)M.dinv32[r0 + i] : M.dinv[r0 + i]; };
        // basic_iteration(b, ...): norm_r0 = two_norm(b); r = b - A*x; rr, rho
        double bb = 0, rr = 0, rho = 0, tmp = 0;
        for (int i = threadIdx.x; i < n; i += BS) {
            const double bi = b[i];
            bb += bi * bi;
            const double v = bi - brow<double>(M, M.val, r0 + i, r0, x);
            r[i] = v;
            rr += v * v;
            if (PCM == BPC_DIAG) rho += v * (dv(i) * v);
        }
        block_sum2<BS>(bb, rr, sh);
        if (PCM == BPC_DIAG) block_sum2<BS>(rho, tmp, sh);
        else rho = rr;
        const double norm_r0 = sqrt(bb);
        double resid = sqrt(fabs(rr)), rho_1 = 0, rmax = resid;
        int iters = 0, error = 0, status = 0, refreshes = LOW ? 1 : 0;
        bool fresh = true, restart = false;
        for (;;) {
            if (LOW && !fresh) {
                const bool est_stop = !isfinite(resid) || bconv(resid, norm_r0, it.rtol, it.atol) ||
                                      iters >= it.max_iter;
                const bool grown = resid > kBatchGrowth * rmax;
                if (est_stop || grown || resid <= delta * rmax) {   // FP64 refresh
                    restart = !isfinite(resid) || grown;
                    rr = 0;
                    rho = 0;
                    for (int i = threadIdx.x; i < n; i += BS) {
                        const double v = b[i] - brow<double>(M, M.val, r0 + i, r0, x);
                        r[i] = v;
                        rr += v * v;
                        if (PCM == BPC_DIAG) rho += v * (dv(i) * v);
                    }
                    block_sum2<BS>(rr, rho, sh);
                    if (PCM != BPC_DIAG) rho = rr;
                    resid = sqrt(fabs(rr));
                    rmax = resid;
                    fresh = true;
                    refreshes++;
                }
            }
            if (!LOW || fresh) {
                if (bconv(resid, norm_r0, it.rtol, it.atol)) break;
            }
            if (iters >= it.max_iter) { error = 1; break; }
            if (!isfinite(resid)) { status = MTL4G_ERR_NONFINITE; break; }
            iters++;
            const bool first = iters <= 1 || restart;
            restart = false;
            if (PCM == BPC_ILU) {   // z = solve(P, r); rho = dot(r, z)
                bilu<BS, LOW>(M, it.sys, r0, r, y, z);
                double rz = 0;
                tmp = 0;
                for (int i = threadIdx.x; i < n; i += BS) rz += r[i] * z[i];
                block_sum2<BS>(rz, tmp, sh);
                rho = rz;
            }
            const double beta = first ? 0.0 : rho / rho_1;
            for (int i = threadIdx.x; i < n; i += BS) {   // p = z + beta*p
                const double zi = PCM == BPC_DIAG ? dv(i) * r[i] : PCM == BPC_ILU ? z[i] : r[i];
                p[i] = first ? zi : zi + beta * p[i];
            }
            __syncthreads();
            double pq = 0;
            tmp = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // q = A*p, p.q
                const double s = brow<double>(M, aval, r0 + i, r0, p);
                q[i] = s;
                pq += p[i] * s;
            }
            block_sum2<BS>(pq, tmp, sh);
            const double alpha = rho / pq;
            rho_1 = rho;
            rr = 0;
            double rhon = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // x += alpha p; r -= alpha q
                x[i] = x[i] + alpha * p[i];
                const double v = r[i] - alpha * q[i];
                r[i] = v;
                rr += v * v;
                if (PCM == BPC_DIAG) rhon += v * (dv(i) * v);
            }
            block_sum2<BS>(rr, rhon, sh);
            resid = sqrt(fabs(rr));
            rho = PCM == BPC_DIAG ? rhon : rr;
            fresh = false;
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = mtl4g_batch_result{iters, error, status, refreshes, resid, norm_r0};
    }
}

// itl::bicgstab (bicgstab.hpp:27-78) with FP64 vectors, including the
// half-step exit and the breakdown tests; MT as in k_batch_cg64, except that
// every refresh restarts the recurrence (rtilde = r, p = r) and is taken
// after a fall of the residual by kBatchBicgDelta.
template <int BS, int PCM, typename MT>
__global__ __launch_bounds__(BS) void k_batch_bicg64(BatchDev M, const MT* __restrict__ aval,
                                                     const mtl4g_batch_item* __restrict__ items, int nitems,
                                                     unsigned* queue, const double* __restrict__ B, double* X,
                                                     double* W, mtl4g_batch_result* out, double delta) {
    constexpr bool LOW = !std::is_same<MT, double>::value;
    (void)delta;   // BiCGSTAB refreshes at kBatchBicgDelta
    __shared__ double sh[2 * (BS / kWave)];
    __shared__ int s_k;
    for (;;) {
        const int k = next_item(queue, &s_k);
        if (k >= nitems) break;
        const mtl4g_batch_item it = items[k];
        const int r0 = M.sys_row[it.sys];
        const int n = M.sys_row[it.sys + 1] - r0;
        const double* b = B + it.b_off;
        double* x = X + it.x_off;
        double* r = W + it.ws_off;
        double* rt = r + n;
        double* p = rt + n;
        double* v = p + n;
        double* sv = v + n;
        double* t = sv + n;
        double* ph = t + n;    // solve(P, p): p itself for the identity
        double* shv = ph + n;  // solve(P, s)
        double* y = shv + n;   // ilu_0 lower-sweep result
        if (PCM == BPC_NONE) {
            ph = p;
            shv = sv;
        }
        auto dv = [&](int i) -> double { return LOW ? (double)M.dinv32[r0 + i] : M.dinv[r0 + i]; };
        auto pc = [&](const double* in, double* o) {
            if (PCM == BPC_DIAG) {
                for (int i = threadIdx.x; i < n; i += BS) o[i] = dv(i) * in[i];
                __syncthreads();
            } else if (PCM == BPC_ILU) {
                bilu<BS, LOW>(M, it.sys, r0, in, y, o);
            }
        };
        // r = b - A*x; rtilde = r
        double bb = 0, rr = 0;
        for (int i = threadIdx.x; i < n; i += BS) {
            const double bi = b[i];
            bb += bi * bi;
            const double vv = bi - brow<double>(M, M.val, r0 + i, r0, x);
            r[i] = vv;
            rt[i] = vv;
            rr += vv * vv;
        }
        block_sum2<BS>(bb, rr, sh);
        const double norm_r0 = sqrt(bb);
        double resid = sqrt(fabs(rr)), rmax = resid;
        double rho_2 = 0, alpha = 0, omega = 0;
        int iters = 0, error = 0, status = 0, refreshes = LOW ? 1 : 0;
        bool fresh = true, restart = false;
        for (;;) {
            if (LOW && !fresh) {
                const bool est_stop = !isfinite(resid) || bconv(resid, norm_r0, it.rtol, it.atol) ||
                                      iters >= it.max_iter;
                const bool grown = resid > kBatchGrowth * rmax;
                if (est_stop || grown || resid <= kBatchBicgDelta * rmax) {
                    // FP64 refresh; the recurrence restarts from it (rtilde = r, p = r)
                    restart = true;
                    rr = 0;
                    double tmp = 0;
                    for (int i = threadIdx.x; i < n; i += BS) {
                        const double vv = b[i] - brow<double>(M, M.val, r0 + i, r0, x);
                        r[i] = vv;
                        rt[i] = vv;
                        rr += vv * vv;
                    }
                    block_sum2<BS>(rr, tmp, sh);
                    resid = sqrt(fabs(rr));
                    rmax = resid;
                    fresh = true;
                    refreshes++;
                }
            }
            if (!LOW || fresh) {
                if (bconv(resid, norm_r0, it.rtol, it.atol)) break;
            }
            if (iters >= it.max_iter) { error = 1; break; }
            if (!isfinite(resid)) { status = MTL4G_ERR_NONFINITE; break; }
            iters++;
            const bool first = iters <= 1 || restart;
            restart = false;
            double rho_1 = 0, tmp = 0;
            for (int i = threadIdx.x; i < n; i += BS) rho_1 += rt[i] * r[i];
            block_sum2<BS>(rho_1, tmp, sh);
            if (rho_1 == 0.0) { status = MTL4G_ERR_BREAKDOWN; break; }
            double beta = 0;
            if (!first) {
                if (omega == 0.0) { status = MTL4G_ERR_BREAKDOWN; break; }
                beta = (rho_1 / rho_2) * (alpha / omega);
            }
            for (int i = threadIdx.x; i < n; i += BS) p[i] = first ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
            __syncthreads();
            pc(p, ph);
            double gamma = 0;
            tmp = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // v = A*phat; gamma = dot(rtilde, v)
                const double s = brow<double>(M, aval, r0 + i, r0, ph);
                v[i] = s;
                gamma += rt[i] * s;
            }
            block_sum2<BS>(gamma, tmp, sh);
            if (gamma == 0.0) { status = MTL4G_ERR_BREAKDOWN; break; }
            alpha = rho_1 / gamma;
            double ss = 0;
            tmp = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // s = r - alpha*v
                const double si = r[i] - alpha * v[i];
                sv[i] = si;
                ss += si * si;
            }
            block_sum2<BS>(ss, tmp, sh);
            const double sres = sqrt(ss);
            const bool half_stop = LOW ? (!isfinite(sres) || bconv(sres, norm_r0, it.rtol, it.atol))
                                       : (bconv(sres, norm_r0, it.rtol, it.atol) || iters >= it.max_iter ||
                                          !isfinite(sres));
            if (half_stop) {   // finished(s): x += alpha*phat (bicgstab.hpp:64-67)
                for (int i = threadIdx.x; i < n; i += BS) x[i] = x[i] + alpha * ph[i];
                __syncthreads();
                resid = sres;
                if (!LOW) {
                    if (!bconv(sres, norm_r0, it.rtol, it.atol)) {
                        if (iters >= it.max_iter) error = 1;
                        else status = MTL4G_ERR_NONFINITE;
                    }
                    break;
                }
                fresh = false;   // the estimated stop asks for the FP64 refresh, which
                restart = true;  // decides; the recurrence then restarts from r
                continue;
            }
            pc(sv, shv);
            double ts = 0, tt = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // t = A*shat; omega = dot(t,s)/dot(t,t)
                const double ti = brow<double>(M, aval, r0 + i, r0, shv);
                t[i] = ti;
                ts += ti * sv[i];
                tt += ti * ti;
            }
            block_sum2<BS>(ts, tt, sh);
            omega = ts / tt;
            rr = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // x += omega*shat + alpha*phat; r = s - omega*t
                x[i] = x[i] + (omega * shv[i] + alpha * ph[i]);
                const double ri = sv[i] - omega * t[i];
                r[i] = ri;
                rr += ri * ri;
            }
            tmp = 0;
            block_sum2<BS>(rr, tmp, sh);
            resid = sqrt(fabs(rr));
            rho_2 = rho_1;
            fresh = false;
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = mtl4g_batch_result{iters, error, status, refreshes, resid, norm_r0};
    }
}

// ---------------------------------------------------------------- FP32 + FP64 refresh
template <int BS, int PCM, typename MT>
__global__ __launch_bounds__(BS) void k_batch_mixed(BatchDev M, const MT* __restrict__ val32,
                                                    const mtl4g_batch_item* __restrict__ items, int nitems,
                                                    unsigned* queue, const double* __restrict__ B, double* X,
                                                    double* W, mtl4g_batch_result* out, double delta) {
    __shared__ double sh[2 * (BS / kWave)];
    __shared__ float shf[2 * (BS / kWave)];
    __shared__ int s_k;
    for (;;) {
        const int k = next_item(queue, &s_k);
        if (k >= nitems) break;
        const mtl4g_batch_item it = items[k];
        const int r0 = M.sys_row[it.sys];
        const int n = M.sys_row[it.sys + 1] - r0;
        const double* b = B + it.b_off;
        double* x = X + it.x_off;
        float* r = reinterpret_cast<float*>(W + it.ws_off);
        float* p = r + n;
        float* q = p + n;
        float* x32 = q + n;
        const float* d = M.dinv32 + r0;
        double bb = 0, tmp = 0;
        for (int i = threadIdx.x; i < n; i += BS) {
            bb += b[i] * b[i];
            x32[i] = 0.0f;
        }
        block_sum2<BS>(bb, tmp, sh);
        const double norm_r0 = sqrt(bb);
        // FP64 refresh: x += x32, x32 = 0, r = b - A*x in FP64, r32 = float(r)
        double resid64 = 0;
        float rho = 0;
        int refreshes = 0;
        auto refresh = [&]() {
            for (int i = threadIdx.x; i < n; i += BS) {
                x[i] = x[i] + (double)x32[i];
                x32[i] = 0.0f;
            }
            __syncthreads();
            double r2 = 0, rz = 0;
            for (int i = threadIdx.x; i < n; i += BS) {
                const double v = b[i] - brow<double>(M, M.val, r0 + i, r0, x);
                const float f = (float)v;
                r[i] = f;
                r2 += v * v;
                rz += PCM == BPC_DIAG ? (double)(f * (d[i] * f)) : (double)(f * f);
            }
            block_sum2<BS>(r2, rz, sh);
            resid64 = sqrt(r2);
            rho = (float)rz;
            refreshes++;
        };
        refresh();
        double rmax = resid64, resid32 = resid64;
        float rho_1 = 0;
        int iters = 0, error = 0, status = 0;
        bool fresh = true;
        for (;;) {
            if (!fresh) {
                const bool est_stop = !isfinite(resid32) || bconv(resid32, norm_r0, it.rtol, it.atol) ||
                                      iters >= it.max_iter;
                if (est_stop || resid32 <= delta * rmax) {
                    refresh();
                    rmax = resid64;
                    fresh = true;
                }
            }
            if (fresh) {
                if (bconv(resid64, norm_r0, it.rtol, it.atol)) break;
                if (!isfinite(resid64)) { status = MTL4G_ERR_NONFINITE; break; }
            }
            if (iters >= it.max_iter) {
                if (!fresh) refresh();
                error = 1;
                break;
            }
            iters++;
            const bool first = iters <= 1;
            const float beta = first ? 0.0f : rho / rho_1;
            for (int i = threadIdx.x; i < n; i += BS) {
                const float zi = PCM == BPC_DIAG ? d[i] * r[i] : r[i];
                p[i] = first ? zi : zi + beta * p[i];
            }
            __syncthreads();
            float pq = 0, tf = 0;
            for (int i = threadIdx.x; i < n; i += BS) {
                const float s = brow<float>(M, val32, r0 + i, r0, p);
                q[i] = s;
                pq += p[i] * s;
            }
            block_sum2<BS>(pq, tf, shf);
            const float alpha = rho / pq;
            rho_1 = rho;
            float rr = 0, rhon = 0;
            for (int i = threadIdx.x; i < n; i += BS) {
                x32[i] = x32[i] + alpha * p[i];
                const float v = r[i] - alpha * q[i];
                r[i] = v;
                rr += v * v;
                if (PCM == BPC_DIAG) rhon += v * (d[i] * v);
            }
            block_sum2<BS>(rr, rhon, shf);
            resid32 = sqrt(fabs((double)rr));
            rho = PCM == BPC_DIAG ? rhon : rr;
            fresh = false;
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = mtl4g_batch_result{iters, error, status, refreshes, resid64, norm_r0};
    }
}

__global__ void k_reset_queue(unsigned* q) { *q = 0u; }

// ---------------------------------------------------------------- host
int64_t batch_ws_elems(int32_t n, int precision) {
    const int64_t pad = ((int64_t)n + 63) / 64 * 64;
    return precision == MTL4G_F64 ? 3 * pad : 2 * pad;   // 3 FP64 or 4 FP32 vectors
}

int64_t batch_ws_elems(int32_t n, int precision, int method, int pc_kind) {
    const int64_t pad = ((int64_t)n + 63) / 64 * 64;
    if (method == MTL4G_BICGSTAB) return 9 * pad;            // r rt p v s t phat shat y
    if (pc_kind == MTL4G_PC_ILU0) return 5 * pad;            // r p q z y
    return batch_ws_elems(n, precision);
}

static uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int batch_ensure_values(mtl4g_batch_s* B, int precision) {
    if (precision != MTL4G_F64 && !B->dinv32.p) {
        std::vector<float> d(B->h_diag.size());
        for (size_t i = 0; i < d.size(); i++) d[i] = (float)(1.0 / B->h_diag[i]);
        M4_TRY(B->dinv32.alloc(d.size()));
        M4_HIP(hipMemcpy(B->dinv32.p, d.data(), d.size() * 4, hipMemcpyHostToDevice));
    }
    if (precision == MTL4G_F32 && !B->val32.p) {
        std::vector<float> v(B->h_val.begin(), B->h_val.end());
        M4_TRY(B->val32.alloc(v.size()));
        M4_HIP(hipMemcpy(B->val32.p, v.data(), v.size() * 4, hipMemcpyHostToDevice));
    }
    if (precision == MTL4G_BF16 && !B->valbf.p) {
        std::vector<uint16_t> v(B->h_val.size());
        for (size_t i = 0; i < v.size(); i++) v[i] = to_bf16((float)B->h_val[i]);
        M4_TRY(B->valbf.alloc(v.size()));
        M4_HIP(hipMemcpy(B->valbf.p, v.data(), v.size() * 2, hipMemcpyHostToDevice));
    }
    return MTL4G_OK;
}

template <int BS>
static int launch_bs(mtl4g_batch_s* Bt, int method, int pcm, int precision, int grid,
                     const mtl4g_batch_item* items, int nitems, const double* b, double* x, double* work,
                     mtl4g_batch_result* out, unsigned* queue, hipStream_t s) {
    BatchDev M{Bt->rowptr.p, Bt->col.p, Bt->val.p, Bt->val32.p, Bt->valbf.p, Bt->dinv.p, Bt->dinv32.p,
               Bt->sys_row.p, Bt->ilu_lo.view(), Bt->ilu_up.view()};
    const double delta = 0.1;
#define M4_BATCH(KERNEL, PCM, MT, VALS)                                                                      \
    hipLaunchKernelGGL((KERNEL<BS, PCM, MT>), dim3(grid), dim3(BS), 0, s, M, VALS, items, nitems, queue, b, x, \
                       work, out, delta)
#define M4_BATCH_PCM(KERNEL, MT, VALS)                                        \
    do {                                                                      \
        if (pcm == BPC_DIAG) M4_BATCH(KERNEL, BPC_DIAG, MT, VALS);            \
        else if (pcm == BPC_ILU) M4_BATCH(KERNEL, BPC_ILU, MT, VALS);         \
        else M4_BATCH(KERNEL, BPC_NONE, MT, VALS);                            \
    } while (0)
#define M4_BATCH_PREC(KERNEL)                                                     \
    do {                                                                          \
        if (precision == MTL4G_F64) M4_BATCH_PCM(KERNEL, double, Bt->val.p);      \
        else if (precision == MTL4G_F32) M4_BATCH_PCM(KERNEL, float, Bt->val32.p); \
        else M4_BATCH_PCM(KERNEL, uint16_t, Bt->valbf.p);                         \
    } while (0)
    if (method == MTL4G_BICGSTAB) {
        M4_BATCH_PREC(k_batch_bicg64);
    } else if (precision == MTL4G_F64 || pcm == BPC_ILU) {
        M4_BATCH_PREC(k_batch_cg64);
    } else if (precision == MTL4G_F32) {   // FP32 recurrence (identity / diagonal)
        if (pcm == BPC_DIAG)
            hipLaunchKernelGGL((k_batch_mixed<BS, BPC_DIAG, float>), dim3(grid), dim3(BS), 0, s, M, Bt->val32.p,
                               items, nitems, queue, b, x, work, out, delta);
        else
            hipLaunchKernelGGL((k_batch_mixed<BS, BPC_NONE, float>), dim3(grid), dim3(BS), 0, s, M, Bt->val32.p,
                               items, nitems, queue, b, x, work, out, delta);
    } else {
        if (pcm == BPC_DIAG)
            hipLaunchKernelGGL((k_batch_mixed<BS, BPC_DIAG, uint16_t>), dim3(grid), dim3(BS), 0, s, M,
                               Bt->valbf.p, items, nitems, queue, b, x, work, out, delta);
        else
            hipLaunchKernelGGL((k_batch_mixed<BS, BPC_NONE, uint16_t>), dim3(grid), dim3(BS), 0, s, M,
                               Bt->valbf.p, items, nitems, queue, b, x, work, out, delta);
    }
#undef M4_BATCH_PREC
#undef M4_BATCH_PCM
#undef M4_BATCH
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

int batch_launch(mtl4g_batch_s* Bt, int method, int pc_kind, int precision, const BatchShape& sh,
                 const mtl4g_batch_item* items, int nitems, const double* b, double* x, double* work,
                 mtl4g_batch_result* out, unsigned* queue, hipStream_t s) {
    if (pc_kind != MTL4G_PC_IDENTITY && pc_kind != MTL4G_PC_DIAGONAL && pc_kind != MTL4G_PC_ILU0) {
        set_error("batched solves support the identity, diagonal and ilu_0 preconditioners");
        return MTL4G_ERR_UNSUPPORTED;
    }
    if (method != MTL4G_CG && method != MTL4G_BICGSTAB) {
        set_error("batched solves run itl::cg or itl::bicgstab");
        return MTL4G_ERR_UNSUPPORTED;
    }
    if (pc_kind == MTL4G_PC_ILU0 && !Bt->has_ilu) {
        set_error("batched ilu_0: factors not built (batch_ensure_ilu0)");
        return MTL4G_ERR_INVALID;
    }
    if (nitems <= 0) return MTL4G_OK;
    M4_TRY(batch_ensure_values(Bt, precision));
    const int pcm = pc_kind == MTL4G_PC_DIAGONAL ? BPC_DIAG : pc_kind == MTL4G_PC_ILU0 ? BPC_ILU : BPC_NONE;
    const int grid = std::max(1, std::min(nitems, Bt->cus * sh.per_cu));
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, s, queue);
    switch (sh.bs) {
    case 256: return launch_bs<256>(Bt, method, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    case 512: return launch_bs<512>(Bt, method, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    case 1024: return launch_bs<1024>(Bt, method, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    }
    set_error("batch workgroup size %d", sh.bs);
    return MTL4G_ERR_INVALID;
}

// Levels and level order of one system's sweep, appended to the batch-wide
// arrays.  rows: local row r's entries rp[r]..rp[r+1] of (col, val), diag
// value per row (unused for the unit sweep); lower: dependencies point to
// smaller rows.
struct SweepBuild {
    std::vector<int32_t> perm, rowptr{0}, col, lev, lev_off{0};
    std::vector<double> val, dia;
};

static void append_sweep(SweepBuild& SB, int32_t r0, int32_t n, const std::vector<int32_t>& rp,
                         const std::vector<int32_t>& ci, const std::vector<double>& va,
                         const std::vector<double>& dg, bool lower) {
    std::vector<int32_t> lev(n, 0);
    int32_t nlev = 1;
    auto level_of = [&](int32_t r) {
        int32_t l = 0;
        for (int32_t j = rp[r]; j < rp[r + 1]; j++) l = std::max(l, lev[ci[j]] + 1);
        lev[r] = l;
        nlev = std::max(nlev, l + 1);
    };
    if (lower)
        for (int32_t r = 0; r < n; r++) level_of(r);
    else
        for (int32_t r = n; r-- > 0;) level_of(r);
    std::vector<int32_t> start(nlev + 1, 0);
    for (int32_t r = 0; r < n; r++) start[lev[r] + 1]++;
    for (int32_t l = 0; l < nlev; l++) start[l + 1] += start[l];
    std::vector<int32_t> order(n), fill(start.begin(), start.end() - 1);
    for (int32_t r = 0; r < n; r++) order[fill[lev[r]]++] = r;   // ascending rows inside a level
    for (int32_t l = 0; l <= nlev; l++) SB.lev.push_back(r0 + start[l]);
    SB.lev_off.push_back((int32_t)SB.lev.size());
    for (int32_t k = 0; k < n; k++) {
        const int32_t r = order[k];
        SB.perm.push_back(r0 + r);
        for (int32_t j = rp[r]; j < rp[r + 1]; j++) {
            SB.col.push_back(r0 + ci[j]);
            SB.val.push_back(va[j]);
        }
        SB.rowptr.push_back((int32_t)SB.col.size());
        SB.dia.push_back(dg.empty() ? 1.0 : dg[r]);
    }
}

static int upload_batch_sweep(const SweepBuild& SB, BatchSweep& D) {
    auto up = [](auto& buf, const auto& v) -> int {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        M4_TRY(buf.alloc(std::max<size_t>(v.size(), 1)));
        if (!v.empty()) M4_HIP(hipMemcpy(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return MTL4G_OK;
    };
    std::vector<float> v32(SB.val.begin(), SB.val.end()), d32(SB.dia.begin(), SB.dia.end());
    M4_TRY(up(D.perm, SB.perm));
    M4_TRY(up(D.rowptr, SB.rowptr));
    M4_TRY(up(D.col, SB.col));
    M4_TRY(up(D.lev, SB.lev));
    M4_TRY(up(D.lev_off, SB.lev_off));
    M4_TRY(up(D.val, SB.val));
    M4_TRY(up(D.dia, SB.dia));
    M4_TRY(up(D.val32, v32));
    M4_TRY(up(D.dia32, d32));
    return MTL4G_OK;
}

int batch_ensure_ilu0(mtl4g_batch_s* B) {
    if (B->has_ilu) return MTL4G_OK;
    SweepBuild lo, up;
    for (int32_t k = 0; k < B->nsys; k++) {
        const int32_t a = B->h_sys_row[k], e = B->h_sys_row[k + 1], n = e - a;
        const int32_t j0 = B->h_rowptr[a], j1 = B->h_rowptr[e];
        std::vector<int32_t> rp(n + 1), ci(j1 - j0);
        for (int32_t r = 0; r <= n; r++) rp[r] = B->h_rowptr[a + r] - j0;
        for (int32_t j = j0; j < j1; j++) ci[j - j0] = B->h_col[j] - a;
        // itl::pc::ilu_0 of the system: the library's device factorization
        mtl4g_matrix A = nullptr;
        mtl4g_pc P = nullptr;
        int st = mtl4g_csr_create(&A, n, n, j1 - j0, rp.data(), ci.data(), B->h_val.data() + j0, 0);
        if (!st) st = mtl4g_pc_create(&P, A, MTL4G_PC_ILU0, MTL4G_F64);
        std::vector<int32_t> lrp(n + 1), urp(n + 1), lci, uci;
        std::vector<double> lva, uva;
        int64_t lnz = 0, unz = 0;
        if (!st) st = mtl4g_pc_factor(P, 1, nullptr, &lnz, nullptr, nullptr, nullptr);
        if (!st) st = mtl4g_pc_factor(P, 2, nullptr, &unz, nullptr, nullptr, nullptr);
        if (!st) {
            lci.resize(std::max<int64_t>(lnz, 1));
            lva.resize(lci.size());
            uci.resize(std::max<int64_t>(unz, 1));
            uva.resize(uci.size());
            st = mtl4g_pc_factor(P, 1, nullptr, nullptr, lrp.data(), lci.data(), lva.data());
            if (!st) st = mtl4g_pc_factor(P, 2, nullptr, nullptr, urp.data(), uci.data(), uva.data());
        }
        if (P) mtl4g_pc_destroy(P);
        if (A) mtl4g_csr_destroy(A);
        if (st) {
            set_error("batched ilu_0 of system %d: %s", k, mtl4g_last_error());
            return st;
        }
        // upper factor rows keep the inverted diagonal first (upper_trisolve v3)
        std::vector<int32_t> urp2(n + 1, 0), uci2;
        std::vector<double> uva2, udg(n);
        for (int32_t r = 0; r < n; r++) {
            udg[r] = uva[urp[r]];
            for (int32_t j = urp[r] + 1; j < urp[r + 1]; j++) {
                uci2.push_back(uci[j]);
                uva2.push_back(uva[j]);
            }
            urp2[r + 1] = (int32_t)uci2.size();
        }
        append_sweep(lo, a, n, lrp, lci, lva, {}, true);
        append_sweep(up, a, n, urp2, uci2, uva2, udg, false);
    }
    M4_TRY(upload_batch_sweep(lo, B->ilu_lo));
    M4_TRY(upload_batch_sweep(up, B->ilu_up));
    B->has_ilu = true;
    return MTL4G_OK;
}

static int prec_index(int precision) {
    return precision == MTL4G_F64 ? 0 : precision == MTL4G_F32 ? 1 : 2;
}

// items for "every system once": b and x at the system's rows, workspace
// slices in order; systems in decreasing estimated cost (nnz * sqrt(n)).
static int prepare_all(mtl4g_batch_s* Bt, int precision, int32_t max_iter, double rtol, double atol,
                       hipStream_t s) {
    std::vector<mtl4g_batch_item> items(Bt->nsys);
    int64_t ws = 0;
    std::vector<int64_t> ws_off(Bt->nsys);
    for (int32_t k = 0; k < Bt->nsys; k++) {
        ws_off[k] = ws;
        ws += batch_ws_elems(Bt->h_sys_row[k + 1] - Bt->h_sys_row[k], MTL4G_F64);
    }
    for (int32_t k = 0; k < Bt->nsys; k++) {
        const int32_t sys = Bt->order[k];
        items[k] = mtl4g_batch_item{sys, max_iter, Bt->h_sys_row[sys], Bt->h_sys_row[sys], ws_off[sys], rtol, atol};
    }
    (void)precision;
    if (Bt->items.n < items.size()) M4_TRY(Bt->items.alloc(items.size()));
    if (Bt->results.n < items.size()) M4_TRY(Bt->results.alloc(items.size()));
    if (Bt->work.n < (size_t)ws) M4_TRY(Bt->work.alloc(ws));
    if (!Bt->queue.p) M4_TRY(Bt->queue.alloc(1));
    M4_HIP(hipMemcpyAsync(Bt->items.p, items.data(), items.size() * sizeof(mtl4g_batch_item),
                          hipMemcpyHostToDevice, s));
    return MTL4G_OK;
}

// Solve-all through a captured graph: recorded once per (preconditioner,
// precision, shape, buffers) and replayed; the queue is reset by a kernel
// inside the graph (no memset node).
static int batch_graph_launch(mtl4g_batch_s* B, int pc_kind, int precision, const double* b, double* x,
                              hipStream_t s) {
    const BatchShape sh = B->shape[prec_index(precision)];
    BatchGraphKey key;
    key.pc_kind = pc_kind;
    key.precision = precision;
    key.bs = sh.bs;
    key.per_cu = sh.per_cu;
    key.b = b;
    key.x = x;
    key.items = B->items.p;
    key.work = B->work.p;
    key.results = B->results.p;
    M4_TRY(batch_ensure_values(B, precision));   // allocations stay outside the capture
    if (!B->gstream) {
        M4_HIP(hipStreamCreateWithFlags(&B->gstream, hipStreamNonBlocking));
        M4_HIP(hipEventCreateWithFlags(&B->gev[0], hipEventDisableTiming));
        M4_HIP(hipEventCreateWithFlags(&B->gev[1], hipEventDisableTiming));
    }
    if (!B->gexec || !(key == B->gkey)) {
        if (B->gexec) {
            (void)hipGraphExecDestroy(B->gexec);
            B->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        M4_HIP(hipStreamBeginCapture(B->gstream, hipStreamCaptureModeThreadLocal));
        const int st = batch_launch(B, MTL4G_CG, pc_kind, precision, sh, B->items.p, B->nsys, b, x, B->work.p,
                                    B->results.p, B->queue.p, B->gstream);
        const hipError_t e = hipStreamEndCapture(B->gstream, &g);
        if (st != MTL4G_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture", __FILE__, __LINE__);
        const hipError_t ei = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            B->gexec = nullptr;
            return hip_status(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        B->gkey = key;
        B->graph_captures++;
    }
    M4_HIP(hipEventRecord(B->gev[0], s));
    M4_HIP(hipStreamWaitEvent(B->gstream, B->gev[0], 0));
    M4_HIP(hipGraphLaunch(B->gexec, B->gstream));
    M4_HIP(hipEventRecord(B->gev[1], B->gstream));
    M4_HIP(hipStreamWaitEvent(s, B->gev[1], 0));
    B->graph_launches++;
    return MTL4G_OK;
}

}  // namespace mtl4g

using namespace mtl4g;

extern "C" {

int mtl4g_batch_create(mtl4g_batch* out, int32_t nsys, const int32_t* sys_rowptr, const int32_t* rowptr,
                       const int32_t* col, const double* val, int where) {
    if (!out || nsys <= 0 || !sys_rowptr || !rowptr || where != 0) {
        set_error("mtl4g_batch_create: bad arguments (host arrays required)");
        return MTL4G_ERR_INVALID;
    }
    *out = nullptr;
    const int32_t nrows = sys_rowptr[nsys];
    const int64_t nnz = rowptr[nrows];
    if (sys_rowptr[0] != 0 || rowptr[0] != 0 || nnz < 0) return MTL4G_ERR_INVALID;
    mtl4g_batch_s* B = new (std::nothrow) mtl4g_batch_s;
    if (!B) return MTL4G_ERR_OOM;
    auto fail = [&](int st) { delete B; return st; };
    (void)hipGetDevice(&B->device);
    (void)hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, B->device);
    if (B->cus <= 0) B->cus = 256;
    B->nsys = nsys;
    B->nrows = nrows;
    B->nnz = nnz;
    B->h_sys_row.assign(sys_rowptr, sys_rowptr + nsys + 1);
    B->h_rowptr.assign(rowptr, rowptr + nrows + 1);
    B->h_col.assign(col, col + nnz);
    B->h_val.assign(val, val + nnz);
    B->h_diag.assign(nrows, 0.0);
    std::vector<double> cost(nsys);
    for (int32_t k = 0; k < nsys; k++) {
        const int32_t a = sys_rowptr[k], e = sys_rowptr[k + 1];
        if (e <= a) {
            set_error("system %d is empty", k);
            return fail(MTL4G_ERR_INVALID);
        }
        B->max_n = std::max(B->max_n, e - a);
        for (int32_t r = a; r < e; r++) {
            if (rowptr[r + 1] < rowptr[r]) return fail(MTL4G_ERR_INVALID);
            for (int32_t j = rowptr[r]; j < rowptr[r + 1]; j++) {
                const int32_t c = col[j];
                if (c < a || c >= e || (j > rowptr[r] && c <= col[j - 1])) {
                    set_error("system %d: column %d of row %d outside the system or unsorted", k, c, r);
                    return fail(MTL4G_ERR_INVALID);
                }
                if (c == r) B->h_diag[r] = val[j];
            }
        }
        cost[k] = (double)(rowptr[e] - rowptr[a]) * std::sqrt((double)(e - a));
    }
    B->order.resize(nsys);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    std::vector<double> dinv(nrows);
    for (int32_t r = 0; r < nrows; r++) dinv[r] = 1.0 / B->h_diag[r];
    int st;
    if ((st = B->sys_row.alloc(nsys + 1)) || (st = B->rowptr.alloc(nrows + 1)) || (st = B->col.alloc(nnz)) ||
        (st = B->val.alloc(nnz)) || (st = B->dinv.alloc(nrows)))
        return fail(st);
    hipError_t e = hipMemcpy(B->sys_row.p, sys_rowptr, (nsys + 1) * 4, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(B->rowptr.p, rowptr, ((size_t)nrows + 1) * 4, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(B->col.p, col, (size_t)nnz * 4, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(B->val.p, val, (size_t)nnz * 8, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(B->dinv.p, dinv.data(), (size_t)nrows * 8, hipMemcpyHostToDevice);
    if (e) return fail(hip_status(e, "hipMemcpy", __FILE__, __LINE__));
    *out = B;
    return MTL4G_OK;
}

int mtl4g_batch_destroy(mtl4g_batch B) {
    delete B;
    return MTL4G_OK;
}

int64_t mtl4g_batch_workspace_elems(int32_t n, int precision) { return batch_ws_elems(n, precision); }

int mtl4g_batch_solve_items(mtl4g_batch B, int pc_kind, int precision, const mtl4g_batch_item* items,
                            int32_t nitems, const double* b, double* x, double* work,
                            mtl4g_batch_result* results, unsigned* queue, void* stream) {
    if (!B || nitems < 0 || (nitems > 0 && (!items || !b || !x || !work || !results || !queue)))
        return MTL4G_ERR_INVALID;
    return batch_launch(B, MTL4G_CG, pc_kind, precision, B->shape[prec_index(precision)], items, nitems, b, x, work,
                        results, queue, S(stream));
}

int mtl4g_batch_graph_info(mtl4g_batch B, int64_t* captures, int64_t* replays) {
    if (!B) return MTL4G_ERR_INVALID;
    if (captures) *captures = B->graph_captures;
    if (replays) *replays = B->graph_launches;
    return MTL4G_OK;
}

int mtl4g_batch_info(mtl4g_batch B, int32_t* nsys, int32_t* nrows, int64_t* nnz) {
    if (!B) return MTL4G_ERR_INVALID;
    if (nsys) *nsys = B->nsys;
    if (nrows) *nrows = B->nrows;
    if (nnz) *nnz = B->nnz;
    return MTL4G_OK;
}

int mtl4g_batch_set_shape(mtl4g_batch B, int precision, int32_t threads, int32_t per_cu) {
    if (!B || (threads != 256 && threads != 512 && threads != 1024) || per_cu < 1 || threads * per_cu > 2048)
        return MTL4G_ERR_INVALID;
    B->shape[prec_index(precision)] = BatchShape{threads, per_cu};
    B->tuned[prec_index(precision)] = true;
    return MTL4G_OK;
}

int mtl4g_batch_get_shape(mtl4g_batch B, int precision, int32_t* threads, int32_t* per_cu) {
    if (!B) return MTL4G_ERR_INVALID;
    const BatchShape& s = B->shape[prec_index(precision)];
    if (threads) *threads = s.bs;
    if (per_cu) *per_cu = s.per_cu;
    return MTL4G_OK;
}

int mtl4g_batch_solve(mtl4g_batch B, int pc_kind, int precision, double* d_x, const double* d_b, double rtol,
                      double atol, int32_t max_iter, mtl4g_batch_result* h_results, void* stream) {
    if (!B || !d_x || !d_b || (precision != MTL4G_F64 && precision != MTL4G_F32 && precision != MTL4G_BF16))
        return MTL4G_ERR_INVALID;
    hipStream_t s = S(stream);
    M4_TRY(prepare_all(B, precision, max_iter, rtol, atol, s));
    M4_TRY(batch_graph_launch(B, pc_kind, precision, d_b, d_x, s));
    if (h_results) {
        // results come back in item order; report them per system
        std::vector<mtl4g_batch_result> res(B->nsys);
        M4_HIP(hipMemcpyAsync(res.data(), B->results.p, res.size() * sizeof(mtl4g_batch_result),
                              hipMemcpyDeviceToHost, s));
        M4_HIP(hipStreamSynchronize(s));
        for (int32_t k = 0; k < B->nsys; k++) h_results[B->order[k]] = res[k];
    } else {
        M4_HIP(hipStreamSynchronize(s));
    }
    return MTL4G_OK;
}

// Autotuner: time the whole batch for every launch shape (x reset from x0
// each time) and keep the fastest for this precision.
int mtl4g_batch_tune(mtl4g_batch B, int pc_kind, int precision, double* d_x, const double* d_x0,
                     const double* d_b, double rtol, int32_t max_iter, double* h_best_ms, void* stream) {
    if (!B || !d_x || !d_x0 || !d_b) return MTL4G_ERR_INVALID;
    hipStream_t s = S(stream);
    M4_TRY(prepare_all(B, precision, max_iter, rtol, 0.0, s));
    const BatchShape cand[] = {{256, 1}, {256, 2}, {256, 4}, {256, 8}, {512, 1}, {512, 2}, {512, 4}, {1024, 1}, {1024, 2}};
    hipEvent_t e0, e1;
    M4_HIP(hipEventCreate(&e0));
    M4_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    BatchShape best_s = B->shape[prec_index(precision)];
    for (const BatchShape& c : cand) {
        M4_HIP(hipMemcpyAsync(d_x, d_x0, (size_t)B->nrows * 8, hipMemcpyDeviceToDevice, s));
        M4_HIP(hipEventRecord(e0, s));
        int st = batch_launch(B, MTL4G_CG, pc_kind, precision, c, B->items.p, B->nsys, d_b, d_x, B->work.p, B->results.p,
                              B->queue.p, s);
        M4_HIP(hipEventRecord(e1, s));
        M4_HIP(hipEventSynchronize(e1));
        if (st) continue;
        float ms = 0;
        M4_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_s = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->shape[prec_index(precision)] = best_s;
    B->tuned[prec_index(precision)] = true;
    if (h_best_ms) *h_best_ms = best;
    return MTL4G_OK;
}

}  // extern "C"
