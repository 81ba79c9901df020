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

namespace mtl4g {

enum { BPC_NONE = 0, BPC_DIAG = 1 };

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

// ---------------------------------------------------------------- FP64
template <int BS, int PCM>
__global__ __launch_bounds__(BS) void k_batch_f64(BatchDev M, const mtl4g_batch_item* __restrict__ items,
                                                  int nitems, unsigned* queue, const double* __restrict__ B,
                                                  double* X, double* W, mtl4g_batch_result* out) {
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
        const double* d = M.dinv + r0;
        // basic_iteration(b, ...): norm_r0 = two_norm(b); r = b - A*x; rr, rho
        double bb = 0, rr = 0, rho = 0, tmp = 0;
        for (int i = threadIdx.x; i < n; i += BS) {
            const double bi = b[i];
            bb += bi * bi;
            const double v = bi - brow<double>(M, M.val, r0 + i, r0, x);
            r[i] = v;
            rr += v * v;
            if (PCM == BPC_DIAG) rho += v * (d[i] * v);
        }
        block_sum2<BS>(bb, rr, sh);
        if (PCM == BPC_DIAG) block_sum2<BS>(rho, tmp, sh);
        else rho = rr;
        const double norm_r0 = sqrt(bb);
        double resid = sqrt(fabs(rr)), rho_1 = 0;
        int iters = 0, error = 0, status = 0;
        for (;;) {
            if (bconv(resid, norm_r0, it.rtol, it.atol)) break;
            if (iters >= it.max_iter) { error = 1; break; }
            if (!isfinite(resid)) { status = MTL4G_ERR_NONFINITE; break; }
            iters++;
            const bool first = iters <= 1;
            const double beta = first ? 0.0 : rho / rho_1;
            for (int i = threadIdx.x; i < n; i += BS) {   // p = z + beta*p
                const double zi = PCM == BPC_DIAG ? d[i] * r[i] : r[i];
                p[i] = first ? zi : zi + beta * p[i];
            }
            __syncthreads();
            double pq = 0;
            tmp = 0;
            for (int i = threadIdx.x; i < n; i += BS) {   // q = A*p, p.q
                const double s = brow<double>(M, M.val, r0 + i, r0, p);
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
                if (PCM == BPC_DIAG) rhon += v * (d[i] * v);
            }
            block_sum2<BS>(rr, rhon, sh);
            resid = sqrt(fabs(rr));
            rho = PCM == BPC_DIAG ? rhon : rr;
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = mtl4g_batch_result{iters, error, status, 0, resid, norm_r0};
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
static int launch_bs(mtl4g_batch_s* Bt, int pcm, int precision, int grid, const mtl4g_batch_item* items,
                     int nitems, const double* b, double* x, double* work, mtl4g_batch_result* out,
                     unsigned* queue, hipStream_t s) {
    BatchDev M{Bt->rowptr.p, Bt->col.p, Bt->val.p, Bt->val32.p, Bt->valbf.p, Bt->dinv.p, Bt->dinv32.p,
               Bt->sys_row.p};
    const double delta = 0.1;
    if (precision == MTL4G_F64) {
        if (pcm == BPC_DIAG)
            hipLaunchKernelGGL((k_batch_f64<BS, BPC_DIAG>), dim3(grid), dim3(BS), 0, s, M, items, nitems, queue, b, x,
                               work, out);
        else
            hipLaunchKernelGGL((k_batch_f64<BS, BPC_NONE>), dim3(grid), dim3(BS), 0, s, M, items, nitems, queue, b, x,
                               work, out);
    } else if (precision == MTL4G_F32) {
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
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

int batch_launch(mtl4g_batch_s* Bt, int pc_kind, int precision, const BatchShape& sh,
                 const mtl4g_batch_item* items, int nitems, const double* b, double* x, double* work,
                 mtl4g_batch_result* out, unsigned* queue, hipStream_t s) {
    if (pc_kind != MTL4G_PC_IDENTITY && pc_kind != MTL4G_PC_DIAGONAL) {
        set_error("batched solves support the identity and diagonal preconditioners");
        return MTL4G_ERR_UNSUPPORTED;
    }
    if (nitems <= 0) return MTL4G_OK;
    M4_TRY(batch_ensure_values(Bt, precision));
    const int pcm = pc_kind == MTL4G_PC_DIAGONAL ? BPC_DIAG : BPC_NONE;
    const int grid = std::max(1, std::min(nitems, Bt->cus * sh.per_cu));
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, s, queue);
    switch (sh.bs) {
    case 256: return launch_bs<256>(Bt, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    case 512: return launch_bs<512>(Bt, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    case 1024: return launch_bs<1024>(Bt, pcm, precision, grid, items, nitems, b, x, work, out, queue, s);
    }
    set_error("batch workgroup size %d", sh.bs);
    return MTL4G_ERR_INVALID;
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
        const int st = batch_launch(B, pc_kind, precision, sh, B->items.p, B->nsys, b, x, B->work.p,
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
    return batch_launch(B, pc_kind, precision, B->shape[prec_index(precision)], items, nitems, b, x, work,
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
        int st = batch_launch(B, pc_kind, precision, c, B->items.p, B->nsys, d_b, d_x, B->work.p, B->results.p,
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
