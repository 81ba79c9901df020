// batched.hpp -- many independent systems solved by one persistent launch.
#pragma once

#include "internal.hpp"

namespace mtl4g {

// One triangular sweep of every system's ILU(0) factor, level scheduled
// inside the system's workgroup.  Position k (the positions of system s are
// [sys_row[s], sys_row[s+1])) handles row perm[k]; its off-diagonal entries
// rowptr[k]..rowptr[k+1] follow in the order MTL4 subtracts them.  Levels of
// system s: positions lev[lev_off[s] + l] .. lev[lev_off[s] + l + 1].
struct BatchSweepDev {
    const int32_t *perm, *rowptr, *col, *lev, *lev_off;
    const double *val, *dia;       // FP64 factor values; dia: stored inverse diagonal (upper)
    const float *val32, *dia32;    // the same rounded to float
};

struct BatchDev {
    const int32_t* rowptr;    // block-diagonal CSR over all systems
    const int32_t* col;       // absolute column indices
    const double* val;
    const float* val32;
    const uint16_t* valbf;
    const double* dinv;       // 1 / A[i][i]
    const float* dinv32;
    const int32_t* sys_row;   // first row of each system, nsys+1
    BatchSweepDev lo, up;     // ILU(0): unit lower sweep, upper sweep (inverse diagonal)
};

// Host side of one batched sweep.
struct BatchSweep {
    DevBuf<int32_t> perm, rowptr, col, lev, lev_off;
    DevBuf<double> val, dia;
    DevBuf<float> val32, dia32;
    BatchSweepDev view() const {
        return BatchSweepDev{perm.p, rowptr.p, col.p, lev.p, lev_off.p, val.p, dia.p, val32.p, dia32.p};
    }
};

// Launch shape picked by the autotuner.
struct BatchShape {
    int bs = 256;            // threads per workgroup
    int per_cu = 2;          // resident workgroups per CU
};

// What a captured solve-all graph was recorded with (kernel arguments).
struct BatchGraphKey {
    int pc_kind = -1, precision = -1, bs = 0, per_cu = 0;
    const void *b = nullptr, *x = nullptr, *items = nullptr, *work = nullptr, *results = nullptr;
    bool operator==(const BatchGraphKey& o) const {
        return pc_kind == o.pc_kind && precision == o.precision && bs == o.bs && per_cu == o.per_cu &&
               b == o.b && x == o.x && items == o.items && work == o.work && results == o.results;
    }
};

}  // namespace mtl4g

struct mtl4g_batch_s {
    int device = 0;
    int32_t nsys = 0, nrows = 0;
    int64_t nnz = 0;
    int32_t max_n = 0;
    std::vector<int32_t> h_sys_row, h_rowptr, h_col;
    std::vector<double> h_val, h_diag;
    mtl4g::DevBuf<int32_t> sys_row, rowptr, col;
    mtl4g::DevBuf<double> val, dinv;
    mtl4g::DevBuf<float> val32, dinv32;
    mtl4g::DevBuf<uint16_t> valbf;
    // ILU(0) of every system (built on demand by batch_ensure_ilu0)
    mtl4g::BatchSweep ilu_lo, ilu_up;
    bool has_ilu = false;
    // solve-all scratch (items in cost order, workspace, queue)
    mtl4g::DevBuf<mtl4g_batch_item> items;
    mtl4g::DevBuf<mtl4g_batch_result> results;
    mtl4g::DevBuf<double> work;
    mtl4g::DevBuf<unsigned> queue;
    std::vector<int32_t> order;   // systems by decreasing estimated cost
    mtl4g::BatchShape shape[3];   // per precision (F64, F32, BF16)
    bool tuned[3] = {false, false, false};
    int cus = 0;
    // solve-all as a captured graph (queue reset + persistent kernel),
    // replayed on a stream of its own ordered after and before the caller's
    hipGraphExec_t gexec = nullptr;
    mtl4g::BatchGraphKey gkey;
    hipStream_t gstream = nullptr;
    hipEvent_t gev[2] = {nullptr, nullptr};
    int64_t graph_launches = 0, graph_captures = 0;
    ~mtl4g_batch_s() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (hipEvent_t e : gev)
            if (e) (void)hipEventDestroy(e);
        if (gstream) (void)hipStreamDestroy(gstream);
    }
};

namespace mtl4g {
// Workspace elements (doubles) one itl::cg item of system size n needs with
// the identity or diagonal preconditioner.
int64_t batch_ws_elems(int32_t n, int precision);
// The same for a method (MTL4G_CG / MTL4G_BICGSTAB) and preconditioner.
int64_t batch_ws_elems(int32_t n, int precision, int method, int pc_kind);
// Launch the persistent kernel over `nitems` items (queue is reset inside):
// method MTL4G_CG or MTL4G_BICGSTAB, preconditioner identity, diagonal or
// ilu_0 (batch_ensure_ilu0 first).
int batch_launch(mtl4g_batch_s* B, int method, int pc_kind, int precision, const BatchShape& sh,
                 const mtl4g_batch_item* d_items, int nitems, const double* d_b, double* d_x,
                 double* d_work, mtl4g_batch_result* d_out, unsigned* d_queue, hipStream_t s);
int batch_ensure_values(mtl4g_batch_s* B, int precision);
// ILU(0) factors of every system (the device factorization of
// itl::pc::ilu_0, one system at a time) laid out for the batched sweeps.
int batch_ensure_ilu0(mtl4g_batch_s* B);
}  // namespace mtl4g
