// precond.hpp -- preconditioner handle and the level-scheduled sweep.
#pragma once

#include <memory>

#include "internal.hpp"

namespace mtl4g {

// Sweep counters live one per 128-byte line: waiting workgroups poll the
// counter of one level, and neighbouring levels' counters must not share
// the line that is being incremented.
constexpr int kCounterStride = 32;
__host__ __device__ inline unsigned* sweep_counter(unsigned* base, int i) {
    return base + (size_t)i * kCounterStride;
}
// counter words of a sweep with nlev levels: ticket, abort flag, finished
// workgroups, then one per level
inline size_t sweep_counter_words(int32_t nlev) { return ((size_t)nlev + 3) * kCounterStride; }

// One triangular sweep  w[r] = (v[r] - sum_j T[r][c_j] * w[c_j]) * dia[r]
// stored in level order: position k of the sweep handles row perm[k]; the
// off-diagonal entries of that row follow in the order MTL4 subtracts them.
struct ChainSweep;

struct Sweep {
    int32_t n = 0, nlev = 0, nblocks = 0;
    int64_t nnz = 0;
    bool unit = false;
    DevBuf<int32_t> perm, rowptr, col, lev_ptr, lev_blk, blk_level;
    DevBuf<int32_t> src, dsrc;   // values gathered from a device factor: entry / diagonal sources
    DevBuf<double> val64, dia64;
    DevBuf<float> val32, dia32;
    DevBuf<uint16_t> val16, dia16;   // bfloat16 storage (MTL4G_BF16)
    DevBuf<unsigned> counters;   // default counters (sweep_counter_words)
    std::vector<int32_t> h_lev_ptr;
    std::unique_ptr<ChainSweep> chain;   // grid-structured factor: the chain pipeline runs instead
    Sweep();
    ~Sweep();
};

// Host description of a sweep before upload: rows in natural order.
struct SweepRows {
    int32_t n = 0;
    bool lower = true;     // dependencies point to smaller (true) or larger rows
    bool unit = false;
    std::vector<int32_t> rowptr, col;   // off-diagonal entries in subtraction order
    std::vector<double> val;
    std::vector<double> dia;            // stored (inverse) diagonal, unused if unit
    // instead of val / dia: positions in a device array of factor values
    // (filled later by fill_sweep); dsrc -1 for a unit diagonal
    std::vector<int32_t> src, dsrc;
};

int upload_sweep(const SweepRows& rows, int value_format, Sweep* S);
// val / dia of a sweep uploaded with src positions: gathered from fval
// (rounded to float for MTL4G_F32 storage)
int fill_sweep(Sweep& S, int value_format, const double* fval, hipStream_t s);

// counters: sweep_counter_words(nlev) zeroed words owned by the caller (NULL =
// the sweep's own); each launch leaves them zeroed again.
// The row arithmetic runs in VT; v is read as IT; w32 (may be null) also
// receives the results rounded to float.
template <typename IT, typename VT>
int launch_sweep(const Sweep& S, int value_format, const IT* v, VT* w, float* w32, const int* skip, int* err,
                 unsigned* counters, hipStream_t s);

}  // namespace mtl4g

struct mtl4g_pc_s {
    int kind = MTL4G_PC_IDENTITY;
    int fmt = MTL4G_F64;
    int32_t n = 0;
    const mtl4g_matrix_s* A = nullptr;
    mtl4g::DevBuf<double> inv_diag64;
    mtl4g::DevBuf<float> inv_diag32;
    mtl4g::DevBuf<uint16_t> inv_diag16;   // bfloat16 storage
    mtl4g::Sweep lower, upper;            // solve: lower then upper
    mtl4g::Sweep adj_lower, adj_upper;    // adjoint solve (built on demand)
    bool has_adjoint = false;
    // host copies of the factors as MTL4 stores them (natural row order);
    // the values of ILU(0)/IC(0) come from the device factor on demand
    std::vector<double> h_inv_diag;
    std::vector<int32_t> L_rowptr, L_col, U_rowptr, U_col;
    std::vector<double> L_val, U_val;
    bool host_values = false;
    mtl4g::DevBuf<int> err;               // device error word of the sweeps
    // device factorization (ILU(0): values in A's CSR order; IC(0): in the
    // order of the upper factor U), kept for mtl4g_pc_refactor
    mtl4g::DevBuf<double> fval, finv;     // factor values, inverse diagonal per row
    mtl4g::DevBuf<int32_t> fpos;          // ILU(0): diagonal position per row; IC(0): A position per U entry
    mtl4g::DevBuf<int32_t> urowptr;       // IC(0): row pointers of U
    mtl4g::DevBuf<unsigned> fcounters;    // level counters of the factorization kernel
    std::vector<int32_t> L_src, U_src;    // positions in fval of the L / U entries (natural order)
    mtl4g::DevBuf<int32_t> dpos;          // diagonal: position of A[i][i] or -1
};

namespace mtl4g {
// z = P^{-1} r on the device.  y0 is caller scratch of n values (ilu/ic
// with double vectors); w64 is caller scratch of 2n doubles (ilu/ic with
// float vectors, whose sweeps run in FP64); skip (may be null) makes every
// kernel a no-op when *skip != 0; counters (may be null) are
// pc_counter_words(P) zeroed words of the caller.
template <typename VT>
int pc_apply(const mtl4g_pc_s* P, const VT* r, VT* z, VT* y0, const int* skip,
             unsigned* counters, hipStream_t s, double* w64 = nullptr);
size_t pc_counter_words(const mtl4g_pc_s* P);
int pc_check_error(const mtl4g_pc_s* P, hipStream_t s);
}  // namespace mtl4g
