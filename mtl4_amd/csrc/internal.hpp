// internal.hpp -- shared host/device pieces of the MI355X sparse Krylov library.
//
// Layout in HBM (one GPU):
//   CSR: int32 rowptr[n+1], int32 col[nnz], double val[nnz] (plus optional
//        float / bfloat16 copies of val for the inner solves of mixed
//        precision), rows sorted by column as in mtl::compressed2D.
//   Row blocks: int32 blk[nb+1] -- the SpMV partition; block b covers rows
//        [blk[b], blk[b+1]) with at most rows_per_block rows and at most
//        nnz_per_block non-zeros, or a single longer row.
//   Solver control: one Ctl struct (scalars of the recurrence, iteration
//        counter, stop flag, reduction tickets) and a partials array.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mtl4g.h"

namespace mtl4g {

// ----------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int hip_status(hipError_t e, const char* what, const char* file, int line);
// MTL4G_ORDER_TREE or MTL4G_ORDER_MTL4 (mtl4g_set_reduction_order)
int reduction_order();

#define M4_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) return ::mtl4g::hip_status(e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define M4_TRY(expr)                \
    do {                            \
        int s_ = (expr);            \
        if (s_ != MTL4G_OK) return s_; \
    } while (0)

constexpr int kWave = 64;      // CDNA wavefront
constexpr int kXcd = 8;        // MI355X: 8 XCDs, workgroups dealt round robin
constexpr int kMaxPartials = 1 << 16;

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device buffer owned by a handle.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        free();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) { p = nullptr; return hip_status(e, "hipMalloc", __FILE__, __LINE__); }
        n = count;
        return MTL4G_OK;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// ----------------------------------------------------------------- device helpers
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// float -> bfloat16 bits, round to nearest even; NaN stays NaN
__host__ __device__ inline uint16_t bf16_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
// the value of bfloat16 bits, on the host
inline double bf16_value(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

template <typename MT>
__device__ __forceinline__ double val_f64(MT v) { return static_cast<double>(v); }

// A stored value (double, float, or bfloat16 bits) in the arithmetic type T.
template <typename T, typename MT>
__device__ __forceinline__ T to_arith(MT v) { return static_cast<T>(v); }
template <>
__device__ __forceinline__ float to_arith<float, uint16_t>(uint16_t v) { return bf16_to_f32(v); }
template <>
__device__ __forceinline__ double to_arith<double, uint16_t>(uint16_t v) { return (double)bf16_to_f32(v); }

template <typename AT, typename MT>
__device__ __forceinline__ AT load_val(const MT* p, int64_t j) { return to_arith<AT>(p[j]); }

template <typename AT, typename MT>
__device__ __forceinline__ AT load_val_nt(const MT* p, int64_t j) {
    return to_arith<AT>(__builtin_nontemporal_load(p + j));
}

// Ordering between workgroups of one kernel.  A value handed to another
// workgroup (a partial sum, a level's or a chain's result, a factor row) is
// written with an agent-scope atomic or agent-scope store, and read with
// agent-scope loads: on gfx950 both are performed at the device-coherent
// level (the sc1 cache policy), never in a per-XCD L2 line that another XCD
// cannot see.  The write that announces such values (a ticket increment, a
// level counter) is issued only after s_waitcnt 0 (wait_memory) has seen
// every earlier write of the wave complete, i.e. acknowledged by that
// level; a workgroup that reads the announcement with an agent-scope load
// therefore reads the values too.  This is the order a release/acquire pair
// at agent scope gives; it is built by hand because an agent-scope release
// on gfx950 also writes back the whole L2 (buffer_wbl2), which cost six
// times a CG iteration.  Readers that poll a value itself (the pending
// patterns of the triangular sweeps) need no announcement: the value and
// its validity are one 8-byte write.  Every use site names this protocol.
//
// Agent-scope loads/stores for values exchanged between workgroups
// inside one kernel (partials, level flags); they bypass non-coherent
// per-XCD cache lines.
template <typename T>
__device__ __forceinline__ T ld_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Butterfly sum over the 64 lanes: every lane ends with the same value and
// the association order is fixed, so results do not depend on timing.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block sum with a fixed order; result valid in every thread.
template <int BS, typename T>
__device__ __forceinline__ T block_sum(T v, T* sh /* BS/64 */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T r = sh[0];
#pragma unroll
    for (int i = 1; i < BS / kWave; i++) r += sh[i];
    return r;
}

// Publish a value for other workgroups of the same kernel: an agent-scope
// atomic exchange is performed at the device-coherent level, so no L2
// write-back fence (buffer_wbl2, which a __threadfence() costs every
// workgroup on a multi-XCD part) is needed.
__device__ __forceinline__ void publish(double* p, double v) {
    (void)__hip_atomic_exchange(reinterpret_cast<unsigned long long*>(p),
                                (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish(float* p, float v) {
    (void)__hip_atomic_exchange(reinterpret_cast<unsigned*>(p), __float_as_uint(v),
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until this wave's memory operations (the publishes above) have
// completed, before the atomic that announces them (the hand-made release
// of the protocol above).
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Last-block pattern: lane 0 of every block publishes its partials, then the
// block that arrives last (ticket nblocks-1) reduces them in index order with
// agent-scope loads and resets the ticket for the next launch (no memset
// needed, graph-replay safe).  Ordering: the inter-workgroup protocol above
// (publish, wait_memory, ticket; the last block reads with agent-scope loads).
template <int BS>
__device__ __forceinline__ bool arrive_last(unsigned* ticket, unsigned nblocks) {
    __shared__ int am_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        wait_memory();
        unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == nblocks - 1);
    }
    __syncthreads();
    if (am_last && threadIdx.x == 0) st_agent(ticket, 0u);
    return am_last != 0;
}

// Sum of partials[0..n) in a fixed order, valid in all threads.  Each lane
// issues U independent agent-scope loads before it adds them, so the last
// workgroup pays one memory round trip per U*BS partials rather than one
// per BS.
template <int BS, typename T>
__device__ __forceinline__ T reduce_partials(const T* partials, int n, int stride, T* sh) {
    constexpr int U = 16;
    T s = 0;
    for (int base = threadIdx.x; base < n; base += BS * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = base + u * BS;
            v[u] = i < n ? ld_agent(partials + (size_t)i * stride) : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) s += v[u];
    }
    return block_sum<BS>(s, sh);
}

// XCD-aware block order: hardware deals workgroup g to XCD g % 8; this maps
// it to logical block (g % 8) * per + g / 8 so each XCD walks a contiguous
// range of rows and the gathers of neighbouring rows share its L2.
__device__ __forceinline__ int xcd_logical_block(int g, int per_xcd) {
    return (g % kXcd) * per_xcd + g / kXcd;
}

// ----------------------------------------------------------------- CSR core
// Row block descriptor: rows [r0, r1), non-zeros [j0, j1).
struct RowBlock {
    int32_t r0, r1, j0, j1;
};

// One workgroup, rows [r0, r1) of a row block whose non-zeros fit the LDS
// tile: all lanes stream val/col coalesced and form the products into LDS,
// then lane t sums row r0+t in column order (the order of
// mat_vec_mult.hpp:656-682, so FP64 results match MTL4 bit for bit).
// Long single rows (nnz > TILE) are summed by the whole workgroup.
// The row bounds and the epilogue's per-row operands (epi.preload) are
// requested together with the value/index stream, so a workgroup waits on
// three dependent memory round trips: descriptor, stream, gathers.
template <int BS, int TILE, typename AT, typename MT, typename XT, typename Epi>
__device__ __forceinline__ AT csr_block_rows(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x, const RowBlock& d,
                                             AT* prod, AT* sh, int& row, const Epi& epi,
                                             typename Epi::Pre& pre) {
    const int t = threadIdx.x;
    row = d.r0 + t;
    const int j0 = d.j0;
    const int nz = d.j1 - d.j0;
    int a = 0, b = 0;
    if (row < d.r1) {
        a = rowptr[row] - j0;
        b = rowptr[row + 1] - j0;
        pre = epi.preload(row);
    }
    if (nz > TILE) {  // one long row: strided partial sums, fixed-order combine
        AT s = 0;
        for (int k = t; k < nz; k += BS)
            s += load_val<AT>(val, (int64_t)j0 + k) * static_cast<AT>(x[col[j0 + k]]);
        s = block_sum<BS>(s, sh);
        return s;   // valid for lane 0 == the row
    }
    // The whole tile is requested at once: U = TILE/BS column indices and
    // values per lane (non-temporal: streamed once, they should not evict the
    // x window from L2), then U independent gathers of x.
    constexpr int U = TILE / BS;
    int c[U];
    AT v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int k = t + u * BS;
        c[u] = k < nz ? __builtin_nontemporal_load(col + j0 + k) : 0;
        v[u] = k < nz ? load_val_nt<AT>(val, (int64_t)j0 + k) : AT(0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int k = t + u * BS;
        if (k < nz) prod[k] = v[u] * static_cast<AT>(x[c[u]]);
    }
    __syncthreads();
    AT s = 0;
    for (int k = a; k < b; k++) s += prod[k];
    return s;
}

}  // namespace mtl4g

// ----------------------------------------------------------------- handles
struct mtl4g_matrix_s {
    int device = 0;
    int32_t n = 0, m = 0;
    int64_t nnz = 0;
    mtl4g::DevBuf<int32_t> rowptr, col;
    mtl4g::DevBuf<double> val;
    mtl4g::DevBuf<float> val32;       // FP32 copy (created on demand)
    mtl4g::DevBuf<uint16_t> valbf;    // BF16 copy (created on demand)
    mtl4g::DevBuf<double> diag;       // A[i][i] (0 where absent)
    mtl4g::DevBuf<mtl4g::RowBlock> desc;   // row-block partition (LDS tile kernel)
    int nb = 0;                            // number of row blocks
    int kernel = MTL4G_SPMV_VECTOR;        // CSR kernel: vector groups or LDS tiles
    int gs = 8;                            // vector kernel: lanes per row
    int vrows = 256;                       // vector kernel: rows per workgroup
    int rows_per_block = 0, nnz_per_block = 0;
    int cus = 256;                         // compute units of the device
    int per_cu = 4;                        // stream kernel: workgroups per CU
    int uv = 8;                            // scalar kernel: values requested per lane-step
    int per_cu_lowp[2] = {0, 0};           // the same for FP32 / BF16 values (0: as FP64)
    bool lowp_tuned[2] = {false, false};
    bool lowp_wide[2] = {false, false};    // FP32 / BF16 stream kernel on the wide partition
    mtl4g::DevBuf<mtl4g::RowBlock> desc_wide;   // blocks of <= 512 rows / 4096 non-zeros
    int nb_wide = 0;
    // sliced ELL (MTL4G_SPMV_SLICED): slices of 64 rows, column-major, padding column -1
    std::vector<int64_t> h_sl_off;          // nsl + 1 entry offsets
    mtl4g::DevBuf<int64_t> sl_off;
    mtl4g::DevBuf<int32_t> sl_col;
    mtl4g::DevBuf<double> sl_val;
    mtl4g::DevBuf<float> sl_val32;
    mtl4g::DevBuf<uint16_t> sl_valbf;
    int nsl = 0;
    // diagonal layout (MTL4G_SPMV_DIA): ndiag values per row, slice-major, absent entries marked
    std::vector<int32_t> h_dia_off;         // column - row of each diagonal, ascending
    mtl4g::DevBuf<int32_t> dia_off;
    mtl4g::DevBuf<double> dia_val;
    mtl4g::DevBuf<float> dia_val32;
    mtl4g::DevBuf<uint16_t> dia_valbf;
    int ndiag = 0;
    // symmetric diagonal layout (MTL4G_SPMV_DIA_SYM): the diagonals with offset >= 0 of a
    // matrix equal to its transpose; dsym_off: all offsets ascending, dsym_map: stored index of |offset|
    std::vector<int32_t> h_dsym_off;
    mtl4g::DevBuf<int32_t> dsym_off, dsym_map;
    mtl4g::DevBuf<double> dsym_val;
    mtl4g::DevBuf<float> dsym_val32;
    mtl4g::DevBuf<uint16_t> dsym_valbf;
    int ndsym = 0, nsym = 0;
    // column panels (MTL4G_SPMV_PANEL): for each panel of pan_w columns, the
    // CSR of the entries in those columns (rows in order, columns ascending),
    // panel after panel; absolute entry indices
    int npan = 0;
    int32_t pan_w = 0;
    int pan_tile = 2048;                    // non-zeros per row block (1024 or 2048)
    std::vector<int32_t> h_pan_src;         // CSR index of each panel entry
    std::vector<int> h_pan_nb;              // row blocks of each panel
    mtl4g::DevBuf<int32_t> pan_rowptr;      // npan * (n + 1)
    mtl4g::DevBuf<int32_t> pan_col;
    mtl4g::DevBuf<double> pan_val;
    mtl4g::DevBuf<float> pan_val32;
    mtl4g::DevBuf<uint16_t> pan_valbf;
    mtl4g::DevBuf<mtl4g::RowBlock> pan_desc;
    int sl_unroll = 8;                      // sliced / diagonal: entries per lane per chunk (4 or 8)
    int sl_unroll_lowp[2] = {8, 8};         // the same for FP32 / BF16 values
    int32_t max_row = 0;
    std::vector<int32_t> h_rowptr, h_col;   // host pattern (factorizations)
    std::vector<double> h_val;
};

namespace mtl4g {
// Row-block partition (host).
int build_row_blocks(mtl4g_matrix_s* A, int rows_per_block, int nnz_per_block);
int ensure_lowp_values(mtl4g_matrix_s* A, int value_format, hipStream_t s);

// SpMV launchers used by several translation units.
int launch_spmv(const mtl4g_matrix_s* A, const double* x, double* y, int assign, hipStream_t s);
int launch_spmv_f32(const mtl4g_matrix_s* A, int fmt, const float* x, float* y, int assign,
                    hipStream_t s);
int blocks_for(int64_t n);   // grid size of the vector kernels

// A roctx range over one C-ABI call, named as MTL4's vampir_trace scope of
// the same operation (interface/vpt.cpp: "crs_cvec_mult" :201, "cg" :353,
// "bicgstab" :355, the other solvers :356-360, the preconditioners
// :114, :313-329).  rocprofv3 --marker-trace records it; without a tool
// attached a push and a pop cost a few tens of nanoseconds.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};
}  // namespace mtl4g
