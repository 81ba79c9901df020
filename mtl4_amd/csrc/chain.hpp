// chain.hpp -- triangular sweeps of grid-structured factors as a pipeline of
// row chains (included by precond.hip).
//
// A triangular factor whose rows depend only on the rows 1, Lc and Lc*Cx
// before them (after them, for an upper factor), with no dependency across
// the edge of a chain of Lc rows or of a plane of Cx chains -- the IC(0) /
// ILU(0) factors of 5- and 7-point stencils in natural order -- is swept
// chain by chain: chain (cx, cy) (rows (cy*Cx + cx)*Lc .. +Lc) handles its
// element e at step e + cx + cy.  A lane owns one chain and keeps its
// previous element in a register; the chains cx-1 and cy-1 hand over their
// element of the previous step through LDS inside a workgroup (a tile of
// 64*TXW chains along x by TY along y, one barrier per step) and through
// memory between workgroups (agent-scope stores, polled reads of the
// reserved pending pattern, issued U steps ahead).  A workgroup therefore
// pays a barrier per step, not a device-wide round trip per level.  Every
// row is computed by one lane with MTL4's subtraction order (ascending
// columns, lower_trisolve.hpp:174-209, upper_trisolve.hpp:161-176), so the
// results are the level-scheduled sweep's bit for bit.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace mtl4g {

constexpr int kChainU = 8;         // steps per unrolled block: loads are issued this far ahead
constexpr int kChainMaxWaves = 8;  // waves per workgroup

struct ChainSweep {
    bool on = false;
    bool lower = true, unit = false;
    int32_t n = 0, Lc = 0, Cx = 0, Cy = 0;
    int ndep = 0;                   // 1..3 offsets: 1, Lc, Lc*Cx
    int txw = 1, ty = 1, tiles_x = 1, tiles_y = 1;
    DevBuf<int32_t> src, dsrc;      // gathered values: positions in the device factor (-1: absent / unit)
    DevBuf<double> val64, dia64;    // val: ndep arrays of n (entry of offset k of row r), absent pattern
    DevBuf<float> val32, dia32;
    DevBuf<uint16_t> val16, dia16;
};

template <typename MT>
struct ChainView {
    const MT* val;   // ndep * n
    const MT* dia;   // n (unused for a unit sweep)
    int32_t n, Lc, Cx, Cy, txw, ty, tiles_x;
};

template <typename T>
__device__ __forceinline__ bool chain_pending(T v);
template <>
__device__ __forceinline__ bool chain_pending<double>(double v) {
    return (uint64_t)__double_as_longlong(v) == 0x7ff5c3c3c3c3c3c3ull;
}

// Buffer (descriptor-based) loads and stores: an offset past the buffer's
// size reads zero and drops a store, so lanes without a row this step run
// the same straight-line code as the others and every load keeps its place
// in the memory counter (vmcnt) -- the loads issued for later steps stay in
// flight across the steps in between.
constexpr int kBufSc1 = 16;   // cache policy: agent (device) scope, as ld_agent / st_agent
typedef unsigned int chain_u2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t chain_rsrc(const void* p, int32_t bytes) {
    // raw buffer: stride 0, 32-bit data format (the fourth descriptor word
    // of gfx9 buffers).  Every input is wave-uniform; readfirstlane keeps the
    // descriptor in scalar registers (a descriptor in vector registers makes
    // the compiler loop over the lanes' values around every access).  The
    // host admits a layout only when its largest buffer fits an int32.
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                             __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
template <int AUX>
__device__ __forceinline__ double chain_ld64(__amdgpu_buffer_rsrc_t r, int32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
}
__device__ __forceinline__ float chain_ld32(__amdgpu_buffer_rsrc_t r, int32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ uint16_t chain_ld16(__amdgpu_buffer_rsrc_t r, int32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
template <typename T>
__device__ __forceinline__ T chain_ld(__amdgpu_buffer_rsrc_t r, int32_t off);
template <>
__device__ __forceinline__ double chain_ld<double>(__amdgpu_buffer_rsrc_t r, int32_t off) { return chain_ld64<0>(r, off); }
template <>
__device__ __forceinline__ float chain_ld<float>(__amdgpu_buffer_rsrc_t r, int32_t off) { return chain_ld32(r, off); }
template <>
__device__ __forceinline__ uint16_t chain_ld<uint16_t>(__amdgpu_buffer_rsrc_t r, int32_t off) { return chain_ld16(r, off); }

// Waits (bounded) until the values another workgroup stores at w[ix]
// (when wx) and w[iy] (when wy) are no longer the pending marker; returns
// them in x and y (the others come back unchanged).  A separate function on
// purpose: the kernel calls it only when a lane of the wave needs it, and
// the kernel's own step keeps no loop (nor the register copies a loop's
// operands need) between a load and its use.
__device__ __noinline__ double2 chain_poll(const double* w, int32_t ix, int32_t iy, bool wx, bool wy, double2 o,
                                           unsigned* abort_flag) {
    unsigned it = 0;
    while (wx || wy) {
        if (++it > (1u << 22) || ((it & 63u) == 0 && ld_agent(abort_flag) != 0u)) {
            st_agent(abort_flag, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (wx) o.x = ld_agent(w + ix);
        if (wy) o.y = ld_agent(w + iy);
        wx = wx && chain_pending(o.x);
        wy = wy && chain_pending(o.y);
    }
    return o;
}

// The pipeline kernel.  LOWER: rows depend on smaller rows (else on larger
// ones: the sweep runs in the mirrored index r' = n-1-r).  NDEP offsets;
// UNIT: no diagonal factor.  v is read as IT, every result is kept in w
// (double) and, when w32 is given, rounded to float there too.  The step
// body is straight-line for every lane: operands of absent entries are
// computed and dropped by a select (the skipped subtraction leaves acc's
// bits as they are), rows outside the chain read and write out of range.
template <typename MT, typename IT, bool LOWER, bool UNIT, int NDEP>
__global__ __launch_bounds__(kChainMaxWaves * kWave) void k_chain(ChainView<MT> C, const IT* __restrict__ v,
                                                                  double* w, float* w32, unsigned* counters,
                                                                  int* err, const int* skip) {
    constexpr int U = kChainU;
    if (skip && *skip) return;
    __shared__ double xch[2][kChainMaxWaves * kWave];   // each lane's element of the last step
    __shared__ int s_tile, s_final;
    if (threadIdx.x == 0)
        s_tile = (int)__hip_atomic_fetch_add(counters, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int tile = s_tile;   // tiles in row-major order: producers come first
    const int tx = tile % C.tiles_x, ty = tile / C.tiles_x;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int wx = wave % C.txw, wy = wave / C.txw;
    const int tw = C.txw * kWave;                       // tile width in chains
    const int gx = wx * kWave + lane;                   // position in the tile's x extent
    const int x0 = tx * tw;                             // first chain of the tile
    const int cx = x0 + gx, cy = ty * C.ty + wy;
    const bool chain = cx < C.Cx && cy < C.Cy;
    const int32_t Lc = C.Lc, plane = C.Lc * C.Cx;
    const int32_t base = (cy * C.Cx + cx) * Lc;         // first row of the chain (mirrored index)
    const int xl = min(C.Cx, x0 + tw) - 1, yl = min(C.Cy, (ty + 1) * C.ty) - 1;
    const int s0 = x0 + ty * C.ty, s1 = (Lc - 1) + xl + yl;
    const bool ext_x = cx == x0 && tx > 0 && NDEP >= 2;      // the chain cx-1 is another tile's
    const bool ext_y = wy == 0 && ty > 0 && NDEP >= 3;       // the chain cy-1 is another tile's
    const int32_t n = C.n;
    constexpr int32_t kOut = 0x7ffffff0;                 // an offset past every buffer
    const __amdgpu_buffer_rsrc_t rv_ = chain_rsrc(v, n * (int32_t)sizeof(IT));
    const __amdgpu_buffer_rsrc_t rval = chain_rsrc(C.val, NDEP * n * (int32_t)sizeof(MT));
    const __amdgpu_buffer_rsrc_t rdia = chain_rsrc(C.dia, n * (int32_t)sizeof(MT));
    const __amdgpu_buffer_rsrc_t rw = chain_rsrc(w, n * 8);
    const __amdgpu_buffer_rsrc_t rw32 = chain_rsrc(w32, n * 4);
    const unsigned* abort_flag = counters + kCounterStride;
    auto row_of = [=](int32_t rm) -> int32_t { return LOWER ? rm : n - 1 - rm; };
    auto dep_row = [=](int32_t r, int32_t d) -> int32_t { return LOWER ? r - d : r + d; };

    // register rings: the operands of the steps s..s+U-1 (slot j: step s+j);
    // the descriptors are passed by value so that they stay scalar
    IT rv[U];
    MT rl[NDEP][U];
    MT rd[U];
    double ex[U], ey[U];
#define M4_CHAIN_ISSUE(ST, J)                                                                              \
    do {                                                                                                  \
        const int e_ = (ST) - cx - cy;                                                                    \
        const bool in_ = chain && e_ >= 0 && e_ < Lc;                                                     \
        const int32_t r_ = row_of(base + (in_ ? e_ : 0));                                                 \
        rv[J] = chain_ld<IT>(rv_, in_ ? r_ * (int32_t)sizeof(IT) : kOut);                                 \
        _Pragma("unroll") for (int k_ = 0; k_ < NDEP; k_++) rl[k_][J] =                                  \
            chain_ld<MT>(rval, in_ ? (k_ * n + r_) * (int32_t)sizeof(MT) : kOut);                         \
        if (!UNIT) rd[J] = chain_ld<MT>(rdia, in_ ? r_ * (int32_t)sizeof(MT) : kOut);                    \
        if (NDEP >= 2) ex[J] = chain_ld64<kBufSc1>(rw, in_ && ext_x ? dep_row(r_, Lc) * 8 : kOut);       \
        if (NDEP >= 3) ey[J] = chain_ld64<kBufSc1>(rw, in_ && ext_y ? dep_row(r_, plane) * 8 : kOut);    \
    } while (0)

#pragma unroll
    for (int j = 0; j < U; j++) M4_CHAIN_ISSUE(s0 + j, j);
    double prev = 0.0;
    int par = 0;
    for (int s = s0; s <= s1; s += U) {
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int st = s + j;
            if (st > s1) break;   // uniform over the workgroup
            const int e = st - cx - cy;
            const bool act = chain && e >= 0 && e < Lc;
            const int32_t r = row_of(base + (act ? e : 0));
            // operands: the previous element, the chain cx-1 (ox) and the
            // chain cy-1 (oy), each of the previous step.  Both selects read
            // LDS unconditionally (the index is clamped into the array).
            const double lx = xch[par ^ 1][wy * tw + max(gx - 1, 0)];
            const double ly = xch[par ^ 1][max(wy - 1, 0) * tw + gx];
            double ox = 0.0, oy = 0.0;
            if (NDEP >= 2) ox = cx > x0 ? lx : ex[j];
            if (NDEP >= 3) oy = wy > 0 ? ly : ey[j];
            bool present[3];
#pragma unroll
            for (int k = 0; k < NDEP; k++) present[k] = act && !dia_absent(rl[k][j]);
            // another tile's value not yet there: poll it (bounded)
            const bool wait_x = NDEP >= 2 && ext_x && present[1] && chain_pending(ox);
            const bool wait_y = NDEP >= 3 && ext_y && present[2] && chain_pending(oy);
            if (__builtin_amdgcn_ballot_w64(wait_x || wait_y)) {
                const double2 o = chain_poll(w, dep_row(r, Lc), NDEP >= 3 ? dep_row(r, plane) : 0, wait_x, wait_y,
                                             make_double2(ox, oy), const_cast<unsigned*>(abort_flag));
                ox = o.x;
                oy = o.y;
            }
            // ascending columns: lower rows subtract offset Lc*Cx, Lc, 1;
            // upper rows 1, Lc, Lc*Cx
            double acc = static_cast<double>(rv[j]);
#pragma unroll
            for (int t = 0; t < NDEP; t++) {
                const int k = LOWER ? NDEP - 1 - t : t;
                const double o = k == 0 ? prev : k == 1 ? ox : oy;
                const double cand = acc - to_arith<double>(rl[k][j]) * o;
                acc = present[k] ? cand : acc;
            }
            if (!UNIT) acc = acc * to_arith<double>(rd[j]);
            acc = chain_pending(acc) ? __longlong_as_double((long long)0x7ffdc3c3c3c3c3c3ull) : acc;
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(chain_u2, acc), rw,
                                                  act ? r * 8 : kOut, 0, kBufSc1);
            if (w32)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)acc), rw32,
                                                      act ? r * 4 : kOut, 0, 0);
            prev = act ? acc : prev;
            xch[par][wy * tw + gx] = acc;
            // a barrier that orders the LDS exchange only: the loads issued
            // for later steps stay in flight across it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            par ^= 1;
            M4_CHAIN_ISSUE(st + U, j);
        }
    }
#undef M4_CHAIN_ISSUE
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned done =
            __hip_atomic_fetch_add(counters + 2 * kCounterStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        s_final = done == (unsigned)gridDim.x;
    }
    __syncthreads();
    if (s_final && threadIdx.x == 0) {   // every other workgroup has finished: reset for the next launch
        if (ld_agent(abort_flag) != 0u) atomicOr(err, 1);
        st_agent(counters, 0u);
        st_agent(const_cast<unsigned*>(abort_flag), 0u);
        st_agent(counters + 2 * kCounterStride, 0u);
    }
}

// value of a stored entry of the chain layout, on the device
template <typename MT>
__global__ void k_fill_chain(int64_t n, int ndep, const int32_t* __restrict__ src, const int32_t* __restrict__ dsrc,
                             const double* __restrict__ fv, MT* __restrict__ val, MT* __restrict__ dia) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ndep * n; i += stride)
        val[i] = src[i] < 0 ? dia_absent_value<MT>() : stored_as<MT>(fv[src[i]]);
    if (dsrc)
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
            dia[i] = dsrc[i] < 0 ? stored_as<MT>(1.0) : stored_as<MT>(fv[dsrc[i]]);
}

}  // namespace mtl4g
