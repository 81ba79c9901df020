// chain.hpp -- triangular sweeps of grid-structured factors as a pipeline of
// row chains (included by precond.hip).
//
// A triangular factor whose rows depend only on the rows 1, Lc and Lc*Cx
// before them (after them, for an upper factor), with no dependency across
// the edge of a chain of Lc rows or of a plane of Cx chains -- the IC(0) /
// ILU(0) factors of 5- and 7-point stencils in natural order -- is swept
// chain by chain: chain (cx, cy) (rows (cy*Cx + cx)*Lc .. +Lc) handles its
// element e at step e + cx + cy.  A lane owns one chain and keeps its
// previous element in a register.  A workgroup is one wave and owns a tile
// of Tx x Ty = 64 chains; inside it the chains cx-1 and cy-1 hand over their
// element of the previous step lane to lane (ds_bpermute: no LDS store, no
// barrier), between tiles through memory (agent-scope stores of the tile's
// edge chains, reads of the reserved pending pattern issued U steps ahead,
// polled when still pending).
//
// Memory access is coalesced although the chains' rows lie Lc apart: the
// factor's values are stored tile by tile in step-major order (slot =
// tile*stile + (step - first step of the tile)*64 + lane), so a step's loads
// read consecutive slots at a scalar offset; the input and output vectors keep their natural
// order and move through LDS in blocks of U steps (8 consecutive rows of a
// chain per 8 lanes).  Every row is computed by one lane with MTL4's
// subtraction order (ascending columns, lower_trisolve.hpp:174-209,
// upper_trisolve.hpp:161-176), so the results are the level-scheduled
// sweep's bit for bit.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace mtl4g {

constexpr int kChainU = 8;     // steps per block: v and w staged per block, values issued this far ahead
constexpr int kChainPad = 8;   // LDS row padding of the staging arrays (bank spread)

struct ChainSweep {
    bool lower = true, unit = false;
    int32_t n = 0, Lc = 0, Cx = 0, Cy = 0;
    int ndep = 0;                   // 1..3 offsets: 1, Lc, Lc*Cx
    int Tx = 64, Ty = 1, tiles_x = 1, tiles_y = 1;
    int64_t stile = 0, nslots = 0;  // slots per tile, slots in all
    DevBuf<int32_t> src, dsrc;      // gathered values: positions in the device factor (-1: absent / unit)
    DevBuf<double> val64, dia64;    // val: ndep arrays of nslots (entry of offset k), absent pattern
    DevBuf<float> val32, dia32;
    DevBuf<uint16_t> val16, dia16;
};

template <typename MT>
struct ChainView {
    const MT* val;   // ndep * nslots, tile-local step-major
    const MT* dia;   // nslots (unused for a unit sweep)
    int32_t n, Lc, Cx, Cy, Tx, Ty, tiles_x, stile, nslots;
};

// Slot of row r of the sweep's own index (rm = r for a lower sweep,
// n-1-r for an upper one) in the tile-local step-major layout.
__host__ __device__ inline int64_t chain_slot(int64_t rm, int64_t Lc, int64_t Cx, int64_t Tx, int64_t Ty,
                                              int64_t tiles_x, int64_t stile) {
    const int64_t e = rm % Lc, q = rm / Lc, cx = q % Cx, cy = q / Cx;
    const int64_t tx = cx / Tx, ty = cy / Ty;
    const int64_t l = (cy - ty * Ty) * Tx + (cx - tx * Tx);
    return (ty * tiles_x + tx) * stile + (e + cx + cy - tx * Tx - ty * Ty) * (Tx * Ty) + l;
}

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
__device__ __forceinline__ double chain_ld64(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff = 0) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, AUX));
}
__device__ __forceinline__ float chain_ld32(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}
__device__ __forceinline__ uint16_t chain_ld16(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, 0);
}
// off: per-lane byte offset (range-checked against the buffer's size);
// soff: a wave-uniform byte offset added to it
template <typename T>
__device__ __forceinline__ T chain_ld(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff = 0);
template <>
__device__ __forceinline__ double chain_ld<double>(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff) {
    return chain_ld64<0>(r, off, soff);
}
template <>
__device__ __forceinline__ float chain_ld<float>(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff) {
    return chain_ld32(r, off, soff);
}
template <>
__device__ __forceinline__ uint16_t chain_ld<uint16_t>(__amdgpu_buffer_rsrc_t r, int32_t off, int32_t soff) {
    return chain_ld16(r, off, soff);
}

// The value v of lane l-1 (lane 0: 0), for every lane: a DPP shift of the
// whole wave by one lane, no LDS.
__device__ __forceinline__ double chain_from_prev_lane(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    // bound_ctrl: lane 0 (no source) reads 0, no initial value to set
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The value v of lane src (src4 = 4*src), for every lane: two ds_bpermute
// (the LDS crossbar, no LDS storage).
__device__ __forceinline__ double chain_from(double v, int src4) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(src4, (int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(src4, (int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The pipeline kernel, one wave per tile.  LOWER: rows depend on smaller
// rows (else on larger ones: the sweep runs in the mirrored index
// r' = n-1-r).  NDEP offsets; UNIT: no diagonal factor.  v is read as IT,
// every result is kept in w (double) and, when w32 is given, rounded to
// float there too.  The step body is straight-line for every lane: operands
// of absent entries are computed and dropped by a select (the skipped
// subtraction leaves acc's bits as they are); lanes without a row this step
// compute on padding.
template <typename MT, typename IT, bool LOWER, bool UNIT, int NDEP>
__global__ __launch_bounds__(kWave) void k_chain(ChainView<MT> C, const IT* __restrict__ v, double* w, float* w32,
                                                 unsigned* counters, int* err, const int* skip) {
    constexpr int U = kChainU;
    constexpr int RW = kWave + kChainPad;   // staging row: one step of the block, all chains
    if (skip && *skip) return;
    __shared__ double vst[U * RW];          // v of the next block: [step][chain]
    __shared__ double wst[U * RW];          // results of the block: [step][chain]
    const int l = threadIdx.x;
    unsigned tk = 0;
    if (l == 0) tk = __hip_atomic_fetch_add(counters, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int tile = (int)__builtin_amdgcn_readfirstlane(tk);   // tiles in row-major order: producers first
    const int tx = tile % C.tiles_x, ty = tile / C.tiles_x;
    const int x0 = tx * C.Tx, y0 = ty * C.Ty;
    const int gx = l % C.Tx, gy = l / C.Tx;
    const int cx = x0 + gx, cy = y0 + gy;
    const bool chain = cx < C.Cx && cy < C.Cy;
    const int32_t Lc = C.Lc, plane = C.Lc * C.Cx, n = C.n;
    const int32_t base = chain ? (cy * C.Cx + cx) * Lc : 0;   // first row of the chain (mirrored index)
    const int e0 = cx + cy;                                   // step of the chain's first element
    const int xl = min(C.Cx, x0 + C.Tx) - 1, yl = min(C.Cy, y0 + C.Ty) - 1;
    const int s0 = x0 + y0, s1 = (Lc - 1) + xl + yl;
    const bool ext_x = NDEP >= 2 && chain && gx == 0 && tx > 0;   // the chain cx-1 is another tile's
    const bool ext_y = NDEP >= 3 && chain && gy == 0 && ty > 0;   // the chain cy-1 is another tile's
    const int srcy = max(l - C.Tx, 0) * 4;
    constexpr int32_t kOut = 0x7ffffff0;                          // an offset past every buffer
    const __amdgpu_buffer_rsrc_t rv_ = chain_rsrc(v, n * (int32_t)sizeof(IT));
    const __amdgpu_buffer_rsrc_t rval = chain_rsrc(C.val, NDEP * C.nslots * (int32_t)sizeof(MT));
    const __amdgpu_buffer_rsrc_t rdia = chain_rsrc(C.dia, C.nslots * (int32_t)sizeof(MT));
    const __amdgpu_buffer_rsrc_t rw = chain_rsrc(w, n * 8);
    // no w32: a descriptor of size 0, every store through it is dropped (no
    // branch around the store keeps the wait counts exact)
    const __amdgpu_buffer_rsrc_t rw32 = chain_rsrc(w32, w32 ? n * 4 : 0);
    const unsigned* abort_flag = counters + kCounterStride;
    auto row_of = [=](int32_t rm) -> int32_t { return LOWER ? rm : n - 1 - rm; };
    // byte offsets of this lane's slot at the tile's first step; a step adds
    // the wave-uniform (step - s0) * 64 * sizeof(MT), clamped to the tile
    const int32_t vo = (tile * C.stile + l) * (int32_t)sizeof(MT);
    // rows of the other tiles' elements: mirrored index of element e of the
    // chain cx-1 (cy-1) is hx + st (hy + st) at step st
    const int32_t hx = base - Lc - e0, hy = base - plane - e0;

    // Block staging: lane l moves element jj = l % 8 of the chains
    // c = l/8 + 8i (i = 0..7) between the natural-order vectors and LDS.
    const int jj = l & 7;
    auto stage_row = [=](int S, int i, bool& ok) -> int32_t {
        const int c = (l >> 3) + i * 8;
        const int ccx = x0 + c % C.Tx, ccy = y0 + c / C.Tx;
        const int e = S + jj - ccx - ccy;
        // bitwise tests, no short-circuit branches: every path issues the
        // same memory operations, which keeps the wait counts exact
        ok = (ccx < C.Cx) & (ccy < C.Cy) & ((unsigned)e < (unsigned)Lc);
        return row_of(ok ? (ccy * C.Cx + ccx) * Lc + e : 0);
    };
    IT vin[U];       // v of the block after the next one, in flight
    double vb[U];    // v of this block (this lane's chain)
    double res[U];   // results of this block (this lane's chain)
    auto stage_issue = [&](int S) {
#pragma unroll
        for (int i = 0; i < U; i++) {
            bool ok;
            const int32_t r = stage_row(S, i, ok);
            vin[i] = chain_ld<IT>(rv_, ok ? r * (int32_t)sizeof(IT) : kOut);
        }
    };
    // vin -> LDS -> vb.  An input equal to the pending pattern is quieted
    // here: a result that is v itself (no subtraction, unit diagonal) must
    // not read as pending, and arithmetic never yields that signalling NaN.
    auto stage_next = [&]() {
#pragma unroll
        for (int i = 0; i < U; i++) {
            const double vi = static_cast<double>(vin[i]);
            vst[jj * RW + (l >> 3) + i * 8] =
                chain_pending(vi) ? __longlong_as_double((long long)0x7ffdc3c3c3c3c3c3ull) : vi;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#pragma unroll
        for (int j = 0; j < U; j++) vb[j] = vst[j * RW + l];
    };

    // register rings: the values of the steps s..s+U-1 (slot j: step s+j),
    // and the other tiles' elements they need
    MT rl[NDEP][U];
    MT rd[U];
    double ex[U], ey[U];
#define M4_CHAIN_ISSUE(ST, J)                                                                              \
    do {                                                                                                  \
        const int32_t so_ = (min((ST), s1) - s0) * (kWave * (int32_t)sizeof(MT));                        \
        _Pragma("unroll") for (int k_ = 0; k_ < NDEP; k_++) rl[k_][J] =                                  \
            chain_ld<MT>(rval, vo + k_ * C.nslots * (int32_t)sizeof(MT), so_);                            \
        if (!UNIT) rd[J] = chain_ld<MT>(rdia, vo, so_);                                                   \
        const bool a_ = (unsigned)((ST) - e0) < (unsigned)Lc;                                             \
        if (NDEP >= 2) ex[J] = chain_ld64<kBufSc1>(rw, a_ && ext_x ? row_of(hx + (ST)) * 8 : kOut);      \
        if (NDEP >= 3) ey[J] = chain_ld64<kBufSc1>(rw, a_ && ext_y ? row_of(hy + (ST)) * 8 : kOut);      \
    } while (0)

    stage_issue(s0);
    stage_next();
    stage_issue(s0 + U);
#pragma unroll
    for (int j = 0; j < U; j++) M4_CHAIN_ISSUE(s0 + j, j);
    double prev = 0.0;   // the chain's element of the last step; the neighbours read it
    // two blocks per iteration: the loads issued in the first block for the
    // second stay inside one iteration, whose wait counts the compiler keeps
    // exact; only those carried into the next iteration may wait longer.
    // Steps past s1 run on padding and store nothing.
    for (int S2 = s0; S2 <= s1; S2 += 2 * U)
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const int S = S2 + half * U;
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int st = S + j;
            // steps past s1 (the last block's tail) run on padding and
            // store nothing: no exit inside the block keeps the wait counts
            // of the loads issued ahead exact
            // operands: the previous element, the chain cx-1 (ox) and the
            // chain cy-1 (oy), each of the previous step; the lane exchange
            // is issued first, it is on every step's dependency chain
            double nx, ny;
            if (NDEP >= 2) nx = chain_from_prev_lane(prev);   // lane l-1 is the chain cx-1 when gx > 0
            if (NDEP >= 3) ny = chain_from(prev, srcy);
            bool present[3];
#pragma unroll
            for (int k = 0; k < NDEP; k++) present[k] = !dia_absent(rl[k][j]);   // padding slots: absent
            // a lane at the tile's edge takes the other tile's value loaded U
            // steps ago.  The row is computed with it at once; a value still
            // pending (tested on the loaded value, off the exchange's chain)
            // is polled (bounded) in a branch that computes the row again,
            // so the only value merged after the branch is the row's result.
            const double ex_j = NDEP >= 2 ? ex[j] : 0.0, ey_j = NDEP >= 3 ? ey[j] : 0.0;
            const bool wait_x = NDEP >= 2 && (ext_x & present[1] & chain_pending(ex_j));
            const bool wait_y = NDEP >= 3 && (ext_y & present[2] & chain_pending(ey_j));
            // ascending columns: lower rows subtract offset Lc*Cx, Lc, 1;
            // upper rows 1, Lc, Lc*Cx
            auto row = [&](double qx, double qy) -> double {
                const double ox = gx > 0 ? nx : qx;
                const double oy = gy > 0 ? ny : qy;
                double a = vb[j];
#pragma unroll
                for (int t = 0; t < NDEP; t++) {
                    const int k = LOWER ? NDEP - 1 - t : t;
                    const double o = k == 0 ? prev : k == 1 ? ox : oy;
                    const double cand = a - to_arith<double>(rl[k][j]) * o;
                    a = present[k] ? cand : a;
                }
                if (!UNIT) a = a * to_arith<double>(rd[j]);
                return a;
            };
            double acc = row(ex_j, ey_j);
            if (__builtin_expect(wait_x | wait_y, 0)) {
                double px = ex_j, py = ey_j;
                bool wx = wait_x, wy = wait_y;
                unsigned it = 0;
                while (wx || wy) {
                    if (++it > (1u << 22) || ((it & 63u) == 0 && ld_agent(abort_flag) != 0u)) {
                        st_agent(const_cast<unsigned*>(abort_flag), 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    if (wx) px = ld_agent(w + row_of(hx + st));
                    if (NDEP >= 3 && wy) py = ld_agent(w + row_of(hy + st));
                    wx = wx && chain_pending(px);
                    wy = wy && chain_pending(py);
                }
                acc = row(px, py);
            }
            res[j] = acc;
            // a lane without a row this step computes on padding; its value
            // is never read: the neighbours that would read it have no row
            // this step either, and its own next row starts a chain
            prev = acc;
            M4_CHAIN_ISSUE(st + U, j);
            // keep the loads where they are issued: moved down to the step
            // that uses them, they would no longer be in flight for U steps
            __builtin_amdgcn_sched_barrier(0);
        }
        // the block's results to w (8 consecutive rows of a chain per 8
        // lanes), the next block's v from LDS, the one after it in flight.
        // The rows are stored at agent scope (through to memory): the edge
        // chains' rows are read by other tiles, which poll each value for
        // the pending pattern with agent-scope loads (the inter-workgroup
        // protocol of internal.hpp; no announcement needed).  One store per
        // row keeps a block's vector-memory operations well under what the
        // wait counter tracks, so the loads issued ahead stay in flight.
#pragma unroll
        for (int j = 0; j < U; j++) wst[j * RW + l] = res[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#pragma unroll
        for (int i = 0; i < U; i++) {
            bool ok;
            const int32_t r = stage_row(S, i, ok);
            const double y = wst[jj * RW + (l >> 3) + i * 8];
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(chain_u2, y), rw, ok ? r * 8 : kOut, 0,
                                                  kBufSc1);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)y), rw32, ok ? r * 4 : kOut, 0,
                                                  0);
        }
        stage_next();
        stage_issue(S + 2 * U);
        __builtin_amdgcn_sched_barrier(0);
    }
#undef M4_CHAIN_ISSUE
    wait_memory();
    if (l == 0) {
        const unsigned done =
            __hip_atomic_fetch_add(counters + 2 * kCounterStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (done == (unsigned)gridDim.x) {   // every other tile has finished: reset for the next launch
            if (ld_agent(abort_flag) != 0u) atomicOr(err, 1);
            st_agent(counters, 0u);
            st_agent(const_cast<unsigned*>(abort_flag), 0u);
            st_agent(counters + 2 * kCounterStride, 0u);
        }
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
