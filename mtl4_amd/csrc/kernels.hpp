// kernels.hpp -- the kernel shapes every operation of the library is built
// from, each parameterized by an epilogue functor:
//
//   CSR kernels (SpMV and everything fused into it: residuals, dots of the
//   product with another vector), all summing each row in column order:
//     k_csr:  one workgroup per row block, products staged in LDS;
//     k_csrp: the same row blocks walked by a persistent grid that keeps the
//             next block's stream in flight;
//     k_csrs: one lane per row;
//     k_csrv: a group of lanes per row, products added through shuffles;
//   k_vec: one workgroup per contiguous chunk of a vector (axpy-type
//          updates fused with their reductions);
//   k_ordered: the final sums of a reduction in MTL4's own order.
//
// Epilogue concept:
//   static constexpr int NP;                       // partial sums produced
//   bool skip() const;                             // uniform early exit
//   void prepare();                                // load scalars once per thread
//   k_csr:  struct Pre; Pre preload(int row) const; // per-row operands, fetched
//           void operator()(int row, AT s, const Pre&, AT* part) const;
//                                                   // with the matrix stream
//   k_vec:  void operator()(int64_t i, AT* part) const;
//   void finish(const AT* totals) const;           // last workgroup, lane 0
// Partial sums are combined in block-index order by the last workgroup to
// arrive, so every reduction is deterministic and needs no memset.
#pragma once

#include "internal.hpp"

namespace mtl4g {

constexpr int kMaxTile = 4096;
constexpr int kFanIn = 128;                                  // workgroups per ticket group
constexpr int kMaxGroups = kMaxPartials / kFanIn;            // 512
constexpr int kTicketStride = 32;                            // one ticket per 128-byte line
constexpr size_t kTicketWords = (size_t)(1 + kMaxGroups) * kTicketStride;
constexpr int kMaxNP = 3;                                    // reductions fused into one launch
constexpr size_t kPartialWords = kMaxNP * (size_t)(kMaxPartials + kMaxGroups);

// Two-level combine of per-workgroup partial sums.  Workgroup b publishes its
// NP partials, then takes the ticket of its group of kFanIn workgroups; the
// last of a group sums the group's partials (index order) and takes the top
// ticket; the last group sums the group sums (index order) and runs
// epi.finish.  Same-address atomics are thus at most kFanIn deep instead of
// one per workgroup, and the result does not depend on arrival order.
// partials: NP*G block values then NP*ngroups group values; tickets: words
// [(1+g)*kTicketStride] per group and [0] on top, all left at zero.
// Ordering: the inter-workgroup protocol of internal.hpp (publish, then
// arrive_last's wait_memory and ticket; agent-scope reads of the partials).
template <int BS, typename AT, typename Epi>
__device__ __forceinline__ void combine_partials(const AT* part, AT* partials, unsigned* tickets,
                                                 int b, int G, const Epi& epi, AT* sh) {
    constexpr int NP = Epi::NP;
    static_assert(NP <= kMaxNP, "partials are sized for kMaxNP reductions per launch");
    const int ngroups = (G + kFanIn - 1) / kFanIn;
    const int g = b / kFanIn;
    const int gsize = min(kFanIn, G - g * kFanIn);
#pragma unroll
    for (int k = 0; k < NP; k++) {
        AT v = block_sum<BS>(part[k], sh);
        if (threadIdx.x == 0) publish(partials + (size_t)k * G + b, v);
    }
    if (!arrive_last<BS>(tickets + (size_t)(1 + g) * kTicketStride, (unsigned)gsize)) return;
    AT* gpart = partials + (size_t)NP * G;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        AT v = reduce_partials<BS>(partials + (size_t)k * G + (size_t)g * kFanIn, gsize, 1, sh);
        if (threadIdx.x == 0) publish(gpart + (size_t)k * ngroups + g, v);
    }
    if (!arrive_last<BS>(tickets, (unsigned)ngroups)) return;
    AT tot[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) tot[k] = reduce_partials<BS>(gpart + (size_t)k * ngroups, ngroups, 1, sh);
    if (threadIdx.x == 0) epi.finish(tot);
}

// ORD: instead of combining partial sums, every row stores its terms in
// terms[k * tstride + row]; k_ordered then adds them in MTL4's order.
template <int BS, int TILE, typename AT, typename MT, typename XT, typename Epi, bool ORD = false>
__global__ __launch_bounds__(BS) void k_csr(const int32_t* __restrict__ rowptr,
                                            const int32_t* __restrict__ col,
                                            const MT* __restrict__ val,
                                            const XT* __restrict__ x,
                                            const RowBlock* __restrict__ desc, int nb, int per_xcd,
                                            Epi epi, AT* partials, unsigned* ticket,
                                            AT* terms, int64_t tstride) {
    const int b = xcd_logical_block(blockIdx.x, per_xcd);
    const RowBlock d = desc[b < nb ? b : nb - 1];   // requested with the stop flag
    if (epi.skip()) return;
    epi.prepare();
    __shared__ AT prod[TILE];
    __shared__ AT sh[BS / kWave];
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);
    if (b < nb) {
        int row;
        typename Epi::Pre pre{};
        AT s = csr_block_rows<BS, TILE, AT>(rowptr, col, val, x, d, prod, sh, row, epi, pre);
        if (row < d.r1) {
            if constexpr (ORD) {
                epi(row, s, pre, part);
#pragma unroll
                for (int k = 0; k < NP; k++) terms[(size_t)k * tstride + row] = part[k];
            } else {
                epi(row, s, pre, part);
            }
        }
    }
    if constexpr (Epi::NP > 0 && !ORD) combine_partials<BS>(part, partials, ticket, b, (int)gridDim.x, epi, sh);
}

// Persistent row-block kernel with the next tile's stream in flight.
// Workgroup g runs on XCD g % 8 and owns every (gridDim/8)-th row block of
// that XCD's contiguous eighth of the matrix, so the XCD's workgroups sweep
// one window of rows together and share its x in L2.  Per tile, a lane
// issues the gathers of x for the current tile, then the descriptor, row
// bounds, epilogue operands and value/index stream of its next tile, and
// only then waits for the gathers: the next tile's HBM reads overlap the
// current tile's gathers, products, row sums and epilogue.  Products go to
// one of two LDS buffers in turn, so one barrier per tile suffices.  Rows
// are summed in column order, as in k_csr; a longer-than-TILE single row is
// summed by the whole workgroup (not overlapped).  RPL rows per lane: lane t
// sums rows t, t + BS, ... of the block (blocks of up to RPL * BS rows).
// Two-element loads of the value/index streams (FP64 and FP32 values): lane
// t requests elements (base + 2p, base + 2p + 1) of pairs p = t + u*BS, with
// base = j0 rounded down to even, so every pair load is aligned; elements
// outside [j0, j1) are loaded (the arrays carry two elements of padding) and
// ignored.  A block then holds at most TILE - 2 non-zeros.
template <typename T>
struct pair_of;
template <>
struct pair_of<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct pair_of<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct pair_of<uint16_t> {
    typedef uint16_t type __attribute__((ext_vector_type(2)));
};
typedef int int_pair __attribute__((ext_vector_type(2)));

// A row's sum starts at epi.init(row) when the epilogue has one (a column
// panel continuing the sums of the panels before it), else at 0.
template <typename E, typename = void>
struct has_init : std::false_type {};
template <typename E>
struct has_init<E, std::void_t<decltype(std::declval<const E&>().init(0))>> : std::true_type {};

// Column panels: every panel but the last stores each row's partial sum
// (started from the previous panel's) in t; the last one starts from t and
// runs the caller's epilogue on the full sum.  The caller's skip() holds for
// every panel.
template <typename AT, typename Epi>
struct EpiPanelCarry {
    static constexpr int NP = 0;
    struct Pre {};
    const AT* tin;   // null for the first panel
    AT* tout;
    Epi inner;
    __device__ bool skip() const { return inner.skip(); }
    __device__ void prepare() {}
    __device__ Pre preload(int) const { return Pre{}; }
    __device__ AT init(int row) const { return tin ? tin[row] : AT(0); }
    __device__ void operator()(int row, AT s, const Pre&, AT*) const { tout[row] = s; }
    __device__ void finish(const AT*) const {}
};
template <typename AT, typename Epi>
struct EpiPanelLast : Epi {
    const AT* tin;
    EpiPanelLast(const Epi& e, const AT* t) : Epi(e), tin(t) {}
    __device__ AT init(int row) const { return tin[row]; }
};

template <int BS, int TILE, typename AT, typename MT, typename XT, typename Epi, int RPL = 1>
__global__ __launch_bounds__(BS) void k_csrp(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x,
                                             const RowBlock* __restrict__ desc, int nb, Epi epi,
                                             AT* partials, unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int U = TILE / BS;
    constexpr bool PAIR = sizeof(MT) >= 4 && U % 2 == 0;
    constexpr int LONG = PAIR ? TILE - 2 : TILE;   // blocks above this are single long rows
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    __shared__ AT prod[2][TILE];
    __shared__ AT sh[BS / kWave];
    const int t = threadIdx.x;
    const int slots = gridDim.x / kXcd;
    const int per = (nb + kXcd - 1) / kXcd;
    const int xcd = blockIdx.x % kXcd;
    const int hi = min(nb, (xcd + 1) * per);
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);

    RowBlock d{};
    int c[U];
    AT v[U];
    int ra[RPL] = {}, re[RPL] = {};
    typename Epi::Pre pre[RPL]{};
    auto fetch = [&](int bb) {
        d = desc[bb];
        const int nz = d.j1 - d.j0;
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            const int row = d.r0 + t + q * BS;
            if (row < d.r1) {
                ra[q] = rowptr[row] - d.j0;
                re[q] = rowptr[row + 1] - d.j0;
                pre[q] = epi.preload(row);
            } else {
                ra[q] = re[q] = 0;
            }
        }
        if constexpr (PAIR) {
            if (nz <= LONG) {
                const int base = d.j0 & ~1;
#pragma unroll
                for (int u = 0; u < U / 2; u++) {
                    const int e0 = min(base + 2 * (t + u * BS), d.j1 & ~1);   // in bounds; unused beyond j1
                    const int_pair cc = __builtin_nontemporal_load(reinterpret_cast<const int_pair*>(col + e0));
                    const typename pair_of<MT>::type vv =
                        __builtin_nontemporal_load(reinterpret_cast<const typename pair_of<MT>::type*>(val + e0));
                    c[2 * u] = cc.x;
                    c[2 * u + 1] = cc.y;
                    v[2 * u] = static_cast<AT>(vv.x);
                    v[2 * u + 1] = static_cast<AT>(vv.y);
                }
            }
        } else if (nz <= TILE) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = t + u * BS;
                c[u] = k < nz ? __builtin_nontemporal_load(col + d.j0 + k) : 0;
                v[u] = k < nz ? load_val_nt<AT>(val, (int64_t)d.j0 + k) : AT(0);
            }
        }
    };

    int b = xcd * per + blockIdx.x / kXcd;
    if (b < hi) fetch(b);
    int buf = 0;
    while (b < hi) {
        const RowBlock dc = d;
        const int nz = dc.j1 - dc.j0;
        const int row = dc.r0 + t;
        int a[RPL], e[RPL];
        typename Epi::Pre pc[RPL];
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            a[q] = ra[q];
            e[q] = re[q];
            pc[q] = pre[q];
        }
        const int bn = b + slots;
        if (nz > LONG) {
            AT s = AT(0);
            for (int k = t; k < nz; k += BS)
                s += load_val<AT>(val, (int64_t)dc.j0 + k) * static_cast<AT>(x[col[dc.j0 + k]]);
            s = block_sum<BS>(s, sh);
            if (bn < hi) fetch(bn);
            if constexpr (has_init<Epi>::value) {
                if (t == 0) s = epi.init(row) + s;
            }
            if (t == 0) epi(row, s, pc[0], part);
        } else {
            // tile-relative index of the element lane t holds in slot u
            const int pbase = (dc.j0 & ~1) - dc.j0;
            auto slot_k = [&](int u) {
                if constexpr (PAIR) return pbase + 2 * (t + (u / 2) * BS) + (u & 1);
                else return t + u * BS;
            };
            AT g[U], vc[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = slot_k(u);
                g[u] = (k >= 0 && k < nz) ? static_cast<AT>(x[c[u]]) : AT(0);
                vc[u] = v[u];
            }
            if (bn < hi) fetch(bn);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = slot_k(u);
                if (k >= 0 && k < nz) prod[buf][k] = vc[u] * g[u];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RPL; q++) {
                AT s = AT(0);
                if constexpr (has_init<Epi>::value) {
                    if (row + q * BS < dc.r1) s = epi.init(row + q * BS);
                }
                for (int k = a[q]; k < e[q]; k++) s += prod[buf][k];
                if (row + q * BS < dc.r1) epi(row + q * BS, s, pc[q], part);
            }
            buf ^= 1;
        }
        b = bn;
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, blockIdx.x, (int)gridDim.x, epi, sh);
}

// k_csrp with DEPTH row blocks requested ahead instead of one.
// Workgroup g runs on XCD g % 8 and owns every (gridDim/8)-th row block of
// that XCD's contiguous eighth of the matrix, so the XCD's workgroups sweep
// one window of rows together and share its x in L2.  A workgroup keeps
// DEPTH row blocks requested ahead: per tile, a lane issues the gathers of x
// for the current tile, then the descriptor, row bounds, epilogue operands
// and value/index stream of the tile DEPTH steps ahead, and only then waits
// for the gathers (the wait counter counts in issue order, so the newer
// requests stay in flight).  Products go to one of two LDS buffers in turn,
// so one barrier per tile suffices.  Rows are summed in column order, as in
// k_csr; a longer-than-TILE single row is summed by the whole workgroup.
template <int BS, int TILE, int DEPTH, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_csrp2(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x,
                                             const RowBlock* __restrict__ desc, int nb, Epi epi,
                                             AT* partials, unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int U = TILE / BS;
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    __shared__ AT prod[2][TILE];
    __shared__ AT sh[BS / kWave];
    const int t = threadIdx.x;
    const int slots = gridDim.x / kXcd;
    const int per = (nb + kXcd - 1) / kXcd;
    const int xcd = blockIdx.x % kXcd;
    const int hi = min(nb, (xcd + 1) * per);
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);

    struct Stage {
        RowBlock d;
        int c[U];
        AT v[U];
        int ra, re;
        typename Epi::Pre pre;
    };
    Stage st[DEPTH];   // requested tiles, st[0] the next to sum
    auto fetch = [&](int bb, Stage& q) {
        q.d = desc[bb];
        const int row = q.d.r0 + t;
        const int nz = q.d.j1 - q.d.j0;
        q.ra = q.re = 0;
        if (row < q.d.r1) {
            q.ra = rowptr[row] - q.d.j0;
            q.re = rowptr[row + 1] - q.d.j0;
            q.pre = epi.preload(row);
        }
        if (nz <= TILE) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = t + u * BS;
                q.c[u] = k < nz ? __builtin_nontemporal_load(col + q.d.j0 + k) : 0;
                q.v[u] = k < nz ? load_val_nt<AT>(val, (int64_t)q.d.j0 + k) : AT(0);
            }
        }
    };

    int b = xcd * per + blockIdx.x / kXcd;
#pragma unroll
    for (int i = 0; i < DEPTH; i++)
        if (b + i * slots < hi) fetch(b + i * slots, st[i]);
    int buf = 0;
    while (b < hi) {
        const Stage cur = st[0];
#pragma unroll
        for (int i = 0; i + 1 < DEPTH; i++) st[i] = st[i + 1];
        const int nz = cur.d.j1 - cur.d.j0;
        const int row = cur.d.r0 + t;
        const int bn = b + DEPTH * slots;   // the tile requested in this step
        if (nz > TILE) {
            AT s = AT(0);
            for (int k = t; k < nz; k += BS)
                s += load_val<AT>(val, (int64_t)cur.d.j0 + k) * static_cast<AT>(x[col[cur.d.j0 + k]]);
            s = block_sum<BS>(s, sh);
            if (bn < hi) fetch(bn, st[DEPTH - 1]);
            if (t == 0) epi(row, s, cur.pre, part);
        } else {
            AT g[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = t + u * BS;
                g[u] = k < nz ? static_cast<AT>(x[cur.c[u]]) : AT(0);
            }
            if (bn < hi) fetch(bn, st[DEPTH - 1]);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = t + u * BS;
                if (k < nz) prod[buf][k] = cur.v[u] * g[u];
            }
            __syncthreads();
            AT s = AT(0);
            for (int k = cur.ra; k < cur.re; k++) s += prod[buf][k];
            if (row < cur.d.r1) epi(row, s, cur.pre, part);
            buf ^= 1;
        }
        b += slots;
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, blockIdx.x, (int)gridDim.x, epi, sh);
}

// Scalar CSR: one lane per row, RPT consecutive rows per lane-pass.  The
// lane requests the first UV values and indices of its row at once (more in
// a loop for longer rows), then the gathers, then sums in column order --
// no LDS, no barrier.  For stencil rows the k-th gathers of neighbouring
// lanes are neighbouring x entries (coalesced), and the strided value and
// index loads of a wavefront cover its rows' contiguous span, reused from
// the vector cache across k.  Rows map to workgroups XCD-aware.
template <int BS, int UV, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_csrs(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x, int nrows, int nb, int per_xcd,
                                             Epi epi, AT* partials, unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    __shared__ AT sh[BS / kWave];
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);
    const int b = xcd_logical_block(blockIdx.x, per_xcd);
    const int row = b * BS + (int)threadIdx.x;
    if (b < nb && row < nrows) {
        const int a = rowptr[row], e = rowptr[row + 1];
        const typename Epi::Pre pre = epi.preload(row);
        AT s = AT(0);
        for (int j0 = a; j0 < e; j0 += UV) {
            int c[UV];
            AT v[UV];
#pragma unroll
            for (int u = 0; u < UV; u++) {
                c[u] = j0 + u < e ? col[j0 + u] : 0;   // cached: neighbouring lanes reuse the lines
                v[u] = j0 + u < e ? load_val<AT>(val, (int64_t)j0 + u) : AT(0);
            }
            AT g[UV];
#pragma unroll
            for (int u = 0; u < UV; u++) g[u] = j0 + u < e ? static_cast<AT>(x[c[u]]) : AT(0);
#pragma unroll
            for (int u = 0; u < UV; u++)
                if (j0 + u < e) s += v[u] * g[u];
        }
        epi(row, s, pre, part);
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, b, (int)gridDim.x, epi, sh);
}

// Vector CSR: a group of GS lanes per row, 64/GS rows per wavefront pass.
// The lanes of a group read consecutive non-zeros of their row, so a
// wavefront's value/index loads cover consecutive rows contiguously; no LDS,
// no workgroup barrier.  The row sum keeps MTL4's order: the group's
// products are added one after the other in column order, each taken from
// its lane with a shuffle (mat_vec_mult.hpp:656-682, bit for bit).
template <int GS, int BS, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_csrv(const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x, int nrows, int rows_per_block,
                                             int nb, int per_xcd, Epi epi, AT* partials,
                                             unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    __shared__ AT sh[BS / kWave];
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    constexpr int GPW = kWave / GS;       // row groups per wavefront
    constexpr int WPB = BS / kWave;       // wavefronts per workgroup
    const int b = xcd_logical_block(blockIdx.x, per_xcd);
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int g = lane / GS, l = lane % GS;
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);
    if (b < nb) {
        const int row_begin = b * rows_per_block;
        const int row_end = min(nrows, row_begin + rows_per_block);
        for (int rb = row_begin + wave * GPW; rb < row_end; rb += WPB * GPW) {
            const int r = rb + g;
            const bool valid = r < row_end;
            int a = 0, e = 0;
            typename Epi::Pre pre{};
            if (valid) {
                a = rowptr[r];
                e = rowptr[r + 1];
                if (l == 0) pre = epi.preload(r);
            }
            AT s = AT(0);
            for (int base = a; __any(base < e); base += GS) {
                const int j = base + l;
                const AT pj = j < e ? load_val_nt<AT>(val, j) * static_cast<AT>(x[__builtin_nontemporal_load(col + j)])
                                    : AT(0);
#pragma unroll
                for (int k = 0; k < GS; k++) {
                    const AT v = __shfl(pj, g * GS + k, kWave);
                    if (base + k < e) s += v;
                }
            }
            if (valid && l == 0) epi(r, s, pre, part);
        }
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, b, (int)gridDim.x, epi, sh);
}

// Sliced ELL: rows in slices of 64 (one wavefront); slice s stores w_s
// entries per row (its longest row), column-major: entry k of row 64s + l
// at soff[s] + 64k + l, padding marked by column -1 and skipped.  A lane
// sums its own row in column order, as every CSR kernel does, so the result
// is the same bit for bit; no LDS, no barrier, no row pointers.  Each
// wavefront walks its slices alone, in chunks of UNR entries, and requests
// the next chunk's values and indices before it waits for the current
// chunk's gathers of x.  XCD g % 8 owns a contiguous eighth of the slices
// and its wavefronts take them in turn, so they share one window of x in L2.
// Slice bookkeeping is wavefront-uniform (scalar registers and loads).  The
// vector loads are unconditional -- entries beyond a slice's width reload
// its last column, a chunk past the end of the array reads the padding the
// arrays carry (UNR * 64 entries) -- so the wait for the gathers never
// waits for the next chunk; two register sets alternate between the
// current and the next chunk.
constexpr int kSellMaxUnroll = 8;                  // entries per lane per chunk: 4 or 8
constexpr int kSellPad = kSellMaxUnroll * kWave;   // padding entries after the arrays

template <int BS, int UNR, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_sell(const int64_t* __restrict__ soff,
                                             const int32_t* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const XT* __restrict__ x, int nrows, int nsl,
                                             Epi epi, AT* partials, unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    constexpr int WPB = BS / kWave;
    __shared__ AT sh[WPB];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int xcd = blockIdx.x % kXcd;
    const int wstep = (int)(gridDim.x / kXcd) * WPB;   // wavefronts per XCD
    const int per = (nsl + kXcd - 1) / kXcd;
    const int hi = min(nsl, (xcd + 1) * per);
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);

    struct Chunk {
        int c[UNR];
        AT v[UNR];
        typename Epi::Pre pre;
    };
    // state of the chunk last requested: slice s, entries k0.. of width w at o
    int s = xcd * per + (int)(blockIdx.x / kXcd) * WPB + wv;
    int64_t o = 0;
    int w = 0, k0 = 0;
    auto fetch = [&](Chunk& ch) {
        ch.pre = epi.preload(min(s * kWave + lane, nrows - 1));
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int slot = max(min(k0 + u, w - 1), 0);
            const int64_t e = o + (int64_t)slot * kWave + lane;
            ch.c[u] = __builtin_nontemporal_load(col + e);
            ch.v[u] = load_val_nt<AT>(val, e);
        }
    };
    auto slice = [&]() {
        const int64_t a = soff[s], b = soff[s + 1];
        o = ((int64_t)__builtin_amdgcn_readfirstlane((int)(a >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((int)a);
        w = __builtin_amdgcn_readfirstlane((int)((b - a) / kWave));
        k0 = 0;
    };
    AT acc = AT(0);
    // one chunk: gathers of cur, then (unless this was the last) the request
    // of the next chunk into nxt, then the products; false at the end
    auto step = [&](const Chunk& cur, Chunk& nxt) -> bool {
        AT g[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) g[u] = static_cast<AT>(x[max(cur.c[u], 0)]);
        const int kc = k0, wc = w, row = s * kWave + lane;
        // wavefront-uniform by construction (scalar branches)
        const bool last = __builtin_amdgcn_readfirstlane(kc + UNR >= wc) != 0;
        const int sn = __builtin_amdgcn_readfirstlane(last ? s + wstep : s);
        auto sum = [&]() {
#pragma unroll
            for (int u = 0; u < UNR; u++)
                if (kc + u < wc && cur.c[u] >= 0) acc += cur.v[u] * g[u];
            if (last) {
                if (row < nrows) epi(row, acc, cur.pre, part);
                acc = AT(0);
            }
        };
        if (__builtin_amdgcn_readfirstlane(sn >= hi)) {
            sum();
            return false;
        }
        s = sn;
        if (last) slice();
        else k0 = kc + UNR;
        fetch(nxt);
        // the products stay behind the requests of the next chunk
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UNR; u++) asm volatile("" : "+v"(g[u])::"memory");
        sum();
        return true;
    };
    if (s < hi) {
        Chunk A, B;
        slice();
        fetch(A);
        for (;;) {
            if (!step(A, B)) break;
            if (!step(B, A)) break;
        }
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, blockIdx.x, (int)gridDim.x, epi, sh);
}

// Diagonal (banded) layout: a matrix whose entries lie on ndiag diagonals
// (column - row = off[d], ascending) stores one value for every row r and
// diagonal d, slice-major -- at ((r / 64) * ndiag + d) * 64 + r % 64, so a
// wavefront's 64 rows read one contiguous block -- with a reserved NaN bit
// pattern where row r has no entry on d (and for the rows past n of the last
// slice).  No column indices are stored or read.  A lane sums its row over
// the diagonals in ascending offset order -- ascending column order, as
// every CSR kernel does -- skipping the absent entries, so the result is
// the same bit for bit.  The walk, chunking and request order are those of
// k_sell (slices of 64 rows per wavefront, UNR diagonals per chunk, the
// next chunk's values in flight while the current chunk's x is gathered).
constexpr uint64_t kDiaAbsent64 = 0x7ff4a5a5a5a5a5a5ull;   // signalling NaN: never produced by arithmetic
constexpr uint32_t kDiaAbsent32 = 0x7fa5a5a5u;             // nor by a conversion (which quiets NaNs)
constexpr uint16_t kDiaAbsent16 = 0x7fa5u;                 // bfloat16: quiet NaNs from rounding carry 0x0040
constexpr int kDiaMax = 32;                                // diagonals at most

__device__ __forceinline__ bool dia_absent(double v) { return (uint64_t)__double_as_longlong(v) == kDiaAbsent64; }
__device__ __forceinline__ bool dia_absent(float v) { return __float_as_uint(v) == kDiaAbsent32; }
__device__ __forceinline__ bool dia_absent(uint16_t v) { return v == kDiaAbsent16; }
template <typename AT>
__device__ __forceinline__ AT dia_value(double v) { return static_cast<AT>(v); }
template <typename AT>
__device__ __forceinline__ AT dia_value(float v) { return static_cast<AT>(v); }
template <typename AT>
__device__ __forceinline__ AT dia_value(uint16_t v) { return static_cast<AT>(bf16_to_f32(v)); }

template <int BS, int UNR, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_dia(const int32_t* __restrict__ doff,
                                            const MT* __restrict__ val,
                                            const XT* __restrict__ x, int nrows, int ncols, int ndiag,
                                            Epi epi, AT* partials, unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    constexpr int WPB = BS / kWave;
    __shared__ AT sh[WPB];
    // the offsets in LDS: read per chunk without waiting on vector memory
    __shared__ int offs[kDiaMax];
    if ((int)threadIdx.x < ndiag) offs[threadIdx.x] = doff[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int xcd = blockIdx.x % kXcd;
    const int wstep = (int)(gridDim.x / kXcd) * WPB;
    const int nsl = (nrows + kWave - 1) / kWave;
    const int per = (nsl + kXcd - 1) / kXcd;
    const int hi = min(nsl, (xcd + 1) * per);
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);

    // Three chunks rotate: two are loading while the third is summed.  A
    // chunk is UNR diagonals of one slice of kWave rows, with its cursor.
    struct Chunk {
        MT v[UNR];
        AT g[UNR];      // x at the chunk's columns
        int off[UNR];   // wavefront-uniform
        int s, k0;      // slice and first diagonal (wavefront-uniform)
        typename Epi::Pre pre;
    };
    int fs = xcd * per + (int)(blockIdx.x / kXcd) * WPB + wv;   // fetch cursor
    int fk = 0;
    auto fetch = [&](Chunk& ch) {
        ch.s = fs;
        ch.k0 = fk;
        if (__builtin_amdgcn_readfirstlane(fs < hi)) {
            const int r = fs * kWave + lane;
            ch.pre = epi.preload(min(r, nrows - 1));
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int d = min(fk + u, ndiag - 1);
                const int o = __builtin_amdgcn_readfirstlane(offs[d]);
                ch.off[u] = o;
                ch.v[u] = __builtin_nontemporal_load(val + ((int64_t)fs * ndiag + d) * kWave + lane);
                ch.g[u] = static_cast<AT>(x[min(max(r + o, 0), ncols - 1)]);
            }
        }
        if (fk + UNR >= ndiag) {
            fk = 0;
            fs += wstep;
        } else {
            fk += UNR;
        }
    };
    AT acc = AT(0);
    // Sums chunk cur (false when it lies past the wave's last slice) after
    // starting the load of nxt, two chunks ahead.
    auto step = [&](const Chunk& cur, Chunk& nxt) -> bool {
        const int cs = __builtin_amdgcn_readfirstlane(cur.s);
        if (cs >= hi) return false;
        fetch(nxt);
        __builtin_amdgcn_sched_barrier(0);
        const int kc = __builtin_amdgcn_readfirstlane(cur.k0), row = cs * kWave + lane;
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (kc + u < ndiag && !dia_absent(cur.v[u])) acc += dia_value<AT>(cur.v[u]) * cur.g[u];
        if (kc + UNR >= ndiag) {
            if (row < nrows) epi(row, acc, cur.pre, part);
            acc = AT(0);
        }
        return true;
    };
    {
        Chunk A, B, C;
        fetch(A);
        fetch(B);
        for (;;) {
            if (!step(A, C)) break;
            if (!step(B, A)) break;
            if (!step(C, B)) break;
        }
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, blockIdx.x, (int)gridDim.x, epi, sh);
}

// Symmetric diagonal layout (MTL4G_SPMV_DIA_SYM): for a matrix equal to its
// transpose bit for bit, only the diagonals with offset >= 0 are stored
// (slice-major as above, nst of them); the entry of row r on diagonal -o is
// the stored entry of row r - o on diagonal +o.  The walk and the summation
// order are k_dia's over the full offset list (ascending columns), so the
// result is the same bit for bit; the stored diagonals are read with
// caching loads, and the mirrored reads of rows a few slices back mostly
// hit the cache, so the value stream from memory is nst of the ndiag
// diagonals.  dmap[d]: the stored diagonal of full diagonal d.
template <typename MT>
__device__ __forceinline__ MT dia_absent_value();
template <>
__device__ __forceinline__ double dia_absent_value<double>() { return __longlong_as_double((long long)kDiaAbsent64); }
template <>
__device__ __forceinline__ float dia_absent_value<float>() { return __uint_as_float(kDiaAbsent32); }
template <>
__device__ __forceinline__ uint16_t dia_absent_value<uint16_t>() { return kDiaAbsent16; }

template <int BS, int UNR, typename AT, typename MT, typename XT, typename Epi>
__global__ __launch_bounds__(BS) void k_dia_sym(const int32_t* __restrict__ doff, const int32_t* __restrict__ dmap,
                                                const MT* __restrict__ val, const XT* __restrict__ x, int nrows,
                                                int ncols, int ndiag, int nst, Epi epi, AT* partials,
                                                unsigned* ticket) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    constexpr int WPB = BS / kWave;
    __shared__ AT sh[WPB];
    __shared__ int offs[kDiaMax], maps[kDiaMax];
    if ((int)threadIdx.x < ndiag) {
        offs[threadIdx.x] = doff[threadIdx.x];
        maps[threadIdx.x] = dmap[threadIdx.x];
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int xcd = blockIdx.x % kXcd;
    const int wstep = (int)(gridDim.x / kXcd) * WPB;
    const int nsl = (nrows + kWave - 1) / kWave;
    const int per = (nsl + kXcd - 1) / kXcd;
    const int hi = min(nsl, (xcd + 1) * per);
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);

    // Three chunks rotate: two are loading while the third is summed.  A
    // chunk is UNR diagonals of one slice of kWave rows, with its cursor.
    struct Chunk {
        MT v[UNR];
        AT g[UNR];      // x at the chunk's columns
        int off[UNR];   // wavefront-uniform
        int s, k0;      // slice and first diagonal (wavefront-uniform)
        typename Epi::Pre pre;
    };
    int fs = xcd * per + (int)(blockIdx.x / kXcd) * WPB + wv;   // fetch cursor
    int fk = 0;
    auto fetch = [&](Chunk& ch) {
        ch.s = fs;
        ch.k0 = fk;
        if (__builtin_amdgcn_readfirstlane(fs < hi)) {
            const int r = fs * kWave + lane;
            ch.pre = epi.preload(min(r, nrows - 1));
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int d = min(fk + u, ndiag - 1);
                const int o = __builtin_amdgcn_readfirstlane(offs[d]);
                const int mi = __builtin_amdgcn_readfirstlane(maps[d]);
                ch.off[u] = o;
                if (o >= 0) {
                    ch.v[u] = val[((int64_t)fs * nst + mi) * kWave + lane];
                } else {   // row r - |o| on the mirrored diagonal; none above row 0
                    const int r2 = r + o;
                    const int rc = max(r2, 0);
                    const MT v = val[((int64_t)(rc / kWave) * nst + mi) * kWave + rc % kWave];
                    ch.v[u] = r2 >= 0 ? v : dia_absent_value<MT>();
                }
                ch.g[u] = static_cast<AT>(x[min(max(r + o, 0), ncols - 1)]);
            }
        }
        if (fk + UNR >= ndiag) {
            fk = 0;
            fs += wstep;
        } else {
            fk += UNR;
        }
    };
    AT acc = AT(0);
    // Sums chunk cur (false when it lies past the wave's last slice) after
    // starting the load of nxt, two chunks ahead.
    auto step = [&](const Chunk& cur, Chunk& nxt) -> bool {
        const int cs = __builtin_amdgcn_readfirstlane(cur.s);
        if (cs >= hi) return false;
        fetch(nxt);
        __builtin_amdgcn_sched_barrier(0);
        const int kc = __builtin_amdgcn_readfirstlane(cur.k0), row = cs * kWave + lane;
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (kc + u < ndiag && !dia_absent(cur.v[u])) acc += dia_value<AT>(cur.v[u]) * cur.g[u];
        if (kc + UNR >= ndiag) {
            if (row < nrows) epi(row, acc, cur.pre, part);
            acc = AT(0);
        }
        return true;
    };
    {
        Chunk A, B, C;
        fetch(A);
        fetch(B);
        for (;;) {
            if (!step(A, C)) break;
            if (!step(B, A)) break;
            if (!step(C, B)) break;
        }
    }
    if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, blockIdx.x, (int)gridDim.x, epi, sh);
}

// Element epilogues may split operator()(i, part) into In load(i) (all
// reads of element i) and store(i, const In&, part) (the arithmetic, writes
// and partial sums); k_vec then batches the loads of several elements.
template <typename E, typename = void>
struct has_split : std::false_type {};
template <typename E>
struct has_split<E, std::void_t<typename E::In>> : std::true_type {};
constexpr int kVecBatch = 4;

template <int BS, typename AT, typename Epi, bool ORD = false>
__global__ __launch_bounds__(BS) void k_vec(int64_t n, Epi epi, AT* partials, unsigned* ticket,
                                            AT* terms) {
    if (epi.skip()) return;
    epi.prepare();
    __shared__ AT sh[BS / kWave];
    constexpr int NP = Epi::NP > 0 ? Epi::NP : 1;
    const unsigned G = gridDim.x;
    const int b = blockIdx.x;
    int64_t chunk = (n + G - 1) / G;
    chunk = (chunk + BS - 1) / BS * BS;
    const int64_t lo = (int64_t)b * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    AT part[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) part[k] = AT(0);
    if constexpr (ORD) {
        for (int64_t i = lo + threadIdx.x; i < hi; i += BS) {
#pragma unroll
            for (int k = 0; k < NP; k++) part[k] = AT(0);
            epi(i, part);
#pragma unroll
            for (int k = 0; k < NP; k++) terms[(size_t)k * n + i] = part[k];
        }
    } else {
        if constexpr (has_split<Epi>::value) {
            // loads of kVecBatch elements first, then their updates in the same
            // element order as the plain loop (partial sums unchanged): the
            // loads of later elements need not wait for earlier stores
            int64_t i0 = lo + threadIdx.x;
            for (; i0 + (kVecBatch - 1) * BS < hi; i0 += kVecBatch * BS) {
                typename Epi::In in[kVecBatch];
#pragma unroll
                for (int u = 0; u < kVecBatch; u++) in[u] = epi.load(i0 + u * BS);
#pragma unroll
                for (int u = 0; u < kVecBatch; u++) epi.store(i0 + u * BS, in[u], part);
            }
            for (; i0 < hi; i0 += BS) epi.store(i0, epi.load(i0), part);
        } else {
            for (int64_t i = lo + threadIdx.x; i < hi; i += BS) epi(i, part);
        }
        if constexpr (Epi::NP > 0) combine_partials<BS>(part, partials, ticket, b, (int)G, epi, sh);
    }
}

// ---- MTL4-order reductions
// MTL4 sums with U accumulators: a[k] holds the terms i = k (mod U) below
// n - n % U in index order, a[0] then also the remainder; the forms below
// combine them.
enum RedOrder {
    RO_DOT4 = 0,        // dot<4>: ((t0+t1)+(t2+t3)) + ((t4+t5)+(t6+t7)), t4..t7 = 0 (operation/dot.hpp:136-147)
    RO_NORM4 = 1,       // reduction<4>: t0+(t1+(t2+t3)) (vector/reduction.hpp:28-82)
    RO_FUSED_DOT = 2,   // dot_index_evaluator: ((t0+t1)+t2)+t3 (fused_expr.hpp:57-77)
    RO_FUSED_NORM = 3,  // reduction_index_evaluator: (t0+t1)+(t2+t3) (reduction_index_evaluator.hpp:22-41)
    RO_NORM8 = 4        // reduction<8> (unary_dot's default, unary_dot.hpp:56-60): t0+(t1+(...+(t6+t7)))
};

__host__ __device__ constexpr int ro_unroll(int order) { return order == RO_NORM8 ? 8 : 4; }

template <typename E, typename = void>
struct has_order : std::false_type {};
template <typename E>
struct has_order<E, std::void_t<decltype(E::order(0))>> : std::true_type {};

template <typename AT>
__device__ __forceinline__ AT mtl4_finish(int order, const AT* a) {
    switch (order) {
    case RO_DOT4: return ((a[0] + a[1]) + (a[2] + a[3])) + ((AT(0) + AT(0)) + (AT(0) + AT(0)));
    case RO_NORM4: return a[0] + (a[1] + (a[2] + a[3]));
    case RO_FUSED_DOT: return ((a[0] + a[1]) + a[2]) + a[3];
    case RO_FUSED_NORM: return (a[0] + a[1]) + (a[2] + a[3]);
    default: return a[0] + (a[1] + (a[2] + (a[3] + (a[4] + (a[5] + (a[6] + a[7]))))));
    }
}

// One wavefront: lane k < U adds terms[k], terms[k+U], ... below n - n % U
// one after the other (loads issued eight ahead), lane 0 then the remainder.
template <typename AT, typename Epi>
__global__ __launch_bounds__(kWave) void k_ordered(int64_t n, const AT* __restrict__ terms, Epi epi) {
    if (epi.skip()) return;
    epi.prepare();
    constexpr int NP = Epi::NP;
    __shared__ AT acc[NP][8];
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        const int U = ro_unroll(Epi::order(k));
        const int64_t sb = n - n % U;
        if (lane < U) {
            const AT* t = terms + (size_t)k * n;
            AT a = AT(0);
            int64_t i = lane;
            for (; i + 7 * U < sb; i += 8 * U) {
                AT v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = t[i + U * u];
#pragma unroll
                for (int u = 0; u < 8; u++) a += v[u];
            }
            for (; i < sb; i += U) a += t[i];
            if (lane == 0)
                for (i = sb; i < n; i++) a += t[i];
            acc[k][lane] = a;
        }
    }
    __syncthreads();
    if (lane == 0) {
        AT tot[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) tot[k] = mtl4_finish(Epi::order(k), acc[k]);
        epi.finish(tot);
    }
}

// True when this launch must run in MTL4's order; an error when it should
// but the caller brought no terms buffer.
template <typename AT, typename Epi>
int want_ordered(const AT* terms, bool& ord) {
    ord = false;
    if constexpr (Epi::NP > 0 && has_order<Epi>::value && std::is_same<AT, double>::value) {
        if (reduction_order() == MTL4G_ORDER_MTL4) {
            if (!terms) {
                set_error("MTL4-order reduction without a terms buffer");
                return MTL4G_ERR_INVALID;
            }
            ord = true;
        }
    }
    return MTL4G_OK;
}

// ---- host launchers
// Grid of a persistent kernel: cus * per_cu workgroups (a multiple of the
// XCD count), no more than the row blocks, and no more per CU than can be
// resident together (queried once per kernel instantiation).
template <auto KERNEL>
int persistent_grid(int cus, int per_cu, int nb) {
    static const int resident = []() {
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, KERNEL, 256, 0) != hipSuccess || nblk < 1)
            nblk = 1;
        return nblk;
    }();
    const int want = cus * std::min(per_cu, resident);
    return std::max(kXcd, std::min(want, (nb + kXcd - 1) / kXcd * kXcd) / kXcd * kXcd);
}

// terms: Epi::NP * A->n elements, used when reductions run in MTL4's order.
template <typename AT, typename MT, typename XT, typename Epi>
int launch_csr(const mtl4g_matrix_s* A, const MT* val, const XT* x, const Epi& epi,
               AT* partials, unsigned* ticket, hipStream_t s, AT* terms = nullptr) {
    if (A->n == 0) return MTL4G_OK;
    bool ord;
    M4_TRY((want_ordered<AT, Epi>(terms, ord)));
    if constexpr (Epi::NP > 0 && has_order<Epi>::value && std::is_same<AT, double>::value) {
        if (ord) {   // the row-block kernel (rows are summed alike in every kernel)
            const int per = (A->nb + kXcd - 1) / kXcd;
            const int tile = A->nnz_per_block;
#define M4_LAUNCHO(T)                                                                          \
    hipLaunchKernelGGL((k_csr<256, T, AT, MT, XT, Epi, true>), dim3(per * kXcd), dim3(256), 0, s, \
                       A->rowptr.p, A->col.p, val, x, A->desc.p, A->nb, per, epi, partials, ticket, \
                       terms, (int64_t)A->n)
            if (tile <= 1024) M4_LAUNCHO(1024);
            else if (tile <= 2048) M4_LAUNCHO(2048);
            else M4_LAUNCHO(4096);
#undef M4_LAUNCHO
            M4_HIP(hipGetLastError());
            hipLaunchKernelGGL((k_ordered<AT, Epi>), dim3(1), dim3(kWave), 0, s, (int64_t)A->n,
                               (const AT*)terms, epi);
            M4_HIP(hipGetLastError());
            return MTL4G_OK;
        }
    }
    if (A->kernel == MTL4G_SPMV_SCALAR) {
        const int nbs = (A->n + 255) / 256;
        const int per = (nbs + kXcd - 1) / kXcd;
        const int grid = per * kXcd;
        if (Epi::NP > 0 && grid > kMaxPartials) {
            set_error("matrix too large for one reduction pass (%d blocks)", grid);
            return MTL4G_ERR_UNSUPPORTED;
        }
        if (A->uv <= 4)
            hipLaunchKernelGGL((k_csrs<256, 4, AT, MT, XT, Epi>), dim3(grid), dim3(256), 0, s, A->rowptr.p,
                               A->col.p, val, x, A->n, nbs, per, epi, partials, ticket);
        else
            hipLaunchKernelGGL((k_csrs<256, 8, AT, MT, XT, Epi>), dim3(grid), dim3(256), 0, s, A->rowptr.p,
                               A->col.p, val, x, A->n, nbs, per, epi, partials, ticket);
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    if (A->kernel == MTL4G_SPMV_VECTOR) {
        const int rpb = A->vrows;
        const int nbv = (A->n + rpb - 1) / rpb;
        const int per = (nbv + kXcd - 1) / kXcd;
        const int grid = per * kXcd;
        if (Epi::NP > 0 && grid > kMaxPartials) {
            set_error("matrix too large for one reduction pass (%d blocks)", grid);
            return MTL4G_ERR_UNSUPPORTED;
        }
#define M4_LAUNCHV(G)                                                                          \
    hipLaunchKernelGGL((k_csrv<G, 256, AT, MT, XT, Epi>), dim3(grid), dim3(256), 0, s,         \
                       A->rowptr.p, A->col.p, val, x, A->n, rpb, nbv, per, epi, partials, ticket)
        switch (A->gs) {
        case 4: M4_LAUNCHV(4); break;
        case 8: M4_LAUNCHV(8); break;
        case 16: M4_LAUNCHV(16); break;
        case 32: M4_LAUNCHV(32); break;
        default: M4_LAUNCHV(64); break;
        }
#undef M4_LAUNCHV
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    if (A->kernel == MTL4G_SPMV_SLICED) {
        // the sliced copy of the value array the caller passed
        const MT* sv = nullptr;
        if constexpr (std::is_same<MT, double>::value) sv = val == A->val.p ? A->sl_val.p : nullptr;
        else if constexpr (std::is_same<MT, float>::value) sv = val == A->val32.p ? A->sl_val32.p : nullptr;
        else sv = val == A->valbf.p ? A->sl_valbf.p : nullptr;
        if (!sv || !A->sl_off.p) {
            set_error("sliced SpMV: no sliced copy of these values");
            return MTL4G_ERR_INVALID;
        }
        int pcu = A->per_cu, unr = A->sl_unroll;
        const int li = sizeof(MT) == 4 ? 0 : 1;
        if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) {
            pcu = A->per_cu_lowp[li];
            unr = A->sl_unroll_lowp[li];
        }
        const int units = (A->nsl + 3) / 4;   // four slices per workgroup pass
#define M4_LAUNCHS(U)                                                                               \
    do {                                                                                            \
        const int g = persistent_grid<k_sell<256, U, AT, MT, XT, Epi>>(A->cus, pcu, units);         \
        hipLaunchKernelGGL((k_sell<256, U, AT, MT, XT, Epi>), dim3(g), dim3(256), 0, s, A->sl_off.p, \
                           A->sl_col.p, sv, x, A->n, A->nsl, epi, partials, ticket);               \
    } while (0)
        if (unr == 4) M4_LAUNCHS(4);
        else M4_LAUNCHS(8);
#undef M4_LAUNCHS
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    if (A->kernel == MTL4G_SPMV_DIA) {
        // the diagonal copy of the value array the caller passed
        const MT* dv = nullptr;
        if constexpr (std::is_same<MT, double>::value) dv = val == A->val.p ? A->dia_val.p : nullptr;
        else if constexpr (std::is_same<MT, float>::value) dv = val == A->val32.p ? A->dia_val32.p : nullptr;
        else dv = val == A->valbf.p ? A->dia_valbf.p : nullptr;
        if (!dv || !A->dia_off.p) {
            set_error("diagonal SpMV: no diagonal copy of these values");
            return MTL4G_ERR_INVALID;
        }
        int pcu = A->per_cu, unr = A->sl_unroll;
        const int li = sizeof(MT) == 4 ? 0 : 1;
        if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) {
            pcu = A->per_cu_lowp[li];
            unr = A->sl_unroll_lowp[li];
        }
        const int units = ((A->n + kWave - 1) / kWave + 3) / 4;
#define M4_LAUNCHD(U)                                                                               \
    do {                                                                                            \
        const int g = persistent_grid<k_dia<256, U, AT, MT, XT, Epi>>(A->cus, pcu, units);          \
        hipLaunchKernelGGL((k_dia<256, U, AT, MT, XT, Epi>), dim3(g), dim3(256), 0, s, A->dia_off.p, dv, \
                           x, A->n, A->m, A->ndiag, epi, partials, ticket);                        \
    } while (0)
        if (unr == 4) M4_LAUNCHD(4);
        else M4_LAUNCHD(8);
#undef M4_LAUNCHD
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    if (A->kernel == MTL4G_SPMV_DIA_SYM) {
        // the symmetric diagonal copy of the value array the caller passed
        const MT* dv = nullptr;
        if constexpr (std::is_same<MT, double>::value) dv = val == A->val.p ? A->dsym_val.p : nullptr;
        else if constexpr (std::is_same<MT, float>::value) dv = val == A->val32.p ? A->dsym_val32.p : nullptr;
        else dv = val == A->valbf.p ? A->dsym_valbf.p : nullptr;
        if (!dv || !A->dsym_off.p) {
            set_error("symmetric diagonal SpMV: no symmetric diagonal copy of these values");
            return MTL4G_ERR_INVALID;
        }
        int pcu = A->per_cu, unr = A->sl_unroll;
        const int li = sizeof(MT) == 4 ? 0 : 1;
        if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) {
            pcu = A->per_cu_lowp[li];
            unr = A->sl_unroll_lowp[li];
        }
        const int units = ((A->n + kWave - 1) / kWave + 3) / 4;
#define M4_LAUNCHY(U)                                                                                      \
    do {                                                                                                   \
        const int g = persistent_grid<k_dia_sym<256, U, AT, MT, XT, Epi>>(A->cus, pcu, units);             \
        hipLaunchKernelGGL((k_dia_sym<256, U, AT, MT, XT, Epi>), dim3(g), dim3(256), 0, s, A->dsym_off.p,  \
                           A->dsym_map.p, dv, x, A->n, A->m, A->ndsym, A->nsym, epi, partials, ticket);    \
    } while (0)
        if (unr == 4) M4_LAUNCHY(4);
        else M4_LAUNCHY(8);
#undef M4_LAUNCHY
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    if (A->kernel == MTL4G_SPMV_PANEL) {
        // the panel copy of the value array the caller passed
        const MT* pv = nullptr;
        if constexpr (std::is_same<MT, double>::value) pv = val == A->val.p ? A->pan_val.p : nullptr;
        else if constexpr (std::is_same<MT, float>::value) pv = val == A->val32.p ? A->pan_val32.p : nullptr;
        else pv = val == A->valbf.p ? A->pan_valbf.p : nullptr;
        if (!pv || !A->npan) {
            set_error("panel SpMV: no panel copy of these values");
            return MTL4G_ERR_INVALID;
        }
        int pcu = A->per_cu;
        const int li = sizeof(MT) == 4 ? 0 : 1;
        if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) pcu = A->per_cu_lowp[li];
        // each call's partial sums live in a stream-ordered scratch vector:
        // calls on other streams never share it
        AT* t = nullptr;
        if (A->npan > 1) M4_HIP(hipMallocAsync((void**)&t, (size_t)A->n * sizeof(AT), s));
        int64_t boff = 0;
        int st = MTL4G_OK;
        for (int p = 0; p < A->npan && !st; p++) {
            const int32_t* rp = A->pan_rowptr.p + (size_t)p * (A->n + 1);
            const RowBlock* d = A->pan_desc.p + boff;
            const int nb = A->h_pan_nb[p];
            boff += nb;
            if (nb == 0) continue;
            if (p + 1 < A->npan) {
                using E = EpiPanelCarry<AT, Epi>;
                const E e{p ? t : nullptr, t, epi};
                const int g = persistent_grid<k_csrp<256, 2048, AT, MT, XT, E>>(A->cus, pcu, nb);
                hipLaunchKernelGGL((k_csrp<256, 2048, AT, MT, XT, E>), dim3(g), dim3(256), 0, s, rp, A->pan_col.p,
                                   pv, x, d, nb, e, (AT*)nullptr, (unsigned*)nullptr);
            } else if (A->npan > 1) {
                using E = EpiPanelLast<AT, Epi>;
                const E e(epi, t);
                const int g = persistent_grid<k_csrp<256, 2048, AT, MT, XT, E>>(A->cus, pcu, nb);
                hipLaunchKernelGGL((k_csrp<256, 2048, AT, MT, XT, E>), dim3(g), dim3(256), 0, s, rp, A->pan_col.p,
                                   pv, x, d, nb, e, partials, ticket);
            } else {
                const int g = persistent_grid<k_csrp<256, 2048, AT, MT, XT, Epi>>(A->cus, pcu, nb);
                hipLaunchKernelGGL((k_csrp<256, 2048, AT, MT, XT, Epi>), dim3(g), dim3(256), 0, s, rp, A->pan_col.p,
                                   pv, x, d, nb, epi, partials, ticket);
            }
            if (hipGetLastError() != hipSuccess) st = MTL4G_ERR_HIP;
        }
        if (t) (void)hipFreeAsync(t, s);
        return st;
    }
    if (A->nb == 0) return MTL4G_OK;
    if (A->kernel == MTL4G_SPMV_STREAM || A->kernel == MTL4G_SPMV_STREAM2) {
        int pcu = A->per_cu;
        const int li = sizeof(MT) == 4 ? 0 : 1;   // FP32 / BF16 values
        if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) pcu = A->per_cu_lowp[li];
        if constexpr (sizeof(MT) < 8) {
            // low-precision values: larger blocks, two rows per lane
            if (A->kernel == MTL4G_SPMV_STREAM && A->lowp_wide[li] && A->nb_wide > 0) {
                const int g = persistent_grid<k_csrp<256, 4096, AT, MT, XT, Epi, 2>>(A->cus, pcu, A->nb_wide);
                hipLaunchKernelGGL((k_csrp<256, 4096, AT, MT, XT, Epi, 2>), dim3(g), dim3(256), 0, s,
                                   A->rowptr.p, A->col.p, val, x, A->desc_wide.p, A->nb_wide, epi, partials,
                                   ticket);
                M4_HIP(hipGetLastError());
                return MTL4G_OK;
            }
        }
        const int tile = A->nnz_per_block;
        // a persistent grid must be resident at once: never more workgroups
        // per CU than this instantiation's registers and LDS allow
#define M4_LAUNCHP(KERNEL)                                                                     \
    do {                                                                                       \
        const int g = persistent_grid<KERNEL>(A->cus, pcu, A->nb);                            \
        hipLaunchKernelGGL((KERNEL), dim3(g), dim3(256), 0, s, A->rowptr.p, A->col.p, val, x,   \
                           A->desc.p, A->nb, epi, partials, ticket);                          \
    } while (0)
        if (A->kernel == MTL4G_SPMV_STREAM2) {
            if (tile <= 1024) M4_LAUNCHP((k_csrp2<256, 1024, 2, AT, MT, XT, Epi>));
            else M4_LAUNCHP((k_csrp2<256, 2048, 2, AT, MT, XT, Epi>));
        } else {
            if (tile <= 1024) M4_LAUNCHP((k_csrp<256, 1024, AT, MT, XT, Epi>));
            else M4_LAUNCHP((k_csrp<256, 2048, AT, MT, XT, Epi>));
        }
#undef M4_LAUNCHP
        M4_HIP(hipGetLastError());
        return MTL4G_OK;
    }
    const int per = (A->nb + kXcd - 1) / kXcd;
    const int grid = per * kXcd;
    if (Epi::NP > 0 && grid > kMaxPartials) {
        set_error("matrix too large for one reduction pass (%d blocks)", grid);
        return MTL4G_ERR_UNSUPPORTED;
    }
    const int tile = A->nnz_per_block;
#define M4_LAUNCH(T)                                                                          \
    hipLaunchKernelGGL((k_csr<256, T, AT, MT, XT, Epi>), dim3(grid), dim3(256), 0, s,         \
                       A->rowptr.p, A->col.p, val, x, A->desc.p, A->nb, per, epi, partials, ticket, \
                       (AT*)nullptr, (int64_t)0)
    if (tile <= 1024) M4_LAUNCH(1024);
    else if (tile <= 2048) M4_LAUNCH(2048);
    else M4_LAUNCH(4096);
#undef M4_LAUNCH
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

template <typename AT, typename Epi>
int launch_vec(int64_t n, const Epi& epi, AT* partials, unsigned* ticket, hipStream_t s,
               AT* terms = nullptr) {
    if (n <= 0) return MTL4G_OK;
    bool ord;
    M4_TRY((want_ordered<AT, Epi>(terms, ord)));
    if constexpr (Epi::NP > 0 && has_order<Epi>::value && std::is_same<AT, double>::value) {
        if (ord) {
            hipLaunchKernelGGL((k_vec<256, AT, Epi, true>), dim3(blocks_for(n)), dim3(256), 0, s, n,
                               epi, partials, ticket, terms);
            M4_HIP(hipGetLastError());
            hipLaunchKernelGGL((k_ordered<AT, Epi>), dim3(1), dim3(kWave), 0, s, n, (const AT*)terms, epi);
            M4_HIP(hipGetLastError());
            return MTL4G_OK;
        }
    }
    hipLaunchKernelGGL((k_vec<256, AT, Epi>), dim3(blocks_for(n)), dim3(256), 0, s, n, epi,
                       partials, ticket, (AT*)nullptr);
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

}  // namespace mtl4g
