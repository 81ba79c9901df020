// precond.hip -- itl::pc::identity / diagonal / ilu_0 / ic_0 on the device.
//
// ILU(0)/IC(0) factor on the GPU in FP64, each row in MTL4's loop order
// (so the stored factors are the ones MTL4 computes), level scheduled like
// the sweeps; the host plans only the pattern.  The triangular sweeps run on
// the GPU, level scheduled.  A sweep is one persistent kernel: workgroups
// take tickets in order, each owns up to kSweepBS rows of one dependency
// level, waits until every workgroup of the previous level has published
// (bounded wait, abort flag), solves its rows with MTL4's per-row
// subtraction order (FP64 results equal MTL4's bit for bit) and publishes.
// The last workgroup of the last level resets the counters, so a sweep
// needs no memset between launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "kernels.hpp"
#include "precond.hpp"

namespace mtl4g {

constexpr int kSweepBS = 128;
constexpr unsigned kSpinLimit = 1u << 20;   // polls before a wait gives up

// ------------------------------------------------------------------ device
template <typename MT, typename VT>
struct SweepView {
    const int32_t *perm, *rowptr, *col, *lev_ptr, *lev_blk, *blk_level;
    const MT *val, *dia;
    int32_t nlev;
    int unit;
};

// A result not yet written reads as this signalling-NaN pattern: every
// output of a sweep is preset to it (k_fill_pending), and a row reads the
// values it depends on until none is pending.  Arithmetic never produces a
// signalling NaN, so the pattern cannot be a result -- except when a unit
// row with no entries copies an input carrying it; that copy is quieted.
constexpr uint64_t kPending64 = 0x7ff5c3c3c3c3c3c3ull;
constexpr uint32_t kPending32 = 0x7f93c3c3u;
__device__ __forceinline__ bool pending(double v) { return (uint64_t)__double_as_longlong(v) == kPending64; }
__device__ __forceinline__ bool pending(float v) { return __float_as_uint(v) == kPending32; }
__device__ __forceinline__ double quiet_if_pending(double v) {
    return pending(v) ? __longlong_as_double((long long)(kPending64 | 0x0008000000000000ull)) : v;
}
__device__ __forceinline__ float quiet_if_pending(float v) {
    return pending(v) ? __uint_as_float(kPending32 | 0x00400000u) : v;
}

template <typename VT>
__global__ void k_fill_pending(int64_t n, VT* w, const int* skip) {
    if (skip && *skip) return;
    VT p;
    if constexpr (sizeof(VT) == 8) p = __longlong_as_double((long long)kPending64);
    else p = __uint_as_float(kPending32);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = p;
}

// Counters: [0] ticket, [1] abort flag, [2] finished workgroups, [3 + l]
// finished workgroups of level l; the last workgroup to finish resets them,
// so a sweep needs no memset between launches.  Workgroups take tickets in
// level order.  A workgroup of level l first waits (one lane polling) until
// level l-2 is complete, then each row reads the values it depends on until
// none is pending: the rows of level l-1 are seen as soon as they are
// stored, without a barrier per level, and only the next level's rows poll
// values.  Every row depended on belongs to a workgroup that took its
// ticket earlier and is running, so every wait ends -- or gives up after
// kSpinLimit polls and raises the abort flag.
// v (input, type IT) and w are distinct arrays; w holds the results in VT,
// the type the row arithmetic runs in, and w32 (may be null) receives each
// result rounded to float as well.
template <int BS, typename MT, typename IT, typename VT>
__global__ __launch_bounds__(BS) void k_sweep(SweepView<MT, VT> S, const IT* __restrict__ v,
                                              VT* w, float* w32, unsigned* counters, int* err,
                                              const int* skip) {
    if (skip && *skip) return;
    __shared__ int s_b, s_ok, s_final;
    if (threadIdx.x == 0)
        s_b = (int)__hip_atomic_fetch_add(sweep_counter(counters, 0), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int b = s_b;
    const int L = S.blk_level[b];
    // The row's own data (position, bounds, right-hand side, diagonal and its
    // first eight entries) does not depend on earlier levels: request it
    // before waiting.
    const int k = S.lev_ptr[L] + (b - S.lev_blk[L]) * BS + (int)threadIdx.x;
    const bool mine = k < S.lev_ptr[L + 1];
    int r = 0, a = 0, e = 0;
    VT rr = VT(0), dk = VT(1);
    int c[8];
    VT m[8];
    if (mine) {
        r = S.perm[k];
        a = S.rowptr[k];
        e = S.rowptr[k + 1];
        rr = static_cast<VT>(v[r]);
        if (!S.unit) dk = to_arith<VT>(S.dia[k]);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool in = a + u < e;
            c[u] = in ? S.col[a + u] : 0;
            m[u] = in ? to_arith<VT>(S.val[a + u]) : VT(0);
        }
    }
    if (threadIdx.x == 0) {
        int ok = 1;
        if (L > 1) {
            const unsigned need = (unsigned)(S.lev_blk[L - 1] - S.lev_blk[L - 2]);
            unsigned it = 0;
            const unsigned* lev = sweep_counter(counters, 3 + L - 2);
            while (ld_agent(lev) < need) {
                if ((++it & 63u) == 0 && ld_agent(sweep_counter(counters, 1)) != 0u) { ok = 0; break; }
                if (it > kSpinLimit) { st_agent(sweep_counter(counters, 1), 1u); ok = 0; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    if (s_ok && mine) {
        bool ok = true;
        // eight entries at a time: the eight values they name (one round
        // trip, repeated for those still pending), then the subtractions in
        // the row's column order
        for (int j0 = a; j0 < e && ok; j0 += 8) {
            if (j0 > a) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const bool in = j0 + u < e;
                    c[u] = in ? S.col[j0 + u] : 0;
                    m[u] = in ? to_arith<VT>(S.val[j0 + u]) : VT(0);
                }
            }
            VT wv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) wv[u] = j0 + u < e ? ld_agent(w + c[u]) : VT(0);
            unsigned it = 0;
            for (;;) {
                bool wait = false;
#pragma unroll
                for (int u = 0; u < 8; u++) wait |= pending(wv[u]);
                if (!wait) break;
                if (++it > kSpinLimit) {
                    st_agent(sweep_counter(counters, 1), 1u);
                    ok = false;
                    break;
                }
                if ((it & 63u) == 0 && ld_agent(sweep_counter(counters, 1)) != 0u) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (pending(wv[u])) wv[u] = ld_agent(w + c[u]);
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (j0 + u < e) rr -= m[u] * wv[u];
        }
        if (!S.unit) rr = rr * dk;
        rr = quiet_if_pending(rr);
        st_agent(w + r, rr);   // the L1 writes through; wait_memory() below waits for it
        if (w32) w32[r] = (float)rr;
    }
    // ordering (the inter-workgroup protocol of internal.hpp): the rows'
    // agent-scope stores complete before the level counter announces them;
    // the next level's rows also poll the values themselves
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(sweep_counter(counters, 3 + L), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned done =
            __hip_atomic_fetch_add(sweep_counter(counters, 2), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        s_final = done == (unsigned)gridDim.x;
    }
    __syncthreads();
    if (s_final) {   // every other workgroup has finished: reset for the next launch
        if (threadIdx.x == 0 && ld_agent(sweep_counter(counters, 1)) != 0u) atomicOr(err, 1);
        for (int i = threadIdx.x; i < S.nlev + 3; i += BS) st_agent(sweep_counter(counters, i), 0u);
    }
}

template <typename VT, typename MT>
__global__ void k_diag_apply(int64_t n, const MT* __restrict__ d, const VT* __restrict__ r,
                             VT* __restrict__ z, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        z[i] = to_arith<VT>(d[i]) * r[i];
}

template <typename VT>
__global__ void k_copy(int64_t n, const VT* __restrict__ a, VT* __restrict__ b, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// ------------------------------------------------------------------ host
static int64_t find_entry(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                          int32_t r, int32_t c) {
    auto b = ci.begin() + rp[r], e = ci.begin() + rp[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? (int64_t)(it - ci.begin()) : -1;
}

static float round_f32(double v) { return (float)v; }

// A factor value as a sweep stores it: double, float, or bfloat16 bits
// rounded from the float (as the matrix's bfloat16 copy is).
template <typename MT>
__host__ __device__ inline MT stored_as(double v) { return static_cast<MT>(v); }
template <>
__host__ __device__ inline uint16_t stored_as<uint16_t>(double v) { return bf16_bits((float)v); }
// the value a stored factor entry has, as a double (host)
static double stored_value(int fmt, double v) {
    if (fmt == MTL4G_F32) return (double)(float)v;
    if (fmt == MTL4G_BF16) return bf16_value(bf16_bits((float)v));
    return v;
}


}  // namespace mtl4g

#include "chain.hpp"

namespace mtl4g {

Sweep::Sweep() = default;
Sweep::~Sweep() = default;

static std::atomic<int> g_sweep_schedule{MTL4G_SWEEP_AUTO};
static int g_chain_tile[2] = {0, 0};   // chain tile (Tx, Ty) for factors created later; 0: chosen per grid

// A sweep whose rows depend only on the rows 1, Lc, Lc*Cx away (towards the
// start for a lower, the end for an upper sweep), never across the edge of a
// chain of Lc rows or of a plane of Cx chains, with every row's entries in
// ascending column order: the chain pipeline's layout.  Anything else keeps
// the level schedule.
static bool chain_layout(const SweepRows& R, int32_t& Lc, int32_t& Cx, int32_t& Cy, int& ndep) {
    const int32_t n = R.n;
    if (n < 2 || g_sweep_schedule.load() != MTL4G_SWEEP_AUTO) return false;
    // every byte offset the kernel forms (3 FP64 values per row at most)
    // stays below its out-of-range marker, 0x7ffffff0
    if ((int64_t)n * 3 * 8 >= 0x7ffffff0) return false;
    int64_t offs[3];
    int no = 0;
    for (int32_t r = 0; r < n; r++)
        for (int32_t j = R.rowptr[r]; j < R.rowptr[r + 1]; j++) {
            const int64_t d = R.lower ? (int64_t)r - R.col[j] : (int64_t)R.col[j] - r;
            if (d <= 0) return false;
            bool seen = false;
            for (int k = 0; k < no; k++) seen |= offs[k] == d;
            if (!seen) {
                if (no == 3) return false;
                offs[no++] = d;
            }
        }
    if (no == 0) return false;
    std::sort(offs, offs + no);
    if (offs[0] != 1) return false;
    if (no == 1) {
        Lc = n, Cx = 1, Cy = 1;
    } else {
        if (offs[1] > n || n % offs[1]) return false;
        Lc = (int32_t)offs[1];
        if (no == 3) {
            if (offs[2] % offs[1] || n % offs[2]) return false;
            Cx = (int32_t)(offs[2] / offs[1]);
            Cy = (int32_t)(n / offs[2]);
        } else {
            Cx = n / Lc;
            Cy = 1;
        }
    }
    ndep = no;
    const int64_t plane = (int64_t)Lc * Cx;
    for (int32_t r = 0; r < n; r++) {
        const int64_t rm = R.lower ? r : (int64_t)n - 1 - r;   // index in the sweep's own direction
        const int64_t e = rm % Lc, cx = (rm / Lc) % Cx, cy = rm / plane;
        int64_t last = 0;
        for (int32_t j = R.rowptr[r]; j < R.rowptr[r + 1]; j++) {
            const int64_t d = R.lower ? (int64_t)r - R.col[j] : (int64_t)R.col[j] - r;
            if ((d == 1 && e < 1) || (d == Lc && no >= 2 && cx < 1) || (d == plane && no >= 3 && cy < 1))
                return false;
            // subtraction order of the stored row: ascending columns
            const int64_t key = R.lower ? -d : d;
            if (j > R.rowptr[r] && key <= last) return false;
            last = key;
        }
    }
    return true;
}

static int build_chain(const SweepRows& R, int fmt, Sweep* S) {
    int32_t Lc = 0, Cx = 0, Cy = 0;
    int ndep = 0;
    S->chain.reset();
    if (!chain_layout(R, Lc, Cx, Cy, ndep)) return MTL4G_OK;
    std::unique_ptr<ChainSweep> C(new ChainSweep);
    const int32_t n = R.n;
    C->lower = R.lower;
    C->unit = R.unit;
    C->n = n;
    C->Lc = Lc;
    C->Cx = Cx;
    C->Cy = Cy;
    C->ndep = ndep;
    // tile: Tx x Ty = 64 chains, one wave.  Square tiles for 3-D grids: the
    // hand-overs between tiles on the way from the first to the last chain
    // number Cx/Tx + Cy/Ty.
    {
        int ty = 1;
        while (ty < 8 && ty < Cy) ty *= 2;
        C->Tx = kWave / ty;
        C->Ty = ty;
    }
    if (g_chain_tile[0] > 0 && g_chain_tile[1] > 0) {   // set by mtl4g_set_chain_tile
        C->Tx = g_chain_tile[0];
        C->Ty = g_chain_tile[1];
    }
    C->tiles_x = (int)(((int64_t)Cx + C->Tx - 1) / C->Tx);
    C->tiles_y = (int)(((int64_t)Cy + C->Ty - 1) / C->Ty);
    const int64_t T = (int64_t)C->Tx * C->Ty;
    C->stile = ((int64_t)Lc + C->Tx + C->Ty - 2) * T;
    C->nslots = C->stile * C->tiles_x * C->tiles_y;
    // every byte offset the kernel forms stays below its out-of-range
    // marker 0x7ffffff0; else the level schedule stays
    if ((C->nslots * ndep + 1024) * 8 >= 0x7ffffff0) return MTL4G_OK;
    const int64_t ns = C->nslots;
    auto slot = [&](int64_t d) { return d == 1 ? 0 : d == Lc ? 1 : 2; };
    const bool gathered = !R.dsrc.empty();
    std::vector<int32_t> src(gathered ? (size_t)ndep * ns : 0, -1), dsrc(gathered ? (size_t)ns : 0, -1);
    std::vector<double> val(gathered ? 0 : (size_t)ndep * ns, 0.0), dia(gathered ? 0 : (size_t)ns, 1.0);
    std::vector<char> present(gathered ? 0 : (size_t)ndep * ns, 0);
    for (int32_t r = 0; r < n; r++) {
        const int64_t rm = R.lower ? r : (int64_t)n - 1 - r;
        const int64_t at = chain_slot(rm, Lc, Cx, C->Tx, C->Ty, C->tiles_x, C->stile);
        if (gathered) dsrc[at] = R.dsrc[r];
        else if (!R.dia.empty()) dia[at] = R.dia[r];
        for (int32_t j = R.rowptr[r]; j < R.rowptr[r + 1]; j++) {
            const int64_t d = R.lower ? (int64_t)r - R.col[j] : (int64_t)R.col[j] - r;
            const size_t k = (size_t)slot(d) * ns + at;
            if (gathered) src[k] = R.src[j];
            else {
                val[k] = R.val[j];
                present[k] = 1;
            }
        }
    }
    auto up = [](auto& buf, const auto& v) -> int {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        M4_TRY(buf.alloc(std::max<size_t>(v.size(), 1)));
        if (!v.empty()) M4_HIP(hipMemcpy(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return MTL4G_OK;
    };
    const size_t nv = (size_t)ndep * ns;
    if (gathered) {   // values come from the device factor (fill_sweep)
        M4_TRY(up(C->src, src));
        M4_TRY(up(C->dsrc, dsrc));
        if (fmt == MTL4G_F64) {
            M4_TRY(C->val64.alloc(nv));
            M4_TRY(C->dia64.alloc(ns));
        } else if (fmt == MTL4G_F32) {
            M4_TRY(C->val32.alloc(nv));
            M4_TRY(C->dia32.alloc(ns));
        } else {
            M4_TRY(C->val16.alloc(nv));
            M4_TRY(C->dia16.alloc(ns));
        }
    } else {
        auto fill = [&](auto& vbuf, auto& dbuf, auto absent, auto conv) -> int {
            using T = decltype(absent);
            std::vector<T> v(nv), dg(ns);
            for (size_t i = 0; i < nv; i++) v[i] = present[i] ? conv(val[i]) : absent;
            for (int64_t i = 0; i < ns; i++) dg[i] = conv(dia[i]);
            M4_TRY(up(vbuf, v));
            return up(dbuf, dg);
        };
        const uint64_t a64 = kDiaAbsent64;
        const uint32_t a32 = kDiaAbsent32;
        double ad;
        float af;
        std::memcpy(&ad, &a64, 8);
        std::memcpy(&af, &a32, 4);
        if (fmt == MTL4G_F64) M4_TRY(fill(C->val64, C->dia64, ad, [](double x) { return x; }));
        else if (fmt == MTL4G_F32) M4_TRY(fill(C->val32, C->dia32, af, [](double x) { return round_f32(x); }));
        else M4_TRY(fill(C->val16, C->dia16, kDiaAbsent16, [](double x) { return stored_as<uint16_t>(x); }));
    }
    S->chain = std::move(C);
    return MTL4G_OK;
}

int fill_chain(ChainSweep& C, int fmt, const double* fval, hipStream_t s) {
    const int64_t n = C.nslots;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n * C.ndep + 1023) / 1024, 2048));
    const int32_t* dsrc = C.unit ? nullptr : C.dsrc.p;
    if (fmt == MTL4G_F64)
        hipLaunchKernelGGL(k_fill_chain<double>, dim3(grid), dim3(256), 0, s, n, C.ndep, C.src.p, dsrc, fval,
                           C.val64.p, C.dia64.p);
    else if (fmt == MTL4G_F32)
        hipLaunchKernelGGL(k_fill_chain<float>, dim3(grid), dim3(256), 0, s, n, C.ndep, C.src.p, dsrc, fval,
                           C.val32.p, C.dia32.p);
    else
        hipLaunchKernelGGL(k_fill_chain<uint16_t>, dim3(grid), dim3(256), 0, s, n, C.ndep, C.src.p, dsrc, fval,
                           C.val16.p, C.dia16.p);
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

template <typename MT, typename IT, bool LOWER, bool UNIT>
static void launch_chain_t(const ChainSweep& C, const MT* val, const MT* dia, const IT* v, double* w, float* w32,
                           unsigned* counters, int* err, const int* skip, hipStream_t s) {
    ChainView<MT> V{val, dia, C.n, C.Lc, C.Cx, C.Cy, C.Tx, C.Ty, C.tiles_x, (int32_t)C.stile, (int32_t)C.nslots};
    const dim3 grid(C.tiles_x * C.tiles_y), block(C.Tx * C.Ty);
    if (C.ndep == 1)
        hipLaunchKernelGGL((k_chain<MT, IT, LOWER, UNIT, 1>), grid, block, 0, s, V, v, w, w32, counters, err, skip);
    else if (C.ndep == 2)
        hipLaunchKernelGGL((k_chain<MT, IT, LOWER, UNIT, 2>), grid, block, 0, s, V, v, w, w32, counters, err, skip);
    else
        hipLaunchKernelGGL((k_chain<MT, IT, LOWER, UNIT, 3>), grid, block, 0, s, V, v, w, w32, counters, err, skip);
}

template <typename MT, typename IT>
static void launch_chain_m(const ChainSweep& C, const MT* val, const MT* dia, const IT* v, double* w, float* w32,
                           unsigned* counters, int* err, const int* skip, hipStream_t s) {
    if (C.lower && C.unit) launch_chain_t<MT, IT, true, true>(C, val, dia, v, w, w32, counters, err, skip, s);
    else if (C.lower) launch_chain_t<MT, IT, true, false>(C, val, dia, v, w, w32, counters, err, skip, s);
    else if (C.unit) launch_chain_t<MT, IT, false, true>(C, val, dia, v, w, w32, counters, err, skip, s);
    else launch_chain_t<MT, IT, false, false>(C, val, dia, v, w, w32, counters, err, skip, s);
}

template <typename IT>
static int launch_chain(const ChainSweep& C, int fmt, const IT* v, double* w, float* w32, unsigned* counters,
                        int* err, const int* skip, hipStream_t s) {
    if (fmt == MTL4G_F64) launch_chain_m<double, IT>(C, C.val64.p, C.dia64.p, v, w, w32, counters, err, skip, s);
    else if (fmt == MTL4G_F32) launch_chain_m<float, IT>(C, C.val32.p, C.dia32.p, v, w, w32, counters, err, skip, s);
    else launch_chain_m<uint16_t, IT>(C, C.val16.p, C.dia16.p, v, w, w32, counters, err, skip, s);
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

int upload_sweep(const SweepRows& R, int fmt, Sweep* S) {
    const int32_t n = R.n;
    // dependency levels
    std::vector<int32_t> lev(n, 0);
    int32_t nlev = n ? 1 : 0;
    auto level_of = [&](int32_t r) {
        int32_t l = 0;
        for (int32_t j = R.rowptr[r]; j < R.rowptr[r + 1]; j++) l = std::max(l, lev[R.col[j]] + 1);
        lev[r] = l;
        nlev = std::max(nlev, l + 1);
    };
    if (R.lower)
        for (int32_t r = 0; r < n; r++) level_of(r);
    else
        for (int32_t r = n; r-- > 0;) level_of(r);
    // rows ordered by level, ascending row inside a level
    std::vector<int32_t> lev_ptr(nlev + 1, 0), perm(n);
    for (int32_t r = 0; r < n; r++) lev_ptr[lev[r] + 1]++;
    for (int32_t l = 0; l < nlev; l++) lev_ptr[l + 1] += lev_ptr[l];
    {
        std::vector<int32_t> fill(lev_ptr.begin(), lev_ptr.end() - 1);
        for (int32_t r = 0; r < n; r++) perm[fill[lev[r]]++] = r;
    }
    std::vector<int32_t> rowptr(n + 1, 0);
    int64_t nnz = R.rowptr[n];
    const bool gathered = !R.dsrc.empty();   // dsrc has one entry per row on that path
    std::vector<int32_t> col(nnz), src(gathered ? nnz : 0), dsrc(gathered ? n : 0);
    std::vector<double> val(gathered ? 0 : nnz), dia(gathered ? 0 : n, 1.0);
    for (int32_t k = 0; k < n; k++) {
        int32_t r = perm[k];
        int32_t len = R.rowptr[r + 1] - R.rowptr[r];
        rowptr[k + 1] = rowptr[k] + len;
        for (int32_t t = 0; t < len; t++) {
            col[rowptr[k] + t] = R.col[R.rowptr[r] + t];
            if (gathered) src[rowptr[k] + t] = R.src[R.rowptr[r] + t];
            else val[rowptr[k] + t] = R.val[R.rowptr[r] + t];
        }
        if (gathered) dsrc[k] = R.unit ? -1 : R.dsrc[r];
        else if (!R.unit) dia[k] = R.dia[r];
    }
    // workgroups: ceil(rows / kSweepBS) per level
    std::vector<int32_t> lev_blk(nlev + 1, 0), blk_level;
    for (int32_t l = 0; l < nlev; l++) {
        int32_t rows = lev_ptr[l + 1] - lev_ptr[l];
        int32_t nb = (rows + kSweepBS - 1) / kSweepBS;
        lev_blk[l + 1] = lev_blk[l] + nb;
        for (int32_t i = 0; i < nb; i++) blk_level.push_back(l);
    }
    S->n = n;
    S->nnz = nnz;
    S->nlev = nlev;
    S->nblocks = lev_blk[nlev];
    S->unit = R.unit;
    S->h_lev_ptr = lev_ptr;
    auto up = [](auto& buf, const auto& vec) -> int {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        M4_TRY(buf.alloc(vec.size()));
        if (!vec.empty()) M4_HIP(hipMemcpy(buf.p, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice));
        return MTL4G_OK;
    };
    M4_TRY(up(S->perm, perm));
    M4_TRY(up(S->rowptr, rowptr));
    M4_TRY(up(S->col, col));
    M4_TRY(up(S->lev_ptr, lev_ptr));
    M4_TRY(up(S->lev_blk, lev_blk));
    M4_TRY(up(S->blk_level, blk_level));
    if (gathered) {   // values come later from the device factor (fill_sweep)
        M4_TRY(up(S->src, src));
        M4_TRY(up(S->dsrc, dsrc));
        if (fmt == MTL4G_F64) {
            M4_TRY(S->val64.alloc(std::max<int64_t>(nnz, 1)));
            M4_TRY(S->dia64.alloc(std::max(n, 1)));
        } else if (fmt == MTL4G_F32) {
            M4_TRY(S->val32.alloc(std::max<int64_t>(nnz, 1)));
            M4_TRY(S->dia32.alloc(std::max(n, 1)));
        } else {
            M4_TRY(S->val16.alloc(std::max<int64_t>(nnz, 1)));
            M4_TRY(S->dia16.alloc(std::max(n, 1)));
        }
    } else if (fmt == MTL4G_F64) {
        M4_TRY(up(S->val64, val));
        M4_TRY(up(S->dia64, dia));
    } else if (fmt == MTL4G_F32) {
        std::vector<float> v32(val.size()), d32(dia.size());
        for (size_t i = 0; i < val.size(); i++) v32[i] = round_f32(val[i]);
        for (size_t i = 0; i < dia.size(); i++) d32[i] = round_f32(dia[i]);
        M4_TRY(up(S->val32, v32));
        M4_TRY(up(S->dia32, d32));
    } else {
        std::vector<uint16_t> v16(val.size()), d16(dia.size());
        for (size_t i = 0; i < val.size(); i++) v16[i] = stored_as<uint16_t>(val[i]);
        for (size_t i = 0; i < dia.size(); i++) d16[i] = stored_as<uint16_t>(dia[i]);
        M4_TRY(up(S->val16, v16));
        M4_TRY(up(S->dia16, d16));
    }
    M4_TRY(build_chain(R, fmt, S));
    M4_TRY(S->counters.alloc(sweep_counter_words(nlev)));
    M4_HIP(hipMemset(S->counters.p, 0, sweep_counter_words(nlev) * sizeof(unsigned)));
    // complete before any stream (including non-blocking ones) uses the counters
    M4_HIP(hipDeviceSynchronize());
    return MTL4G_OK;
}

// ------------------------------------------------------------------ device factorization
// ILU(0) and IC(0) factor row by row in the lower sweep's level order, with
// the sweep's ticket and level-counter protocol: a row starts once every
// row its strictly lower entries name is final.  Values other rows write
// are read with agent-scope loads and written with agent-scope stores.
struct FactorView {
    const int32_t *perm, *lev_ptr, *lev_blk, *blk_level;
    const int32_t *lrp, *lcol, *lsrc;   // IC(0): lower sweep rows (level order), U positions
    int32_t nlev;
};

// ilu_0_factorizer::factorize (itl/pc/ilu_0.hpp:45-82), row r: each
// strictly lower entry k in column order becomes a_rk * inv_dia[k], then
// every later entry j of the row loses a_rk * a_kj (a_kj looked up in row
// k by column, 0 where row k has none); inv_dia[r] = 1 / a_rr.
__device__ __forceinline__ void ilu0_row(int r, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                         const int32_t* __restrict__ dpos, double* lu, double* finv) {
    const int a = rp[r], e = rp[r + 1];
    for (int kc = a; kc < e; kc++) {
        const int k = ci[kc];
        if (k == r) break;
        const double ark = ld_agent(lu + kc) * ld_agent(finv + k);
        st_agent(lu + kc, ark);
        const int ka = rp[k], ke = rp[k + 1];
        for (int jc = kc + 1; jc < e; jc++) {
            const int j = ci[jc];
            int lo = ka, hi = ke;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            const double akj = (lo < ke && ci[lo] == j) ? ld_agent(lu + lo) : 0.0;
            st_agent(lu + jc, ld_agent(lu + jc) - ark * akj);
        }
    }
    st_agent(finv + r, 1.0 / ld_agent(lu + dpos[r]));
}

// ic_0::factorizer::factorize (itl/pc/ic_0.hpp:135-186) gathered by row:
// U[r][r] loses d*d for every d = U[k][r] * inv_dia[k], k ascending (the
// order the row-k steps reach it); inv_dia[r] = 1 / sqrt(U[r][r]).
__device__ __forceinline__ void ic0_row(int q, int r, const FactorView& F, const int32_t* __restrict__ urp,
                                        double* u, double* finv) {
    double dg = ld_agent(u + urp[r]);
    for (int t = F.lrp[q]; t < F.lrp[q + 1]; t++) {
        const double d = ld_agent(u + F.lsrc[t]) * ld_agent(finv + F.lcol[t]);
        dg = dg - d * d;
    }
    st_agent(finv + r, 1.0 / sqrt(dg));
}

// pos: ILU(0) diagonal positions in A; IC(0) row pointers of U
template <int BS, bool IC>
__global__ __launch_bounds__(BS) void k_factor(FactorView F, const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ ci, const int32_t* __restrict__ pos,
                                               double* fv, double* finv, unsigned* counters, int* err) {
    __shared__ int s_b, s_ok, s_final;
    if (threadIdx.x == 0)
        s_b = (int)__hip_atomic_fetch_add(sweep_counter(counters, 0), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int b = s_b;
    const int L = F.blk_level[b];
    const int q = F.lev_ptr[L] + (b - F.lev_blk[L]) * BS + (int)threadIdx.x;
    const bool mine = q < F.lev_ptr[L + 1];
    if (threadIdx.x == 0) {
        int ok = 1;
        if (L > 0) {
            const unsigned need = (unsigned)(F.lev_blk[L] - F.lev_blk[L - 1]);
            unsigned it = 0;
            const unsigned* lev = sweep_counter(counters, 2 + L - 1);
            while (ld_agent(lev) < need) {
                if ((++it & 63u) == 0 && ld_agent(sweep_counter(counters, 1)) != 0u) { ok = 0; break; }
                if (it > kSpinLimit) { st_agent(sweep_counter(counters, 1), 1u); ok = 0; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    if (s_ok && mine) {
        const int r = F.perm[q];
        if (IC) ic0_row(q, r, F, pos, fv, finv);
        else ilu0_row(r, rp, ci, pos, fv, finv);
    }
    // ordering (the inter-workgroup protocol of internal.hpp): the level's
    // agent-scope row writes complete before its counter announces them to
    // the next level, which reads the rows with agent-scope loads
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned need = (unsigned)(F.lev_blk[L + 1] - F.lev_blk[L]);
        const unsigned done =
            __hip_atomic_fetch_add(sweep_counter(counters, 2 + L), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        s_final = (L == F.nlev - 1 && done == need);
    }
    __syncthreads();
    if (s_final) {   // every other workgroup has published: reset for the next launch
        if (threadIdx.x == 0 && ld_agent(sweep_counter(counters, 1)) != 0u) atomicOr(err, 1);
        for (int i = threadIdx.x; i < F.nlev + 2; i += BS) st_agent(sweep_counter(counters, i), 0u);
    }
}

// U's entries from A: u[p] = a[src[p]]
__global__ void k_gather_values(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ a,
                                double* __restrict__ u) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        u[i] = a[src[i]];
}

// ILU(0) invert_diagonal (ilu_0.hpp:78-79): a_rr <- 1/a_rr, the inv_dia of the row
__global__ void k_ilu0_finish(int32_t n, const int32_t* __restrict__ dpos, const double* __restrict__ finv,
                              double* __restrict__ lu) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        lu[dpos[r]] = finv[r];
}

// IC(0) rows of U: U[r][r] <- inv_dia[r], U[r][i] <- U[r][i] * inv_dia[r]  (ic_0.hpp:143-150)
__global__ void k_ic0_finish(int32_t n, const int32_t* __restrict__ urp, const double* __restrict__ finv,
                             double* __restrict__ u) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double f = finv[r];
        const int a = urp[r], e = urp[r + 1];
        u[a] = f;
        for (int p = a + 1; p < e; p++) u[p] = u[p] * f;
    }
}

// sweep values from the factor: val[k] = fv[src[k]], dia[k] = fv[dsrc[k]] (1 for a unit diagonal)
template <typename MT>
__global__ void k_fill_sweep(int64_t nnz, int32_t n, const int32_t* __restrict__ src,
                             const int32_t* __restrict__ dsrc, const double* __restrict__ fv,
                             MT* __restrict__ val, MT* __restrict__ dia) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += stride)
        val[i] = stored_as<MT>(fv[src[i]]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
        dia[i] = dsrc[i] < 0 ? stored_as<MT>(1.0) : stored_as<MT>(fv[dsrc[i]]);
}

static int grid_for(int64_t n) {
    int64_t g = (n + 1023) / 1024;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

template <typename IT, typename VT>
int launch_sweep(const Sweep& S, int fmt, const IT* v, VT* w, float* w32, const int* skip, int* err,
                 unsigned* counters, hipStream_t s) {
    if (S.n == 0) return MTL4G_OK;
    if (!counters) counters = S.counters.p;
    hipLaunchKernelGGL(k_fill_pending<VT>, dim3(grid_for(S.n)), dim3(256), 0, s, (int64_t)S.n, w, skip);
    if (S.chain) {
        static_assert(std::is_same<VT, double>::value, "the chain pipeline keeps its results in FP64");
        return launch_chain<IT>(*S.chain, fmt, v, w, w32, counters, err, skip, s);
    }
    if (fmt == MTL4G_F64) {
        SweepView<double, VT> V{S.perm.p, S.rowptr.p, S.col.p, S.lev_ptr.p, S.lev_blk.p,
                                S.blk_level.p, S.val64.p, S.dia64.p, S.nlev, S.unit ? 1 : 0};
        hipLaunchKernelGGL((k_sweep<kSweepBS, double, IT, VT>), dim3(S.nblocks), dim3(kSweepBS), 0, s,
                           V, v, w, w32, counters, err, skip);
    } else if (fmt == MTL4G_F32) {
        SweepView<float, VT> V{S.perm.p, S.rowptr.p, S.col.p, S.lev_ptr.p, S.lev_blk.p,
                               S.blk_level.p, S.val32.p, S.dia32.p, S.nlev, S.unit ? 1 : 0};
        hipLaunchKernelGGL((k_sweep<kSweepBS, float, IT, VT>), dim3(S.nblocks), dim3(kSweepBS), 0, s,
                           V, v, w, w32, counters, err, skip);
    } else {
        SweepView<uint16_t, VT> V{S.perm.p, S.rowptr.p, S.col.p, S.lev_ptr.p, S.lev_blk.p,
                                  S.blk_level.p, S.val16.p, S.dia16.p, S.nlev, S.unit ? 1 : 0};
        hipLaunchKernelGGL((k_sweep<kSweepBS, uint16_t, IT, VT>), dim3(S.nblocks), dim3(kSweepBS), 0, s,
                           V, v, w, w32, counters, err, skip);
    }
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

template int launch_sweep<double, double>(const Sweep&, int, const double*, double*, float*, const int*, int*,
                                          unsigned*, hipStream_t);
template int launch_sweep<float, double>(const Sweep&, int, const float*, double*, float*, const int*, int*,
                                         unsigned*, hipStream_t);


size_t pc_counter_words(const mtl4g_pc_s* P) {
    return sweep_counter_words(P->lower.nlev) + sweep_counter_words(P->upper.nlev);
}

template <typename VT>
int pc_apply(const mtl4g_pc_s* P, const VT* r, VT* z, VT* y0, const int* skip,
             unsigned* counters, hipStream_t s, double* w64) {
    const int64_t n = P->n;
    if (n == 0) return MTL4G_OK;
    switch (P->kind) {
    case MTL4G_PC_IDENTITY:
        hipLaunchKernelGGL((k_copy<VT>), dim3(grid_for(n)), dim3(256), 0, s, n, r, z, skip);
        break;
    case MTL4G_PC_DIAGONAL:
        if (P->fmt == MTL4G_F64)
            hipLaunchKernelGGL((k_diag_apply<VT, double>), dim3(grid_for(n)), dim3(256), 0, s, n,
                               P->inv_diag64.p, r, z, skip);
        else if (P->fmt == MTL4G_F32)
            hipLaunchKernelGGL((k_diag_apply<VT, float>), dim3(grid_for(n)), dim3(256), 0, s, n,
                               P->inv_diag32.p, r, z, skip);
        else
            hipLaunchKernelGGL((k_diag_apply<VT, uint16_t>), dim3(grid_for(n)), dim3(256), 0, s, n,
                               P->inv_diag16.p, r, z, skip);
        break;
    default: {
        unsigned* cu = counters ? counters + sweep_counter_words(P->lower.nlev) : nullptr;
        if constexpr (sizeof(VT) == 8) {
            M4_TRY((launch_sweep<double, double>(P->lower, P->fmt, r, y0, nullptr, skip, P->err.p, counters, s)));
            M4_TRY((launch_sweep<double, double>(P->upper, P->fmt, y0, z, nullptr, skip, P->err.p, cu, s)));
        } else {
            // float vectors: each sweep's chain runs in FP64 (w64: 2n doubles),
            // only the result is rounded to float -- a rounding error made
            // inside a chain would be amplified along it, and a preconditioner
            // that is not one fixed linear map stalls the Krylov recurrence
            if (!w64) {
                set_error("float preconditioner solve without its FP64 sweep buffer");
                return MTL4G_ERR_INVALID;
            }
            (void)y0;
            M4_TRY((launch_sweep<float, double>(P->lower, P->fmt, r, w64, nullptr, skip, P->err.p, counters, s)));
            M4_TRY((launch_sweep<double, double>(P->upper, P->fmt, w64, w64 + n, z, skip, P->err.p, cu, s)));
        }
        return MTL4G_OK;
    }
    }
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

template int pc_apply<double>(const mtl4g_pc_s*, const double*, double*, double*, const int*, unsigned*, hipStream_t,
                              double*);
template int pc_apply<float>(const mtl4g_pc_s*, const float*, float*, float*, const int*, unsigned*, hipStream_t,
                             double*);

int pc_check_error(const mtl4g_pc_s* P, hipStream_t s) {
    if (!P || !P->err.p) return MTL4G_OK;
    int h = 0;
    M4_HIP(hipMemcpyAsync(&h, P->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    if (h) {
        M4_HIP(hipMemsetAsync(P->err.p, 0, sizeof(int), s));
        set_error("triangular sweep: a bounded level wait expired");
        return MTL4G_ERR_TIMEOUT;
    }
    return MTL4G_OK;
}

// ---- factorizations: the pattern is planned on the host, the values are
// computed on the device (k_factor) in MTL4's loop order per row, so the
// stored factors are the ones MTL4 computes bit for bit.

int fill_sweep(Sweep& S, int fmt, const double* fval, hipStream_t s) {
    if (S.n == 0) return MTL4G_OK;
    if (S.chain) M4_TRY(fill_chain(*S.chain, fmt, fval, s));
    const int grid = grid_for(std::max<int64_t>(S.nnz, S.n));
    if (fmt == MTL4G_F64)
        hipLaunchKernelGGL(k_fill_sweep<double>, dim3(grid), dim3(256), 0, s, S.nnz, S.n, S.src.p, S.dsrc.p, fval,
                           S.val64.p, S.dia64.p);
    else if (fmt == MTL4G_F32)
        hipLaunchKernelGGL(k_fill_sweep<float>, dim3(grid), dim3(256), 0, s, S.nnz, S.n, S.src.p, S.dsrc.p, fval,
                           S.val32.p, S.dia32.p);
    else
        hipLaunchKernelGGL(k_fill_sweep<uint16_t>, dim3(grid), dim3(256), 0, s, S.nnz, S.n, S.src.p, S.dsrc.p, fval,
                           S.val16.p, S.dia16.p);
    M4_HIP(hipGetLastError());
    return MTL4G_OK;
}

// ilu_0 (itl/pc/ilu_0.hpp:45-82): L = strictly lower part (unit diagonal),
// U = diagonal (inverted) and upper part, both in A's positions; every row
// needs its diagonal entry.
static int plan_ilu0(mtl4g_pc_s* P, SweepRows& lo, SweepRows& up, std::vector<int32_t>& dpos) {
    const mtl4g_matrix_s* A = P->A;
    const int32_t n = A->n;
    const auto& rp = A->h_rowptr;
    const auto& ci = A->h_col;
    dpos.assign(n, 0);
    for (int32_t i = 0; i < n; i++) {
        const int64_t d = find_entry(rp, ci, i, i);
        if (d < 0) {
            set_error("ilu_0: row %d has no diagonal entry", i);
            return MTL4G_ERR_MISSING_DIAG;
        }
        dpos[i] = (int32_t)d;
    }
    P->L_rowptr.assign(n + 1, 0);
    P->U_rowptr.assign(n + 1, 0);
    for (int32_t r = 0; r < n; r++) {
        for (int32_t j = rp[r]; j < rp[r + 1]; j++) {
            if (ci[j] < r) { P->L_col.push_back(ci[j]); P->L_src.push_back(j); }
            else { P->U_col.push_back(ci[j]); P->U_src.push_back(j); }
        }
        P->L_rowptr[r + 1] = (int32_t)P->L_col.size();
        P->U_rowptr[r + 1] = (int32_t)P->U_col.size();
    }
    lo.n = up.n = n;
    lo.lower = true;
    lo.unit = true;
    lo.rowptr = P->L_rowptr;
    lo.col = P->L_col;
    lo.src = P->L_src;
    lo.dsrc.assign(n, -1);
    up.lower = false;
    up.unit = false;
    up.rowptr.assign(n + 1, 0);
    up.dsrc.assign(n, 0);
    for (int32_t r = 0; r < n; r++) {   // inverse diagonal first (upper_trisolve v3)
        up.dsrc[r] = P->U_src[P->U_rowptr[r]];
        for (int32_t j = P->U_rowptr[r] + 1; j < P->U_rowptr[r + 1]; j++) {
            up.col.push_back(P->U_col[j]);
            up.src.push_back(P->U_src[j]);
        }
        up.rowptr[r + 1] = (int32_t)up.col.size();
    }
    return MTL4G_OK;
}

// ic_0 (itl/pc/ic_0.hpp:135-186) on U = upper(A), L = trans(U) (:57) with
// the inverse diagonal last; values in U's order.  usrc: A position of
// each U entry.
static int plan_ic0(mtl4g_pc_s* P, SweepRows& lo, SweepRows& up, std::vector<int32_t>& usrc) {
    const mtl4g_matrix_s* A = P->A;
    const int32_t n = A->n;
    auto& urp = P->U_rowptr;
    auto& uci = P->U_col;
    urp.assign(n + 1, 0);
    for (int32_t r = 0; r < n; r++) {
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++)
            if (A->h_col[j] >= r) { uci.push_back(A->h_col[j]); usrc.push_back(j); }
        urp[r + 1] = (int32_t)uci.size();
    }
    for (int32_t k = 0; k < n; k++)
        if (urp[k] == urp[k + 1] || uci[urp[k]] != k) {
            set_error("ic_0: row %d has no diagonal entry", k);
            return MTL4G_ERR_MISSING_DIAG;
        }
    P->U_src.resize(uci.size());
    for (size_t p = 0; p < uci.size(); p++) P->U_src[p] = (int32_t)p;
    auto& lrp = P->L_rowptr;
    lrp.assign(n + 1, 0);
    for (int32_t c : uci) lrp[c + 1]++;
    for (int32_t r = 0; r < n; r++) lrp[r + 1] += lrp[r];
    P->L_col.resize(uci.size());
    P->L_src.resize(uci.size());
    std::vector<int32_t> fill(lrp.begin(), lrp.end() - 1);
    for (int32_t r = 0; r < n; r++)
        for (int32_t j = urp[r]; j < urp[r + 1]; j++) {
            const int32_t d = fill[uci[j]]++;
            P->L_col[d] = r;
            P->L_src[d] = j;
        }
    lo.n = up.n = n;
    lo.lower = true;
    lo.unit = false;
    lo.rowptr.assign(n + 1, 0);
    lo.dsrc.assign(n, 0);
    for (int32_t r = 0; r < n; r++) {   // inverse diagonal last (lower_trisolve v6)
        const int32_t a = lrp[r], e = lrp[r + 1] - 1;
        lo.dsrc[r] = P->L_src[e];
        for (int32_t j = a; j < e; j++) {
            lo.col.push_back(P->L_col[j]);
            lo.src.push_back(P->L_src[j]);
        }
        lo.rowptr[r + 1] = (int32_t)lo.col.size();
    }
    up.lower = false;
    up.unit = false;
    up.rowptr.assign(n + 1, 0);
    up.dsrc.assign(n, 0);
    for (int32_t r = 0; r < n; r++) {   // inverse diagonal first (upper_trisolve v3)
        up.dsrc[r] = urp[r];
        for (int32_t j = urp[r] + 1; j < urp[r + 1]; j++) {
            up.col.push_back(uci[j]);
            up.src.push_back(j);
        }
        up.rowptr[r + 1] = (int32_t)up.col.size();
    }
    return MTL4G_OK;
}

// Factor values from A's current device values, then the sweeps' values.
static int factor_device(mtl4g_pc_s* P, hipStream_t s) {
    const mtl4g_matrix_s* A = P->A;
    const int32_t n = P->n;
    if (n == 0) return MTL4G_OK;
    const bool ic = P->kind == MTL4G_PC_IC0;
    const int64_t nf = ic ? (int64_t)P->U_col.size() : A->nnz;
    if (ic)
        hipLaunchKernelGGL(k_gather_values, dim3(grid_for(nf)), dim3(256), 0, s, nf, P->fpos.p, A->val.p, P->fval.p);
    else
        M4_HIP(hipMemcpyAsync(P->fval.p, A->val.p, (size_t)nf * sizeof(double), hipMemcpyDeviceToDevice, s));
    M4_HIP(hipGetLastError());
    const Sweep& Lo = P->lower;
    FactorView F{Lo.perm.p, Lo.lev_ptr.p, Lo.lev_blk.p, Lo.blk_level.p, Lo.rowptr.p, Lo.col.p, Lo.src.p, Lo.nlev};
    if (ic) {
        hipLaunchKernelGGL((k_factor<kSweepBS, true>), dim3(Lo.nblocks), dim3(kSweepBS), 0, s, F, A->rowptr.p,
                           A->col.p, P->urowptr.p, P->fval.p, P->finv.p, P->fcounters.p, P->err.p);
        hipLaunchKernelGGL(k_ic0_finish, dim3(grid_for(n)), dim3(256), 0, s, n, P->urowptr.p, P->finv.p, P->fval.p);
    } else {
        hipLaunchKernelGGL((k_factor<kSweepBS, false>), dim3(Lo.nblocks), dim3(kSweepBS), 0, s, F, A->rowptr.p,
                           A->col.p, P->fpos.p, P->fval.p, P->finv.p, P->fcounters.p, P->err.p);
        hipLaunchKernelGGL(k_ilu0_finish, dim3(grid_for(n)), dim3(256), 0, s, n, P->fpos.p, P->finv.p, P->fval.p);
    }
    M4_HIP(hipGetLastError());
    M4_TRY(fill_sweep(P->lower, P->fmt, P->fval.p, s));
    M4_TRY(fill_sweep(P->upper, P->fmt, P->fval.p, s));
    P->host_values = false;
    P->has_adjoint = false;
    return MTL4G_OK;
}

static int setup_factor(mtl4g_pc_s* P) {
    SweepRows lo, up;
    std::vector<int32_t> pos;
    const bool ic = P->kind == MTL4G_PC_IC0;
    M4_TRY(ic ? plan_ic0(P, lo, up, pos) : plan_ilu0(P, lo, up, pos));
    M4_TRY(upload_sweep(lo, P->fmt, &P->lower));
    M4_TRY(upload_sweep(up, P->fmt, &P->upper));
    auto up32 = [](DevBuf<int32_t>& buf, const std::vector<int32_t>& v) -> int {
        M4_TRY(buf.alloc(v.size()));
        if (!v.empty()) M4_HIP(hipMemcpy(buf.p, v.data(), v.size() * 4, hipMemcpyHostToDevice));
        return MTL4G_OK;
    };
    M4_TRY(up32(P->fpos, pos));
    if (ic) M4_TRY(up32(P->urowptr, P->U_rowptr));
    M4_TRY(P->fval.alloc(ic ? P->U_col.size() : (size_t)P->A->nnz));
    M4_TRY(P->finv.alloc((size_t)P->n));
    const size_t words = ((size_t)P->lower.nlev + 2) * kCounterStride;
    M4_TRY(P->fcounters.alloc(words));
    M4_HIP(hipMemset(P->fcounters.p, 0, words * sizeof(unsigned)));
    M4_HIP(hipDeviceSynchronize());
    M4_TRY(factor_device(P, nullptr));
    M4_HIP(hipDeviceSynchronize());
    return pc_check_error(P, nullptr);
}

// Host copies of the factor values (mtl4g_pc_factor, the adjoint sweeps),
// rounded as the sweeps store them.
static int ensure_host_values(mtl4g_pc_s* P) {
    if (P->host_values || (P->kind != MTL4G_PC_ILU0 && P->kind != MTL4G_PC_IC0)) return MTL4G_OK;
    std::vector<double> fv(P->fval.n);
    M4_HIP(hipDeviceSynchronize());
    if (!fv.empty()) M4_HIP(hipMemcpy(fv.data(), P->fval.p, fv.size() * 8, hipMemcpyDeviceToHost));
    auto store = [&](double v) { return stored_value(P->fmt, v); };
    P->L_val.resize(P->L_src.size());
    P->U_val.resize(P->U_src.size());
    for (size_t j = 0; j < P->L_src.size(); j++) P->L_val[j] = store(fv[P->L_src[j]]);
    for (size_t j = 0; j < P->U_src.size(); j++) P->U_val[j] = store(fv[P->U_src[j]]);
    P->host_values = true;
    return MTL4G_OK;
}

// Sweep rows from a stored factor.
//  kind 0: unit lower on strict L           (lower_trisolve v5, :174-189)
//  kind 1: lower, inverse diagonal last      (lower_trisolve v6, :193-209)
//  kind 2: upper, inverse diagonal first     (upper_trisolve v3, :161-176)
static SweepRows rows_from(int32_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                           const std::vector<double>& va, int kind) {
    SweepRows R;
    R.n = n;
    R.lower = kind != 2;
    R.unit = kind == 0;
    R.rowptr.assign(n + 1, 0);
    R.dia.assign(n, 1.0);
    for (int32_t r = 0; r < n; r++) {
        int32_t a = rp[r], b = rp[r + 1];
        if (kind == 1) { R.dia[r] = va[b - 1]; b--; }
        if (kind == 2) { R.dia[r] = va[a]; a++; }
        for (int32_t j = a; j < b; j++) { R.col.push_back(ci[j]); R.val.push_back(va[j]); }
        R.rowptr[r + 1] = (int32_t)R.col.size();
    }
    return R;
}

// Adjoint sweeps of ilu (ilu.hpp:105-115) on the transposed views: the
// column-oriented generic sweeps (lower_trisolve.hpp:154-170 on trans(U),
// upper_trisolve.hpp:128-147 on trans(L)) subtract in the order the columns
// are visited: ascending for the lower, descending for the upper one.
static int build_adjoint(mtl4g_pc_s* P) {
    const int32_t n = P->n;
    M4_TRY(ensure_host_values(P));
    auto transpose = [n](const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                         const std::vector<double>& va, std::vector<int32_t>& trp,
                         std::vector<int32_t>& tci, std::vector<double>& tva) {
        trp.assign(n + 1, 0);
        for (int32_t c : ci) trp[c + 1]++;
        for (int32_t r = 0; r < n; r++) trp[r + 1] += trp[r];
        tci.resize(ci.size());
        tva.resize(ci.size());
        std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
        for (int32_t r = 0; r < n; r++)
            for (int32_t j = rp[r]; j < rp[r + 1]; j++) {
                int32_t d = fill[ci[j]]++;
                tci[d] = r;
                tva[d] = va[j];
            }
    };
    std::vector<int32_t> trp, tci;
    std::vector<double> tva;
    transpose(P->U_rowptr, P->U_col, P->U_val, trp, tci, tva);   // trans(U): diagonal last
    M4_TRY(upload_sweep(rows_from(n, trp, tci, tva, 1), P->fmt, &P->adj_lower));
    transpose(P->L_rowptr, P->L_col, P->L_val, trp, tci, tva);   // trans(L): strictly upper
    SweepRows R = rows_from(n, trp, tci, tva, 0);
    R.lower = false;
    for (int32_t r = 0; r < n; r++) {   // descending subtraction order
        std::reverse(R.col.begin() + R.rowptr[r], R.col.begin() + R.rowptr[r + 1]);
        std::reverse(R.val.begin() + R.rowptr[r], R.val.begin() + R.rowptr[r + 1]);
    }
    M4_TRY(upload_sweep(R, P->fmt, &P->adj_upper));
    P->has_adjoint = true;
    return MTL4G_OK;
}

}  // namespace mtl4g

namespace mtl4g {
// FP32 / BF16 inverse diagonal of a diagonal preconditioner: rounded from
// h_inv_diag, which then keeps the stored values.
static int upload_inv_diag(mtl4g_pc_s* P, int32_t n) {
    if (P->fmt == MTL4G_F32) {
        std::vector<float> f(P->h_inv_diag.begin(), P->h_inv_diag.end());
        for (size_t i = 0; i < f.size(); i++) P->h_inv_diag[i] = f[i];
        if (!P->inv_diag32.p) M4_TRY(P->inv_diag32.alloc(std::max(n, 1)));
        if (n) M4_HIP(hipMemcpy(P->inv_diag32.p, f.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    } else {
        std::vector<uint16_t> h(n);
        for (int32_t i = 0; i < n; i++) {
            h[i] = stored_as<uint16_t>(P->h_inv_diag[i]);
            P->h_inv_diag[i] = bf16_value(h[i]);
        }
        if (!P->inv_diag16.p) M4_TRY(P->inv_diag16.alloc(std::max(n, 1)));
        if (n) M4_HIP(hipMemcpy(P->inv_diag16.p, h.data(), (size_t)n * 2, hipMemcpyHostToDevice));
    }
    return MTL4G_OK;
}
}  // namespace mtl4g

using namespace mtl4g;

extern "C" {

int mtl4g_pc_create(mtl4g_pc* out, mtl4g_matrix A, int kind, int fmt) {
    if (!out || !A || kind < MTL4G_PC_IDENTITY || kind > MTL4G_PC_IC0 ||
        (fmt != MTL4G_F64 && fmt != MTL4G_F32 && fmt != MTL4G_BF16)) {
        set_error("mtl4g_pc_create: bad arguments");
        return MTL4G_ERR_INVALID;
    }
    *out = nullptr;
    if (A->n != A->m) {
        set_error("preconditioner needs a square matrix");
        return MTL4G_ERR_NOT_SQUARE;
    }
    TraceRange tr(kind == MTL4G_PC_ILU0 ? "ilu_0::factorize" : kind == MTL4G_PC_IC0 ? "ic_0::factorize"
                  : kind == MTL4G_PC_DIAGONAL ? "diagonal::setup" : "identity::setup");
    mtl4g_pc_s* P = new (std::nothrow) mtl4g_pc_s;
    if (!P) return MTL4G_ERR_OOM;
    P->kind = kind;
    P->fmt = fmt;
    P->n = A->n;
    P->A = A;
    int st = P->err.alloc(1);
    if (!st && (hipMemset(P->err.p, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess))
        st = MTL4G_ERR_HIP;   // cleared before any (non-blocking) stream reads it
    if (!st && kind == MTL4G_PC_DIAGONAL) {
        // diagonal(A): inv_diag[i] = reciprocal(A[i][i]) (diagonal.hpp:37-45)
        P->h_inv_diag.assign(A->n, 0.0);
        for (int32_t r = 0; r < A->n; r++) {
            int64_t d = find_entry(A->h_rowptr, A->h_col, r, r);
            P->h_inv_diag[r] = 1.0 / (d >= 0 ? A->h_val[d] : 0.0);
        }
        if (fmt == MTL4G_F64) {
            st = P->inv_diag64.alloc(A->n);
            if (!st && A->n)
                st = hipMemcpy(P->inv_diag64.p, P->h_inv_diag.data(), A->n * 8, hipMemcpyHostToDevice) ? MTL4G_ERR_HIP : 0;
        } else {
            st = upload_inv_diag(P, A->n);
        }
    } else if (!st && (kind == MTL4G_PC_ILU0 || kind == MTL4G_PC_IC0)) {
        st = setup_factor(P);
    }
    if (st) {
        delete P;
        return st;
    }
    *out = P;
    return MTL4G_OK;
}

int mtl4g_pc_destroy(mtl4g_pc P) {
    delete P;
    return MTL4G_OK;
}

static int pc_solve_impl(mtl4g_pc P, const double* x, double* y, void* stream, bool adjoint) {
    if (!P || (P->n && (!x || !y))) return MTL4G_ERR_INVALID;
    static const char* const kSolve[4][2] = {{"identity::solve", "identity::adjoint_solve"},
                                             {"diagonal::solve", "diagonal::adjoint_solve"},
                                             {"ilu_0::solve", "ilu_0::adjoint_solve"},
                                             {"ic_0::solve", "ic_0::adjoint_solve"}};
    TraceRange tr(kSolve[std::min(std::max(P->kind, 0), 3)][adjoint ? 1 : 0]);
    hipStream_t s = S(stream);
    if (P->n == 0) return MTL4G_OK;
    if (P->kind == MTL4G_PC_ILU0 || P->kind == MTL4G_PC_IC0) {
        if (adjoint && P->kind == MTL4G_PC_ILU0 && !P->has_adjoint) M4_TRY(build_adjoint(P));
        double* y0 = nullptr;
        M4_HIP(hipMallocAsync((void**)&y0, (size_t)P->n * sizeof(double), s));
        int st;
        if (adjoint && P->kind == MTL4G_PC_ILU0) {
            st = launch_sweep<double, double>(P->adj_lower, P->fmt, x, y0, nullptr, nullptr, P->err.p, nullptr, s);
            if (!st) st = launch_sweep<double, double>(P->adj_upper, P->fmt, y0, y, nullptr, nullptr, P->err.p, nullptr, s);
        } else {
            st = pc_apply<double>(P, x, y, y0, nullptr, nullptr, s);
        }
        (void)hipFreeAsync(y0, s);
        M4_TRY(st);
        return pc_check_error(P, s);
    }
    return pc_apply<double>(P, x, y, nullptr, nullptr, nullptr, s);
}

int mtl4g_pc_solve(mtl4g_pc P, const double* x, double* y, void* stream) {
    return pc_solve_impl(P, x, y, stream, false);
}

int mtl4g_pc_adjoint_solve(mtl4g_pc P, const double* x, double* y, void* stream) {
    return pc_solve_impl(P, x, y, stream, true);
}

int mtl4g_pc_factor(mtl4g_pc P, int which, int32_t* n_out, int64_t* nnz_out, int32_t* rowptr,
                    int32_t* col, double* val) {
    if (!P || which < 0 || which > 2) return MTL4G_ERR_INVALID;
    if (n_out) *n_out = P->n;
    if (which == 0) {
        if (nnz_out) *nnz_out = (int64_t)P->h_inv_diag.size();
        if (val && !P->h_inv_diag.empty())
            std::memcpy(val, P->h_inv_diag.data(), P->h_inv_diag.size() * 8);
        return MTL4G_OK;
    }
    M4_TRY(ensure_host_values(P));
    const auto& rp = which == 1 ? P->L_rowptr : P->U_rowptr;
    const auto& ci = which == 1 ? P->L_col : P->U_col;
    const auto& va = which == 1 ? P->L_val : P->U_val;
    if (rp.empty()) {
        if (nnz_out) *nnz_out = -1;
        return MTL4G_OK;
    }
    if (nnz_out) *nnz_out = (int64_t)ci.size();
    if (rowptr) std::memcpy(rowptr, rp.data(), rp.size() * 4);
    if (col && !ci.empty()) std::memcpy(col, ci.data(), ci.size() * 4);
    if (val && !va.empty()) std::memcpy(val, va.data(), va.size() * 8);
    return MTL4G_OK;
}

int mtl4g_pc_refactor(mtl4g_pc P, void* stream) {
    if (!P || !P->A || P->A->n != P->n) return MTL4G_ERR_INVALID;
    TraceRange tr(P->kind == MTL4G_PC_ILU0 ? "ilu_0::factorize" : P->kind == MTL4G_PC_IC0 ? "ic_0::factorize"
                  : P->kind == MTL4G_PC_DIAGONAL ? "diagonal::setup" : "identity::setup");
    hipStream_t s = S(stream);
    if (P->kind == MTL4G_PC_ILU0 || P->kind == MTL4G_PC_IC0) {
        M4_TRY(factor_device(P, s));
        return pc_check_error(P, s);
    }
    if (P->kind == MTL4G_PC_DIAGONAL) {   // n reciprocals, as at creation (diagonal.hpp:37-45)
        const mtl4g_matrix_s* A = P->A;
        for (int32_t r = 0; r < A->n; r++) {
            const int64_t d = find_entry(A->h_rowptr, A->h_col, r, r);
            P->h_inv_diag[r] = 1.0 / (d >= 0 ? A->h_val[d] : 0.0);
        }
        M4_HIP(hipStreamSynchronize(s));
        if (P->fmt == MTL4G_F64) {
            if (A->n) M4_HIP(hipMemcpy(P->inv_diag64.p, P->h_inv_diag.data(), (size_t)A->n * 8, hipMemcpyHostToDevice));
        } else {
            M4_TRY(upload_inv_diag(P, A->n));
        }
    }
    return MTL4G_OK;
}

int mtl4g_set_sweep_schedule(int schedule) {
    if (schedule != MTL4G_SWEEP_AUTO && schedule != MTL4G_SWEEP_LEVELS) return MTL4G_ERR_INVALID;
    g_sweep_schedule.store(schedule);
    return MTL4G_OK;
}

int mtl4g_set_chain_tile(int32_t tx, int32_t ty) {
    if (tx == 0 && ty == 0) {
        g_chain_tile[0] = g_chain_tile[1] = 0;
        return MTL4G_OK;
    }
    if (tx < 1 || ty < 1 || (int64_t)tx * ty != kWave) {
        set_error("mtl4g_set_chain_tile: Tx*Ty must be %d (one wave)", kWave);
        return MTL4G_ERR_INVALID;
    }
    g_chain_tile[0] = tx;
    g_chain_tile[1] = ty;
    return MTL4G_OK;
}

int mtl4g_pc_chain_tile(mtl4g_pc P, int32_t* lower_tx, int32_t* lower_ty, int32_t* upper_tx, int32_t* upper_ty) {
    if (!P) return MTL4G_ERR_INVALID;
    const ChainSweep* lo = P->lower.chain.get();
    const ChainSweep* hi = P->upper.chain.get();
    if (lower_tx) *lower_tx = lo ? lo->Tx : 0;
    if (lower_ty) *lower_ty = lo ? lo->Ty : 0;
    if (upper_tx) *upper_tx = hi ? hi->Tx : 0;
    if (upper_ty) *upper_ty = hi ? hi->Ty : 0;
    return MTL4G_OK;
}

int mtl4g_pc_sweeps(mtl4g_pc P, int* lower, int* upper) {
    if (!P) return MTL4G_ERR_INVALID;
    if (lower) *lower = P->lower.chain ? MTL4G_SWEEP_AUTO : MTL4G_SWEEP_LEVELS;
    if (upper) *upper = P->upper.chain ? MTL4G_SWEEP_AUTO : MTL4G_SWEEP_LEVELS;
    return MTL4G_OK;
}

int mtl4g_pc_levels(mtl4g_pc P, int32_t* lower_levels, int32_t* upper_levels) {
    if (!P) return MTL4G_ERR_INVALID;
    if (lower_levels) *lower_levels = P->lower.nlev;
    if (upper_levels) *upper_levels = P->upper.nlev;
    return MTL4G_OK;
}

}  // extern "C"
