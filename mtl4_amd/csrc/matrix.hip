// matrix.hip -- compressed2D on the device, CSR x dense_vector kernels and
// the dense_vector reductions/updates of the C ABI.
//
// SpMV (mat_vec_mult.hpp:613-683) is HBM bound: per non-zero it streams
// 4 B of column index and V B of value and gathers x[col] (mostly from
// L2/MALL).  Each workgroup owns a row block sized to an LDS tile of
// products: every lane streams val/col coalesced (4 independent loads in
// flight per lane), forms val*x[col] into LDS, then lane t sums its row in
// column order.  Row lengths are binned by construction (a block holds as
// many rows as fit the tile), and blocks are mapped XCD-contiguously.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include "kernels.hpp"

namespace mtl4g {

static thread_local std::string g_last_error;
static std::atomic<int> g_reduction_order{MTL4G_ORDER_TREE};

int reduction_order() { return g_reduction_order.load(std::memory_order_relaxed); }

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int hip_status(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? MTL4G_ERR_OOM : MTL4G_ERR_HIP;
}

int blocks_for(int64_t n) {
    int64_t g = (n + 256 * 8 - 1) / (256 * 8);
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

// ------------------------------------------------------------------ partition
// Row blocks of at most rows_per_block rows and nnz_per_block non-zeros (a
// longer row gets a block of its own), in row order.
static std::vector<RowBlock> partition_rows(const mtl4g_matrix_s* A, int rows_per_block, int nnz_per_block) {
    const std::vector<int32_t>& rp = A->h_rowptr;
    std::vector<RowBlock> out;
    if (nnz_per_block >= 64) nnz_per_block -= 2;   // headroom for aligned pair loads
    int32_t r = 0;
    while (r < A->n) {
        const int32_t r0 = r;
        int64_t nz = 0;
        while (r < A->n && r - r0 < rows_per_block && nz + (rp[r + 1] - rp[r]) <= nnz_per_block) {
            nz += rp[r + 1] - rp[r];
            r++;
        }
        if (r == r0) r++;
        out.push_back(RowBlock{r0, r, rp[r0], rp[r]});
    }
    return out;
}

int build_row_blocks(mtl4g_matrix_s* A, int rows_per_block, int nnz_per_block) {
    if (rows_per_block <= 0) rows_per_block = 256;
    if (nnz_per_block <= 0) nnz_per_block = 2048;
    if (rows_per_block > 256 || nnz_per_block > kMaxTile) {
        set_error("row block limits: rows_per_block <= 256, nnz_per_block <= %d", kMaxTile);
        return MTL4G_ERR_INVALID;
    }
    std::vector<int32_t> blk;
    blk.reserve(A->n / 64 + 2);
    const std::vector<int32_t>& rp = A->h_rowptr;
    // two non-zeros of headroom for the stream kernel's aligned pair loads
    const int cap = nnz_per_block >= 64 ? nnz_per_block - 2 : nnz_per_block;
    int32_t r = 0;
    blk.push_back(0);
    while (r < A->n) {
        int32_t r0 = r;
        int64_t nz = 0;
        while (r < A->n && r - r0 < rows_per_block && nz + (rp[r + 1] - rp[r]) <= cap) {
            nz += rp[r + 1] - rp[r];
            r++;
        }
        if (r == r0) r++;   // a row longer than the tile gets its own block
        blk.push_back(r);
    }
    A->nb = (int)blk.size() - 1;
    A->rows_per_block = rows_per_block;
    A->nnz_per_block = nnz_per_block;
    std::vector<RowBlock> desc(std::max(A->nb, 1));
    for (int b = 0; b < A->nb; b++)
        desc[b] = RowBlock{blk[b], blk[b + 1], rp[blk[b]], rp[blk[b + 1]]};
    M4_TRY(A->desc.alloc(desc.size()));
    M4_HIP(hipMemcpy(A->desc.p, desc.data(), desc.size() * sizeof(RowBlock), hipMemcpyHostToDevice));
    return MTL4G_OK;
}

static uint16_t f32_to_bf16(float f) { return bf16_bits(f); }

static int tune_lowp(mtl4g_matrix_s* A, int fmt, hipStream_t s);

// ---- sliced ELL copy of the matrix
// Entries of the sliced layout: slices of 64 rows, each padded to its
// longest row.
static int64_t sliced_entries(const mtl4g_matrix_s* A) {
    int64_t tot = 0;
    for (int32_t s0 = 0; s0 < A->n; s0 += kWave) {
        int32_t w = 0;
        for (int32_t r = s0; r < std::min(A->n, s0 + kWave); r++) w = std::max(w, A->h_rowptr[r + 1] - A->h_rowptr[r]);
        tot += (int64_t)w * kWave;
    }
    return tot;
}

// out[position of CSR entry j in the sliced layout] = src[j], pad elsewhere
template <typename T, typename S, typename F>
static void sliced_fill(const mtl4g_matrix_s* A, const S* src, T pad, F conv, std::vector<T>& out) {
    out.assign((size_t)A->h_sl_off.back() + kSellPad, pad);   // the kernel may read the padding
    for (int32_t s = 0; s < A->nsl; s++)
        for (int l = 0; l < kWave; l++) {
            const int32_t r = s * kWave + l;
            if (r >= A->n) break;
            for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++)
                out[(size_t)(A->h_sl_off[s] + (int64_t)(j - A->h_rowptr[r]) * kWave + l)] = conv(src[j]);
        }
}

template <typename T>
static int upload(mtl4g::DevBuf<T>& d, const std::vector<T>& h) {
    M4_TRY(d.alloc(h.size()));
    if (!h.empty()) M4_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return MTL4G_OK;
}

static void free_sliced(mtl4g_matrix_s* A) {
    A->h_sl_off.clear();
    A->sl_off.free();
    A->sl_col.free();
    A->sl_val.free();
    A->sl_val32.free();
    A->sl_valbf.free();
    A->nsl = 0;
}

static int sliced_lowp_values(mtl4g_matrix_s* A, int fmt) {
    if (A->h_sl_off.empty()) return MTL4G_OK;
    if (fmt == MTL4G_F32 && !A->sl_val32.p) {
        std::vector<float> v;
        sliced_fill(A, A->h_val.data(), 0.0f, [](double d) { return (float)d; }, v);
        M4_TRY(upload(A->sl_val32, v));
    }
    if (fmt == MTL4G_BF16 && !A->sl_valbf.p) {
        std::vector<uint16_t> v;
        sliced_fill(A, A->h_val.data(), (uint16_t)0, [](double d) { return f32_to_bf16((float)d); }, v);
        M4_TRY(upload(A->sl_valbf, v));
    }
    return MTL4G_OK;
}

// Build the sliced copy (pattern and FP64 values; low-precision values for
// the formats already in use).  Refused when padding exceeds half the nnz.
static int build_sliced(mtl4g_matrix_s* A) {
    if (!A->h_sl_off.empty()) return MTL4G_OK;
    const int64_t ent = sliced_entries(A);
    if (ent > A->nnz + A->nnz / 2 + kWave) {
        set_error("sliced ELL: %lld padded entries for %lld non-zeros", (long long)ent, (long long)A->nnz);
        return MTL4G_ERR_UNSUPPORTED;
    }
    A->nsl = (A->n + kWave - 1) / kWave;
    A->h_sl_off.assign((size_t)A->nsl + 1, 0);
    for (int32_t s = 0; s < A->nsl; s++) {
        int32_t w = 0;
        for (int32_t r = s * kWave; r < std::min(A->n, (s + 1) * kWave); r++)
            w = std::max(w, A->h_rowptr[r + 1] - A->h_rowptr[r]);
        A->h_sl_off[s + 1] = A->h_sl_off[s] + (int64_t)w * kWave;
    }
    int st;
    std::vector<int32_t> c;
    sliced_fill(A, A->h_col.data(), (int32_t)-1, [](int32_t j) { return j; }, c);
    std::vector<double> v;
    sliced_fill(A, A->h_val.data(), 0.0, [](double d) { return d; }, v);
    if ((st = upload(A->sl_off, A->h_sl_off)) || (st = upload(A->sl_col, c)) || (st = upload(A->sl_val, v)) ||
        (A->val32.p && (st = sliced_lowp_values(A, MTL4G_F32))) ||
        (A->valbf.p && (st = sliced_lowp_values(A, MTL4G_BF16)))) {
        free_sliced(A);
        return st;
    }
    return MTL4G_OK;
}

// ---- column panels
static void free_panels(mtl4g_matrix_s* A) {
    A->npan = 0;
    A->pan_w = 0;
    A->h_pan_src.clear();
    A->h_pan_nb.clear();
    A->pan_rowptr.free();
    A->pan_col.free();
    A->pan_val.free();
    A->pan_val32.free();
    A->pan_valbf.free();
    A->pan_desc.free();
}

// the panel copy of the values: FP64 always, FP32 / BF16 for the formats in use
static int panel_values(mtl4g_matrix_s* A) {
    const size_t ne = A->h_pan_src.size();
    std::vector<double> v(ne + 2, 0.0);   // padding for the aligned pair loads
    for (size_t i = 0; i < ne; i++) v[i] = A->h_val[A->h_pan_src[i]];
    M4_TRY(upload(A->pan_val, v));
    A->pan_val32.free();
    A->pan_valbf.free();
    if (A->val32.p) {
        std::vector<float> f(v.size());
        for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
        M4_TRY(upload(A->pan_val32, f));
    }
    if (A->valbf.p) {
        std::vector<uint16_t> h(v.size());
        for (size_t i = 0; i < v.size(); i++) h[i] = f32_to_bf16((float)v[i]);
        M4_TRY(upload(A->pan_valbf, h));
    }
    return MTL4G_OK;
}

static int panel_lowp_values(mtl4g_matrix_s* A, int fmt) {
    if (!A->npan) return MTL4G_OK;
    const size_t ne = A->h_pan_src.size();
    if (fmt == MTL4G_F32 && !A->pan_val32.p) {
        std::vector<float> f(ne + 2, 0.0f);
        for (size_t i = 0; i < ne; i++) f[i] = (float)A->h_val[A->h_pan_src[i]];
        M4_TRY(upload(A->pan_val32, f));
    }
    if (fmt == MTL4G_BF16 && !A->pan_valbf.p) {
        std::vector<uint16_t> h(ne + 2, 0);
        for (size_t i = 0; i < ne; i++) h[i] = f32_to_bf16((float)A->h_val[A->h_pan_src[i]]);
        M4_TRY(upload(A->pan_valbf, h));
    }
    return MTL4G_OK;
}

// Split the columns into panels of w columns: per panel the CSR of its
// entries (absolute indices into the panel-major arrays) and its row blocks
// of at most 256 rows and `tile` - 2 non-zeros.  Every row's entries must
// be in ascending column order, so that summing panel after panel is
// summing in column order.
static int build_panels(mtl4g_matrix_s* A, int32_t w) {
    if (w <= 0) w = 262144;
    if (A->npan && A->pan_w == w) return MTL4G_OK;
    free_panels(A);
    const int32_t n = A->n;
    const std::vector<int32_t>& rp = A->h_rowptr;
    const std::vector<int32_t>& ci = A->h_col;
    for (int32_t r = 0; r < n; r++)
        for (int32_t j = rp[r] + 1; j < rp[r + 1]; j++)
            if (ci[j] <= ci[j - 1]) {
                set_error("column panels: row %d's columns are not ascending", r);
                return MTL4G_ERR_UNSUPPORTED;
            }
    const int npan = (int)std::max<int64_t>(1, ((int64_t)A->m + w - 1) / w);
    // entries per (panel, row), then absolute row pointers panel after panel
    std::vector<int32_t> prp((size_t)npan * (n + 1), 0);
    for (int32_t r = 0; r < n; r++)
        for (int32_t j = rp[r]; j < rp[r + 1]; j++) prp[(size_t)(ci[j] / w) * (n + 1) + r + 1]++;
    int64_t off = 0;
    for (int p = 0; p < npan; p++) {
        int32_t* q = prp.data() + (size_t)p * (n + 1);
        q[0] = (int32_t)off;
        for (int32_t r = 0; r < n; r++) q[r + 1] += q[r];
        off = q[n];
    }
    std::vector<int32_t> cur(prp);
    std::vector<int32_t> pc((size_t)A->nnz + 2, 0);   // padding for the aligned pair loads
    A->h_pan_src.assign((size_t)A->nnz, 0);
    for (int32_t r = 0; r < n; r++)
        for (int32_t j = rp[r]; j < rp[r + 1]; j++) {
            const int p = ci[j] / w;
            const int32_t at = cur[(size_t)p * (n + 1) + r]++;
            pc[at] = ci[j];
            A->h_pan_src[at] = j;
        }
    int32_t maxblk = 0;
    for (int p = 0; p < npan; p++) {
        const int32_t* q = prp.data() + (size_t)p * (n + 1);
        for (int32_t r = 0; r < n; r++) maxblk = std::max(maxblk, q[r + 1] - q[r]);
    }
    A->pan_tile = 2048;
    std::vector<RowBlock> desc;
    A->h_pan_nb.assign(npan, 0);
    for (int p = 0; p < npan; p++) {
        const int32_t* q = prp.data() + (size_t)p * (n + 1);
        const int cap = A->pan_tile - 2;   // headroom for the aligned pair loads
        int32_t r = 0;
        while (r < n) {
            const int32_t r0 = r;
            int64_t nz = 0;
            while (r < n && r - r0 < 256 && nz + (q[r + 1] - q[r]) <= cap) {
                nz += q[r + 1] - q[r];
                r++;
            }
            if (r == r0) r++;
            desc.push_back(RowBlock{r0, r, q[r0], q[r]});
            A->h_pan_nb[p]++;
        }
    }
    int st;
    if ((st = upload(A->pan_rowptr, prp)) || (st = upload(A->pan_col, pc)) || (st = upload(A->pan_desc, desc)) ||
        (st = panel_values(A))) {
        free_panels(A);
        return st;
    }
    A->npan = npan;
    A->pan_w = w;
    return MTL4G_OK;
}

// ---- diagonal copy of the matrix
// Distinct diagonals (column - row) in ascending order, or empty when there
// are more than kDiaMax of them.
static std::vector<int32_t> dia_offsets(const mtl4g_matrix_s* A) {
    std::vector<int32_t> off;
    for (int32_t r = 0; r < A->n; r++)
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++) {
            const int32_t o = A->h_col[j] - r;
            auto it = std::lower_bound(off.begin(), off.end(), o);
            if (it != off.end() && *it == o) continue;
            if ((int)off.size() == kDiaMax) return {};
            off.insert(it, o);
        }
    return off;
}

// Slice-major diagonal layout: the entry of row r on diagonal d at
// ((r / 64) * ndiag + d) * 64 + r % 64 (a wavefront's rows read one
// contiguous block per slice); absent entries and the rows past n of the
// last slice hold the reserved pattern.
template <typename T, typename F>
static void dia_fill(const mtl4g_matrix_s* A, T absent, F conv, std::vector<T>& out) {
    const int nd = (int)A->h_dia_off.size();
    const size_t nsl = ((size_t)A->n + kWave - 1) / kWave;
    out.assign(nsl * nd * kWave, absent);
    for (int32_t r = 0; r < A->n; r++) {
        int d = 0;
        const size_t base = (size_t)(r / kWave) * nd * kWave + r % kWave;
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++) {
            while (A->h_dia_off[d] != A->h_col[j] - r) d++;   // columns ascend, so do the diagonals
            out[base + (size_t)d * kWave] = conv(A->h_val[j]);
        }
    }
}

static void free_dia(mtl4g_matrix_s* A) {
    A->h_dia_off.clear();
    A->dia_off.free();
    A->dia_val.free();
    A->dia_val32.free();
    A->dia_valbf.free();
    A->ndiag = 0;
}

template <typename T>
static bool bits_equal(T a, T b) {
    return std::memcmp(&a, &b, sizeof(T)) == 0;
}

static double absent64() {
    double d;
    std::memcpy(&d, &kDiaAbsent64, 8);
    return d;
}
static float absent32() {
    float f;
    std::memcpy(&f, &kDiaAbsent32, 4);
    return f;
}

static int dia_values(mtl4g_matrix_s* A, int fmt) {
    if (A->h_dia_off.empty()) return MTL4G_OK;
    if (fmt == MTL4G_F64 && !A->dia_val.p) {
        std::vector<double> v;
        dia_fill(A, absent64(), [](double d) { return d; }, v);
        M4_TRY(upload(A->dia_val, v));
    }
    if (fmt == MTL4G_F32 && !A->dia_val32.p) {
        std::vector<float> v;
        dia_fill(A, absent32(), [](double d) { return (float)d; }, v);
        M4_TRY(upload(A->dia_val32, v));
    }
    if (fmt == MTL4G_BF16 && !A->dia_valbf.p) {
        std::vector<uint16_t> v;
        dia_fill(A, kDiaAbsent16, [](double d) { return f32_to_bf16((float)d); }, v);
        M4_TRY(upload(A->dia_valbf, v));
    }
    return MTL4G_OK;
}

// Entries per slot of the diagonal layout: refused above kDiaMax diagonals,
// below 4/5 fill, or when a value carries the reserved absent-entry bits.
static int build_dia(mtl4g_matrix_s* A) {
    if (!A->h_dia_off.empty()) return MTL4G_OK;
    std::vector<int32_t> off = dia_offsets(A);
    if (A->n == 0 || off.empty() || (double)off.size() * A->n * 0.8 > (double)A->nnz) {
        set_error("diagonal layout: %zu diagonals for %lld non-zeros in %d rows (at most %d, filled 4/5)",
                  off.size(), (long long)A->nnz, A->n, kDiaMax);
        return MTL4G_ERR_UNSUPPORTED;
    }
    for (double v : A->h_val)
        if (bits_equal(v, absent64())) {
            set_error("diagonal layout: a value has the reserved absent-entry bits");
            return MTL4G_ERR_UNSUPPORTED;
        }
    A->h_dia_off = off;
    A->ndiag = (int)off.size();
    int st;
    if ((st = upload(A->dia_off, off)) || (st = dia_values(A, MTL4G_F64)) ||
        (A->val32.p && (st = dia_values(A, MTL4G_F32))) || (A->valbf.p && (st = dia_values(A, MTL4G_BF16)))) {
        free_dia(A);
        return st;
    }
    return MTL4G_OK;
}

// ---- symmetric diagonal copy (MTL4G_SPMV_DIA_SYM)
// Diagonals with offset >= 0 of a square matrix whose every entry (r, c)
// has a mirror (c, r) with the same bits; slice-major as dia_fill, rows'
// entries with column >= row only.
template <typename T, typename F>
static void dsym_fill(const mtl4g_matrix_s* A, const std::vector<int32_t>& up, T absent, F conv,
                      std::vector<T>& out) {
    const int nd = (int)up.size();
    const size_t nsl = ((size_t)A->n + kWave - 1) / kWave;
    out.assign(nsl * nd * kWave, absent);
    for (int32_t r = 0; r < A->n; r++) {
        int d = 0;
        const size_t base = (size_t)(r / kWave) * nd * kWave + r % kWave;
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++) {
            const int32_t o = A->h_col[j] - r;
            if (o < 0) continue;
            while (up[d] != o) d++;
            out[base + (size_t)d * kWave] = conv(A->h_val[j]);
        }
    }
}

static std::vector<int32_t> dsym_stored(const mtl4g_matrix_s* A) {
    std::vector<int32_t> up;
    for (int32_t o : A->h_dsym_off)
        if (o >= 0) up.push_back(o);
    return up;
}

static void free_dsym(mtl4g_matrix_s* A) {
    A->h_dsym_off.clear();
    A->dsym_off.free();
    A->dsym_map.free();
    A->dsym_val.free();
    A->dsym_val32.free();
    A->dsym_valbf.free();
    A->ndsym = A->nsym = 0;
}

static int dsym_values(mtl4g_matrix_s* A, int fmt) {
    if (A->h_dsym_off.empty()) return MTL4G_OK;
    const std::vector<int32_t> up = dsym_stored(A);
    if (fmt == MTL4G_F64 && !A->dsym_val.p) {
        std::vector<double> v;
        dsym_fill(A, up, absent64(), [](double d) { return d; }, v);
        M4_TRY(upload(A->dsym_val, v));
    }
    if (fmt == MTL4G_F32 && !A->dsym_val32.p) {
        std::vector<float> v;
        dsym_fill(A, up, absent32(), [](double d) { return (float)d; }, v);
        M4_TRY(upload(A->dsym_val32, v));
    }
    if (fmt == MTL4G_BF16 && !A->dsym_valbf.p) {
        std::vector<uint16_t> v;
        dsym_fill(A, up, kDiaAbsent16, [](double d) { return f32_to_bf16((float)d); }, v);
        M4_TRY(upload(A->dsym_valbf, v));
    }
    return MTL4G_OK;
}

// The diagonal layout's conditions, plus a square matrix equal to its
// transpose bit for bit (every entry has its mirror with the same value).
static int build_dsym(mtl4g_matrix_s* A) {
    if (!A->h_dsym_off.empty()) return MTL4G_OK;
    std::vector<int32_t> off = dia_offsets(A);
    if (A->n == 0 || A->n != A->m || off.empty() || (double)off.size() * A->n * 0.8 > (double)A->nnz) {
        set_error("symmetric diagonal layout: not a square matrix on at most %d well-filled diagonals", kDiaMax);
        return MTL4G_ERR_UNSUPPORTED;
    }
    for (int32_t r = 0; r < A->n; r++)
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++) {
            const int32_t c = A->h_col[j];
            const double v = A->h_val[j];
            if (bits_equal(v, absent64())) {
                set_error("symmetric diagonal layout: a value has the reserved absent-entry bits");
                return MTL4G_ERR_UNSUPPORTED;
            }
            const auto b = A->h_col.begin() + A->h_rowptr[c], e = A->h_col.begin() + A->h_rowptr[c + 1];
            const auto it = std::lower_bound(b, e, r);
            if (it == e || *it != r || !bits_equal(A->h_val[it - A->h_col.begin()], v)) {
                set_error("symmetric diagonal layout: the matrix is not equal to its transpose");
                return MTL4G_ERR_UNSUPPORTED;
            }
        }
    A->h_dsym_off = off;
    const std::vector<int32_t> up = dsym_stored(A);
    std::vector<int32_t> map(off.size());
    for (size_t d = 0; d < off.size(); d++)
        map[d] = (int32_t)(std::lower_bound(up.begin(), up.end(), std::abs(off[d])) - up.begin());
    A->ndsym = (int)off.size();
    A->nsym = (int)up.size();
    int st;
    if ((st = upload(A->dsym_off, off)) || (st = upload(A->dsym_map, map)) || (st = dsym_values(A, MTL4G_F64)) ||
        (A->val32.p && (st = dsym_values(A, MTL4G_F32))) || (A->valbf.p && (st = dsym_values(A, MTL4G_BF16)))) {
        free_dsym(A);
        return st;
    }
    return MTL4G_OK;
}

int ensure_lowp_values(mtl4g_matrix_s* A, int fmt, hipStream_t s) {
    if (fmt == MTL4G_F32 || fmt == MTL4G_BF16) {
        M4_TRY(sliced_lowp_values(A, fmt));
        M4_TRY(panel_lowp_values(A, fmt));
        M4_TRY(dia_values(A, fmt));
        M4_TRY(dsym_values(A, fmt));
    }
    if (fmt == MTL4G_F32 && !A->val32.p) {
        std::vector<float> v(A->h_val.begin(), A->h_val.end());
        M4_TRY(A->val32.alloc(v.size() + 2));   // padding for aligned pair loads
        M4_HIP(hipMemcpy(A->val32.p, v.data(), v.size() * 4, hipMemcpyHostToDevice));
    }
    if (fmt == MTL4G_BF16 && !A->valbf.p) {
        std::vector<uint16_t> v(A->h_val.size());
        for (size_t i = 0; i < v.size(); i++) v[i] = f32_to_bf16((float)A->h_val[i]);
        M4_TRY(A->valbf.alloc(v.size() + 2));
        M4_HIP(hipMemcpy(A->valbf.p, v.data(), v.size() * 2, hipMemcpyHostToDevice));
    }
    if ((fmt == MTL4G_F32 || fmt == MTL4G_BF16) && !A->lowp_tuned[fmt - 1]) M4_TRY(tune_lowp(A, fmt, s));
    return MTL4G_OK;
}

// ------------------------------------------------------------------ kernels
template <typename AT, int MODE>
struct EpiAssign {
    static constexpr int NP = 0;
    struct Pre {
        AT y;
    };
    AT* y;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{MODE == MTL4G_ASSIGN ? AT(0) : y[row]}; }
    __device__ void operator()(int row, AT s, const Pre& pre, AT*) const {
        if (MODE == MTL4G_ASSIGN) y[row] = s;
        else if (MODE == MTL4G_PLUS) y[row] = pre.y + s;
        else y[row] = pre.y - s;
    }
    __device__ void finish(const AT*) const {}
};

// (lazy(q)= A*p) || (lazy(alpha)= lazy_dot(p,q)), result to *out
struct EpiSpmvDot {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_FUSED_DOT; }
    struct Pre {
        double p;
    };
    const double* p;
    double* q;
    double* out;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ Pre preload(int row) const { return Pre{p[row]}; }
    __device__ void operator()(int row, double s, const Pre& pre, double* part) const {
        q[row] = s;
        part[0] += pre.p * s;
    }
    __device__ void finish(const double* tot) const { *out = tot[0]; }
};

int launch_spmv(const mtl4g_matrix_s* A, const double* x, double* y, int assign, hipStream_t s) {
    switch (assign) {
    case MTL4G_ASSIGN: return launch_csr<double>(A, A->val.p, x, EpiAssign<double, 0>{y}, (double*)nullptr, nullptr, s);
    case MTL4G_PLUS: return launch_csr<double>(A, A->val.p, x, EpiAssign<double, 1>{y}, (double*)nullptr, nullptr, s);
    case MTL4G_MINUS: return launch_csr<double>(A, A->val.p, x, EpiAssign<double, 2>{y}, (double*)nullptr, nullptr, s);
    }
    set_error("assign mode %d", assign);
    return MTL4G_ERR_INVALID;
}

template <int MODE>
static int spmv32_mode(const mtl4g_matrix_s* A, int fmt, const float* x, float* y, hipStream_t s) {
    if (fmt == MTL4G_F32)
        return launch_csr<float>(A, A->val32.p, x, EpiAssign<float, MODE>{y}, (float*)nullptr, nullptr, s);
    return launch_csr<float>(A, A->valbf.p, x, EpiAssign<float, MODE>{y}, (float*)nullptr, nullptr, s);
}

// Workgroups per CU of the stream kernel on the FP32 / BF16 values: fewer
// bytes per non-zero and fewer registers than FP64, so the best persistent
// grid can differ.  Timed on the first low-precision use of the matrix.
static int tune_lowp(mtl4g_matrix_s* A, int fmt, hipStream_t s) {
    A->lowp_tuned[fmt - 1] = true;
    A->per_cu_lowp[fmt - 1] = 0;
    A->lowp_wide[fmt - 1] = false;
    if ((A->kernel != MTL4G_SPMV_STREAM && A->kernel != MTL4G_SPMV_STREAM2 && A->kernel != MTL4G_SPMV_SLICED &&
         A->kernel != MTL4G_SPMV_DIA && A->kernel != MTL4G_SPMV_DIA_SYM) ||
        A->n == 0 || A->m == 0)
        return MTL4G_OK;
    if (A->kernel == MTL4G_SPMV_STREAM && !A->desc_wide.p) {   // wide partition for the candidates below
        std::vector<RowBlock> w = partition_rows(A, 512, 4096);
        M4_TRY(A->desc_wide.alloc(std::max<size_t>(w.size(), 1)));
        if (!w.empty())
            M4_HIP(hipMemcpy(A->desc_wide.p, w.data(), w.size() * sizeof(RowBlock), hipMemcpyHostToDevice));
        A->nb_wide = (int)w.size();
    }
    float *x = nullptr, *y = nullptr;
    M4_HIP(hipMallocAsync((void**)&x, (size_t)A->m * 4, s));
    M4_HIP(hipMallocAsync((void**)&y, (size_t)A->n * 4, s));
    M4_HIP(hipMemsetAsync(x, 0, (size_t)A->m * 4, s));
    hipEvent_t e0, e1;
    M4_HIP(hipEventCreate(&e0));
    M4_HIP(hipEventCreate(&e1));
    double best = 1e30;
    int win = 0;
    bool win_wide = false;
    struct Cand {
        int f;
        bool wide;
        int unr;
    };
    std::vector<Cand> cands;
    int win_unr = A->sl_unroll;
    if (A->kernel == MTL4G_SPMV_SLICED || A->kernel == MTL4G_SPMV_DIA || A->kernel == MTL4G_SPMV_DIA_SYM) {
        for (int u : {4, 8})
            for (int f : {2, 4, 6, 8}) cands.push_back({f, false, u});
    } else {
        for (int f = 2; f <= 6; f++) cands.push_back({f, false, 8});
    }
    if (A->kernel == MTL4G_SPMV_STREAM && A->nb_wide > 0)
        for (int f = 2; f <= 4; f++) cands.push_back({f, true, 8});
    for (const Cand& cd : cands) {
        const int f = cd.f;
        A->per_cu_lowp[fmt - 1] = f;
        A->lowp_wide[fmt - 1] = cd.wide;
        A->sl_unroll_lowp[fmt - 1] = cd.unr;
        int st = launch_spmv_f32(A, fmt, x, y, MTL4G_ASSIGN, s);
        if (!st) st = launch_spmv_f32(A, fmt, x, y, MTL4G_ASSIGN, s);
        if (st) continue;
        const int reps = 10;
        M4_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) M4_TRY(launch_spmv_f32(A, fmt, x, y, MTL4G_ASSIGN, s));
        M4_HIP(hipEventRecord(e1, s));
        M4_HIP(hipEventSynchronize(e1));
        float ms = 0;
        M4_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            win = f;
            win_wide = cd.wide;
            win_unr = cd.unr;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFreeAsync(x, s);
    (void)hipFreeAsync(y, s);
    A->per_cu_lowp[fmt - 1] = win;
    A->lowp_wide[fmt - 1] = win_wide;
    A->sl_unroll_lowp[fmt - 1] = win_unr;
    return MTL4G_OK;
}

// Workgroups of the persistent grid y = A*x launches with on these values
// (the persistent kernels clamp the requested per-CU count to what one CU
// holds of that instantiation); 0 for the one-block-per-tile kernels.
template <typename AT, typename MT>
static int assign_grid(const mtl4g_matrix_s* A) {
    using Epi = EpiAssign<AT, 0>;
    int pcu = A->per_cu, unr = A->sl_unroll;
    const int li = sizeof(MT) == 4 ? 0 : 1;
    if (sizeof(MT) < 8 && A->per_cu_lowp[li] > 0) {
        pcu = A->per_cu_lowp[li];
        unr = A->sl_unroll_lowp[li];
    }
    const int units = ((A->n + kWave - 1) / kWave + 3) / 4;
    switch (A->kernel) {
    case MTL4G_SPMV_SLICED:
        return unr == 4 ? persistent_grid<k_sell<256, 4, AT, MT, AT, Epi>>(A->cus, pcu, units)
                        : persistent_grid<k_sell<256, 8, AT, MT, AT, Epi>>(A->cus, pcu, units);
    case MTL4G_SPMV_DIA:
        return unr == 4 ? persistent_grid<k_dia<256, 4, AT, MT, AT, Epi>>(A->cus, pcu, units)
                        : persistent_grid<k_dia<256, 8, AT, MT, AT, Epi>>(A->cus, pcu, units);
    case MTL4G_SPMV_DIA_SYM:
        return unr == 4 ? persistent_grid<k_dia_sym<256, 4, AT, MT, AT, Epi>>(A->cus, pcu, units)
                        : persistent_grid<k_dia_sym<256, 8, AT, MT, AT, Epi>>(A->cus, pcu, units);
    case MTL4G_SPMV_STREAM:
        if (sizeof(MT) < 8 && A->lowp_wide[li] && A->nb_wide > 0)
            return persistent_grid<k_csrp<256, 4096, AT, MT, AT, Epi, 2>>(A->cus, pcu, A->nb_wide);
        return A->nnz_per_block <= 1024 ? persistent_grid<k_csrp<256, 1024, AT, MT, AT, Epi>>(A->cus, pcu, A->nb)
                                        : persistent_grid<k_csrp<256, 2048, AT, MT, AT, Epi>>(A->cus, pcu, A->nb);
    case MTL4G_SPMV_STREAM2:
        return A->nnz_per_block <= 1024
                   ? persistent_grid<k_csrp2<256, 1024, 2, AT, MT, AT, Epi>>(A->cus, pcu, A->nb)
                   : persistent_grid<k_csrp2<256, 2048, 2, AT, MT, AT, Epi>>(A->cus, pcu, A->nb);
    default:
        return 0;
    }
}

int launch_spmv_f32(const mtl4g_matrix_s* A, int fmt, const float* x, float* y, int assign,
                    hipStream_t s) {
    switch (assign) {
    case MTL4G_ASSIGN: return spmv32_mode<0>(A, fmt, x, y, s);
    case MTL4G_PLUS: return spmv32_mode<1>(A, fmt, x, y, s);
    case MTL4G_MINUS: return spmv32_mode<2>(A, fmt, x, y, s);
    }
    set_error("assign mode %d", assign);
    return MTL4G_ERR_INVALID;
}

// ---- vector kernels: block b owns a contiguous chunk; fixed partial layout
struct EpiDot {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_DOT4; }
    const double *a, *b;
    double* out;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const { part[0] += a[i] * b[i]; }
    __device__ void finish(const double* tot) const { *out = tot[0]; }
};

// unary_dot(v) (reduction<8>) or, with SQRT, two_norm(v) (reduction<4>)
template <int SQRT>
struct EpiUnaryDot {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return SQRT ? RO_NORM4 : RO_NORM8; }
    const double* a;
    double* out;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const { part[0] += a[i] * a[i]; }
    __device__ void finish(const double* tot) const { *out = SQRT ? sqrt(tot[0]) : tot[0]; }
};

struct EpiAxpy {
    static constexpr int NP = 0;
    double alpha;
    const double* v;
    double* w;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double*) const { w[i] = w[i] + alpha * v[i]; }
    __device__ void finish(const double*) const {}
};

struct EpiAxpyUnary {
    static constexpr int NP = 1;
    __host__ __device__ static constexpr int order(int) { return RO_FUSED_NORM; }
    double alpha;
    const double* q;
    double* r;
    double* out;
    __device__ bool skip() const { return false; }
    __device__ void prepare() {}
    __device__ void operator()(int64_t i, double* part) const {
        double v = r[i] - alpha * q[i];
        r[i] = v;
        part[0] += v * v;
    }
    __device__ void finish(const double* tot) const { *out = tot[0]; }
};

// Stream-ordered scratch for one reduction: tickets, result, partials and,
// when reductions run in MTL4's order, n terms.
struct Scratch {
    static constexpr size_t kTicketBytes = kTicketWords * sizeof(unsigned);
    char* base = nullptr;
    hipStream_t s = nullptr;
    bool ordered = false;
    unsigned* ticket() { return reinterpret_cast<unsigned*>(base); }
    double* result() { return reinterpret_cast<double*>(base + kTicketBytes); }
    double* partials() { return reinterpret_cast<double*>(base + kTicketBytes + 256); }
    double* terms() { return ordered ? partials() + kPartialWords : nullptr; }
    int get(hipStream_t st, int64_t n) {
        s = st;
        ordered = reduction_order() == MTL4G_ORDER_MTL4;
        const size_t words = kPartialWords + (ordered ? (size_t)n : 0);
        M4_HIP(hipMallocAsync((void**)&base, kTicketBytes + 256 + sizeof(double) * words, st));
        M4_HIP(hipMemsetAsync(base, 0, kTicketBytes + 256, st));
        return MTL4G_OK;
    }
    ~Scratch() {
        if (base) (void)hipFreeAsync(base, s);
    }
};

template <typename Epi>
static int run_vec(int64_t n, const Epi& epi, Scratch* sc, hipStream_t s) {
    return launch_vec<double>(n, epi, sc ? sc->partials() : nullptr, sc ? sc->ticket() : nullptr, s,
                              sc ? sc->terms() : nullptr);
}

static int fetch(Scratch& sc, double* h, hipStream_t s) {
    M4_HIP(hipMemcpyAsync(h, sc.result(), sizeof(double), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    return MTL4G_OK;
}

// ---- fixed scratch for the entry point that allocates nothing
// Layout as Scratch: tickets, a 256-byte result slot, partials, then room for
// the terms of an MTL4-order reduction of up to kNoAllocTermRows rows.  Each
// device holds its own copy of the array from module load to process end.
constexpr size_t kNoAllocTermRows = (size_t)1 << 20;
constexpr size_t kNoAllocBytes =
    Scratch::kTicketBytes + 256 + sizeof(double) * (kPartialWords + kNoAllocTermRows);
static __device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices];   // device address of `scratch`, per device

// The device address of this device's `scratch`, looked up once per device
// with that device current.  Host code passes this pointer, never the
// symbol's own name, to kernels and copies.
static int noalloc_scratch(char** out) {
    int dev = 0;
    M4_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d devices the fixed scratch serves", dev, kMaxDevices);
        return MTL4G_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        M4_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the fixed scratch");
            return MTL4G_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return MTL4G_OK;
}

}  // namespace mtl4g

using namespace mtl4g;

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* mtl4g_status_string(int st) {
    switch (st) {
    case MTL4G_OK: return "ok";
    case MTL4G_MAX_ITER: return "too many iterations";
    case MTL4G_ERR_INVALID: return "invalid argument";
    case MTL4G_ERR_NOT_SQUARE: return "matrix not square";
    case MTL4G_ERR_SIZE: return "incompatible size";
    case MTL4G_ERR_MISSING_DIAG: return "missing diagonal";
    case MTL4G_ERR_BREAKDOWN: return "unexpected orthogonality";
    case MTL4G_ERR_NONFINITE: return "non-finite value";
    case MTL4G_ERR_OOM: return "out of device memory";
    case MTL4G_ERR_HIP: return "HIP runtime error";
    case MTL4G_ERR_UNSUPPORTED: return "unsupported";
    case MTL4G_ERR_TIMEOUT: return "device wait timed out";
    }
    return "unknown status";
}

const char* mtl4g_last_error(void) { return g_last_error.c_str(); }

int mtl4g_version(void) { return 100; }

int mtl4g_set_reduction_order(int order) {
    if (order != MTL4G_ORDER_TREE && order != MTL4G_ORDER_MTL4) {
        set_error("reduction order %d", order);
        return MTL4G_ERR_INVALID;
    }
    mtl4g::g_reduction_order.store(order, std::memory_order_relaxed);
    return MTL4G_OK;
}

int mtl4g_get_reduction_order(void) { return mtl4g::reduction_order(); }

int mtl4g_set_device(int device) {
    M4_HIP(hipSetDevice(device));
    return MTL4G_OK;
}

int mtl4g_malloc(void** p, size_t bytes) {
    if (!p) return MTL4G_ERR_INVALID;
    M4_HIP(hipMalloc(p, bytes ? bytes : 1));
    return MTL4G_OK;
}

int mtl4g_free(void* p) {
    if (p) M4_HIP(hipFree(p));
    return MTL4G_OK;
}

int mtl4g_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
    if (bytes == 0) return MTL4G_OK;
    if (!dst || !src || kind < 0 || kind > 2) return MTL4G_ERR_INVALID;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                            : kind == 1 ? hipMemcpyDeviceToHost
                                        : hipMemcpyDeviceToDevice;
    M4_HIP(hipMemcpyAsync(dst, src, bytes, k, S(stream)));
    if (kind != 2) M4_HIP(hipStreamSynchronize(S(stream)));
    return MTL4G_OK;
}

int mtl4g_memset(void* p, int value, size_t bytes, void* stream) {
    if (bytes == 0) return MTL4G_OK;
    if (!p) return MTL4G_ERR_INVALID;
    M4_HIP(hipMemsetAsync(p, value, bytes, S(stream)));
    return MTL4G_OK;
}

int mtl4g_stream_synchronize(void* stream) {
    M4_HIP(hipStreamSynchronize(S(stream)));
    return MTL4G_OK;
}

int mtl4g_csr_create(mtl4g_matrix* out, int32_t nrows, int32_t ncols, int64_t nnz,
                     const int32_t* rowptr, const int32_t* col, const double* val, int where) {
    if (!out || nrows < 0 || ncols < 0 || nnz < 0 || (nrows > 0 && !rowptr) ||
        (nnz > 0 && (!col || !val))) {
        set_error("mtl4g_csr_create: bad arguments");
        return MTL4G_ERR_INVALID;
    }
    if (nnz >= (int64_t)INT32_MAX) {
        set_error("mtl4g_csr_create: nnz %lld exceeds int32 indexing", (long long)nnz);
        return MTL4G_ERR_UNSUPPORTED;
    }
    *out = nullptr;
    mtl4g_matrix_s* A = new (std::nothrow) mtl4g_matrix_s;
    if (!A) return MTL4G_ERR_OOM;
    auto fail = [&](int st) { delete A; return st; };
    (void)hipGetDevice(&A->device);
    (void)hipDeviceGetAttribute(&A->cus, hipDeviceAttributeMultiprocessorCount, A->device);
    if (A->cus <= 0) A->cus = 256;
    A->n = nrows;
    A->m = ncols;
    A->nnz = nnz;
    A->h_rowptr.resize((size_t)nrows + 1);
    A->h_col.resize((size_t)nnz);
    A->h_val.resize((size_t)nnz);
    if (where == 1) {
        hipError_t e = hipMemcpy(A->h_rowptr.data(), rowptr, ((size_t)nrows + 1) * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && nnz)
            e = hipMemcpy(A->h_col.data(), col, (size_t)nnz * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess && nnz)
            e = hipMemcpy(A->h_val.data(), val, (size_t)nnz * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(hip_status(e, "hipMemcpy", __FILE__, __LINE__));
    } else {
        std::memcpy(A->h_rowptr.data(), rowptr, ((size_t)nrows + 1) * 4);
        if (nnz) {
            std::memcpy(A->h_col.data(), col, (size_t)nnz * 4);
            std::memcpy(A->h_val.data(), val, (size_t)nnz * 8);
        }
    }
    // validate: monotone row pointers, sorted in-range columns
    if (A->h_rowptr[0] != 0 || A->h_rowptr[nrows] != nnz) {
        set_error("mtl4g_csr_create: rowptr[0] must be 0 and rowptr[n] == nnz");
        return fail(MTL4G_ERR_INVALID);
    }
    std::vector<double> diag((size_t)nrows, 0.0);
    for (int32_t r = 0; r < nrows; r++) {
        int32_t a = A->h_rowptr[r], b = A->h_rowptr[r + 1];
        if (b < a) {
            set_error("mtl4g_csr_create: rowptr not monotone at row %d", r);
            return fail(MTL4G_ERR_INVALID);
        }
        A->max_row = std::max(A->max_row, b - a);
        for (int32_t j = a; j < b; j++) {
            int32_t c = A->h_col[j];
            if (c < 0 || c >= ncols || (j > a && c <= A->h_col[j - 1])) {
                set_error("mtl4g_csr_create: column %d of row %d out of range or unsorted", c, r);
                return fail(MTL4G_ERR_INVALID);
            }
            if (c == r) diag[r] = A->h_val[j];
        }
    }
    int st;
    // col and val carry two elements of padding for the aligned pair loads
    if ((st = A->rowptr.alloc((size_t)nrows + 1)) || (st = A->col.alloc((size_t)nnz + 2)) ||
        (st = A->val.alloc((size_t)nnz + 2)) || (st = A->diag.alloc((size_t)nrows)))
        return fail(st);
    hipError_t e = hipMemcpy(A->rowptr.p, A->h_rowptr.data(), ((size_t)nrows + 1) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(A->col.p, A->h_col.data(), (size_t)nnz * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(A->val.p, A->h_val.data(), (size_t)nnz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && nrows) e = hipMemcpy(A->diag.p, diag.data(), (size_t)nrows * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(hip_status(e, "hipMemcpy", __FILE__, __LINE__));
    if ((st = build_row_blocks(A, 0, 0))) return fail(st);
    // row-length binning: lanes per row from the average row length
    const double avg = nrows ? (double)nnz / nrows : 0.0;
    A->gs = avg <= 4 ? 4 : avg <= 8 ? 8 : avg <= 16 ? 16 : avg <= 32 ? 32 : 64;
    A->vrows = std::max(4, 2048 / A->gs);
    // default kernel, chosen from the structure alone (so it is the same on
    // every run; mtl4g_csr_autotune may refine it): the diagonal layout for
    // matrices on few well-filled diagonals, the sliced layout where slices
    // of 64 rows pad little, else the CSR stream kernel
    A->sl_unroll = A->sl_unroll_lowp[0] = A->sl_unroll_lowp[1] = 8;
    if (nnz > 0 && build_dia(A) == MTL4G_OK) {
        A->kernel = MTL4G_SPMV_DIA;
        A->per_cu = 4;
    } else if (nnz > 0 && sliced_entries(A) <= nnz + nnz / 8 + kWave && build_sliced(A) == MTL4G_OK) {
        A->kernel = MTL4G_SPMV_SLICED;
        A->per_cu = 2;
    } else {
        A->kernel = MTL4G_SPMV_STREAM;
        A->per_cu = 4;
    }
    *out = A;
    return MTL4G_OK;
}

int mtl4g_csr_destroy(mtl4g_matrix A) {
    delete A;
    return MTL4G_OK;
}

int mtl4g_csr_info(mtl4g_matrix A, int32_t* nrows, int32_t* ncols, int64_t* nnz) {
    if (!A) return MTL4G_ERR_INVALID;
    if (nrows) *nrows = A->n;
    if (ncols) *ncols = A->m;
    if (nnz) *nnz = A->nnz;
    return MTL4G_OK;
}

int mtl4g_csr_set_values(mtl4g_matrix A, const double* val, int where) {
    if (!A || (A->nnz && !val)) return MTL4G_ERR_INVALID;
    if (where == 1) {
        M4_HIP(hipMemcpy(A->h_val.data(), val, (size_t)A->nnz * 8, hipMemcpyDeviceToHost));
    } else if (A->nnz) {
        std::memcpy(A->h_val.data(), val, (size_t)A->nnz * 8);
    }
    std::vector<double> diag((size_t)A->n, 0.0);
    for (int32_t r = 0; r < A->n; r++)
        for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++)
            if (A->h_col[j] == r) diag[r] = A->h_val[j];
    if (A->nnz) M4_HIP(hipMemcpy(A->val.p, A->h_val.data(), (size_t)A->nnz * 8, hipMemcpyHostToDevice));
    if (A->n) M4_HIP(hipMemcpy(A->diag.p, diag.data(), (size_t)A->n * 8, hipMemcpyHostToDevice));
    A->val32.free();
    A->valbf.free();
    A->sl_val32.free();
    A->sl_valbf.free();
    if (A->npan) M4_TRY(panel_values(A));
    if (!A->h_sl_off.empty()) {
        std::vector<double> v;
        sliced_fill(A, A->h_val.data(), 0.0, [](double d) { return d; }, v);
        M4_TRY(upload(A->sl_val, v));
    }
    if (!A->h_dsym_off.empty()) {
        const bool sym = A->kernel == MTL4G_SPMV_DIA_SYM;
        free_dsym(A);
        if (sym && build_dsym(A)) {   // no longer symmetric: the plain diagonal layout, else the stream kernel
            A->kernel = MTL4G_SPMV_DIA;
            if (!A->h_dia_off.empty()) free_dia(A);
            if (build_dia(A)) {
                A->kernel = MTL4G_SPMV_STREAM;
                if (!A->desc.p) M4_TRY(build_row_blocks(A, 0, 0));
            }
        }
    }
    if (!A->h_dia_off.empty()) {
        const bool dia = A->kernel == MTL4G_SPMV_DIA;
        free_dia(A);
        if (dia) {
            const int st = build_dia(A);
            if (st) {   // a value now carries the reserved bits: back to the CSR stream kernel
                A->kernel = MTL4G_SPMV_STREAM;
                if (!A->desc.p) M4_TRY(build_row_blocks(A, 0, 0));
            }
        }
    }
    return MTL4G_OK;
}

int mtl4g_csr_tune(mtl4g_matrix A, int32_t rows_per_block, int32_t nnz_per_block) {
    if (!A) return MTL4G_ERR_INVALID;
    M4_TRY(build_row_blocks(A, rows_per_block, nnz_per_block));
    A->kernel = MTL4G_SPMV_TILE;
    A->lowp_tuned[0] = A->lowp_tuned[1] = false;
    A->per_cu_lowp[0] = A->per_cu_lowp[1] = 0;
    A->lowp_wide[0] = A->lowp_wide[1] = false;
    return MTL4G_OK;
}

int mtl4g_csr_set_kernel(mtl4g_matrix A, int kernel, int32_t group, int32_t rows_per_block) {
    if (!A) return MTL4G_ERR_INVALID;
    if (kernel == MTL4G_SPMV_TILE) return mtl4g_csr_tune(A, rows_per_block, 0);
    if (kernel == MTL4G_SPMV_STREAM || kernel == MTL4G_SPMV_STREAM2) {
        if (group == 0) group = A->per_cu;
        if (group < 1 || group > 8) {
            set_error("stream kernel: 1 to 8 workgroups per compute unit");
            return MTL4G_ERR_INVALID;
        }
        if (rows_per_block > 0 || !A->desc.p) M4_TRY(build_row_blocks(A, rows_per_block, 0));
        A->kernel = kernel;
        A->per_cu = group;
        A->lowp_tuned[0] = A->lowp_tuned[1] = false;
        A->per_cu_lowp[0] = A->per_cu_lowp[1] = 0;
        A->lowp_wide[0] = A->lowp_wide[1] = false;
        return MTL4G_OK;
    }
    if (kernel == MTL4G_SPMV_SLICED || kernel == MTL4G_SPMV_DIA || kernel == MTL4G_SPMV_DIA_SYM) {
        if (group == 0) group = A->per_cu;
        if (rows_per_block == 0) rows_per_block = 8;
        if (group < 1 || group > 8 || (rows_per_block != 4 && rows_per_block != 8)) {
            set_error("sliced / diagonal kernel: 1 to 8 workgroups per compute unit, 4 or 8 entries per lane and chunk");
            return MTL4G_ERR_INVALID;
        }
        M4_TRY(kernel == MTL4G_SPMV_SLICED ? build_sliced(A) : kernel == MTL4G_SPMV_DIA ? build_dia(A) : build_dsym(A));
        A->kernel = kernel;
        A->per_cu = group;
        A->sl_unroll = rows_per_block;
        A->sl_unroll_lowp[0] = A->sl_unroll_lowp[1] = rows_per_block;
        A->lowp_tuned[0] = A->lowp_tuned[1] = false;
        A->per_cu_lowp[0] = A->per_cu_lowp[1] = 0;
        A->lowp_wide[0] = A->lowp_wide[1] = false;
        return MTL4G_OK;
    }
    if (kernel == MTL4G_SPMV_PANEL) {
        if (group == 0) group = A->per_cu;
        if (group < 1 || group > 8 || rows_per_block < 0) {
            set_error("panel kernel: 1 to 8 workgroups per compute unit, a positive panel width");
            return MTL4G_ERR_INVALID;
        }
        M4_TRY(build_panels(A, rows_per_block));
        A->kernel = kernel;
        A->per_cu = group;
        A->lowp_tuned[0] = A->lowp_tuned[1] = false;
        A->per_cu_lowp[0] = A->per_cu_lowp[1] = 0;
        A->lowp_wide[0] = A->lowp_wide[1] = false;
        return MTL4G_OK;
    }
    if (kernel == MTL4G_SPMV_SCALAR) {
        if (group == 0) group = A->uv;
        if (group != 4 && group != 8) {
            set_error("scalar CSR: 4 or 8 values per step");
            return MTL4G_ERR_INVALID;
        }
        A->kernel = MTL4G_SPMV_SCALAR;
        A->uv = group;
        A->lowp_tuned[0] = A->lowp_tuned[1] = false;
        return MTL4G_OK;
    }
    if (kernel != MTL4G_SPMV_VECTOR) return MTL4G_ERR_INVALID;
    if (group == 0) group = A->gs;
    if (group != 4 && group != 8 && group != 16 && group != 32 && group != 64) {
        set_error("vector CSR group must be 4, 8, 16, 32 or 64 lanes");
        return MTL4G_ERR_INVALID;
    }
    if (rows_per_block <= 0) rows_per_block = std::max(4, 2048 / group);
    A->kernel = MTL4G_SPMV_VECTOR;
    A->gs = group;
    A->vrows = rows_per_block;
    return MTL4G_OK;
}

int mtl4g_csr_grid(mtl4g_matrix A, int value_format, int32_t* workgroups) {
    if (!A || !workgroups || value_format < MTL4G_F64 || value_format > MTL4G_BF16) return MTL4G_ERR_INVALID;
    *workgroups = value_format == MTL4G_F64   ? assign_grid<double, double>(A)
                  : value_format == MTL4G_F32 ? assign_grid<float, float>(A)
                                              : assign_grid<float, uint16_t>(A);
    return MTL4G_OK;
}

int mtl4g_csr_get_kernel(mtl4g_matrix A, int* kernel, int32_t* group, int32_t* rows_per_block) {
    if (!A) return MTL4G_ERR_INVALID;
    if (kernel) *kernel = A->kernel;
    if (group)
        *group = A->kernel == MTL4G_SPMV_VECTOR   ? A->gs
                 : A->kernel == MTL4G_SPMV_STREAM || A->kernel == MTL4G_SPMV_STREAM2 ||
                           A->kernel == MTL4G_SPMV_SLICED || A->kernel == MTL4G_SPMV_DIA ||
                           A->kernel == MTL4G_SPMV_DIA_SYM || A->kernel == MTL4G_SPMV_PANEL
                     ? A->per_cu
                 : A->kernel == MTL4G_SPMV_SCALAR ? A->uv
                                                  : 0;
    if (rows_per_block)
        *rows_per_block = A->kernel == MTL4G_SPMV_VECTOR   ? A->vrows
                          : A->kernel == MTL4G_SPMV_SLICED || A->kernel == MTL4G_SPMV_DIA ||
                                    A->kernel == MTL4G_SPMV_DIA_SYM
                              ? A->sl_unroll
                          : A->kernel == MTL4G_SPMV_PANEL ? A->pan_w
                                                           : A->rows_per_block;
    return MTL4G_OK;
}

int mtl4g_csr_autotune(mtl4g_matrix A, double* h_best_us, void* stream) {
    if (!A) return MTL4G_ERR_INVALID;
    hipStream_t s = S(stream);
    if (A->n == 0) return MTL4G_OK;
    struct Cand {
        int kernel, a, b, rows, nnz;   // stream kernel: a = workgroups per CU on (rows, nnz) blocks
    };
    std::vector<Cand> cands = {{MTL4G_SPMV_TILE, 256, 2048, 0, 0}, {MTL4G_SPMV_TILE, 128, 1024, 0, 0}};
    // persistent grids stay within what one CU holds at once (LDS: two
    // product buffers of TILE doubles per workgroup)
    for (int f : {2, 3, 4}) cands.push_back({MTL4G_SPMV_STREAM, f, 0, 256, 2048});
    for (int f : {4, 6, 8}) cands.push_back({MTL4G_SPMV_STREAM, f, 0, 128, 1024});
    for (int f : {2, 3, 4}) cands.push_back({MTL4G_SPMV_STREAM2, f, 0, 256, 2048});
    for (int f : {4, 6}) cands.push_back({MTL4G_SPMV_STREAM2, f, 0, 128, 1024});
    for (int u : {4, 8}) cands.push_back({MTL4G_SPMV_SCALAR, u, 0, 0, 0});
    // sliced ELL where slices of 64 rows pad little (a copy of the matrix)
    const bool had_sliced = !A->h_sl_off.empty();
    const bool try_sliced = had_sliced || sliced_entries(A) <= A->nnz + A->nnz / 8 + kWave;
    if (try_sliced)
        for (int u : {4, 8})
            for (int f : {2, 4, 6, 8}) cands.push_back({MTL4G_SPMV_SLICED, f, u, 0, 0});
    // the diagonal layout where the entries lie on few, well-filled diagonals
    const bool had_dia = !A->h_dia_off.empty();
    if (had_dia || build_dia(A) == MTL4G_OK)
        for (int u : {4, 8})
            for (int f : {2, 4, 6, 8}) cands.push_back({MTL4G_SPMV_DIA, f, u, 0, 0});
    // its symmetric half where the matrix equals its transpose
    const bool had_dsym = !A->h_dsym_off.empty();
    if (had_dsym || build_dsym(A) == MTL4G_OK)
        for (int u : {4, 8})
            for (int f : {2, 4, 6, 8}) cands.push_back({MTL4G_SPMV_DIA_SYM, f, u, 0, 0});
    // column panels where x is larger than two L2s and most gathers land far
    // from the row's own column (sampled every 61st row)
    const bool had_panels = A->npan > 0;
    if ((int64_t)A->m * 8 > (8 << 20) && A->n == A->m) {
        int64_t far = 0, tot = 0;
        for (int32_t r = 0; r < A->n; r += 61)
            for (int32_t j = A->h_rowptr[r]; j < A->h_rowptr[r + 1]; j++, tot++)
                far += std::abs((int64_t)A->h_col[j] - r) * 8 > (2 << 20);
        if (had_panels || (tot > 0 && far * 2 > tot))
            for (int w : {262144, 524288, 1048576})
                for (int f : {2, 3, 4, 6}) cands.push_back({MTL4G_SPMV_PANEL, f, w, 0, 0});
    }
    // wide vector groups (16+ lanes) only pay for rows of hundreds of non-zeros
    const double avg = A->n ? (double)A->nnz / A->n : 0.0;
    for (int g : {4, 8, 16, 32, 64}) {
        if (g >= 16 && avg < 4.0 * g) continue;
        for (int f : {1, 2, 4}) cands.push_back({MTL4G_SPMV_VECTOR, g, std::max(4, f * 1024 / g), 0, 0});
    }
    double *x = nullptr, *y = nullptr;
    M4_HIP(hipMallocAsync((void**)&x, (size_t)A->m * 8 + 8, s));
    M4_HIP(hipMallocAsync((void**)&y, (size_t)A->n * 8 + 8, s));
    M4_HIP(hipMemsetAsync(x, 0, (size_t)A->m * 8, s));
    hipEvent_t e0, e1;
    M4_HIP(hipEventCreate(&e0));
    M4_HIP(hipEventCreate(&e1));
    double best = 1e30;
    Cand win = cands[0];
    for (const Cand& c : cands) {
        int st = c.kernel == MTL4G_SPMV_TILE ? mtl4g_csr_tune(A, c.a, c.b)
                 : c.kernel == MTL4G_SPMV_STREAM || c.kernel == MTL4G_SPMV_STREAM2
                     ? (mtl4g_csr_tune(A, c.rows, c.nnz) ||
                                                    mtl4g_csr_set_kernel(A, c.kernel, c.a, 0))
                                                 : mtl4g_csr_set_kernel(A, c.kernel, c.a, c.b);
        if (st) continue;
        if (launch_spmv(A, x, y, MTL4G_ASSIGN, s) || launch_spmv(A, x, y, MTL4G_ASSIGN, s)) continue;   // warm-up
        const int reps = 10;
        M4_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) M4_TRY(launch_spmv(A, x, y, MTL4G_ASSIGN, s));
        M4_HIP(hipEventRecord(e1, s));
        M4_HIP(hipEventSynchronize(e1));
        float ms = 0;
        M4_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1000.0 * ms / reps;
        if (us < best) {
            best = us;
            win = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFreeAsync(x, s);
    (void)hipFreeAsync(y, s);
    if (win.kernel != MTL4G_SPMV_SLICED && !had_sliced) free_sliced(A);
    if (win.kernel != MTL4G_SPMV_DIA && !had_dia) free_dia(A);
    if (win.kernel != MTL4G_SPMV_DIA_SYM && !had_dsym) free_dsym(A);
    if (win.kernel != MTL4G_SPMV_PANEL && !had_panels) free_panels(A);
    if (win.kernel == MTL4G_SPMV_STREAM || win.kernel == MTL4G_SPMV_STREAM2)
        M4_TRY(mtl4g_csr_tune(A, win.rows, win.nnz));
    M4_TRY(win.kernel == MTL4G_SPMV_TILE ? mtl4g_csr_tune(A, win.a, win.b)
                                         : mtl4g_csr_set_kernel(A, win.kernel, win.a, win.b));
    if (h_best_us) *h_best_us = best;
    return MTL4G_OK;
}

int mtl4g_spmv(mtl4g_matrix A, const double* v, double* w, int assign, void* stream) {
    if (!A || (A->n && !w) || (A->m && !v)) return MTL4G_ERR_INVALID;
    TraceRange tr("crs_cvec_mult");
    return launch_spmv(A, v, w, assign, S(stream));
}

int mtl4g_csr_prepare(mtl4g_matrix A, int fmt, void* stream) {
    if (!A || (fmt != MTL4G_F64 && fmt != MTL4G_F32 && fmt != MTL4G_BF16)) return MTL4G_ERR_INVALID;
    if (fmt == MTL4G_F64) return MTL4G_OK;
    return ensure_lowp_values(A, fmt, S(stream));
}

int mtl4g_spmv_f32(mtl4g_matrix A, int fmt, const float* v, float* w, int assign, void* stream) {
    if (!A || (A->n && !w) || (A->m && !v) || (fmt != MTL4G_F32 && fmt != MTL4G_BF16))
        return MTL4G_ERR_INVALID;
    TraceRange tr("crs_cvec_mult");
    M4_TRY(ensure_lowp_values(A, fmt, S(stream)));
    return launch_spmv_f32(A, fmt, v, w, assign, S(stream));
}

int mtl4g_spmv_dot(mtl4g_matrix A, const double* p, double* q, double* h_dot, void* stream) {
    if (!A || !h_dot || A->n != A->m) return A && A->n != A->m ? MTL4G_ERR_NOT_SQUARE : MTL4G_ERR_INVALID;
    TraceRange tr("crs_cvec_mult");
    hipStream_t s = S(stream);
    if (A->n == 0) { *h_dot = 0; return MTL4G_OK; }
    Scratch sc;
    M4_TRY(sc.get(s, A->n));
    M4_TRY(launch_csr<double>(A, A->val.p, p, EpiSpmvDot{p, q, sc.result()}, sc.partials(), sc.ticket(), s,
                              sc.terms()));
    return fetch(sc, h_dot, s);
}

// mtl4g_spmv_dot with its tickets, partials and terms in the fixed scratch:
// no device memory is allocated or freed.  Calls on one device share the
// scratch, so calls that overlap on one device (other streams, other
// threads) race; the tickets and the result slot are cleared on every call.
int mtl4g_spmv_dot_noalloc(mtl4g_matrix A, const double* p, double* q, double* h_dot, void* stream) {
    if (!A || !h_dot || A->n != A->m) return A && A->n != A->m ? MTL4G_ERR_NOT_SQUARE : MTL4G_ERR_INVALID;
    if (A->n && (!p || !q)) return MTL4G_ERR_INVALID;
    TraceRange tr("crs_cvec_mult");
    hipStream_t s = S(stream);
    if (A->n == 0) { *h_dot = 0; return MTL4G_OK; }
    const bool ordered = reduction_order() == MTL4G_ORDER_MTL4;
    if (ordered && (size_t)A->n > kNoAllocTermRows) {
        set_error("spmv_dot_noalloc: %d rows exceed the fixed scratch's %zu rows for MTL4-order reductions",
                  A->n, kNoAllocTermRows);
        return MTL4G_ERR_UNSUPPORTED;
    }
    char* base = nullptr;
    M4_TRY(noalloc_scratch(&base));
    unsigned* ticket = reinterpret_cast<unsigned*>(base);
    double* result = reinterpret_cast<double*>(base + Scratch::kTicketBytes);
    double* partials = reinterpret_cast<double*>(base + Scratch::kTicketBytes + 256);
    double* terms = ordered ? partials + kPartialWords : nullptr;
    M4_HIP(hipMemsetAsync(base, 0, Scratch::kTicketBytes + 256, s));
    M4_TRY(launch_csr<double>(A, A->val.p, p, EpiSpmvDot{p, q, result}, partials, ticket, s, terms));
    M4_HIP(hipMemcpyAsync(h_dot, result, sizeof(double), hipMemcpyDeviceToHost, s));
    M4_HIP(hipStreamSynchronize(s));
    return MTL4G_OK;
}

int mtl4g_dot(int64_t n, const double* v, const double* w, double* h, void* stream) {
    if (!h || n < 0) return MTL4G_ERR_INVALID;
    if (n == 0) { *h = 0; return MTL4G_OK; }
    hipStream_t s = S(stream);
    Scratch sc;
    M4_TRY(sc.get(s, n));
    M4_TRY(run_vec(n, EpiDot{v, w, sc.result()}, &sc, s));
    return fetch(sc, h, s);
}

int mtl4g_unary_dot(int64_t n, const double* v, double* h, void* stream) {
    if (!h || n < 0) return MTL4G_ERR_INVALID;
    if (n == 0) { *h = 0; return MTL4G_OK; }
    hipStream_t s = S(stream);
    Scratch sc;
    M4_TRY(sc.get(s, n));
    M4_TRY(run_vec(n, EpiUnaryDot<0>{v, sc.result()}, &sc, s));
    return fetch(sc, h, s);
}

int mtl4g_two_norm(int64_t n, const double* v, double* h, void* stream) {
    if (!h || n < 0) return MTL4G_ERR_INVALID;
    if (n == 0) { *h = 0; return MTL4G_OK; }
    hipStream_t s = S(stream);
    Scratch sc;
    M4_TRY(sc.get(s, n));
    M4_TRY(run_vec(n, EpiUnaryDot<1>{v, sc.result()}, &sc, s));
    return fetch(sc, h, s);
}

int mtl4g_axpy(int64_t n, double alpha, const double* v, double* w, void* stream) {
    if (n < 0) return MTL4G_ERR_INVALID;
    return run_vec(n, EpiAxpy{alpha, v, w}, nullptr, S(stream));
}

int mtl4g_axpy_unary_dot(int64_t n, double alpha, const double* q, double* r, double* h,
                         void* stream) {
    if (!h || n < 0) return MTL4G_ERR_INVALID;
    if (n == 0) { *h = 0; return MTL4G_OK; }
    hipStream_t s = S(stream);
    Scratch sc;
    M4_TRY(sc.get(s, n));
    M4_TRY(run_vec(n, EpiAxpyUnary{alpha, q, r, sc.result()}, &sc, s));
    return fetch(sc, h, s);
}

}  // extern "C"
