// dropin.hpp -- MTL4's interface for the sparse Krylov path, executed on
// MI355X through the C ABI of mtl4g.h.
//
// Source compatible with the MTL4 calls of this path:
//   mtl::compressed2D<double> A(m, n);  laplacian_setup(A, m, n);
//   mtl::mat::inserter<Matrix> ins(A);  ins[r][c] << v;
//   mtl::compressed2D<double> A(m, n, nnz, Major, Minor, Entries);
//   mtl::dense_vector<double> x(n), b(n, 1.0);  b = A * x;  w += A * v;
//   itl::pc::{identity,diagonal,ilu_0,ic_0}<Matrix[, Value]> P(A);  y = solve(P, x);
//   itl::basic_iteration / cyclic_iteration / noisy_iteration<double> iter(b, max, rtol[, atol]);
//   int err = itl::cg(A, x, b, P, iter);  itl::bicgstab(A, x, b, P, iter);
//   mtl::io::matrix_market_istream(name) >> A;  A = mtl::io::matrix_market(name);
// Host containers keep MTL4's storage (row-major CSR with std::size_t
// indices by default, contiguous vectors); operations on the path run on the
// device, with the matrix mirrored once into HBM and vectors mirrored on
// demand.  Errors surface as MTL4's exception types.
#pragma once

#include <mtl4g.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstddef>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace mtl {

// ---------------------------------------------------------------- exceptions
// mtl/utility/exception.hpp:95-236
struct runtime_error : std::runtime_error {
    explicit runtime_error(const char* s = "Run-time error") : std::runtime_error(s) {}
};
struct logic_error : std::logic_error {
    explicit logic_error(const char* s = "Logic error") : std::logic_error(s) {}
};
struct index_out_of_range : logic_error {
    explicit index_out_of_range(const char* s = "Index out of range") : logic_error(s) {}
};
struct incompatible_size : logic_error {
    explicit incompatible_size(const char* s = "Arguments have incompatible size.") : logic_error(s) {}
};
struct matrix_not_square : logic_error {
    explicit matrix_not_square(const char* s = "Matrix must be square for this operation.") : logic_error(s) {}
};
struct missing_diagonal : runtime_error {
    explicit missing_diagonal(const char* s = "Diagonal entry missing or not where it belongs to.")
        : runtime_error(s) {}
};
struct device_error : runtime_error {
    explicit device_error(const char* s) : runtime_error(s) {}
};
struct io_error : std::runtime_error {
    explicit io_error(const char* s = "I/O error") : std::runtime_error(s) {}
};
struct file_not_found : io_error {
    explicit file_not_found(const char* s = "File not found") : io_error(s) {}
};

#ifndef MTL_THROW_IF
#define MTL_THROW_IF(Test, Exception) \
    do {                              \
        if (Test) throw Exception;    \
    } while (0)
#endif

namespace detail {
[[noreturn]] inline void throw_status(int st);
inline int check(int st) {
    if (st < 0) throw_status(st);
    return st;
}
}  // namespace detail

// ---------------------------------------------------------------- tags, parameters
struct row_major {};
struct col_major {};
namespace index {
struct c_index {};
}
namespace non_fixed {
struct dimensions {};
}
namespace mat {
// mtl::mat::parameters (matrix/parameter.hpp:29-44); only row-major is on the path
template <typename Orientation = row_major, typename Index = index::c_index,
          typename Dimensions = non_fixed::dimensions, bool OnStack = false,
          typename SizeType = std::size_t>
struct parameters {
    typedef Orientation orientation;
    typedef SizeType size_type;
};
typedef parameters<row_major, index::c_index, non_fixed::dimensions, false, unsigned> unsigned_parameters;
}  // namespace mat
namespace vec {
template <typename Orientation = col_major>
struct parameters {};
}  // namespace vec

// ---------------------------------------------------------------- device storage
namespace detail {
struct device_block {
    void* p = nullptr;
    std::size_t bytes = 0;
    device_block() = default;
    device_block(const device_block&) = delete;
    device_block& operator=(const device_block&) = delete;
    ~device_block() { release(); }
    void reserve(std::size_t b) {
        if (b <= bytes && p) return;
        release();
        check(mtl4g_malloc(&p, b ? b : 1));
        bytes = b;
    }
    void release() {
        if (p) mtl4g_free(p);
        p = nullptr;
        bytes = 0;
    }
};
}  // namespace detail

// ---------------------------------------------------------------- dense_vector
namespace vec {

template <typename Value, typename Parameters = parameters<>>
class dense_vector;

template <typename V>
struct comma_init {
    V& v;
    std::size_t i;
    template <typename T>
    comma_init& operator,(const T& x) {
        MTL_THROW_IF(i >= v.size(), index_out_of_range("too many values in comma initialization"));
        v[i++] = x;
        return *this;
    }
};

// dense_vector<Value>: contiguous host storage plus an HBM mirror that is
// refreshed on demand (vector/dense_vector.hpp:50-269).
template <typename Value, typename Parameters>
class dense_vector {
  public:
    typedef Value value_type;
    typedef std::size_t size_type;
    typedef dense_vector self;

    dense_vector() {}
    explicit dense_vector(size_type n) : h_(n, Value(0)) {}
    dense_vector(size_type n, const Value& v) : h_(n, v) {}
    // non-owning view of external memory (dense_vector.hpp:122-124)
    dense_vector(size_type n, Value* ext) : ext_(ext), n_ext_(n) {}
    template <std::size_t N>
    explicit dense_vector(const Value (&a)[N]) : h_(a, a + N) {}
    dense_vector(std::initializer_list<Value> l) : h_(l) {}
    dense_vector(const dense_vector& o) {
        const Value* p = o.host_read_();   // the live copy may be on the device
        h_.assign(p, p + o.size());
    }
    dense_vector(dense_vector&& o) noexcept { swap_(o); }
    template <typename Expr, typename = decltype(std::declval<const Expr&>().assign_to(std::declval<dense_vector&>()))>
    dense_vector(const Expr& e) {
        e.assign_to(*this);
    }

    size_type size() const { return ext_ ? n_ext_ : h_.size(); }
    void change_dim(size_type n) {
        if (ext_) {
            MTL_THROW_IF(n != n_ext_, incompatible_size("cannot resize a vector on external memory"));
            return;
        }
        sync_host_();
        h_.resize(n);
        dev_ok_ = false;
    }
    void checked_change_resource(const dense_vector& o) { checked_change_resource_size(o.size()); }
    void checked_change_resource_size(size_type n) {
        if (size() == 0) change_dim(n);
        MTL_THROW_IF(size() != n, incompatible_size());
    }

    Value& operator[](size_type i) {
        sync_host_();
        dev_ok_ = false;
        return begin_()[i];
    }
    const Value& operator[](size_type i) const {
        sync_host_();
        return begin_()[i];
    }
    Value& operator()(size_type i) { return (*this)[i]; }
    const Value& operator()(size_type i) const { return (*this)[i]; }

    dense_vector& operator=(const dense_vector& o) {
        if (this == &o) return *this;
        if (size() == 0) change_dim(o.size());
        MTL_THROW_IF(size() != o.size(), incompatible_size());
        if (o.dev_ok_ && !o.host_ok_) {   // copy on the device
            Value* d = device_write_();
            mtl::detail::check(mtl4g_memcpy(d, o.dev_.p, size() * sizeof(Value), 2, nullptr));
        } else {
            o.sync_host_();
            std::copy(o.begin_(), o.begin_() + o.size(), begin_());
            host_ok_ = true;
            dev_ok_ = false;
        }
        return *this;
    }
    dense_vector& operator=(dense_vector&& o) noexcept {
        if (ext_) {
            *this = static_cast<const dense_vector&>(o);
            return *this;
        }
        swap_(o);
        return *this;
    }
    comma_init<dense_vector> operator=(const Value& v) {
        std::fill(host_write_(), host_write_() + size(), v);
        return comma_init<dense_vector>{*this, 1};
    }
    template <typename Expr, typename = decltype(std::declval<const Expr&>().assign_to(std::declval<dense_vector&>()))>
    dense_vector& operator=(const Expr& e) {
        e.assign_to(*this);
        return *this;
    }
    template <typename Expr, typename = decltype(std::declval<const Expr&>().plus_to(std::declval<dense_vector&>()))>
    dense_vector& operator+=(const Expr& e) {
        e.plus_to(*this);
        return *this;
    }
    template <typename Expr, typename = decltype(std::declval<const Expr&>().minus_to(std::declval<dense_vector&>()))>
    dense_vector& operator-=(const Expr& e) {
        e.minus_to(*this);
        return *this;
    }
    dense_vector& operator*=(const Value& a) {
        Value* p = host_write_();
        for (size_type i = 0; i < size(); i++) p[i] *= a;
        return *this;
    }

    // ---- device mirror (internal)
    const Value* device_read_() const {
        if (!dev_ok_) {
            const_cast<detail::device_block&>(dev_).reserve(size() * sizeof(Value));
            if (size())
                mtl::detail::check(mtl4g_memcpy(dev_.p, begin_(), size() * sizeof(Value), 0, nullptr));
            dev_ok_ = true;
        }
        return static_cast<const Value*>(dev_.p);
    }
    Value* device_readwrite_() {
        device_read_();
        host_ok_ = false;
        return static_cast<Value*>(dev_.p);
    }
    Value* device_write_() {
        dev_.reserve(size() * sizeof(Value));
        dev_ok_ = true;
        host_ok_ = false;
        return static_cast<Value*>(dev_.p);
    }
    Value* host_write_() {
        sync_host_();
        dev_ok_ = false;
        return begin_();
    }
    const Value* host_read_() const {
        sync_host_();
        return begin_();
    }

  private:
    void sync_host_() const {
        if (!host_ok_) {
            if (size())
                mtl::detail::check(mtl4g_memcpy(const_cast<Value*>(begin_()), dev_.p,
                                                size() * sizeof(Value), 1, nullptr));
            host_ok_ = true;
        }
    }
    Value* begin_() { return ext_ ? ext_ : h_.data(); }
    const Value* begin_() const { return ext_ ? ext_ : h_.data(); }
    void swap_(dense_vector& o) {
        std::swap(h_, o.h_);
        std::swap(ext_, o.ext_);
        std::swap(n_ext_, o.n_ext_);
        std::swap(dev_.p, o.dev_.p);
        std::swap(dev_.bytes, o.dev_.bytes);
        std::swap(host_ok_, o.host_ok_);
        std::swap(dev_ok_, o.dev_ok_);
    }

    std::vector<Value> h_;
    Value* ext_ = nullptr;
    size_type n_ext_ = 0;
    detail::device_block dev_;
    mutable bool host_ok_ = true;
    mutable bool dev_ok_ = false;
};

template <typename V, typename P>
inline std::size_t size(const dense_vector<V, P>& v) { return v.size(); }
template <typename V, typename P>
inline std::size_t num_rows(const dense_vector<V, P>& v) { return v.size(); }

template <typename V, typename P>
inline std::ostream& operator<<(std::ostream& os, const dense_vector<V, P>& v) {
    os << "{" << v.size() << "}[";
    for (std::size_t i = 0; i < v.size(); i++) os << v[i] << (i + 1 < v.size() ? "," : "");
    return os << "]";
}

// ---- element-wise expressions, evaluated on the host (not on the path)
template <typename E>
struct vec_expr_base {};

template <typename V, typename P>
inline const V* expr_data(const dense_vector<V, P>& v) { return v.host_read_(); }

template <typename L, typename R, int Op>   // 0: l+r, 1: l-r
struct vec_vec_expr {
    const L& l;
    const R& r;
    std::size_t size() const { return l.size(); }
    template <typename Out>
    void assign_to(Out& w) const {
        MTL_THROW_IF(l.size() != r.size(), incompatible_size());
        w.checked_change_resource_size(l.size());
        std::vector<double> a = materialize(l), b = materialize(r);
        auto* p = w.host_write_();
        for (std::size_t i = 0; i < l.size(); i++) p[i] = Op == 0 ? a[i] + b[i] : a[i] - b[i];
    }
    template <typename Out>
    void plus_to(Out& w) const { apply_(w, 1.0); }
    template <typename Out>
    void minus_to(Out& w) const { apply_(w, -1.0); }
    template <typename Out>
    void apply_(Out& w, double sign) const {
        std::vector<double> a = materialize(l), b = materialize(r);
        auto* p = w.host_write_();
        for (std::size_t i = 0; i < l.size(); i++) p[i] += sign * (Op == 0 ? a[i] + b[i] : a[i] - b[i]);
    }
};

template <typename S, typename R>
struct scal_vec_expr {
    S a;
    const R& r;
    std::size_t size() const { return r.size(); }
    template <typename Out>
    void assign_to(Out& w) const {
        w.checked_change_resource_size(r.size());
        std::vector<double> b = materialize(r);
        auto* p = w.host_write_();
        for (std::size_t i = 0; i < r.size(); i++) p[i] = a * b[i];
    }
    template <typename Out>
    void plus_to(Out& w) const {
        std::vector<double> b = materialize(r);
        auto* p = w.host_write_();
        for (std::size_t i = 0; i < r.size(); i++) p[i] += a * b[i];
    }
    template <typename Out>
    void minus_to(Out& w) const {
        std::vector<double> b = materialize(r);
        auto* p = w.host_write_();
        for (std::size_t i = 0; i < r.size(); i++) p[i] -= a * b[i];
    }
};

template <typename V, typename P>
inline std::vector<double> materialize(const dense_vector<V, P>& v) {
    const V* p = v.host_read_();
    return std::vector<double>(p, p + v.size());
}
template <typename E>
inline auto materialize(const E& e) -> decltype(e.size(), std::vector<double>()) {
    dense_vector<double> t(e.size());
    e.assign_to(t);
    const double* p = t.host_read_();
    return std::vector<double>(p, p + t.size());
}

template <typename T>
struct is_vec_operand : std::false_type {};
template <typename V, typename P>
struct is_vec_operand<dense_vector<V, P>> : std::true_type {};
template <typename L, typename R, int Op>
struct is_vec_operand<vec_vec_expr<L, R, Op>> : std::true_type {};
template <typename S, typename R>
struct is_vec_operand<scal_vec_expr<S, R>> : std::true_type {};

template <typename L, typename R,
          typename = std::enable_if_t<is_vec_operand<L>::value && is_vec_operand<R>::value>>
inline vec_vec_expr<L, R, 0> operator+(const L& l, const R& r) { return {l, r}; }
template <typename L, typename R,
          typename = std::enable_if_t<is_vec_operand<L>::value && is_vec_operand<R>::value>>
inline vec_vec_expr<L, R, 1> operator-(const L& l, const R& r) { return {l, r}; }
template <typename S, typename R,
          typename = std::enable_if_t<std::is_arithmetic<S>::value && is_vec_operand<R>::value>>
inline scal_vec_expr<double, R> operator*(const S& a, const R& r) { return {double(a), r}; }

template <typename V, typename P>
inline void iota(dense_vector<V, P>& v, const V offset = V(0)) {
    V* p = v.host_write_();
    for (std::size_t i = 0; i < v.size(); i++) p[i] = V(i) + offset;
}

}  // namespace vec

using vec::dense_vector;
using vec::iota;
using vec::size;

// ---------------------------------------------------------------- compressed2D
namespace mat {

template <typename Value, typename Parameters = parameters<>>
class compressed2D;

template <typename Value>
struct update_store {
    void operator()(Value& x, const Value& y) const { x = y; }
};
template <typename Value>
struct update_plus {
    void operator()(Value& x, const Value& y) const { x += y; }
};

// Row-major CSR (matrix/compressed2D.hpp:258-653) with a device mirror.
template <typename Value, typename Parameters>
class compressed2D {
    static_assert(std::is_same<Value, double>::value,
                  "the device path of compressed2D holds double values");
    static_assert(std::is_same<typename Parameters::orientation, row_major>::value,
                  "the device path of compressed2D is row-major");

  public:
    typedef Value value_type;
    typedef typename Parameters::size_type size_type;
    typedef Parameters parameters;
    typedef compressed2D self;

    compressed2D() : starts_(1, 0) {}
    compressed2D(size_type m, size_type n) : nr_(m), nc_(n), starts_(m + 1, 0) {}
    // Raw CSR constructor: copies the arrays (compressed2D.hpp:355-364)
    compressed2D(size_type m, size_type n, size_type nnz, size_type* Major, size_type* Minor,
                 Value* Entries)
        : nr_(m), nc_(n), starts_(Major, Major + m + 1), indices_(Minor, Minor + nnz),
          data_(Entries, Entries + nnz) {}
    compressed2D(const compressed2D& o)
        : nr_(o.nr_), nc_(o.nc_), starts_(o.starts_), indices_(o.indices_), data_(o.data_) {}
    // From a file description, e.g. compressed2D<double> A(mtl::io::matrix_market(name))
    template <typename File,
              typename = decltype(std::declval<const File&>().read_to_(std::declval<compressed2D&>()))>
    explicit compressed2D(const File& f) : starts_(1, 0) {
        f.read_to_(*this);
    }
    template <typename File,
              typename = decltype(std::declval<const File&>().read_to_(std::declval<compressed2D&>()))>
    compressed2D& operator=(const File& f) {
        f.read_to_(*this);
        return *this;
    }
    compressed2D& operator=(const compressed2D& o) {
        if (this != &o) {
            nr_ = o.nr_;
            nc_ = o.nc_;
            starts_ = o.starts_;
            indices_ = o.indices_;
            data_ = o.data_;
            invalidate_();
        }
        return *this;
    }
    ~compressed2D() { invalidate_(); }

    void change_dim(size_type m, size_type n) {
        nr_ = m;
        nc_ = n;
        starts_.assign(m + 1, 0);
        indices_.clear();
        data_.clear();
        invalidate_();
    }
    // A = a: a*I (the diagonal set to a, everything else removed)
    compressed2D& operator=(const Value& a) {
        size_type k = std::min(nr_, nc_);
        indices_.clear();
        data_.clear();
        starts_.assign(nr_ + 1, 0);
        for (size_type r = 0; r < nr_; r++) {
            if (r < k && a != Value(0)) {
                indices_.push_back(r);
                data_.push_back(a);
            }
            starts_[r + 1] = indices_.size();
        }
        invalidate_();
        return *this;
    }

    size_type nnz() const { return indices_.size(); }
    size_type num_rows() const { return nr_; }
    size_type num_cols() const { return nc_; }

    // A(r, c), A[r][c]: search in the row, zero when absent (compressed2D.hpp:537-547)
    Value operator()(size_type r, size_type c) const {
        MTL_THROW_IF(r >= nr_ || c >= nc_, index_out_of_range());
        auto b = indices_.begin() + starts_[r], e = indices_.begin() + starts_[r + 1];
        auto it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? data_[it - indices_.begin()] : Value(0);
    }
    struct row_ref {
        const compressed2D& A;
        size_type r;
        Value operator[](size_type c) const { return A(r, c); }
    };
    row_ref operator[](size_type r) const { return row_ref{*this, r}; }

    const size_type* ref_major() const { return starts_.data(); }
    const size_type* ref_minor() const { return indices_.data(); }
    const Value* address_data() const { return data_.data(); }
    const std::vector<size_type>& starts() const { return starts_; }
    const std::vector<size_type>& indices() const { return indices_; }
    const std::vector<Value>& data() const { return data_; }

    // (internal) replace the whole pattern
    void set_csr_(std::vector<size_type>&& s, std::vector<size_type>&& i, std::vector<Value>&& v) {
        starts_ = std::move(s);
        indices_ = std::move(i);
        data_ = std::move(v);
        invalidate_();
    }

    // (internal) HBM mirror, built on first use and after every change
    mtl4g_matrix device_() const {
        if (!dev_) {
            std::vector<int32_t> rp(starts_.begin(), starts_.end()), ci(indices_.begin(), indices_.end());
            mtl4g_matrix h = nullptr;
            mtl::detail::check(mtl4g_csr_create(&h, (int32_t)nr_, (int32_t)nc_, (int64_t)nnz(), rp.data(),
                                                ci.empty() ? nullptr : ci.data(),
                                                data_.empty() ? nullptr : data_.data(), 0));
            dev_ = std::shared_ptr<mtl4g_matrix_s>(h, [](mtl4g_matrix p) { mtl4g_csr_destroy(p); });
        }
        return dev_.get();
    }

  private:
    void invalidate_() { dev_.reset(); }

    size_type nr_ = 0, nc_ = 0;
    std::vector<size_type> starts_, indices_;
    std::vector<Value> data_;
    mutable std::shared_ptr<mtl4g_matrix_s> dev_;
};

template <typename V, typename P>
inline typename compressed2D<V, P>::size_type num_rows(const compressed2D<V, P>& A) { return A.num_rows(); }
template <typename V, typename P>
inline typename compressed2D<V, P>::size_type num_cols(const compressed2D<V, P>& A) { return A.num_cols(); }
template <typename V, typename P>
inline typename compressed2D<V, P>::size_type size(const compressed2D<V, P>& A) {
    return A.num_rows() * A.num_cols();
}
template <typename V, typename P>
inline void set_to_zero(compressed2D<V, P>& A) { A.change_dim(A.num_rows(), A.num_cols()); }

// mtl::mat::inserter (matrix/inserter.hpp:30-48): entries collected, then the
// CSR rebuilt (rows sorted by column, duplicates combined by Updater) when
// the inserter is destroyed, as compressed2D_inserter::finish does.
template <typename Matrix, typename Updater = update_store<typename Matrix::value_type>>
class inserter {
    typedef typename Matrix::value_type Value;
    typedef typename Matrix::size_type size_type;

  public:
    explicit inserter(Matrix& A, size_type = 5) : A_(A) {
        for (size_type r = 0; r < A.num_rows(); r++)
            for (size_type j = A.ref_major()[r]; j < A.ref_major()[r + 1]; j++)
                e_[{r, A.ref_minor()[j]}] = A.address_data()[j];
    }
    ~inserter() {
        std::vector<size_type> s(A_.num_rows() + 1, 0), i;
        std::vector<Value> v;
        i.reserve(e_.size());
        v.reserve(e_.size());
        for (auto& kv : e_) {
            s[kv.first.first + 1]++;
            i.push_back(kv.first.second);
            v.push_back(kv.second);
        }
        for (size_type r = 0; r < A_.num_rows(); r++) s[r + 1] += s[r];
        A_.set_csr_(std::move(s), std::move(i), std::move(v));
    }
    struct entry {
        inserter& ins;
        size_type r, c;
        template <typename T>
        entry& operator<<(const T& val) {
            MTL_THROW_IF(r >= ins.A_.num_rows() || c >= ins.A_.num_cols(), index_out_of_range());
            auto it = ins.e_.find({r, c});
            if (it == ins.e_.end()) ins.e_[{r, c}] = Value(val);
            else Updater()(it->second, Value(val));
            return *this;
        }
    };
    struct row_proxy {
        inserter& ins;
        size_type r;
        entry operator[](size_type c) { return entry{ins, r, c}; }
    };
    entry operator()(size_type r, size_type c) { return entry{*this, r, c}; }
    row_proxy operator[](size_type r) { return row_proxy{*this, r}; }

  private:
    Matrix& A_;
    std::map<std::pair<size_type, size_type>, Value> e_;
};

// laplacian_setup(A, m, n) (matrix/laplacian_setup.hpp:25-44)
template <typename Matrix>
inline void laplacian_setup(Matrix& A, unsigned m, unsigned n) {
    typedef typename Matrix::size_type size_type;
    A.change_dim(m * n, m * n);
    std::vector<size_type> s(1, 0), idx;
    std::vector<double> v;
    for (unsigned i = 0; i < m; i++)
        for (unsigned j = 0; j < n; j++) {
            unsigned row = i * n + j;
            if (i > 0) { idx.push_back(row - n); v.push_back(-1.0); }
            if (j > 0) { idx.push_back(row - 1); v.push_back(-1.0); }
            idx.push_back(row); v.push_back(4.0);
            if (j < n - 1) { idx.push_back(row + 1); v.push_back(-1.0); }
            if (i < m - 1) { idx.push_back(row + n); v.push_back(-1.0); }
            s.push_back(idx.size());
        }
    A.set_csr_(std::move(s), std::move(idx), std::move(v));
}

// trans(A): explicit transposed copy
template <typename V, typename P>
inline compressed2D<V, P> trans(const compressed2D<V, P>& A) {
    typedef typename compressed2D<V, P>::size_type size_type;
    compressed2D<V, P> T(A.num_cols(), A.num_rows());
    std::vector<size_type> s(A.num_cols() + 1, 0), i(A.nnz());
    std::vector<V> v(A.nnz());
    for (size_type j = 0; j < A.nnz(); j++) s[A.ref_minor()[j] + 1]++;
    for (size_type c = 0; c < A.num_cols(); c++) s[c + 1] += s[c];
    std::vector<size_type> fill(s.begin(), s.end() - 1);
    for (size_type r = 0; r < A.num_rows(); r++)
        for (size_type j = A.ref_major()[r]; j < A.ref_major()[r + 1]; j++) {
            size_type d = fill[A.ref_minor()[j]]++;
            i[d] = r;
            v[d] = A.address_data()[j];
        }
    T.set_csr_(std::move(s), std::move(i), std::move(v));
    return T;
}

template <typename V, typename P>
inline std::ostream& operator<<(std::ostream& os, const compressed2D<V, P>& A) {
    for (typename compressed2D<V, P>::size_type r = 0; r < A.num_rows(); r++) {
        os << "[";
        for (typename compressed2D<V, P>::size_type c = 0; c < A.num_cols(); c++)
            os << (c ? " " : "") << A(r, c);
        os << "]\n";
    }
    return os;
}

// ---- A * v (operation/operators.hpp:38-45 -> smat_cvec_mult on the device)
template <typename Matrix, typename Vector>
struct mat_cvec_times_expr {
    const Matrix& A;
    const Vector& v;
    std::size_t size() const { return A.num_rows(); }
    template <typename Out>
    void run_(Out& w, int assign) const {
        MTL_THROW_IF(v.size() != A.num_cols(), incompatible_size());
        if (assign == MTL4G_ASSIGN) w.checked_change_resource_size(A.num_rows());
        MTL_THROW_IF(w.size() != A.num_rows(), incompatible_size());
        mtl4g_matrix h = A.device_();
        const double* dv = v.device_read_();
        double* dw = assign == MTL4G_ASSIGN ? w.device_write_() : w.device_readwrite_();
        mtl::detail::check(mtl4g_spmv(h, dv, dw, assign, nullptr));
    }
    template <typename Out>
    void assign_to(Out& w) const { run_(w, MTL4G_ASSIGN); }
    template <typename Out>
    void plus_to(Out& w) const { run_(w, MTL4G_PLUS); }
    template <typename Out>
    void minus_to(Out& w) const { run_(w, MTL4G_MINUS); }
};

template <typename V, typename P, typename VV, typename VP>
inline mat_cvec_times_expr<compressed2D<V, P>, vec::dense_vector<VV, VP>>
operator*(const compressed2D<V, P>& A, const vec::dense_vector<VV, VP>& v) {
    return {A, v};
}

}  // namespace mat

using mat::compressed2D;
using mat::laplacian_setup;
using mat::num_cols;
using mat::num_rows;
using mat::set_to_zero;
using mat::trans;

// ---------------------------------------------------------------- Matrix Market IO
// io/matrix_market.hpp: header "%%MatrixMarket matrix <coordinate|array>
// <real|integer|pattern|complex> <general|symmetric|skew-symmetric|hermitian>",
// '%' comment lines, then "rows cols [entries]".  Entries go through an
// inserter (later duplicates overwrite); off-diagonal entries of symmetric,
// skew-symmetric and Hermitian files are mirrored (negated for skew);
// pattern entries are stored as 0; complex values cannot go into a real
// matrix; array files are read column by column.
namespace io {

class matrix_market_istream {
  public:
    explicit matrix_market_istream(const char* p) : own_(new std::ifstream(p)), in_(*own_) { check_(p); }
    explicit matrix_market_istream(const std::string& s) : matrix_market_istream(s.c_str()) {}
    explicit matrix_market_istream(std::istream& s = std::cin) : in_(s) { check_(nullptr); }

    template <typename V, typename P>
    matrix_market_istream& operator>>(mat::compressed2D<V, P>& A) {
        read_(A);
        return *this;
    }
    void close() {
        if (own_) own_->close();
    }

  private:
    enum symmetry_t { general, symmetric, skew, hermitian };

    void check_(const char* name) {
        if (in_.good()) return;
        std::string msg("matrix_market_istream: Error in input stream!\n");
        if (name) msg += std::string("Probably file ") + name + " not found.\n";
        throw file_not_found(msg.c_str());
    }
    static std::string lower_(std::string s) {
        for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
        return s;
    }

    template <typename V, typename P>
    void read_(mat::compressed2D<V, P>& A) {
        typedef typename mat::compressed2D<V, P>::size_type size_type;
        std::string marker, type, layout, field, sym;
        in_ >> marker >> type >> layout >> field >> sym;
        MTL_THROW_IF(marker != "%%MatrixMarket", runtime_error("File not in Matrix Market format"));
        MTL_THROW_IF(type != "matrix", runtime_error("Try to read matrix from non-matrix file"));
        layout = lower_(layout);
        sym = lower_(sym);
        symmetry_t symm;
        if (sym == "general") symm = general;
        else if (sym == "symmetric") symm = symmetric;
        else if (sym == "skew-symmetric") symm = skew;
        else if (sym == "hermitian") symm = hermitian;
        else throw runtime_error("Unknown symmetry in Matrix Market file");
        MTL_THROW_IF(layout != "coordinate" && layout != "array",
                     runtime_error("Unknown sparsity in Matrix Market file"));
        for (;;) {   // comment lines
            char first = 0;
            in_ >> first;
            if (first != '%') {
                in_.putback(first);
                break;
            }
            std::string rest;
            std::getline(in_, rest);
        }
        std::size_t nr = 0, nc = 0, nnz = 0;
        in_ >> nr >> nc;
        if (layout == "coordinate") in_ >> nnz;
        int kind;   // 0 real, 1 integer, 2 pattern
        if (field == "real") kind = 0;
        else if (field == "integer") kind = 1;
        else if (field == "pattern") kind = 2;
        else if (field == "complex") throw runtime_error("Cannot convert complex value in real\n");
        else throw runtime_error("Unknown tag for matrix value type in file");
        A.change_dim((size_type)nr, (size_type)nc);
        mat::inserter<mat::compressed2D<V, P>> ins(A);
        auto value = [&]() -> V {
            if (kind == 0) {
                double v = 0;
                in_ >> v;
                return V(v);
            }
            if (kind == 1) {
                long v = 0;
                in_ >> v;
                return V(v);
            }
            return V(0);
        };
        auto put = [&](std::size_t r, std::size_t c, V v) {
            ins[(size_type)r][(size_type)c] << v;
            if (r == c) return;
            if (symm == symmetric || symm == hermitian) ins[(size_type)c][(size_type)r] << v;
            else if (symm == skew) ins[(size_type)c][(size_type)r] << -v;
        };
        if (layout == "coordinate") {
            for (std::size_t k = 0; k < nnz; k++) {
                std::size_t r = 0, c = 0;
                in_ >> r >> c;
                if (!in_) break;
                put(r - 1, c - 1, value());
            }
        } else {
            for (std::size_t c = 0; c < nc; c++)
                for (std::size_t r = 0; r < nr; r++) put(r, c, value());
        }
    }

    std::unique_ptr<std::ifstream> own_;
    std::istream& in_;
};

// Coordinate real general, 17 significant digits (values read back exactly).
class matrix_market_ostream {
  public:
    explicit matrix_market_ostream(const char* p) : own_(new std::ofstream(p)), out_(*own_) {}
    explicit matrix_market_ostream(const std::string& s) : matrix_market_ostream(s.c_str()) {}
    explicit matrix_market_ostream(std::ostream& s = std::cout) : out_(s) {}

    template <typename V, typename P>
    matrix_market_ostream& operator<<(const mat::compressed2D<V, P>& A) {
        out_ << "%%MatrixMarket matrix coordinate real general\n";
        out_ << A.num_rows() << " " << A.num_cols() << " " << A.nnz() << "\n";
        const std::streamsize prec = out_.precision(17);
        for (std::size_t r = 0; r < (std::size_t)A.num_rows(); r++)
            for (std::size_t j = A.ref_major()[r]; j < (std::size_t)A.ref_major()[r + 1]; j++)
                out_ << r + 1 << " " << A.ref_minor()[j] + 1 << " " << A.address_data()[j] << "\n";
        out_.precision(prec);
        return *this;
    }
    void close() {
        if (own_) own_->close();
    }

  private:
    std::unique_ptr<std::ofstream> own_;
    std::ostream& out_;
};

// mtl::io::matrix_market(name): read with `compressed2D<double> A(matrix_market(name))`
// or `A = matrix_market(name)`, write with `matrix_market(name) = A`
// (io/matrix_file.hpp:21-40).
class matrix_market {
  public:
    explicit matrix_market(const std::string& fname) : fname_(fname) {}
    explicit matrix_market(const char* fname) : fname_(fname) {}
    template <typename V, typename P>
    matrix_market& operator=(const mat::compressed2D<V, P>& A) {
        matrix_market_ostream out(fname_);
        out << A;
        return *this;
    }
    template <typename V, typename P>
    void read_to_(mat::compressed2D<V, P>& A) const {
        matrix_market_istream in(fname_);
        in >> A;
    }
    const std::string& file_name() const { return fname_; }

  private:
    std::string fname_;
};

}  // namespace io

namespace vec {
template <typename M, typename V>
struct is_vec_operand<mat::mat_cvec_times_expr<M, V>> : std::true_type {};
}  // namespace vec

// ---------------------------------------------------------------- reductions
// dot (operation/dot.hpp:226-232), unary_dot, two_norm (two_norm.hpp:52-57)
template <typename V1, typename P1, typename V2, typename P2>
inline double dot(const dense_vector<V1, P1>& v, const dense_vector<V2, P2>& w) {
    MTL_THROW_IF(v.size() != w.size(), incompatible_size());
    double r = 0;
    detail::check(mtl4g_dot((int64_t)v.size(), v.device_read_(), w.device_read_(), &r, nullptr));
    return r;
}
template <typename V, typename P>
inline double unary_dot(const dense_vector<V, P>& v) {
    double r = 0;
    detail::check(mtl4g_unary_dot((int64_t)v.size(), v.device_read_(), &r, nullptr));
    return r;
}
template <typename V, typename P>
inline double two_norm(const dense_vector<V, P>& v) {
    double r = 0;
    detail::check(mtl4g_two_norm((int64_t)v.size(), v.device_read_(), &r, nullptr));
    return r;
}
template <typename E, typename = decltype(std::declval<const E&>().assign_to(std::declval<dense_vector<double>&>()))>
inline double two_norm(const E& e) {
    dense_vector<double> t(e);
    return two_norm(t);
}

}  // namespace mtl

// ---------------------------------------------------------------- ITL
namespace itl {

// itl/utility/exception.hpp:22-31
struct search_space_exhaustion : mtl::runtime_error {
    explicit search_space_exhaustion(const char* s = "Search space exhausted.") : mtl::runtime_error(s) {}
};
struct unexpected_orthogonality : mtl::runtime_error {
    explicit unexpected_orthogonality(const char* s = "Vectors unexpectedly orthogonal")
        : mtl::runtime_error(s) {}
};

namespace detail {
struct iteration_access;
}

// itl::basic_iteration (iteration/basic_iteration.hpp:23-164)
template <class Real>
class basic_iteration {
  public:
    typedef basic_iteration self;
    typedef Real real;

    template <class Vector>
    basic_iteration(const Vector& r0, int max_iter_, Real t, Real a = Real(0))
        : error(0), i(0), my_norm_r0(std::abs(mtl::two_norm(r0))), max_iter(max_iter_), rtol_(t),
          atol_(a), resid_(0), is_finished(false), my_quite(false), my_suppress(false) {}
    basic_iteration(Real nb, int max_iter_, Real t, Real a = Real(0))
        : error(0), i(0), my_norm_r0(nb), max_iter(max_iter_), rtol_(t), atol_(a), resid_(0),
          is_finished(false), my_quite(false), my_suppress(false) {}
    virtual ~basic_iteration() {}

    bool check_max() {
        if (i >= max_iter) error = 1, is_finished = true, err_msg = "Too many iterations.";
        return is_finished;
    }
    bool finished(const Real& r) {
        if (converged(r)) return is_finished = true;
        return check_max();
    }
    template <class Vector>
    bool finished(const Vector& r) {
        if (converged(mtl::two_norm(r))) return is_finished = true;
        return check_max();
    }
    bool finished() const { return is_finished; }
    template <class T>
    int terminate(const T& r) {
        finished(r);
        return error;
    }

    self& operator++() { ++i; return *this; }
    self& operator+=(int n) { i += n; return *this; }
    bool first() const { return i <= 1; }
    virtual operator int() const { return error; }
    virtual int error_code() const { return error; }
    bool is_converged() const { return is_finished && error == 0; }
    int iterations() const { return i; }
    int max_iterations() const { return max_iter; }
    void set_max_iterations(int m) { max_iter = m; }
    void restart() { i = 0; }
    Real resid() const { return resid_; }
    Real relresid() const { return resid_ / my_norm_r0; }
    Real normb() const { return my_norm_r0; }
    Real norm_r0() const { return my_norm_r0; }
    void set_norm_r0(Real r0) { my_norm_r0 = r0; }
    Real tol() const { return rtol_; }
    Real atol() const { return atol_; }
    int fail(int err_code) { error = err_code; return error_code(); }
    int fail(int err_code, const std::string& msg) { error = err_code; err_msg = msg; return error_code(); }
    void set(Real v) { my_norm_r0 = v; }
    void set_quite(bool q) { my_quite = q; }
    bool is_quite() const { return my_quite; }
    void suppress_resume(bool s) { my_suppress = s; }
    bool resume_suppressed() const { return my_suppress; }

  protected:
    friend struct detail::iteration_access;
    bool converged(const Real& r) { resid_ = r; return converged(); }
    bool converged() const {
        if (my_norm_r0 == 0) return resid_ <= atol_;
        return resid_ <= rtol_ * my_norm_r0 || resid_ <= atol_;
    }
    // residuals of the device solve, one per finished() call
    // arg: l of BiCGSTAB(l), restart of GMRES (0: gmres_full), unused otherwise
    virtual void replay_(const std::vector<double>&, int /*method*/, int /*arg*/, int /*i0*/) {}
    virtual bool wants_history_() const { return false; }

    int error, i;
    Real my_norm_r0;
    int max_iter;
    Real rtol_, atol_, resid_;
    std::string err_msg;
    bool is_finished, my_quite, my_suppress;
};

// itl::cyclic_iteration (iteration/cyclic_iteration.hpp:23-100)
template <class Real, class OStream = std::ostream>
class cyclic_iteration : public basic_iteration<Real> {
    typedef basic_iteration<Real> super;

    void print_resid() {
        if (!this->my_quite && this->i % cycle == 0)
            if (multi_print || this->i != last_print) {
                out << "iteration " << this->i << ": resid " << this->resid() << std::endl;
                last_print = this->i;
            }
    }

  public:
    template <class Vector>
    cyclic_iteration(const Vector& r0, int max_iter_, Real tol_, Real atol_ = Real(0), int cycle_ = 100,
                     OStream& out = std::cout)
        : super(r0, max_iter_, tol_, atol_), cycle(cycle_), last_print(-1), multi_print(false), out(out) {}
    cyclic_iteration(Real r0, int max_iter_, Real tol_, Real atol_ = Real(0), int cycle_ = 100,
                     OStream& out = std::cout)
        : super(r0, max_iter_, tol_, atol_), cycle(cycle_), last_print(-1), multi_print(false), out(out) {}

    bool finished() { return super::finished(); }
    template <typename T>
    bool finished(const T& r) {
        bool ret = super::finished(r);
        print_resid();
        return ret;
    }
    cyclic_iteration& operator++() { ++this->i; return *this; }
    cyclic_iteration& operator+=(int n) { this->i += n; return *this; }
    operator int() const { return error_code(); }
    bool is_multi_print() const { return multi_print; }
    void set_multi_print(bool m) { multi_print = m; }
    void set_cycle(int c) { cycle = c; }
    int error_code() const {
        if (!this->my_suppress)
            out << "finished! error code = " << this->error << '\n'
                << this->iterations() << " iterations\n"
                << this->resid() << " is actual final residual. \n"
                << this->relresid() << " is actual relative tolerance achieved. \n"
                << "Relative tol: " << this->rtol_ << "  Absolute tol: " << this->atol_ << '\n'
                << "Convergence:  " << std::pow(this->relresid(), 1.0 / double(this->iterations()))
                << std::endl;
        return this->error;
    }

  protected:
    // the device ran the loop; print what finished() would have printed
    void replay_(const std::vector<double>& hist, int method, int arg, int i0) override {
        int i = i0;
        for (std::size_t k = 0; k < hist.size(); k++) {
            if (method == MTL4G_TFQMR        ? k > 1        // finished(rt) before the loop, then tau
                : method == MTL4G_BICGSTAB_2 ? k % 3 == 1   // loop head, then r_hat[0] and r_hat[1]
                : method == MTL4G_GMRES      ? false        // set per cycle below
                : method == MTL4G_BICGSTAB_ELL ? k % (arg + 1) == 1   // loop head, then r_hat[0..l-1]
                                             : (k > 0 && (method != MTL4G_BICGSTAB || k % 2 == 1)))   // cg, cgs: one per iteration
                i++;
            if (method == MTL4G_GMRES && k % 2 == 1)   // end of a cycle of `restart` Arnoldi steps
                i = std::min(i + (arg > 0 ? arg : this->max_iter), this->max_iter);
            this->i = i;
            this->resid_ = hist[k];
            print_resid();
        }
    }
    bool wants_history_() const override { return !this->my_quite; }

    int cycle, last_print;
    bool multi_print;
    OStream& out;
};

// itl::noisy_iteration (iteration/noisy_iteration.hpp:21-30)
template <class Real, class OStream = std::ostream>
class noisy_iteration : public cyclic_iteration<Real, OStream> {
  public:
    template <class Vector>
    noisy_iteration(const Vector& r0, int max_iter_, Real tol_, Real atol_ = Real(0), OStream& out = std::cout)
        : cyclic_iteration<Real, OStream>(r0, max_iter_, tol_, atol_, 1, out) {}
};

// itl::default_iteration (iteration/default_iteration.hpp:22-28): quiet cyclic
template <class Real>
class default_iteration : public cyclic_iteration<Real> {
  public:
    template <class Vector>
    default_iteration(const Vector& r0, int max_iter_, Real tol_, Real atol_ = Real(0))
        : cyclic_iteration<Real>(r0, max_iter_, tol_, atol_) {
        this->set_quite(true);
        this->suppress_resume(true);
    }
    default_iteration(Real r0 = 1, int max_iter_ = 1000, Real tol_ = 1e-6, Real atol_ = Real(0))
        : cyclic_iteration<Real>(r0, max_iter_, tol_, atol_) {
        this->set_quite(true);
        this->suppress_resume(true);
    }
};

namespace detail {
struct iteration_access {
    template <class Real>
    static mtl4g_iteration state(const basic_iteration<Real>& it) {
        mtl4g_iteration s;
        s.norm_r0 = it.my_norm_r0;
        s.max_iter = it.max_iter;
        s.rtol = it.rtol_;
        s.atol = it.atol_;
        s.i = it.i;
        s.error = 0;
        s.resid = 0;
        return s;
    }
    template <class Real>
    static bool wants_history(const basic_iteration<Real>& it) { return it.wants_history_(); }
    template <class Real>
    static void update(basic_iteration<Real>& it, const mtl4g_iteration& s, const std::vector<double>& hist,
                       int method, int arg, int i0) {
        if (!hist.empty()) it.replay_(hist, method, arg, i0);
        it.i = s.i;
        it.resid_ = s.resid;
        it.error = s.error;
        it.is_finished = true;
        if (s.error == 1) it.err_msg = "Too many iterations.";
    }
};
}  // namespace detail

// ---------------------------------------------------------------- preconditioners
namespace pc {

template <typename Matrix, typename Value>
class device_pc {
  public:
    typedef Value value_type;
    typedef typename Matrix::size_type size_type;

    device_pc(const Matrix& A, int kind) : n_(A.num_rows()) {
        MTL_THROW_IF(A.num_rows() != A.num_cols(), mtl::matrix_not_square());
        mtl4g_pc h = nullptr;
        const int fmt = std::is_same<Value, float>::value ? MTL4G_F32 : MTL4G_F64;
        mtl::detail::check(mtl4g_pc_create(&h, A.device_(), kind, fmt));
        h_ = std::shared_ptr<mtl4g_pc_s>(h, [](mtl4g_pc p) { mtl4g_pc_destroy(p); });
        A_keep_ = std::make_shared<Matrix>(A);   // the factors were built from this A
    }
    template <typename VectorIn, typename VectorOut>
    void solve(const VectorIn& x, VectorOut& y) const {
        MTL_THROW_IF(x.size() != n_, mtl::incompatible_size());
        y.checked_change_resource_size(x.size());
        mtl::detail::check(mtl4g_pc_solve(h_.get(), x.device_read_(), y.device_write_(), nullptr));
    }
    template <typename Vector>
    Vector solve(const Vector& x) const {
        Vector y(x.size());
        solve(x, y);
        return y;
    }
    template <typename VectorIn, typename VectorOut>
    void adjoint_solve(const VectorIn& x, VectorOut& y) const {
        MTL_THROW_IF(x.size() != n_, mtl::incompatible_size());
        y.checked_change_resource_size(x.size());
        mtl::detail::check(mtl4g_pc_adjoint_solve(h_.get(), x.device_read_(), y.device_write_(), nullptr));
    }
    template <typename Vector>
    Vector adjoint_solve(const Vector& x) const {
        Vector y(x.size());
        adjoint_solve(x, y);
        return y;
    }
    mtl4g_pc handle_() const { return h_.get(); }

  protected:
    mtl::compressed2D<double> factor_(int which) const {
        int32_t n = 0;
        int64_t nnz = 0;
        mtl::detail::check(mtl4g_pc_factor(h_.get(), which, &n, &nnz, nullptr, nullptr, nullptr));
        MTL_THROW_IF(nnz < 0, mtl::logic_error("factor not present"));
        std::vector<int32_t> rp(n + 1), ci(nnz);
        std::vector<double> va(nnz);
        mtl::detail::check(mtl4g_pc_factor(h_.get(), which, nullptr, nullptr, rp.data(), ci.data(), va.data()));
        std::vector<std::size_t> s(rp.begin(), rp.end()), i(ci.begin(), ci.end());
        mtl::compressed2D<double> F(n, n);
        F.set_csr_(std::move(s), std::move(i), std::move(va));
        return F;
    }
    size_type n_;
    std::shared_ptr<mtl4g_pc_s> h_;
    std::shared_ptr<Matrix> A_keep_;
};

// itl::pc::identity (pc/identity.hpp:24-71)
template <typename Matrix, typename Value = double>
class identity : public device_pc<Matrix, Value> {
  public:
    explicit identity(const Matrix& A) : device_pc<Matrix, Value>(A, MTL4G_PC_IDENTITY) {}
};
// itl::pc::diagonal (pc/diagonal.hpp:28-103)
template <typename Matrix, typename Value = typename Matrix::value_type>
class diagonal : public device_pc<Matrix, Value> {
  public:
    explicit diagonal(const Matrix& A) : device_pc<Matrix, Value>(A, MTL4G_PC_DIAGONAL) {}
};
// itl::pc::ilu_0 (pc/ilu_0.hpp:34-92, pc/ilu.hpp:32-133)
template <typename Matrix, typename Value = typename Matrix::value_type>
class ilu_0 : public device_pc<Matrix, Value> {
  public:
    explicit ilu_0(const Matrix& A) : device_pc<Matrix, Value>(A, MTL4G_PC_ILU0) {}
    mtl::compressed2D<double> get_L() const { return this->factor_(1); }
    mtl::compressed2D<double> get_U() const { return this->factor_(2); }
};
// itl::pc::ic_0 (pc/ic_0.hpp:39-194)
template <typename Matrix, typename Value = typename Matrix::value_type>
class ic_0 : public device_pc<Matrix, Value> {
  public:
    explicit ic_0(const Matrix& A) : device_pc<Matrix, Value>(A, MTL4G_PC_IC0) {}
    mtl::compressed2D<double> get_L() const { return this->factor_(1); }
    mtl::compressed2D<double> get_U() const { return this->factor_(2); }
};

// is_identity / static_is_identity (pc/is_identity.hpp:21-37)
template <typename PC>
bool is_identity(const PC&) { return false; }
template <typename Matrix, typename Value>
bool is_identity(const identity<Matrix, Value>&) { return true; }
template <typename PC>
struct static_is_identity : std::false_type {};
template <typename Matrix, typename Value>
struct static_is_identity<identity<Matrix, Value>> : std::true_type {};

// solve(P, x) / adjoint_solve(P, x) (pc/solver.hpp:23-52): the result vector
template <typename PC, typename Vector>
inline Vector solve(const PC& P, const Vector& x) { return P.solve(x); }
template <typename PC, typename Vector>
inline Vector adjoint_solve(const PC& P, const Vector& x) { return P.adjoint_solve(x); }

}  // namespace pc

// ---------------------------------------------------------------- Krylov solvers
namespace detail {
template <typename PC>
inline mtl4g_pc pc_handle(const PC& P) { return P.handle_(); }

// method: MTL4G_*; arg: l of BiCGSTAB(l), restart of GMRES (0: gmres_full)
template <typename Matrix, typename VX, typename VB, typename Iter>
inline int run(int method, const Matrix& A, VX& x, const VB& b, mtl4g_pc P, Iter& iter,
               mtl4g_pc R = nullptr, int arg = 0) {
    MTL_THROW_IF(A.num_rows() != A.num_cols(), mtl::matrix_not_square());
    MTL_THROW_IF(x.size() != A.num_cols() || b.size() != A.num_rows(), mtl::incompatible_size());
    mtl4g_iteration s = iteration_access::state(iter);
    const int i0 = s.i;
    const bool want = iteration_access::wants_history(iter);
    // finished() calls per iteration, at most
    const int per = method == MTL4G_GMRES ? 2 : method == MTL4G_BICGSTAB_ELL ? arg + 1 : 3;
    const int cap = want ? per * std::max(s.max_iter - s.i, 0) + 4 : 0;
    std::vector<double> hist(cap);
    mtl4g_matrix h = A.device_();
    const double* db = b.device_read_();
    double* dx = x.device_readwrite_();
    double* hp = cap ? hist.data() : nullptr;
    int st = method == MTL4G_CG      ? mtl4g_cg(h, P, dx, db, &s, hp, cap, nullptr)
             : method == MTL4G_CGS   ? mtl4g_cgs(h, P, dx, db, &s, hp, cap, nullptr)
             : method == MTL4G_TFQMR ? mtl4g_tfqmr(h, P, R, dx, db, &s, hp, cap, nullptr)
             : method == MTL4G_BICGSTAB_2 ? mtl4g_bicgstab_2(h, P, dx, db, &s, hp, cap, nullptr)
             : method == MTL4G_GMRES ? mtl4g_gmres(h, P, R, dx, db, arg, &s, hp, cap, nullptr)
             : method == MTL4G_BICGSTAB_ELL ? mtl4g_bicgstab_ell(h, P, R, dx, db, arg, &s, hp, cap, nullptr)
                                     : mtl4g_bicgstab(h, P, dx, db, &s, hp, cap, nullptr);
    std::vector<double> used;
    for (double v : hist)
        if (v == v) used.push_back(v);
    iteration_access::update(iter, s, used, method, arg, i0);
    if (st == MTL4G_ERR_BREAKDOWN) throw itl::unexpected_orthogonality();
    mtl::detail::check(st);
    return s.error;
}
}  // namespace detail

// itl::cg without preconditioner (krylov/cg.hpp:38-71)
template <typename LinearOperator, typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
int cg(const LinearOperator& A, HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) {
    detail::run(MTL4G_CG, A, x, b, nullptr, iter);
    return iter;
}
// itl::cg with preconditioner (krylov/cg.hpp:74-111); identity takes the form above
template <typename LinearOperator, typename HilbertSpaceX, typename HilbertSpaceB, typename Preconditioner,
          typename Iteration>
int cg(const LinearOperator& A, HilbertSpaceX& x, const HilbertSpaceB& b, const Preconditioner& L,
       Iteration& iter) {
    if (pc::is_identity(L)) return cg(A, x, b, iter);
    detail::run(MTL4G_CG, A, x, b, detail::pc_handle(L), iter);
    return iter;
}
// right preconditioner ignored (krylov/cg.hpp:113-120)
template <typename LinearOperator, typename HilbertSpaceX, typename HilbertSpaceB, typename Preconditioner,
          typename RightPreconditioner, typename Iteration>
int cg(const LinearOperator& A, HilbertSpaceX& x, const HilbertSpaceB& b, const Preconditioner& L,
       const RightPreconditioner&, Iteration& iter) {
    return cg(A, x, b, L, iter);
}

// itl::bicgstab (krylov/bicgstab.hpp:27-78)
template <class LinearOperator, class HilbertSpaceX, class HilbertSpaceB, class Preconditioner, class Iteration>
int bicgstab(const LinearOperator& A, HilbertSpaceX& x, const HilbertSpaceB& b, const Preconditioner& M,
             Iteration& iter) {
    detail::run(MTL4G_BICGSTAB, A, x, b, pc::is_identity(M) ? nullptr : detail::pc_handle(M), iter);
    return iter;
}

// itl::cgs (krylov/cgs.hpp:24-62)
template <typename LinearOperator, typename Vector, typename Preconditioner, typename Iteration>
int cgs(const LinearOperator& A, Vector& x, const Vector& b, const Preconditioner& M, Iteration& iter) {
    detail::run(MTL4G_CGS, A, x, b, pc::is_identity(M) ? nullptr : detail::pc_handle(M), iter);
    return iter;
}

// itl::tfqmr (krylov/tfqmr.hpp:32-96): left and right preconditioners
template <typename Matrix, typename Vector, typename LeftPreconditioner, typename RightPreconditioner,
          typename Iteration>
int tfqmr(const Matrix& A, Vector& x, const Vector& b, const LeftPreconditioner& L, const RightPreconditioner& R,
          Iteration& iter) {
    if (b.size() == 0) throw mtl::logic_error("empty rhs vector");
    detail::run(MTL4G_TFQMR, A, x, b, pc::is_identity(L) ? nullptr : detail::pc_handle(L), iter,
                pc::is_identity(R) ? nullptr : detail::pc_handle(R));
    return iter;
}

// itl::bicgstab_2 (krylov/bicgstab_2.hpp:37-130): BiCGSTAB(l), l = 2, left
// preconditioner; omega == 0 returns iter.fail(3)
template <typename LinearOperator, typename Vector, typename Preconditioner, typename Iteration>
int bicgstab_2(const LinearOperator& A, Vector& x, const Vector& b, const Preconditioner& L, Iteration& iter) {
    if (b.size() == 0) throw mtl::logic_error("empty rhs vector");
    detail::run(MTL4G_BICGSTAB_2, A, x, b, pc::is_identity(L) ? nullptr : detail::pc_handle(L), iter);
    return iter;
}

// itl::bicgstab_ell (krylov/bicgstab_ell.hpp:37-134): BiCGSTAB(l), left and
// right preconditioners, 1 <= l <= 8 here
template <typename LinearOperator, typename Vector, typename LeftPreconditioner, typename RightPreconditioner,
          typename Iteration>
int bicgstab_ell(const LinearOperator& A, Vector& x, const Vector& b, const LeftPreconditioner& L,
                 const RightPreconditioner& R, Iteration& iter, std::size_t l) {
    if (b.size() == 0) throw mtl::logic_error("empty rhs vector");
    MTL_THROW_IF(l < 1 || l > 8, mtl::logic_error("bicgstab_ell: l must be in 1..8"));
    detail::run(MTL4G_BICGSTAB_ELL, A, x, b, pc::is_identity(L) ? nullptr : detail::pc_handle(L), iter,
                pc::is_identity(R) ? nullptr : detail::pc_handle(R), int(l));
    return iter;
}

// itl::gmres_full (krylov/gmres.hpp:37-119) and itl::gmres with restart
// (:122-137); at most 256 Krylov vectors per cycle here
template <typename Matrix, typename Vector, typename LeftPreconditioner, typename RightPreconditioner,
          typename Iteration>
int gmres_full(const Matrix& A, Vector& x, const Vector& b, LeftPreconditioner& L, RightPreconditioner& R,
               Iteration& iter) {
    if (b.size() == 0) throw mtl::logic_error("empty rhs vector");
    detail::run(MTL4G_GMRES, A, x, b, pc::is_identity(L) ? nullptr : detail::pc_handle(L), iter,
                pc::is_identity(R) ? nullptr : detail::pc_handle(R), 0);
    return iter;
}
template <typename Matrix, typename Vector, typename LeftPreconditioner, typename RightPreconditioner,
          typename Iteration>
int gmres(const Matrix& A, Vector& x, const Vector& b, LeftPreconditioner& L, RightPreconditioner& R,
          Iteration& iter, std::size_t restart) {
    if (b.size() == 0) throw mtl::logic_error("empty rhs vector");
    MTL_THROW_IF(restart < 1, mtl::logic_error("gmres: restart must be positive"));
    detail::run(MTL4G_GMRES, A, x, b, pc::is_identity(L) ? nullptr : detail::pc_handle(L), iter,
                pc::is_identity(R) ? nullptr : detail::pc_handle(R), int(restart));
    return iter;
}

// itl::base_solver, cg_solver, bicgstab_solver, cgs_solver, tfqmr_solver,
// bicgstab_2_solver, bicgstab_ell_solver (bicgstab_ell.hpp:137-167),
// gmres_solver (gmres.hpp:139-176) (krylov/base_solver.hpp:24-84, cg.hpp:122-156,
// bicgstab.hpp:80-112, cgs.hpp:64-95, tfqmr.hpp:98-127, bicgstab_2.hpp:132-165)
template <typename Solver, typename LinearOperator>
struct base_solver {
    typedef double magnitude_type;
    typedef cyclic_iteration<magnitude_type> iteration_type;

    explicit base_solver(const LinearOperator& A, iteration_type iter = default_iteration<magnitude_type>())
        : A(A), my_iteration(iter), level(1) {
        my_iteration.set_quite(true);
    }
    template <typename HilbertSpaceB, typename HilbertSpaceX>
    int step(HilbertSpaceX& x, const HilbertSpaceB& b) const {
        basic_iteration<double> iter(b, 1, 0, 0);
        return static_cast<Solver const*>(this)->solve(x, b, iter);
    }
    template <typename HilbertSpaceB, typename HilbertSpaceX>
    int operator()(HilbertSpaceX& x, const HilbertSpaceB& b) {
        my_iteration.restart();
        if (mtl::two_norm(x) == 0) my_iteration.set_norm_r0(mtl::two_norm(b));
        else {
            HilbertSpaceB r(b);
            r -= A * x;
            my_iteration.set_norm_r0(mtl::two_norm(r));
        }
        return static_cast<Solver const*>(this)->solve(x, b, my_iteration);
    }
    void set_log_level(unsigned l) {
        level = l;
        my_iteration.set_quite(l < 2);
        my_iteration.suppress_resume(l == 0);
    }
    unsigned log_level() const { return level; }
    iteration_type iteration() const { return my_iteration; }
    iteration_type& iteration_ref() { return my_iteration; }

  protected:
    const LinearOperator& A;
    iteration_type my_iteration;
    unsigned level;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class cg_solver : public base_solver<cg_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<cg_solver, LinearOperator> base;

  public:
    explicit cg_solver(const LinearOperator& A) : base(A), L(A) {}
    cg_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return cg(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicgstab_solver
    : public base_solver<bicgstab_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_solver, LinearOperator> base;

  public:
    explicit bicgstab_solver(const LinearOperator& A) : base(A), L(A) {}
    bicgstab_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return bicgstab(this->A, x, b, L, iter);
    }

  private:
    Preconditioner L;
};

// right preconditioner ignored, with a warning (cgs.hpp:64-95)
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class cgs_solver
    : public base_solver<cgs_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<cgs_solver, LinearOperator> base;

  public:
    explicit cgs_solver(const LinearOperator& A) : base(A), L(A) { warn(); }
    cgs_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L) { warn(); }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return cgs(this->A, x, b, L, iter);
    }

  private:
    static void warn() {
        if (!std::is_same<RightPreconditioner, pc::identity<LinearOperator>>::value)
            std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    Preconditioner L;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class tfqmr_solver
    : public base_solver<tfqmr_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<tfqmr_solver, LinearOperator> base;

  public:
    explicit tfqmr_solver(const LinearOperator& A) : base(A), L(A), R(A) {}
    tfqmr_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L), R(A) {}
    tfqmr_solver(const LinearOperator& A, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return tfqmr(this->A, x, b, L, R, iter);
    }

  private:
    Preconditioner L;
    RightPreconditioner R;
};

template <typename LinearOperator, typename Preconditioner, typename RightPreconditioner>
class bicgstab_ell_solver
    : public base_solver<bicgstab_ell_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_ell_solver, LinearOperator> base;

  public:
    explicit bicgstab_ell_solver(const LinearOperator& A, std::size_t l = 8) : base(A), l(l), L(A), R(A) {}
    bicgstab_ell_solver(const LinearOperator& A, std::size_t l, const Preconditioner& L)
        : base(A), l(l), L(L), R(A) {}
    bicgstab_ell_solver(const LinearOperator& A, std::size_t l, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), l(l), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return bicgstab_ell(this->A, x, b, L, R, iter, l);
    }

  private:
    std::size_t l;
    Preconditioner L;
    RightPreconditioner R;
};

template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class gmres_solver
    : public base_solver<gmres_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<gmres_solver, LinearOperator> base;

  public:
    explicit gmres_solver(const LinearOperator& A, std::size_t restart = 8) : base(A), restart(restart), L(A), R(A) {}
    gmres_solver(const LinearOperator& A, std::size_t restart, const Preconditioner& L)
        : base(A), restart(restart), L(L), R(A) {}
    gmres_solver(const LinearOperator& A, std::size_t restart, const Preconditioner& L, const RightPreconditioner& R)
        : base(A), restart(restart), L(L), R(R) {}
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return gmres(this->A, x, b, L, R, iter, restart);
    }

  private:
    std::size_t restart;
    Preconditioner L;
    RightPreconditioner R;
};

// right preconditioner ignored, with a warning (bicgstab_2.hpp:132-165)
template <typename LinearOperator, typename Preconditioner = pc::identity<LinearOperator>,
          typename RightPreconditioner = pc::identity<LinearOperator>>
class bicgstab_2_solver
    : public base_solver<bicgstab_2_solver<LinearOperator, Preconditioner, RightPreconditioner>, LinearOperator> {
    typedef base_solver<bicgstab_2_solver, LinearOperator> base;

  public:
    explicit bicgstab_2_solver(const LinearOperator& A) : base(A), L(A) { warn(); }
    bicgstab_2_solver(const LinearOperator& A, const Preconditioner& L) : base(A), L(L) { warn(); }
    template <typename HilbertSpaceX, typename HilbertSpaceB, typename Iteration>
    int solve(HilbertSpaceX& x, const HilbertSpaceB& b, Iteration& iter) const {
        return bicgstab_2(this->A, x, b, L, iter);
    }

  private:
    static void warn() {
        if (!std::is_same<RightPreconditioner, pc::identity<LinearOperator>>::value)
            std::cerr << "Right Preconditioner ignored!" << std::endl;
    }
    Preconditioner L;
};

}  // namespace itl

// ---------------------------------------------------------------- status -> exception
namespace mtl {
namespace detail {
[[noreturn]] inline void throw_status(int st) {
    std::string msg = std::string(mtl4g_status_string(st)) + ": " + mtl4g_last_error();
    switch (st) {
    case MTL4G_ERR_NOT_SQUARE: throw matrix_not_square(msg.c_str());
    case MTL4G_ERR_SIZE: throw incompatible_size(msg.c_str());
    case MTL4G_ERR_MISSING_DIAG: throw missing_diagonal(msg.c_str());
    case MTL4G_ERR_BREAKDOWN: throw itl::unexpected_orthogonality(msg.c_str());
    case MTL4G_ERR_INVALID: throw logic_error(msg.c_str());
    default: throw device_error(msg.c_str());
    }
}
}  // namespace detail
}  // namespace mtl
