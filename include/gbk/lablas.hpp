// lablas.hpp — C++ drop-in front end over the gfx950 library (include/gbk.h).
//
// Mirrors the reference's public API (afanasyev-ilya/GB_Kun src/cpp_graphblas/,
// src/cblas_wrappers/): lablas::Matrix<T>, Vector<T>, Descriptor with the
// GrB_DESC_* presets, the functors, monoids and semirings of types.hpp, and
// the operations mxv / vxm / eWiseAdd / eWiseMult / apply / assign /
// assignScatter / reduce / extract with the NULL-accumulator overloads, plus
// the graph drivers.  Template signatures, argument order, defaults and error
// codes follow the reference; each call lowers to one extern "C" entry point.
//
// Semirings and operators are compile-time types here as in the reference;
// each carries the (monoid, multiplier) or operator identifiers the C ABI
// takes, so a call site that compiles against the reference compiles here.
// A Vector<T> / Matrix<T> owns HBM; host views (get_at, print, iteration)
// copy back.  Where the reference throws a `const char*`, so does this
// front end, with the same text.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../gbk.h"

#ifndef GBK_LABLAS_NO_GLOBAL_ENUMS
// Enumerations of src/cpp_graphblas/types.hpp:10-88, same values.
enum Storage { GrB_UNKNOWN, GrB_SPARSE, GrB_DENSE };
enum Desc_field { GrB_MASK, GrB_OUTPUT, GrB_INP0, GrB_INP1, GrB_MODE, GrB_TA, GrB_TB, GrB_NT, GrB_MXVMODE, GrB_TOL,
                  GrB_BACKEND, GrB_MXMMODE, GrB_NDESCFIELD };
enum Desc_value { GrB_SCMP = 0, GrB_REPLACE = 1, GrB_COMP = 2, GrB_TRAN = 3, GrB_STRUCTURE = 4, GrB_STR_COMP = 5,
                  GrB_DEFAULT = 6, GrB_FIXEDROW, GrB_PUSHPULL = 10, GrB_PUSHONLY = 11, GrB_PULLONLY = 12,
                  GrB_SEQUENTIAL = 13, GrB_CUDA = 14, GrB_8 = 8, GrB_16 = 16, GrB_32 = 32, GrB_64 = 64, GrB_128 = 128,
                  GrB_256 = 256, GrB_512 = 512, GrB_1024 = 1024, GrB_IJK = 20, GrB_IKJ = 21, GrB_IKJ_MASKED = 22,
                  GrB_IJK_DOUBLE_SORT = 23 };
typedef enum { GrB_SUCCESS = 0, GrB_NO_VALUE = 1, GrB_UNINITIALIZED_OBJECT = 2, GrB_INVALID_OBJECT = 3,
               GrB_NULL_POINTER = 4, GrB_INVALID_VALUE = 5, GrB_INVALID_INDEX = 6, GrB_DOMAIN_MISMATCH = 7,
               GrB_DIMENSION_MISMATCH = 8, GrB_OUTPUT_NOT_EMPTY = 9, GrB_OUT_OF_MEMORY = 10,
               GrB_INSUFFICIENT_SPACE = 11, GrB_INDEX_OUT_OF_BOUNDS = 12, GrB_PANIC = 13 } LA_Info;
// MatrixStorageFormat, src/helpers/cmd_parser/parser_options.h:15-24.
enum MatrixStorageFormat { CSR = 0, COO = 1, CSR_SEG = 3, LAV = 4, VECT_GROUP_CSR = 5, SELL_C = 6, SORTED_CSR = 7 };
typedef long long Index;
#endif

#define NULL_TYPE long int

namespace lablas {

// ------------------------------------------------------------- type map --
template <typename T> struct gbk_type_of;
template <> struct gbk_type_of<float> { static constexpr int value = GBK_FP32; };
template <> struct gbk_type_of<double> { static constexpr int value = GBK_FP64; };
template <> struct gbk_type_of<int> { static constexpr int value = GBK_INT32; };
template <> struct gbk_type_of<long long> { static constexpr int value = GBK_INT64; };
template <> struct gbk_type_of<long> { static constexpr int value = GBK_INT64; };
template <> struct gbk_type_of<bool> { static constexpr int value = GBK_BOOL; };

namespace detail {
inline LA_Info code(int rc) {
    if (rc != GBK_SUCCESS) {
        std::string msg = gbk_last_error();
        // the reference throws these (spmv.h:130, spmspv.h:201-204, build.hpp:110)
        if (msg.find("unsupported matrix storage format") != std::string::npos) throw "unsupported matrix storage format in VSpM";
    }
    return (LA_Info)rc;
}
inline gbk_desc desc_of(int mask) { gbk_desc d; d.mask = mask; d.output = GBK_DEFAULT; return d; }
}  // namespace detail

// -------------------------------------------- functors (types.hpp:93-209) --
// Each binary functor computes on the host like the reference's and names
// the C ABI operator it lowers to.
template <typename D1, typename D2 = D1> struct Identity {
    inline D2 operator()(const D1 input) const { return input; }
};
#define GBK_BINOP(NAME, ID, EXPR)                                                                          \
    template <typename T_in1, typename T_in2 = T_in1, typename T_out = T_in1> struct NAME {                 \
        static constexpr int gbk_op = ID;                                                                  \
        inline T_out operator()(const T_in1 lhs, const T_in2 rhs) const { return (T_out)(EXPR); }          \
    };
GBK_BINOP(plus, GBK_ACC_PLUS, lhs + rhs)
GBK_BINOP(minus, GBK_ACC_MINUS, lhs - rhs)
GBK_BINOP(multiplies, GBK_ACC_TIMES, lhs * rhs)
GBK_BINOP(first, GBK_ACC_FIRST, lhs)
GBK_BINOP(second, GBK_ACC_NONE, rhs)
GBK_BINOP(minimum, GBK_ACC_MIN, (rhs < lhs) ? rhs : lhs)
GBK_BINOP(maximum, GBK_ACC_MAX, (rhs > lhs) ? rhs : lhs)
GBK_BINOP(not_equal_to, GBK_ACC_NE, lhs != rhs)
#undef GBK_BINOP
template <typename T_in1 = bool, typename T_in2 = bool, typename T_out = bool> struct logical_and {
    static constexpr int gbk_op = GBK_ACC_LAND;
    inline T_out operator()(const T_in1 l, const T_in2 r) const { return l && r; }
};
template <typename T_in1 = bool, typename T_in2 = bool, typename T_out = bool> struct logical_or {
    static constexpr int gbk_op = GBK_ACC_LOR;
    inline T_out operator()(const T_in1 l, const T_in2 r) const { return l || r; }
};
template <typename T_in1, typename T_in2 = T_in1, typename T_out = bool> struct less {
    static constexpr int gbk_op = GBK_ACC_LESS;
    inline T_out operator()(T_in1 l, T_in2 r) const { return l < r; }
};
template <typename T_in1, typename T_in2 = T_in1, typename T_out = T_in1> struct GrB_ONEB_T {
    inline T_out operator()(const T_in1, const T_in2) const { return (T_out)1; }
};
template <typename T> struct div {
    static constexpr int gbk_op = GBK_ACC_DIV;
    inline T operator()(const T l, const T r) const { return l / r; }
};
template <typename T_out> struct abs {
    static constexpr int gbk_unary = 1;
    inline T_out operator()(const T_out a) const { return std::abs(a); }
};

// ------------------------------------------- monoids (types.hpp:211-232) --
#define GBK_MONOID(NAME, ADD, IDENT)                                                         \
    template <typename T_out> struct NAME {                                                  \
        static constexpr int gbk_add = ADD;                                                  \
        inline T_out identity() const { return static_cast<T_out>(IDENT); }                  \
    };
GBK_MONOID(PlusMonoid, GBK_ADD_PLUS, 0)
GBK_MONOID(LogicalOrMonoid, GBK_ADD_LOR, false)
GBK_MONOID(MinimumMonoid, GBK_ADD_MIN, std::numeric_limits<T_out>::max())
GBK_MONOID(FirstWinsMonoid, GBK_ADD_FIRST, 0)
GBK_MONOID(SecondWinsMonoid, GBK_ADD_SECOND, 0)
GBK_MONOID(CustomLessMonoid, GBK_ADD_LESS, std::numeric_limits<T_out>::max())
GBK_MONOID(FirstMin, GBK_ADD_FIRSTMIN, 0)
#undef GBK_MONOID

// ----------------------------------------- semirings (types.hpp:234-284) --
#define GBK_SEMIRING(NAME, ADD, MUL, IDENT, MONOP)                                            \
    template <typename T_in1, typename T_in2 = T_in1, typename T_out = T_in1> struct NAME {    \
        typedef T_out result_type;                                                            \
        typedef T_out T_out_type;                                                             \
        static constexpr int gbk_add = ADD;                                                   \
        static constexpr int gbk_mul = MUL;                                                   \
        static constexpr int gbk_monoid_op = MONOP;                                           \
        inline T_out identity() const { return static_cast<T_out>(IDENT); }                   \
    };
GBK_SEMIRING(PlusMultipliesSemiring, GBK_ADD_PLUS, GBK_MUL_TIMES, 0, GBK_ACC_PLUS)
GBK_SEMIRING(LogicalOrAndSemiring, GBK_ADD_LOR, GBK_MUL_LAND, false, GBK_ACC_LOR)
GBK_SEMIRING(FirstWinsSemiring, GBK_ADD_FIRST, GBK_MUL_TIMES, 0, GBK_ACC_FIRST)
GBK_SEMIRING(FirstMinSemiring, GBK_ADD_FIRSTMIN, GBK_MUL_TIMES, 0, GBK_ACC_MIN)
GBK_SEMIRING(PlusSecondSemiring, GBK_ADD_PLUS, GBK_MUL_SECOND, 0, GBK_ACC_PLUS)
GBK_SEMIRING(StructuralBool, GBK_ADD_LOR, GBK_MUL_ONEB, false, GBK_ACC_LOR)
GBK_SEMIRING(MinimumPlusSemiring, GBK_ADD_MIN, GBK_MUL_PLUS, std::numeric_limits<T_out>::max(), GBK_ACC_MIN)
GBK_SEMIRING(CustomLessPlusSemiring, GBK_ADD_LESS, GBK_MUL_PLUS, std::numeric_limits<T_out>::max(), GBK_ACC_LESS)
GBK_SEMIRING(MinimumSelectSecondSemiring, GBK_ADD_MIN, GBK_MUL_SECOND, std::numeric_limits<T_out>::max(), GBK_ACC_MIN)
GBK_SEMIRING(MinimumNotEqualToSemiring, GBK_ADD_MIN, GBK_MUL_NE, std::numeric_limits<T_out>::max(), GBK_ACC_MIN)
GBK_SEMIRING(PlusOneSemiring, GBK_ADD_PLUS, GBK_MUL_ONEB, 0, GBK_ACC_PLUS)
#undef GBK_SEMIRING

namespace detail {
// element-wise operator of an eWise call: a semiring's add or mul, or a functor
template <typename S, typename = void> struct has_monoid : std::false_type {};
template <typename S> struct has_monoid<S, std::void_t<decltype(S::gbk_monoid_op)>> : std::true_type {};
template <typename S, typename = void> struct has_mul : std::false_type {};
template <typename S> struct has_mul<S, std::void_t<decltype(S::gbk_mul)>> : std::true_type {};
template <typename Op> constexpr int add_binop() {
    if constexpr (has_monoid<Op>::value) return Op::gbk_monoid_op;
    else return Op::gbk_op;
}
template <typename Op> constexpr int mul_binop() {
    if constexpr (has_mul<Op>::value) {
        constexpr int m = Op::gbk_mul;
        return m == GBK_MUL_TIMES ? GBK_ACC_TIMES : m == GBK_MUL_LAND ? GBK_ACC_LAND : m == GBK_MUL_FIRST ? GBK_ACC_FIRST
             : m == GBK_MUL_SECOND ? GBK_ACC_NONE : m == GBK_MUL_PLUS ? GBK_ACC_PLUS : m == GBK_MUL_NE ? GBK_ACC_NE
             : m == GBK_MUL_MINUS ? GBK_ACC_MINUS : GBK_ACC_DIV;
    } else {
        return Op::gbk_op;
    }
}
template <typename Acc> constexpr int accum_id() { return Acc::gbk_op; }
}  // namespace detail

// ------------------------------------------------ Descriptor (descriptor.hpp) --
class Descriptor {
public:
    Descriptor() { for (auto& v : desc_) v = GrB_DEFAULT; }
    explicit Descriptor(const std::vector<std::pair<Desc_field, Desc_value>>& init) : Descriptor() {
        for (auto& p : init) desc_[p.first] = p.second;
    }
    LA_Info set(Desc_field f, Desc_value v) { desc_[f] = v; return GrB_SUCCESS; }
    LA_Info get(Desc_field f, Desc_value* v) const { *v = desc_[f]; return GrB_SUCCESS; }
    Desc_value get(Desc_field f) const { return desc_[f]; }
    // descriptor.h:31-45: a non-default field returns to DEFAULT; a default
    // one becomes GrB_SCMP (mask / output / inp0 by field number) or GrB_TRAN.
    LA_Info toggle(Desc_field field) {
        int f = (int)field;
        if (f < 4) {
            if (desc_[field] != GrB_DEFAULT) desc_[field] = GrB_DEFAULT;
            else desc_[field] = (f > 2) ? GrB_TRAN : (Desc_value)f;
        }
        return GrB_SUCCESS;
    }
    gbk_desc c_desc() const {
        gbk_desc d;
        d.mask = desc_[GrB_MASK];
        d.output = desc_[GrB_OUTPUT];
        return d;
    }
    Descriptor* get_descriptor() { return this; }

private:
    Desc_value desc_[GrB_NDESCFIELD];
};

inline Descriptor GrB_DESC_RSC({{GrB_OUTPUT, GrB_REPLACE}, {GrB_MASK, GrB_STR_COMP}});
inline Descriptor GrB_DESC_C({{GrB_MASK, GrB_COMP}});
inline Descriptor GrB_DESC_S({{GrB_MASK, GrB_STRUCTURE}});
inline Descriptor GrB_DESC_SC({{GrB_MASK, GrB_STR_COMP}});

// ---------------------------------------------------- Vector (vector.hpp) --
template <typename T> class Vector {
public:
    using ValueType = T;
    Vector() : Vector(0) {}
    explicit Vector(Index nsize) : n_(nsize) {
        if (gbk_vector_new(&h_, (gbk_type)gbk_type_of<T>::value, nsize) != GBK_SUCCESS) throw std::runtime_error(gbk_last_error());
    }
    ~Vector() { gbk_vector_free(h_); }
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;

    LA_Info build(const std::vector<Index>* indices, const std::vector<T>* values, Index nvals) {
        if (indices == nullptr || values == nullptr) return GrB_NULL_POINTER;
        if (nvals == 0) return GrB_INVALID_VALUE;
        std::vector<int64_t> ids(indices->begin(), indices->begin() + nvals);
        return detail::code(gbk_vector_build_sparse(h_, ids.data(), values->data(), nvals));
    }
    LA_Info build(const std::vector<T>* values, Index nvals) {
        if (values == nullptr) return GrB_NULL_POINTER;
        return build(values->data(), nvals);
    }
    LA_Info build(const T* values, Index nvals) {
        if (values == nullptr) return GrB_NULL_POINTER;
        if (nvals == 0) return GrB_INVALID_VALUE;
        return detail::code(gbk_vector_build_dense(h_, values, nvals));
    }
    LA_Info fill(T val) { return detail::code(gbk_vector_fill(h_, (double)val)); }
    LA_Info clear() { return detail::code(gbk_vector_clear(h_)); }
    LA_Info set_element(T val, Index pos) { return detail::code(gbk_vector_set_element(h_, (double)val, pos)); }
    LA_Info get_nvals(Index* nv) const {
        int64_t v = 0;
        LA_Info rc = detail::code(gbk_vector_nvals(h_, &v));
        *nv = v;
        return rc;
    }
    Index nvals() const { Index v = 0; get_nvals(&v); return v; }
    Index size() const { return n_; }
    void swap(Vector* other) { gbk_vector_swap(h_, other->h_); std::swap(n_, other->n_); }
    LA_Info fillAscending(Index nvals) { return detail::code(gbk_vector_fill_ascending(h_, nvals)); }
    LA_Info dup(const Vector<T>* rhs) { return detail::code(gbk_vector_dup(h_, rhs->h_)); }
    // host view (copies the dense form back)
    std::vector<T> to_host() const {
        std::vector<T> out(n_);
        if (n_) detail::code(gbk_vector_to_host(h_, out.data()));
        return out;
    }
    T get_at(Index i) const {
        if (i < 0 || i >= n_) throw "Error: out of range in backend::vector";
        return to_host()[i];
    }
    bool is_dense() const { int s = 0; gbk_vector_storage(h_, &s); return s == GBK_STORAGE_DENSE; }
    bool is_sparse() const { return !is_dense(); }
    void print() const {
        auto v = to_host();
        for (Index i = 0; i < n_; ++i) printf("[%lld]:%g ", (long long)i, (double)v[i]);
        printf("\n");
    }
    void set_name(const std::string& s) { name_ = s; }
    gbk_vector handle() const { return h_; }

private:
    gbk_vector h_ = nullptr;
    Index n_ = 0;
    std::string name_;
};

// ---------------------------------------------------- Matrix (matrix.hpp) --
template <typename T> class Matrix {
public:
    Matrix() {
        if (gbk_matrix_new(&h_, (gbk_type)gbk_type_of<T>::value) != GBK_SUCCESS) throw std::runtime_error(gbk_last_error());
    }
    Matrix(Index, Index) : Matrix() {}
    ~Matrix() { gbk_matrix_free(h_); }
    Matrix(const Matrix&) = delete;
    Matrix& operator=(const Matrix&) = delete;

    LA_Info set_preferred_matrix_format(const MatrixStorageFormat f) {
        if (f != CSR && f != CSR_SEG && f != LAV && f != SELL_C) throw "Error: unsupported format in Matrix<T>::build";
        return detail::code(gbk_matrix_set_format(h_, (int)f, 0, 0));
    }
    template <typename BinaryOpT>
    LA_Info build(const std::vector<Index>* rows, const std::vector<Index>* cols, const std::vector<T>* values,
                  Index nvals, BinaryOpT, char* = nullptr) {
        if (rows == nullptr || cols == nullptr || values == nullptr) return GrB_NULL_POINTER;
        if (rows->empty() && cols->empty() && values->empty()) return GrB_NO_VALUE;
        if (rows->size() != cols->size()) return GrB_DIMENSION_MISMATCH;
        (void)nvals;  // the reference builds from the vectors' sizes too (matrix.hpp:87-89)
        std::vector<int64_t> r(rows->begin(), rows->end()), c(cols->begin(), cols->end());
        return detail::code(gbk_matrix_build_coo(h_, r.data(), c.data(), values->data(), (int64_t)r.size()));
    }
    LA_Info build_from_csr_arrays(const Index* row_ptrs, const Index* col_ids, const T* values, Index nrows, Index nvals) {
        if (row_ptrs == nullptr || col_ids == nullptr || values == nullptr) return GrB_NULL_POINTER;
        return detail::code(gbk_matrix_build_csr(h_, (const int64_t*)row_ptrs, (const int64_t*)col_ids, values, nrows, nvals));
    }
    // matrix.hpp:107-110 / read_file.hpp:556-615: .mtx or .mtxbin by suffix;
    // the reference throws on an unreadable file, so does this
    void init_from_mtx(const std::string& mtx_name) {
        if (gbk_matrix_read_mtx(h_, mtx_name.c_str()) != GBK_SUCCESS) throw gbk_last_error();
    }
    LA_Info get_nrows(Index* n) const { int64_t v; LA_Info rc = detail::code(gbk_matrix_nrows(h_, &v)); *n = v; return rc; }
    LA_Info get_ncols(Index* n) const { return get_nrows(n); }
    LA_Info get_nvals(Index* n) const { int64_t v; LA_Info rc = detail::code(gbk_matrix_nvals(h_, &v)); *n = v; return rc; }
    Index nrows() const { Index v = 0; get_nrows(&v); return v; }
    Index ncols() const { return nrows(); }
    Index get_nnz() const { Index v = 0; get_nvals(&v); return v; }
    std::vector<Index> get_rowdegrees() const { return degrees(0); }
    std::vector<Index> get_coldegrees() const { return degrees(1); }
    gbk_matrix handle() const { return h_; }

private:
    std::vector<Index> degrees(int t) const {
        std::vector<int64_t> d(nrows());
        gbk_matrix_degrees(h_, t, d.data());
        return std::vector<Index>(d.begin(), d.end());
    }
    gbk_matrix h_ = nullptr;
};

// ----------------------------------------------- operations (operations.hpp) --
namespace detail {
template <typename V> gbk_vector h(const V* v) { return v ? v->handle() : nullptr; }
template <typename SemiringT> gbk_semiring sr() {
    gbk_semiring s;
    s.add = SemiringT::gbk_add;
    s.mul = SemiringT::gbk_mul;
    return s;
}
}  // namespace detail

template <typename W, typename M, typename a, typename U, typename BinaryOpTAccum, typename SemiringT>
LA_Info mxv(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, SemiringT, const Matrix<a>* A, const Vector<U>* u,
            Descriptor* desc) {
    if (w == nullptr || u == nullptr || A == nullptr || desc == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc->c_desc();
    return detail::code(gbk_mxv(w->handle(), detail::h(mask), detail::accum_id<BinaryOpTAccum>(), detail::sr<SemiringT>(),
                                A->handle(), u->handle(), &d));
}
template <typename W, typename M, typename a, typename U, typename SemiringT>
LA_Info mxv(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, SemiringT op, const Matrix<a>* A, const Vector<U>* u,
            Descriptor* desc) {
    return mxv(w, mask, second<U, W, U>(), op, A, u, desc);
}

template <typename W, typename M, typename a, typename U, typename BinaryOpTAccum, typename SemiringT>
LA_Info vxm(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, SemiringT, const Vector<U>* u, const Matrix<a>* A,
            Descriptor* desc) {
    if (w == nullptr || u == nullptr || A == nullptr || desc == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc->c_desc();
    return detail::code(gbk_vxm(w->handle(), detail::h(mask), detail::accum_id<BinaryOpTAccum>(), detail::sr<SemiringT>(),
                                u->handle(), A->handle(), &d));
}
template <typename W, typename M, typename a, typename U, typename SemiringT>
LA_Info vxm(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, SemiringT op, const Vector<U>* u, const Matrix<a>* A,
            Descriptor* desc) {
    return vxm(w, mask, second<U, W, U>(), op, u, A, desc);
}

// mxm (operations.hpp:411-437): with a mask C gets the mask's entries; a NULL
// mask gives the unmasked product (entries whose fold is 0 dropped).
template <typename c, typename m, typename a, typename b, typename BinaryOpTAccum, typename SemiringT>
LA_Info mxm(Matrix<c>* C, const Matrix<m>* mask, BinaryOpTAccum, SemiringT, const Matrix<a>* A, const Matrix<b>* B,
            Descriptor* desc) {
    if (C == nullptr || A == nullptr || B == nullptr || desc == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc->c_desc();
    const int rc = gbk_mxm(C->handle(), mask ? mask->handle() : nullptr, detail::accum_id<BinaryOpTAccum>(),
                           detail::sr<SemiringT>(), A->handle(), B->handle(), &d);
    if (rc != GBK_SUCCESS && (d.mask == GBK_COMP || d.mask == GBK_STR_COMP))
        throw "Error: complementary mask is not supported yet";
    return detail::code(rc);
}
template <typename c, typename m, typename a, typename b, typename SemiringT>
LA_Info mxm(Matrix<c>* C, const Matrix<m>* mask, NULL_TYPE, SemiringT op, const Matrix<a>* A, const Matrix<b>* B,
            Descriptor* desc) {
    return mxm(C, mask, second<c, a, b>(), op, A, B, desc);
}

template <typename W, typename M, typename U, typename V, typename BinaryOpTAccum, typename BinaryOpT>
LA_Info eWiseAdd(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, BinaryOpT, const Vector<U>* u, const Vector<V>* v,
                 Descriptor* desc) {
    if (w == nullptr || u == nullptr || v == nullptr) return GrB_UNINITIALIZED_OBJECT;
    if (w->size() != u->size() || w->size() != v->size()) return GrB_DIMENSION_MISMATCH;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_ewise_add(w->handle(), detail::h(mask), GBK_ACC_NONE, detail::add_binop<BinaryOpT>(),
                                      u->handle(), v->handle(), &d));
}
template <typename W, typename M, typename U, typename V, typename BinaryOpT>
LA_Info eWiseAdd(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, BinaryOpT op, const Vector<U>* u, const Vector<V>* v,
                 Descriptor* desc) {
    return eWiseAdd(w, mask, second<U, W, U>(), op, u, v, desc);
}

template <typename W, typename M, typename U, typename V, typename BinaryOpTAccum, typename BinaryOpT>
LA_Info eWiseMult(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, BinaryOpT, const Vector<U>* u,
                  const Vector<V>* v, Descriptor* desc) {
    if (w == nullptr || u == nullptr || v == nullptr) return GrB_UNINITIALIZED_OBJECT;
    if (w->size() != u->size() || w->size() != v->size()) return GrB_DIMENSION_MISMATCH;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_ewise_mult(w->handle(), detail::h(mask), GBK_ACC_NONE, detail::mul_binop<BinaryOpT>(),
                                       u->handle(), v->handle(), &d));
}
template <typename W, typename M, typename U, typename V, typename BinaryOpT>
LA_Info eWiseMult(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, BinaryOpT op, const Vector<U>* u, const Vector<V>* v,
                  Descriptor* desc) {
    return eWiseMult(w, mask, second<U, W, U>(), op, u, v, desc);
}

// w[i] = op(u[i], val) (operations.hpp apply with a trailing scalar)
template <typename W, typename M, typename U, typename T, typename BinaryOpTAccum, typename BinaryOpT>
LA_Info apply(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, BinaryOpT, const Vector<U>* u, const T val,
              Descriptor* desc) {
    if (w == nullptr || u == nullptr) return GrB_UNINITIALIZED_OBJECT;
    constexpr int op = BinaryOpT::gbk_op;
    int unary = op == GBK_ACC_DIV ? 2 : op == GBK_ACC_TIMES ? 3 : op == GBK_ACC_PLUS ? 6 : op == GBK_ACC_MINUS ? 7 : 0;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_apply(w->handle(), detail::h(mask), GBK_ACC_NONE, unary, (double)val, u->handle(), &d));
}
// w[i] = op(u[i]) (unary)
template <typename W, typename M, typename U, typename BinaryOpTAccum, typename UnaryOpT>
LA_Info apply(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, UnaryOpT, const Vector<U>* u, Descriptor* desc) {
    if (w == nullptr || u == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_apply(w->handle(), detail::h(mask), GBK_ACC_NONE, UnaryOpT::gbk_unary, 0.0, u->handle(), &d));
}

// w<mask> = value over all indices (GrB_ALL), operations.h:83-112
template <typename W, typename M, typename U, typename BinaryOpTAccum>
LA_Info assign(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, U value, const Index* indices, const Index,
               Descriptor* desc) {
    if (w == nullptr) return GrB_UNINITIALIZED_OBJECT;
    if (indices != nullptr) return GrB_INVALID_VALUE;  // index lists: use assignScatter
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_assign_scalar(w->handle(), detail::h(mask), GBK_ACC_NONE, (double)value, &d));
}
template <typename W, typename M, typename U, typename BinaryOpTAccum>
LA_Info assign(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, Vector<U>* u, const Index* indices, const Index,
               Descriptor* desc) {
    if (w == nullptr || u == nullptr) return GrB_UNINITIALIZED_OBJECT;
    if (indices != nullptr) return GrB_INVALID_VALUE;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_assign_vector(w->handle(), detail::h(mask), detail::accum_id<BinaryOpTAccum>(), u->handle(), &d));
}
template <typename W, typename M, typename U>
LA_Info assign(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, Vector<U>* u, const Index* indices, const Index n,
               Descriptor* desc) {
    return assign(w, mask, second<U, W, U>(), u, indices, n, desc);
}

template <typename W, typename M, typename U, typename I, typename BinaryOpTAccum>
LA_Info assignScatter(Vector<W>* w, const Vector<M>* mask, BinaryOpTAccum, Vector<U>* u, const Vector<I>* indices,
                      const Index nindices, Descriptor* desc) {
    if (w == nullptr) return GrB_UNINITIALIZED_OBJECT;
    if (w->size() != u->size()) return GrB_DIMENSION_MISMATCH;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_assign_scatter(w->handle(), detail::h(mask), detail::accum_id<BinaryOpTAccum>(), u->handle(),
                                           indices->handle(), nindices, &d));
}
template <typename W, typename M, typename U, typename I>
LA_Info assignScatter(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, Vector<U>* u, const Vector<I>* indices,
                      const Index nindices, Descriptor* desc) {
    return assignScatter(w, mask, second<U, W, U>(), u, indices, nindices, desc);
}

// *val = accum(*val, reduce(u)), operations.h:345-376
template <typename T, typename U, typename BinaryOpTAccum, typename MonoidT>
LA_Info reduce(T* val, BinaryOpTAccum, MonoidT, const Vector<U>* u, Descriptor* desc) {
    if (u == nullptr || val == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    if constexpr (std::is_integral<T>::value && sizeof(T) == 8 && !std::is_floating_point<U>::value) {
        int64_t r = (int64_t)*val;
        LA_Info rc = detail::code(gbk_reduce_i64(&r, detail::accum_id<BinaryOpTAccum>(), MonoidT::gbk_add, u->handle(), &d));
        *val = (T)r;
        return rc;
    } else {
        double r = (double)*val;
        LA_Info rc = detail::code(gbk_reduce(&r, detail::accum_id<BinaryOpTAccum>(), MonoidT::gbk_add, u->handle(), &d));
        *val = (T)r;
        return rc;
    }
}
template <typename T, typename U, typename MonoidT>
LA_Info reduce(T* val, NULL_TYPE, MonoidT op, const Vector<U>* u, Descriptor* desc) {
    return reduce(val, second<T, T, T>(), op, u, desc);
}

// w[i] = accum(w[i], u[indices[i]]), operations.h:400-430
template <typename W, typename M, typename U, typename I, typename BinaryOpT>
LA_Info extract(Vector<W>* w, const Vector<M>* mask, BinaryOpT, const Vector<U>* u, const Vector<I>* indices,
                Descriptor* desc) {
    if (w == nullptr || u == nullptr) return GrB_UNINITIALIZED_OBJECT;
    gbk_desc d = desc ? desc->c_desc() : detail::desc_of(GBK_DEFAULT);
    return detail::code(gbk_extract(w->handle(), detail::h(mask), detail::accum_id<BinaryOpT>(), u->handle(),
                                    detail::h(indices), &d));
}
template <typename W, typename M, typename U, typename I>
LA_Info extract(Vector<W>* w, const Vector<M>* mask, NULL_TYPE, const Vector<U>* u, const Vector<I>* indices,
                Descriptor* desc) {
    return extract(w, mask, second<U, W, U>(), u, indices, desc);
}

// ------------------------------------------------------- graph drivers --
namespace algorithm {

// bfs_blast (algorithms/bfs/bfs.hpp:94-142): levels 1-based, 0 unreached.
template <typename T> void bfs_blast(Vector<T>* v, const Matrix<T>* A, Index s, Descriptor* desc) {
    (void)desc;
    Index n = A->nrows();
    std::vector<int32_t> lv(n);
    detail::code(gbk_bfs(A->handle(), s, lv.data(), nullptr));
    std::vector<T> out(lv.begin(), lv.end());
    v->build(out.data(), n);
}

// sssp_bellman_ford_blast (algorithms/sssp/sssp_blast.hpp:8-72) for float
// weights, and the same over int weights (INT_MAX unreachable).
template <typename T> void sssp_bellman_ford_blast(Vector<T>* v, const Matrix<T>* A, Index s, Descriptor* desc) {
    (void)desc;
    Index n = A->nrows();
    std::vector<T> d(n);
    detail::code(gbk_sssp(A->handle(), s, d.data(), nullptr));
    v->build(d.data(), n);
}

// cc (algorithms/cc/cc.hpp:12-119): minimum vertex id per component.
template <typename T> void cc(Vector<T>* v, const Matrix<T>* A, int seed, Descriptor* desc) {
    (void)seed;
    (void)desc;
    Index n = A->nrows();
    std::vector<int32_t> lab(n);
    detail::code(gbk_cc(A->handle(), lab.data(), nullptr));
    std::vector<T> out(lab.begin(), lab.end());
    v->build(out.data(), n);
}

// LAGraph_page_rank_sinks (algorithms/pr/pr.hpp:20-153); tol > 0 stops at
// L1 change < tol (the reference always runs itermax iterations).
inline void page_rank_sinks(Vector<float>* centrality, const Matrix<float>* A, int* iters, int itermax = 100,
                            double damping = 0.85, double tol = 0.0) {
    Index n = A->nrows();
    std::vector<float> r(n);
    detail::code(gbk_pagerank(A->handle(), r.data(), itermax, damping, tol, iters));
    centrality->build(r.data(), n);
}

// LAGr_TriangleCount (algorithms/tc/tc.hpp:80-291) on a symmetric matrix;
// method numbering as LAGraph_TriangleCount_Method (0 = Burkhardt).
template <typename T> LA_Info tricount(uint64_t* ntriangles, const Matrix<T>* A, int method = 0) {
    if (ntriangles == nullptr || A == nullptr) return GrB_NULL_POINTER;
    return detail::code(gbk_tricount(A->handle(), method, ntriangles));
}

}  // namespace algorithm
}  // namespace lablas

// cblas-style forwards (src/cblas_wrappers/operations.hpp:112-151)
#define GrB_ALL nullptr
#define GrB_NULL ((long int)0)
template <typename... Args> LA_Info GrB_mxv(Args... a) { return lablas::mxv(a...); }
template <typename... Args> LA_Info GrB_vxm(Args... a) { return lablas::vxm(a...); }
template <typename... Args> LA_Info GrB_mxm(Args... a) { return lablas::mxm(a...); }
