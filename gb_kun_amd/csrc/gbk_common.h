// Common definitions of the gfx950 library: error plumbing, element types and
// the semiring / accumulator arithmetic shared by host and device code.
//
// The arithmetic restates the reference functors in
// src/cpp_graphblas/types.hpp:93-284 for a single element type T:
//   monoids  plus / logical_or / minimum / first / second / less / FirstMin
//   mul ops  multiplies / logical_and / first / second / plus / GrB_ONEB_T /
//            not_equal_to / minus / div
// Boolean results are stored as T(1) / T(0), as the reference's functors do
// when they return bool into T_out.  Deliberate differences:
//   * integer `plus` used as a multiplier (min.plus) saturates instead of
//     wrapping, so INT_MAX + w stays INT_MAX (the reference wraps);
//   * integer division by zero returns 0 instead of trapping.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>

#include "../../include/gbk.h"

#define GBK_HD __host__ __device__ __forceinline__

// add(acc, mul(a, x)) is two roundings in the reference (a separate multiply
// and add); keep it so on the device instead of fusing it into an FMA, so a
// kernel that folds a row in storage order reproduces the reference's fp32
// bits exactly.
#pragma clang fp contract(off)

namespace gbk {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define GBK_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::gbk::fail(GBK_PANIC, "%s failed: %s (%s:%d)", #call,                 \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define GBK_TRY(call)                       \
    do {                                    \
        int rc_ = (call);                   \
        if (rc_ != GBK_SUCCESS) return rc_; \
    } while (0)

#define GBK_LAUNCH_CHECK()                                                                 \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return ::gbk::fail(GBK_PANIC, "kernel launch failed: %s (%s:%d)",             \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

// ----------------------------------------------------------------- types --
typedef uint8_t boolean;  // GBK_BOOL storage

template <typename T> struct Lim;
template <> struct Lim<float> {
    static GBK_HD float max() { return FLT_MAX; }
    static GBK_HD float lowest() { return -FLT_MAX; }
};
template <> struct Lim<double> {
    static GBK_HD double max() { return DBL_MAX; }
    static GBK_HD double lowest() { return -DBL_MAX; }
};
template <> struct Lim<int32_t> {
    static GBK_HD int32_t max() { return INT32_MAX; }
    static GBK_HD int32_t lowest() { return INT32_MIN; }
};
template <> struct Lim<int64_t> {
    static GBK_HD int64_t max() { return INT64_MAX; }
    static GBK_HD int64_t lowest() { return INT64_MIN; }
};
template <> struct Lim<boolean> {
    static GBK_HD boolean max() { return 1; }
    static GBK_HD boolean lowest() { return 0; }
};

inline size_t type_size(int t) {
    switch (t) {
        case GBK_FP32: return 4;
        case GBK_FP64: return 8;
        case GBK_INT32: return 4;
        case GBK_INT64: return 8;
        case GBK_BOOL: return 1;
    }
    return 0;
}
inline bool valid_type(int t) { return t >= GBK_FP32 && t <= GBK_BOOL; }

template <typename T> GBK_HD T tbool(bool b) { return b ? T(1) : T(0); }
template <typename T> GBK_HD bool truth(T v) { return v != T(0); }
// C++ conversion U -> W, with bool targets taking (u != 0) as a C++ bool would.
template <typename W, typename U> GBK_HD W cvt(U u) {
    if constexpr (std::is_same<W, uint8_t>::value && !std::is_same<U, uint8_t>::value) return u != U(0) ? 1 : 0;
    else return (W)u;
}

template <typename T> constexpr bool is_int_v = std::is_integral<T>::value && !std::is_same<T, boolean>::value;
template <typename T> constexpr bool is_bool_v = std::is_same<T, boolean>::value;

// Wrapping integer add / sub / mul (the reference's plain int arithmetic,
// defined here through unsigned types).
template <typename T> GBK_HD T wrap_add(T a, T b) {
    if constexpr (is_int_v<T>) {
        typedef typename std::make_unsigned<T>::type U;
        return (T)((U)a + (U)b);
    } else if constexpr (is_bool_v<T>) {
        return tbool<T>(a || b);
    } else {
        return a + b;
    }
}
template <typename T> GBK_HD T wrap_sub(T a, T b) {
    if constexpr (is_int_v<T>) {
        typedef typename std::make_unsigned<T>::type U;
        return (T)((U)a - (U)b);
    } else if constexpr (is_bool_v<T>) {
        return tbool<T>(a != b);
    } else {
        return a - b;
    }
}
template <typename T> GBK_HD T wrap_mul(T a, T b) {
    if constexpr (is_int_v<T>) {
        typedef typename std::make_unsigned<T>::type U;
        return (T)((U)a * (U)b);
    } else if constexpr (is_bool_v<T>) {
        return tbool<T>(a && b);
    } else {
        return a * b;
    }
}
template <typename T> GBK_HD T sat_add(T a, T b) {
    if constexpr (std::is_same<T, int32_t>::value) {
        long long s = (long long)a + (long long)b;
        return s > INT32_MAX ? INT32_MAX : (s < INT32_MIN ? INT32_MIN : (T)s);
    } else if constexpr (std::is_same<T, int64_t>::value) {
        T r;
        if (__builtin_add_overflow(a, b, &r)) return a > 0 ? INT64_MAX : INT64_MIN;
        return r;
    } else {
        return wrap_add(a, b);
    }
}
template <typename T> GBK_HD T safe_div(T a, T b) {
    if constexpr (is_int_v<T>) {
        if (b == 0) return 0;
        if (std::is_same<T, int32_t>::value && b == -1) return wrap_sub<T>(0, a);
        if (std::is_same<T, int64_t>::value && b == -1) return wrap_sub<T>(0, a);
        return a / b;
    } else if constexpr (is_bool_v<T>) {
        return b ? a : 0;
    } else {
        return a / b;
    }
}

// ------------------------------------------------------------- semirings --
template <int ADD, typename T> GBK_HD T add_identity() {
    if constexpr (ADD == GBK_ADD_MIN || ADD == GBK_ADD_LESS) return Lim<T>::max();
    else if constexpr (ADD == GBK_ADD_MAX) return Lim<T>::lowest();
    else return T(0);
}

template <int ADD, typename T> GBK_HD T add_op(T l, T r) {
    if constexpr (ADD == GBK_ADD_PLUS) return wrap_add(l, r);
    else if constexpr (ADD == GBK_ADD_LOR) return tbool<T>(truth(l) || truth(r));
    else if constexpr (ADD == GBK_ADD_MIN || ADD == GBK_ADD_FIRSTMIN) return (r < l) ? r : l;
    else if constexpr (ADD == GBK_ADD_MAX) return (r > l) ? r : l;
    else if constexpr (ADD == GBK_ADD_FIRST) return l;
    else if constexpr (ADD == GBK_ADD_SECOND) return r;
    else return tbool<T>(l < r);  // GBK_ADD_LESS
}

template <int MUL, typename T> GBK_HD T mul_op(T a, T x) {
    if constexpr (MUL == GBK_MUL_TIMES) return wrap_mul(a, x);
    else if constexpr (MUL == GBK_MUL_LAND) return tbool<T>(truth(a) && truth(x));
    else if constexpr (MUL == GBK_MUL_FIRST) return a;
    else if constexpr (MUL == GBK_MUL_SECOND) return x;
    else if constexpr (MUL == GBK_MUL_PLUS) return sat_add(a, x);
    else if constexpr (MUL == GBK_MUL_ONEB) return T(1);
    else if constexpr (MUL == GBK_MUL_NE) return tbool<T>(a != x);
    else if constexpr (MUL == GBK_MUL_MINUS) return wrap_sub(a, x);
    else return safe_div(a, x);  // GBK_MUL_DIV
}

// Runtime binary operator (accumulators, element-wise ops).  `op` is uniform
// across a launch, so the switch is a scalar branch.
template <typename T> GBK_HD T binop(int op, T l, T r) {
    switch (op) {
        case GBK_ACC_NONE: return r;
        case GBK_ACC_FIRST: return l;
        case GBK_ACC_PLUS: return wrap_add(l, r);
        case GBK_ACC_MINUS: return wrap_sub(l, r);
        case GBK_ACC_TIMES: return wrap_mul(l, r);
        case GBK_ACC_DIV: return safe_div(l, r);
        case GBK_ACC_MIN: return (r < l) ? r : l;
        case GBK_ACC_MAX: return (r > l) ? r : l;
        case GBK_ACC_LOR: return tbool<T>(truth(l) || truth(r));
        case GBK_ACC_LAND: return tbool<T>(truth(l) && truth(r));
        case GBK_ACC_LESS: return tbool<T>(l < r);
        case GBK_ACC_NE: return tbool<T>(l != r);
    }
    return r;
}

// Host-side identity of a runtime monoid, as a double (used for fills).
template <typename T> T identity_rt(int add) {
    switch (add) {
        case GBK_ADD_MIN: case GBK_ADD_LESS: return Lim<T>::max();
        case GBK_ADD_MAX: return Lim<T>::lowest();
        default: return T(0);
    }
}

// ------------------------------------------------------------- dispatch --
template <int V> using IC = std::integral_constant<int, V>;

// Calls f(IC<ADD>, IC<MUL>) for the semirings the reference registers
// (types.hpp:274-284); other pairs are rejected.
template <typename F> int dispatch_semiring(int add, int mul, F&& f) {
#define GBK_SR(A, M) if (add == A && mul == M) return f(IC<A>{}, IC<M>{});
    GBK_SR(GBK_ADD_PLUS, GBK_MUL_TIMES)
    GBK_SR(GBK_ADD_LOR, GBK_MUL_LAND)
    GBK_SR(GBK_ADD_FIRST, GBK_MUL_TIMES)
    GBK_SR(GBK_ADD_FIRSTMIN, GBK_MUL_TIMES)
    GBK_SR(GBK_ADD_PLUS, GBK_MUL_SECOND)
    GBK_SR(GBK_ADD_LOR, GBK_MUL_ONEB)
    GBK_SR(GBK_ADD_MIN, GBK_MUL_PLUS)
    GBK_SR(GBK_ADD_LESS, GBK_MUL_PLUS)
    GBK_SR(GBK_ADD_MIN, GBK_MUL_SECOND)
    GBK_SR(GBK_ADD_MIN, GBK_MUL_NE)
    GBK_SR(GBK_ADD_PLUS, GBK_MUL_ONEB)
#undef GBK_SR
    return fail(GBK_INVALID_VALUE, "semiring (add=%d, mul=%d) is not one of the registered semirings", add, mul);
}

// Calls f(T{}) for the element type.
template <typename F> int dispatch_type(int type, F&& f) {
    switch (type) {
        case GBK_FP32: return f(float{});
        case GBK_FP64: return f(double{});
        case GBK_INT32: return f(int32_t{});
        case GBK_INT64: return f(int64_t{});
        case GBK_BOOL: return f(boolean{});
    }
    return fail(GBK_DOMAIN_MISMATCH, "unknown element type %d", type);
}

// Converts a host double to T the way a C++ cast would (bool: != 0).
template <typename T> T from_double(double v) {
    if constexpr (is_bool_v<T>) return v != 0.0 ? 1 : 0;
    else return (T)v;
}

}  // namespace gbk
