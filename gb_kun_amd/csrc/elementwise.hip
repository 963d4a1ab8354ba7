// Element-wise operations the drivers use: assign, assignScatter, eWiseAdd,
// eWiseMult, apply, reduce, extract (src/backend/operations/operations.h:15-147,
// 216-430, with the mask loops of generic_operations.h:11-214 and
// indexed_operations.h:12-161).
//
// Reference semantics kept:
//   * assign (value) ignores the accumulator (operations.h:30-32) and makes
//     w dense, then converts it by the 30% rule;
//   * eWiseAdd / eWiseMult / apply ignore the accumulator and the descriptor
//     and leave w dense (no conversion);
//   * masks in assign: sparse mask + DEFAULT/STRUCTURE -> listed ids,
//     STR_COMP -> entries whose mask value is 0; dense mask + STR_COMP ->
//     mask == 0, DEFAULT/STRUCTURE -> mask != 0; COMP and SCMP select
//     nothing (generic_operations.h:11-66); in eWise/apply every mask is a
//     plain value mask (generic_operations.h:117-158);
//   * reduce folds the dense values, or the stored values when sparse, then
//     applies the accumulator to the scalar (operations.h:345-376).
// Deliberate difference: a masked vector assign with a sparse mask reads
// u[id] (GraphBLAS semantics); the reference reads u at the position inside
// the mask's id list (generic_operations.h:24-33 calls lambda(idx, i)), which
// is a defect (it skews PageRank's sink sum).  Same for extract with a sparse
// mask and no index vector.
#include <vector>

#include "gbk_objects.h"

namespace gbk {

// Which elements an operation touches.
struct Sel {
    int kind = 0;  // 0 all, 1 dense predicate, 2 id list, 3 none
    const void* m = nullptr;
    int mtype = 0;
    int pass = 1;  // dense: element passes when (m != 0) == pass
    const int32_t* ids = nullptr;
    int64_t count = 0;
};

__device__ __forceinline__ bool mtest(const void* m, int mt, int64_t i) {
    switch (mt) {
        case GBK_FP32: return ((const float*)m)[i] != 0.0f;
        case GBK_FP64: return ((const double*)m)[i] != 0.0;
        case GBK_INT32: return ((const int32_t*)m)[i] != 0;
        case GBK_INT64: return ((const int64_t*)m)[i] != 0;
        default: return ((const uint8_t*)m)[i] != 0;
    }
}

// Element index of work item t under the selector, or -1 to skip.
__device__ __forceinline__ int64_t sel_index(const Sel& s, int64_t t) {
    if (s.kind == 2) return s.ids[t];
    if (s.kind == 1) return (mtest(s.m, s.mtype, t) == (s.pass != 0)) ? t : -1;
    return t;
}

static int64_t sel_items(const Sel& s, int64_t n) {
    if (s.kind == 2) return s.count;
    if (s.kind == 3) return 0;
    return n;
}

// generic_dense_vector_op_assign's mask rule.
static int assign_sel(gbk_vector mask, const gbk_desc* desc, Sel* s) {
    *s = Sel();
    if (!mask) return GBK_SUCCESS;
    const int f = desc ? desc->mask : GBK_DEFAULT;
    if (mask->storage == GBK_STORAGE_SPARSE) {
        if (f == GBK_DEFAULT || f == GBK_STRUCTURE) {
            s->kind = 2;
            s->ids = mask->ids;
            s->count = mask->nvals;
        } else if (f == GBK_STR_COMP) {
            GBK_TRY(vec_force_dense(mask));  // the reference's getDense() converts the mask
            s->kind = 1;
            s->m = mask->dense;
            s->mtype = mask->type;
            s->pass = 0;
        } else {
            s->kind = 3;
        }
        return GBK_SUCCESS;
    }
    if (f == GBK_STR_COMP || f == GBK_DEFAULT || f == GBK_STRUCTURE) {
        s->kind = 1;
        s->m = mask->dense;
        s->mtype = mask->type;
        s->pass = (f == GBK_STR_COMP) ? 0 : 1;
    } else {
        s->kind = 3;
    }
    return GBK_SUCCESS;
}

// generic_dense_vector_op's mask rule (eWise*, apply).
static void value_sel(gbk_vector mask, Sel* s) {
    *s = Sel();
    if (!mask) return;
    if (mask->storage == GBK_STORAGE_SPARSE) {
        s->kind = 2;
        s->ids = mask->ids;
        s->count = mask->nvals;
    } else {
        s->kind = 1;
        s->m = mask->dense;
        s->mtype = mask->type;
        s->pass = 1;
    }
}

static dim3 grid1(int64_t n) {
    int64_t b = (n + 255) / 256;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// --------------------------------------------------------------- kernels --
template <typename W>
__global__ void k_assign_value(W* __restrict__ w, int64_t items, Sel s, W value) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    int64_t i = sel_index(s, t);
    if (i >= 0) w[i] = value;
}

template <typename W, typename U>
__global__ void k_assign_vec(W* __restrict__ w, const U* __restrict__ u, int64_t items, Sel s, int accum) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    int64_t i = sel_index(s, t);
    if (i >= 0) w[i] = binop<W>(accum, w[i], cvt<W>(u[i]));
}

template <typename W, typename U, typename I>
__global__ void k_assign_scatter(W* __restrict__ w, const U* __restrict__ u, const I* __restrict__ idx, int64_t count,
                                 int64_t n, const void* mask, int mtype, int accum) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int64_t i = (int64_t)idx[t];
    if (i < 0 || i >= n) return;
    if (mask && !mtest(mask, mtype, i)) return;
    w[i] = binop<W>(accum, w[i], cvt<W>(u[t]));
}

template <typename W, typename U, typename I>
__global__ void k_extract(W* __restrict__ w, const U* __restrict__ u, const I* __restrict__ idx, int64_t count,
                          int64_t n, const void* mask, int mtype, int accum) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int64_t j = idx ? (int64_t)idx[t] : t;
    if (j < 0 || j >= n) return;
    if (mask && !mtest(mask, mtype, idx ? j : t)) return;
    w[t] = binop<W>(accum, w[t], cvt<W>(u[j]));
}

template <typename W, typename U>
__global__ void k_ewise(W* __restrict__ w, const U* __restrict__ u, const U* __restrict__ v, int64_t items, Sel s,
                        int op) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    int64_t i = sel_index(s, t);
    if (i >= 0) w[i] = cvt<W>(binop<U>(op, u[i], v[i]));
}

enum { UN_IDENTITY = 0, UN_ABS = 1, UN_DIV_SCALAR = 2, UN_TIMES_SCALAR = 3, UN_AINV = 4, UN_SCALAR_DIV = 5,
       UN_PLUS_SCALAR = 6, UN_MINUS_SCALAR = 7 };

template <typename W, typename U>
__global__ void k_apply(W* __restrict__ w, const U* __restrict__ u, int64_t items, Sel s, int unary, U scalar) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    int64_t i = sel_index(s, t);
    if (i < 0) return;
    U x = u[i], r;
    switch (unary) {
        case UN_ABS: r = (x < U(0)) ? wrap_sub<U>(U(0), x) : x; break;
        case UN_DIV_SCALAR: r = safe_div<U>(x, scalar); break;
        case UN_TIMES_SCALAR: r = wrap_mul<U>(x, scalar); break;
        case UN_AINV: r = wrap_sub<U>(U(0), x); break;
        case UN_SCALAR_DIV: r = safe_div<U>(scalar, x); break;
        case UN_PLUS_SCALAR: r = wrap_add<U>(x, scalar); break;
        case UN_MINUS_SCALAR: r = wrap_sub<U>(x, scalar); break;
        default: r = x;
    }
    w[i] = cvt<W>(r);
}

// Block reduction of a monoid into partial[blockIdx.x]; values are widened
// to double (float/double/bool/int32 exact in double) or int64 (INT64).
template <typename T, typename R>
__global__ __launch_bounds__(256) void k_reduce(const T* __restrict__ v, int64_t n, int monoid, R* __restrict__ partial) {
    __shared__ R s[256];
    R acc = (monoid == GBK_ADD_MIN) ? (R)Lim<T>::max() : (monoid == GBK_ADD_MAX) ? (R)Lim<T>::lowest() : (R)0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        R x = (R)v[i];
        switch (monoid) {
            case GBK_ADD_MIN: acc = (x < acc) ? x : acc; break;
            case GBK_ADD_MAX: acc = (x > acc) ? x : acc; break;
            case GBK_ADD_LOR: acc = (acc != R(0) || x != R(0)) ? R(1) : R(0); break;
            default: acc += x;
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            R a = s[threadIdx.x], b = s[threadIdx.x + off];
            switch (monoid) {
                case GBK_ADD_MIN: a = (b < a) ? b : a; break;
                case GBK_ADD_MAX: a = (b > a) ? b : a; break;
                case GBK_ADD_LOR: a = (a != R(0) || b != R(0)) ? R(1) : R(0); break;
                default: a += b;
            }
            s[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

template <typename R> static R combine_host(int monoid, R a, R b) {
    switch (monoid) {
        case GBK_ADD_MIN: return (b < a) ? b : a;
        case GBK_ADD_MAX: return (b > a) ? b : a;
        case GBK_ADD_LOR: return (a != R(0) || b != R(0)) ? R(1) : R(0);
        default: return a + b;
    }
}

template <typename T, typename R> static int reduce_typed(const T* v, int64_t n, int monoid, R* out) {
    const int blocks = 512;
    void* p;
    GBK_TRY(scratch(4, sizeof(R) * blocks, &p));
    R part[blocks];
    R acc = (monoid == GBK_ADD_MIN) ? (R)Lim<T>::max() : (monoid == GBK_ADD_MAX) ? (R)Lim<T>::lowest() : (R)0;
    if (n > 0) {
        hipLaunchKernelGGL((k_reduce<T, R>), dim3(blocks), dim3(256), 0, stream(), v, n, monoid, (R*)p);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipMemcpyAsync(part, p, sizeof(R) * blocks, hipMemcpyDeviceToHost, stream()));
        GBK_HIP(hipStreamSynchronize(stream()));
        for (int b = 0; b < blocks; ++b) acc = combine_host(monoid, acc, part[b]);
    }
    *out = acc;
    return GBK_SUCCESS;
}

static int check_monoid(int monoid) {
    if (monoid != GBK_ADD_PLUS && monoid != GBK_ADD_MIN && monoid != GBK_ADD_MAX && monoid != GBK_ADD_LOR)
        return fail(GBK_INVALID_VALUE, "reduce supports the plus, min, max and lor monoids");
    return GBK_SUCCESS;
}

template <typename R> static int reduce_vector(gbk_vector u, int monoid, R* out) {
    GBK_TRY(check_monoid(monoid));
    const bool dense = u->storage == GBK_STORAGE_DENSE;
    const void* v = dense ? u->dense : u->svals;
    const int64_t cnt = dense ? u->n : u->nvals;
    return dispatch_type(u->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return reduce_typed<T, R>((const T*)v, cnt, monoid, out);
    });
}

static int check_pair(gbk_vector w, gbk_vector u) {
    if (!w || !u) return fail(GBK_UNINITIALIZED_OBJECT, "uninitialized vector");
    if (w->n != u->n) return fail(GBK_DIMENSION_MISMATCH, "vector sizes differ");
    return GBK_SUCCESS;
}

// Dispatches f(W{}, U{}) over the two element types.
template <typename F> static int dispatch_two(int wt, int ut, F&& f) {
    return dispatch_type(wt, [&](auto wv) -> int {
        return dispatch_type(ut, [&](auto uv) -> int { return f(wv, uv); });
    });
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_assign_scalar(gbk_vector w, gbk_vector mask, int accum, double value, const gbk_desc* desc) {
    if (!w) return fail(GBK_UNINITIALIZED_OBJECT, "uninitialized vector");
    GBK_TRY(ensure_init());
    (void)accum;  // the value assign ignores the accumulator (operations.h:30-32)
    if (mask && mask->n != w->n) return fail(GBK_DIMENSION_MISMATCH, "mask size differs");
    GBK_TRY(vec_force_dense(w));
    Sel s;
    GBK_TRY(assign_sel(mask, desc, &s));
    const int64_t items = sel_items(s, w->n);
    if (items > 0) {
        int rc = dispatch_type(w->type, [&](auto tv) -> int {
            typedef decltype(tv) W;
            hipLaunchKernelGGL(k_assign_value<W>, grid1(items), dim3(256), 0, stream(), (W*)w->dense, items, s,
                               from_double<W>(value));
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        });
        GBK_TRY(rc);
    }
    return vec_convert_if_required(w);
}

extern "C" int gbk_assign_vector(gbk_vector w, gbk_vector mask, int accum, gbk_vector u, const gbk_desc* desc) {
    GBK_TRY(check_pair(w, u));
    GBK_TRY(ensure_init());
    if (mask && mask->n != w->n) return fail(GBK_DIMENSION_MISMATCH, "mask size differs");
    GBK_TRY(vec_force_dense(w));
    GBK_TRY(vec_force_dense(u));
    Sel s;
    GBK_TRY(assign_sel(mask, desc, &s));
    const int64_t items = sel_items(s, w->n);
    if (items > 0) {
        int rc = dispatch_two(w->type, u->type, [&](auto wv, auto uv) -> int {
            typedef decltype(wv) W;
            typedef decltype(uv) U;
            hipLaunchKernelGGL((k_assign_vec<W, U>), grid1(items), dim3(256), 0, stream(), (W*)w->dense,
                               (const U*)u->dense, items, s, accum);
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        });
        GBK_TRY(rc);
    }
    return vec_convert_if_required(w);
}

extern "C" int gbk_assign_scatter(gbk_vector w, gbk_vector mask, int accum, gbk_vector u, gbk_vector indices,
                                  int64_t nindices, const gbk_desc* desc) {
    GBK_TRY(check_pair(w, u));
    if (!indices) return fail(GBK_UNINITIALIZED_OBJECT, "null index vector");
    GBK_TRY(ensure_init());
    (void)desc;
    if (nindices < 0 || nindices > indices->n || nindices > u->n) return fail(GBK_INVALID_VALUE, "bad index count");
    GBK_TRY(vec_force_dense(w));
    GBK_TRY(vec_force_dense(u));
    GBK_TRY(vec_force_dense(indices));
    if (mask) GBK_TRY(vec_force_dense(mask));
    if (nindices == 0) return vec_convert_if_required(w);
    int rc = dispatch_two(w->type, u->type, [&](auto wv, auto uv) -> int {
        typedef decltype(wv) W;
        typedef decltype(uv) U;
        return dispatch_type(indices->type, [&](auto iv) -> int {
            typedef decltype(iv) I;
            hipLaunchKernelGGL((k_assign_scatter<W, U, I>), grid1(nindices), dim3(256), 0, stream(), (W*)w->dense,
                               (const U*)u->dense, (const I*)indices->dense, nindices, w->n,
                               mask ? (const void*)mask->dense : nullptr, mask ? mask->type : 0, accum);
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        });
    });
    GBK_TRY(rc);
    return vec_convert_if_required(w);
}

static int ewise(gbk_vector w, gbk_vector mask, int op, gbk_vector u, gbk_vector v) {
    GBK_TRY(check_pair(w, u));
    GBK_TRY(check_pair(w, v));
    GBK_TRY(ensure_init());
    if (u->type != v->type) return fail(GBK_DOMAIN_MISMATCH, "eWise operands must share a type");
    if (op < GBK_ACC_NONE || op > GBK_ACC_NE) return fail(GBK_INVALID_VALUE, "unknown binary operator %d", op);
    if (mask && mask->n != w->n) return fail(GBK_DIMENSION_MISMATCH, "mask size differs");
    GBK_TRY(vec_force_dense(u));
    GBK_TRY(vec_force_dense(v));
    GBK_TRY(vec_force_dense(w));
    Sel s;
    value_sel(mask, &s);
    const int64_t items = sel_items(s, w->n);
    if (items == 0) return GBK_SUCCESS;
    return dispatch_two(w->type, u->type, [&](auto wv, auto uv) -> int {
        typedef decltype(wv) W;
        typedef decltype(uv) U;
        hipLaunchKernelGGL((k_ewise<W, U>), grid1(items), dim3(256), 0, stream(), (W*)w->dense, (const U*)u->dense,
                           (const U*)v->dense, items, s, op);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    });
}

extern "C" int gbk_ewise_add(gbk_vector w, gbk_vector mask, int accum, int op, gbk_vector u, gbk_vector v,
                             const gbk_desc* desc) {
    (void)accum;
    (void)desc;
    return ewise(w, mask, op, u, v);
}

extern "C" int gbk_ewise_mult(gbk_vector w, gbk_vector mask, int accum, int op, gbk_vector u, gbk_vector v,
                              const gbk_desc* desc) {
    (void)accum;
    (void)desc;
    return ewise(w, mask, op, u, v);
}

extern "C" int gbk_apply(gbk_vector w, gbk_vector mask, int accum, int unary, double scalar, gbk_vector u,
                         const gbk_desc* desc) {
    (void)accum;
    (void)desc;
    GBK_TRY(check_pair(w, u));
    GBK_TRY(ensure_init());
    if (unary < UN_IDENTITY || unary > UN_MINUS_SCALAR) return fail(GBK_INVALID_VALUE, "unknown unary operator %d", unary);
    if (mask && mask->n != w->n) return fail(GBK_DIMENSION_MISMATCH, "mask size differs");
    GBK_TRY(vec_force_dense(u));
    GBK_TRY(vec_force_dense(w));
    Sel s;
    value_sel(mask, &s);
    const int64_t items = sel_items(s, w->n);
    if (items == 0) return GBK_SUCCESS;
    return dispatch_two(w->type, u->type, [&](auto wv, auto uv) -> int {
        typedef decltype(wv) W;
        typedef decltype(uv) U;
        hipLaunchKernelGGL((k_apply<W, U>), grid1(items), dim3(256), 0, stream(), (W*)w->dense, (const U*)u->dense,
                           items, s, unary, from_double<U>(scalar));
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    });
}

extern "C" int gbk_reduce(double* result, int accum, int monoid, gbk_vector u, const gbk_desc* desc) {
    (void)desc;
    if (!result) return fail(GBK_NULL_POINTER, "null result");
    if (!u) return fail(GBK_UNINITIALIZED_OBJECT, "uninitialized vector");
    GBK_TRY(ensure_init());
    double r = 0;
    GBK_TRY(reduce_vector<double>(u, monoid, &r));
    *result = binop<double>(accum, *result, r);
    return GBK_SUCCESS;
}

extern "C" int gbk_reduce_i64(int64_t* result, int accum, int monoid, gbk_vector u, const gbk_desc* desc) {
    (void)desc;
    if (!result) return fail(GBK_NULL_POINTER, "null result");
    if (!u) return fail(GBK_UNINITIALIZED_OBJECT, "uninitialized vector");
    GBK_TRY(ensure_init());
    if (u->type == GBK_FP32 || u->type == GBK_FP64) return fail(GBK_DOMAIN_MISMATCH, "integer reduce of a floating vector");
    long long r = 0;
    GBK_TRY(reduce_vector<long long>(u, monoid, &r));
    *result = binop<int64_t>(accum, *result, (int64_t)r);
    return GBK_SUCCESS;
}

extern "C" int gbk_extract(gbk_vector w, gbk_vector mask, int accum, gbk_vector u, gbk_vector indices,
                           const gbk_desc* desc) {
    (void)desc;
    GBK_TRY(check_pair(w, u));
    GBK_TRY(ensure_init());
    if (mask && mask->n != w->n) return fail(GBK_DIMENSION_MISMATCH, "mask size differs");
    GBK_TRY(vec_force_dense(w));
    GBK_TRY(vec_force_dense(u));
    if (mask) GBK_TRY(vec_force_dense(mask));
    if (indices) GBK_TRY(vec_force_dense(indices));
    const int64_t count = indices ? std::min(indices->n, w->n) : w->n;
    if (count > 0) {
        int rc = dispatch_two(w->type, u->type, [&](auto wv, auto uv) -> int {
            typedef decltype(wv) W;
            typedef decltype(uv) U;
            if (!indices) {
                hipLaunchKernelGGL((k_extract<W, U, int32_t>), grid1(count), dim3(256), 0, stream(), (W*)w->dense,
                                   (const U*)u->dense, (const int32_t*)nullptr, count, u->n,
                                   mask ? (const void*)mask->dense : nullptr, mask ? mask->type : 0, accum);
                GBK_LAUNCH_CHECK();
                return GBK_SUCCESS;
            }
            return dispatch_type(indices->type, [&](auto iv) -> int {
                typedef decltype(iv) I;
                hipLaunchKernelGGL((k_extract<W, U, I>), grid1(count), dim3(256), 0, stream(), (W*)w->dense,
                                   (const U*)u->dense, (const I*)indices->dense, count, u->n,
                                   mask ? (const void*)mask->dense : nullptr, mask ? mask->type : 0, accum);
                GBK_LAUNCH_CHECK();
                return GBK_SUCCESS;
            });
        });
        GBK_TRY(rc);
    }
    return vec_convert_if_required(w);
}
