// gbk_vector: a device vector with the reference's sparse/dense switch.
//
// Replaces backend::Vector<T> (src/backend/vector/vector.h:22-276) with its
// DenseVector (dense_vector.h/.hpp) and SparseVector (sparse_vector.h/.hpp):
//   * a vector starts sparse and empty;
//   * a dense vector's nvals counts nonzeros (dense_vector.hpp:44-52);
//   * dense -> sparse keeps the nonzeros with ascending ids
//     (sparse_vector.hpp:36-92); sparse -> dense zero-fills and scatters
//     (dense_vector.hpp:87-102);
//   * convert_if_required: dense iff nvals > (Index)(0.3 * n) (vector.h:117-129,
//     SPARSE_VECTOR_THRESHOLD in la_backend.h:11).
// Both forms are kept allocated (n entries each), as the reference does.
#include <cmath>
#include <cstring>
#include <vector>

#include "gbk_objects.h"

namespace gbk {

static const double kSparseThreshold = 0.3;

int vec_nvals(gbk_vector v, int64_t* nvals) {
    if (v->storage == GBK_STORAGE_SPARSE) {
        *nvals = v->nvals;
        return GBK_SUCCESS;
    }
    return count_nonzero(v->type, v->dense, v->n, nvals);
}

int vec_force_dense(gbk_vector v) {
    if (v->storage == GBK_STORAGE_DENSE) return GBK_SUCCESS;
    GBK_TRY(fill_value(v->type, v->dense, v->n, 0.0));
    GBK_TRY(scatter_sparse(v->type, v->ids, v->svals, v->nvals, v->dense));
    v->storage = GBK_STORAGE_DENSE;
    return GBK_SUCCESS;
}

int vec_force_sparse(gbk_vector v) {
    if (v->storage == GBK_STORAGE_SPARSE) return GBK_SUCCESS;
    int64_t cnt = 0;
    GBK_TRY(compact_nonzero(v->type, v->dense, v->n, v->ids, v->svals, &cnt));
    v->nvals = cnt;
    v->storage = GBK_STORAGE_SPARSE;
    return GBK_SUCCESS;
}

int vec_convert_if_required(gbk_vector v) {
    int64_t nv = 0;
    GBK_TRY(vec_nvals(v, &nv));
    if (nv > (int64_t)((double)v->n * kSparseThreshold)) return vec_force_dense(v);
    if (v->storage == GBK_STORAGE_SPARSE) return GBK_SUCCESS;
    return vec_force_sparse(v);
}

}  // namespace gbk

using namespace gbk;

#define GBK_CHECK_VEC(v)                                                  \
    do {                                                                  \
        if (!(v)) return fail(GBK_UNINITIALIZED_OBJECT, "null vector");   \
        GBK_TRY(ensure_init());                                           \
    } while (0)

extern "C" int gbk_vector_new(gbk_vector* v, gbk_type type, int64_t n) {
    if (!v) return fail(GBK_NULL_POINTER, "null vector handle pointer");
    if (!valid_type(type)) return fail(GBK_DOMAIN_MISMATCH, "unknown type %d", (int)type);
    if (n < 0 || n >= (int64_t)INT32_MAX) return fail(GBK_INVALID_VALUE, "vector size %lld out of range", (long long)n);
    GBK_TRY(ensure_init());
    gbk_vector x = new gbk_vector_s();
    x->type = type;
    x->n = n;
    const size_t vs = type_size(type);
    int rc = dev_alloc(&x->dense, vs * (n > 0 ? n : 1));
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&x->ids, sizeof(int32_t) * (n > 0 ? n : 1));
    if (rc == GBK_SUCCESS) rc = dev_alloc(&x->svals, vs * (n > 0 ? n : 1));
    if (rc != GBK_SUCCESS) {
        dev_free(x->dense);
        dev_free(x->ids);
        dev_free(x->svals);
        delete x;
        return rc;
    }
    x->storage = GBK_STORAGE_SPARSE;
    x->nvals = 0;
    *v = x;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_free(gbk_vector v) {
    if (!v) return GBK_SUCCESS;
    (void)hipStreamSynchronize(stream());
    dev_free(v->dense);
    dev_free(v->ids);
    dev_free(v->svals);
    delete v;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_size(gbk_vector v, int64_t* n) {
    if (!v || !n) return fail(GBK_NULL_POINTER, "null argument");
    *n = v->n;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_type(gbk_vector v, int* type) {
    if (!v || !type) return fail(GBK_NULL_POINTER, "null argument");
    *type = v->type;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_nvals(gbk_vector v, int64_t* nvals) {
    GBK_CHECK_VEC(v);
    if (!nvals) return fail(GBK_NULL_POINTER, "null argument");
    return vec_nvals(v, nvals);
}

extern "C" int gbk_vector_storage(gbk_vector v, int* storage) {
    if (!v || !storage) return fail(GBK_NULL_POINTER, "null argument");
    *storage = v->storage;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_build_dense(gbk_vector v, const void* vals, int64_t n) {
    GBK_CHECK_VEC(v);
    if (!vals) return fail(GBK_NULL_POINTER, "null values");
    if (n == 0) return fail(GBK_INVALID_VALUE, "nvals == 0");
    if (n > v->n) return fail(GBK_INDEX_OUT_OF_BOUNDS, "%lld values for a vector of size %lld", (long long)n, (long long)v->n);
    const size_t vs = type_size(v->type);
    if (n < v->n) GBK_TRY(fill_value(v->type, v->dense, v->n, 0.0));
    GBK_HIP(hipMemcpyAsync(v->dense, vals, vs * n, hipMemcpyHostToDevice, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    v->storage = GBK_STORAGE_DENSE;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_build_sparse(gbk_vector v, const int64_t* ids, const void* vals, int64_t nvals) {
    GBK_CHECK_VEC(v);
    if (!ids || !vals) return fail(GBK_NULL_POINTER, "null ids or values");
    if (nvals == 0) return fail(GBK_INVALID_VALUE, "nvals == 0");
    if (nvals > v->n) return fail(GBK_INDEX_OUT_OF_BOUNDS, "too many entries");
    std::vector<int32_t> id32(nvals);
    for (int64_t i = 0; i < nvals; ++i) {
        if (ids[i] < 0 || ids[i] >= v->n) return fail(GBK_INVALID_INDEX, "index %lld out of range", (long long)ids[i]);
        id32[i] = (int32_t)ids[i];
    }
    const size_t vs = type_size(v->type);
    GBK_HIP(hipMemcpyAsync(v->ids, id32.data(), sizeof(int32_t) * nvals, hipMemcpyHostToDevice, stream()));
    GBK_HIP(hipMemcpyAsync(v->svals, vals, vs * nvals, hipMemcpyHostToDevice, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    v->nvals = nvals;
    v->storage = GBK_STORAGE_SPARSE;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_fill(gbk_vector v, double value) {
    GBK_CHECK_VEC(v);
    if (value == 0.0) {  // set_constant(0): empty sparse vector (vector.h:39-42)
        v->storage = GBK_STORAGE_SPARSE;
        v->nvals = 0;
        return GBK_SUCCESS;
    }
    GBK_TRY(fill_value(v->type, v->dense, v->n, value));
    v->storage = GBK_STORAGE_DENSE;
    return GBK_SUCCESS;
}

template <typename T> __global__ void k_fill_i64(T* d, int64_t n, int64_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (T)v;
}

extern "C" int gbk_vector_fill_int(gbk_vector v, int64_t value) {
    GBK_CHECK_VEC(v);
    if (value == 0) {
        v->storage = GBK_STORAGE_SPARSE;
        v->nvals = 0;
        return GBK_SUCCESS;
    }
    if (v->n > 0) {
        int rc = dispatch_type(v->type, [&](auto tv) -> int {
            typedef decltype(tv) T;
            hipLaunchKernelGGL(k_fill_i64<T>, dim3((unsigned)((v->n + 255) / 256)), dim3(256), 0, stream(), (T*)v->dense,
                               v->n, value);
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        });
        GBK_TRY(rc);
    }
    v->storage = GBK_STORAGE_DENSE;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_set_element(gbk_vector v, double value, int64_t pos) {
    GBK_CHECK_VEC(v);
    if (pos < 0 || pos >= v->n) return fail(GBK_INVALID_INDEX, "position %lld out of range", (long long)pos);
    return dispatch_type(v->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        T val = from_double<T>(value);
        hipStream_t st = stream();
        if (v->storage == GBK_STORAGE_DENSE) {
            GBK_HIP(hipMemcpyAsync((T*)v->dense + pos, &val, sizeof(T), hipMemcpyHostToDevice, st));
            GBK_HIP(hipStreamSynchronize(st));
            return GBK_SUCCESS;
        }
        // sparse: replace an existing entry or append (sparse_vector.hpp:5-19)
        std::vector<int32_t> ids(v->nvals);
        if (v->nvals) GBK_HIP(hipMemcpyAsync(ids.data(), v->ids, sizeof(int32_t) * v->nvals, hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        int64_t at = v->nvals;
        for (int64_t i = 0; i < v->nvals; ++i)
            if (ids[i] == pos) { at = i; break; }
        int32_t p32 = (int32_t)pos;
        if (at == v->nvals) {
            GBK_HIP(hipMemcpyAsync(v->ids + at, &p32, sizeof(int32_t), hipMemcpyHostToDevice, st));
            v->nvals++;
        }
        GBK_HIP(hipMemcpyAsync((T*)v->svals + at, &val, sizeof(T), hipMemcpyHostToDevice, st));
        GBK_HIP(hipStreamSynchronize(st));
        return GBK_SUCCESS;
    });
}

template <typename T> __global__ void k_iota(T* d, int64_t n, int64_t limit) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (i < limit) ? (T)i : T(0);
}

extern "C" int gbk_vector_fill_ascending(gbk_vector v, int64_t nvals) {
    GBK_CHECK_VEC(v);
    GBK_TRY(vec_force_dense(v));
    if (v->n == 0) return GBK_SUCCESS;
    return dispatch_type(v->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        hipLaunchKernelGGL(k_iota<T>, dim3((unsigned)((v->n + 255) / 256)), dim3(256), 0, stream(), (T*)v->dense, v->n,
                           nvals);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    });
}

extern "C" int gbk_vector_dup(gbk_vector dst, gbk_vector src) {
    GBK_CHECK_VEC(dst);
    if (!src) return fail(GBK_UNINITIALIZED_OBJECT, "null vector");
    if (dst->n != src->n || dst->type != src->type)
        return fail(GBK_DIMENSION_MISMATCH, "dup between vectors of different size or type");
    const size_t vs = type_size(src->type);
    hipStream_t st = stream();
    if (src->storage == GBK_STORAGE_DENSE) {
        if (src->n) GBK_HIP(hipMemcpyAsync(dst->dense, src->dense, vs * src->n, hipMemcpyDeviceToDevice, st));
        dst->storage = GBK_STORAGE_DENSE;
    } else {
        if (src->nvals) {
            GBK_HIP(hipMemcpyAsync(dst->ids, src->ids, sizeof(int32_t) * src->nvals, hipMemcpyDeviceToDevice, st));
            GBK_HIP(hipMemcpyAsync(dst->svals, src->svals, vs * src->nvals, hipMemcpyDeviceToDevice, st));
        }
        dst->nvals = src->nvals;
        dst->storage = GBK_STORAGE_SPARSE;
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_swap(gbk_vector a, gbk_vector b) {
    if (!a || !b) return fail(GBK_UNINITIALIZED_OBJECT, "null vector");
    std::swap(*a, *b);
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_clear(gbk_vector v) {
    if (!v) return fail(GBK_UNINITIALIZED_OBJECT, "null vector");
    v->storage = GBK_STORAGE_SPARSE;
    v->nvals = 0;
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_force_dense(gbk_vector v) {
    GBK_CHECK_VEC(v);
    return vec_force_dense(v);
}
extern "C" int gbk_vector_force_sparse(gbk_vector v) {
    GBK_CHECK_VEC(v);
    return vec_force_sparse(v);
}
extern "C" int gbk_vector_convert_if_required(gbk_vector v) {
    GBK_CHECK_VEC(v);
    return vec_convert_if_required(v);
}

extern "C" int gbk_vector_to_host(gbk_vector v, void* vals) {
    GBK_CHECK_VEC(v);
    if (!vals) return fail(GBK_NULL_POINTER, "null output");
    GBK_TRY(vec_force_dense(v));
    if (v->n) GBK_HIP(hipMemcpyAsync(vals, v->dense, type_size(v->type) * v->n, hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_sparse_to_host(gbk_vector v, int64_t* ids, void* vals, int64_t* nvals) {
    GBK_CHECK_VEC(v);
    if (!nvals) return fail(GBK_NULL_POINTER, "null nvals");
    GBK_TRY(vec_force_sparse(v));
    *nvals = v->nvals;
    if (v->nvals) {
        std::vector<int32_t> id32(v->nvals);
        GBK_HIP(hipMemcpyAsync(id32.data(), v->ids, sizeof(int32_t) * v->nvals, hipMemcpyDeviceToHost, stream()));
        if (vals) GBK_HIP(hipMemcpyAsync(vals, v->svals, type_size(v->type) * v->nvals, hipMemcpyDeviceToHost, stream()));
        GBK_HIP(hipStreamSynchronize(stream()));
        if (ids)
            for (int64_t i = 0; i < v->nvals; ++i) ids[i] = id32[i];
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_vector_device_dense(gbk_vector v, void** vals) {
    GBK_CHECK_VEC(v);
    if (!vals) return fail(GBK_NULL_POINTER, "null output");
    GBK_TRY(vec_force_dense(v));
    *vals = v->dense;
    return GBK_SUCCESS;
}
