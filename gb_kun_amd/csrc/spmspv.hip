// Push SpMSpV (sparse x): y = x (+.*) A over the rows of A named by x.
//
// Reference: backend::SpMSpV + apply_mask (src/backend/spmspv/spmspv.h:17-158)
// and the kernels in spmspv_atomics.h:5-170.  The kernel family is picked, as
// in the reference, by probing the monoid with add(3, 5):
//   8 (plus)      -> y[d] += mul(a, x)  (atomic add)
//   1 (or, less)  -> y[d] = 1           (structural: ignores a and x)
//   other         -> y[d] = add(y[d], mul(a, x))  (CAS loop; atomicMin-like)
// y is first set to the monoid identity everywhere.  Without a mask the
// accumulator is ignored (the reference only applies it in apply_mask).
// With a mask, the old y is combined as apply_mask does:
//   not STR_COMP, dense mask : y = m != 0 ? accum(old, y) : 0
//   not STR_COMP, sparse mask: listed rows accum(old, y), others keep y
//   STR_COMP, dense mask     : y = m == 0 ? accum(old, y) : 0
//   STR_COMP, sparse mask    : listed rows 0, others accum(old, y)
//
// GPU form: the frontier's row lengths are prefix-summed, and the expansion
// is edge-balanced: every thread owns one edge and finds its frontier entry
// in the prefix sums (edge_owner: one global search per block, then a search
// in an LDS window), so hub vertices in the frontier are spread over many
// waves.
#include "gbk_objects.h"

namespace gbk {

template <typename T> __device__ __forceinline__ T atomic_combine_add(T* p, T v) {
    if constexpr (std::is_same<T, int64_t>::value) {
        return (T)atomicAdd((unsigned long long*)p, (unsigned long long)v);
    } else {
        return atomicAdd(p, v);
    }
}

// Generic read-modify-write with a compare-and-swap loop; skips the write
// when the combined value equals the stored one.
template <typename T, int ADD> __device__ __forceinline__ void atomic_combine_cas(T* p, T v) {
    if constexpr (sizeof(T) == 4) {
        unsigned int* a = (unsigned int*)p;
        unsigned int old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            T o;
            __builtin_memcpy(&o, &old, 4);
            T nv = add_op<ADD, T>(o, v);
            unsigned int nb;
            __builtin_memcpy(&nb, &nv, 4);
            if (nb == old) return;
            unsigned int prev = atomicCAS(a, old, nb);
            if (prev == old) return;
            old = prev;
        }
    } else if constexpr (sizeof(T) == 8) {
        unsigned long long* a = (unsigned long long*)p;
        unsigned long long old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            T o;
            __builtin_memcpy(&o, &old, 8);
            T nv = add_op<ADD, T>(o, v);
            unsigned long long nb;
            __builtin_memcpy(&nb, &nv, 8);
            if (nb == old) return;
            unsigned long long prev = atomicCAS(a, old, nb);
            if (prev == old) return;
            old = prev;
        }
    } else {
        // one byte inside an aligned 32-bit word
        uintptr_t addr = (uintptr_t)p;
        unsigned int* a = (unsigned int*)(addr & ~(uintptr_t)3);
        const int sh = (int)(addr & 3) * 8;
        unsigned int old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            T o = (T)((old >> sh) & 0xFFu);
            T nv = add_op<ADD, T>(o, v);
            unsigned int nw = (old & ~(0xFFu << sh)) | ((unsigned int)(uint8_t)nv << sh);
            if (nw == old) return;
            unsigned int prev = atomicCAS(a, old, nw);
            if (prev == old) return;
            old = prev;
        }
    }
}

enum { PUSH_ADD = 0, PUSH_OR = 1, PUSH_CAS = 2 };

template <typename T, int ADD> constexpr int push_kind() {
    // add(3, 5) computed in T, as the reference probes it
    if constexpr (ADD == GBK_ADD_PLUS) return is_bool_v<T> ? PUSH_OR : PUSH_ADD;
    else if constexpr (ADD == GBK_ADD_LOR || ADD == GBK_ADD_LESS) return PUSH_OR;
    else return PUSH_CAS;
}

__global__ void k_frontier_degrees(const int32_t* __restrict__ rp, const int32_t* __restrict__ ids, int64_t nv,
                                   int32_t* __restrict__ deg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) {
        int u = ids[i];
        deg[i] = rp[u + 1] - rp[u];
    }
}

template <typename T, int ADD, int MUL>
__global__ __launch_bounds__(256) void k_push(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                              const T* __restrict__ val, const int32_t* __restrict__ x_ids,
                                              const T* __restrict__ x_vals, int64_t nv,
                                              const int32_t* __restrict__ offs, int64_t nedges, T* __restrict__ y) {
    __shared__ int32_t s_offs[EO_WIN + 1];
    __shared__ int64_t s_i0;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // frontier entry i with offs[i] <= e < offs[i+1]
    const int64_t i = edge_owner(offs, nv, e, s_offs, &s_i0);
    if (e >= nedges) return;
    const int u = x_ids[i];
    const int64_t k = rp[u] + (e - offs[i]);
    const int d = col[k];
    constexpr int kind = push_kind<T, ADD>();
    if constexpr (kind == PUSH_OR) {
        y[d] = T(1);
    } else {
        const T v = mul_op<MUL, T>(val[k], x_vals[i]);
        if constexpr (kind == PUSH_ADD) atomic_combine_add(y + d, v);
        else atomic_combine_cas<T, ADD>(y + d, v);
    }
}

// apply_mask for a dense mask: y = pass ? accum(y_old, t) : 0.
template <typename T>
__global__ void k_mask_dense(T* __restrict__ y, const T* __restrict__ t, int64_t n, const void* __restrict__ mask,
                             int mtype, int pass_on_true, int accum) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool m;
    switch (mtype) {
        case GBK_FP32: m = ((const float*)mask)[i] != 0.0f; break;
        case GBK_FP64: m = ((const double*)mask)[i] != 0.0; break;
        case GBK_INT32: m = ((const int32_t*)mask)[i] != 0; break;
        case GBK_INT64: m = ((const int64_t*)mask)[i] != 0; break;
        default: m = ((const uint8_t*)mask)[i] != 0; break;
    }
    y[i] = (m == (pass_on_true != 0)) ? binop<T>(accum, y[i], t[i]) : T(0);
}

// sparse mask, not complemented: t[ids] = accum(y_old[ids], t[ids]).
template <typename T>
__global__ void k_mask_listed(const T* __restrict__ y, T* __restrict__ t, const int32_t* __restrict__ ids, int64_t nv,
                              int accum) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) {
        int r = ids[i];
        t[r] = binop<T>(accum, y[r], t[r]);
    }
}

__global__ void k_flag_ids(uint8_t* __restrict__ flag, const int32_t* __restrict__ ids, int64_t nv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) flag[ids[i]] = 1;
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

int build_flag_mask(gbk_vector mask, uint8_t* flag) {
    GBK_TRY(fill_value(GBK_BOOL, flag, mask->n, 0.0));
    if (mask->nvals > 0) {
        hipLaunchKernelGGL(k_flag_ids, grid_for(mask->nvals), dim3(256), 0, stream(), flag, mask->ids, mask->nvals);
        GBK_LAUNCH_CHECK();
    }
    return GBK_SUCCESS;
}

int spmspv(gbk_matrix A, bool use_csr, const int32_t* x_ids, const void* x_vals, int64_t x_nvals, void* y, int add,
           int mul, int accum, gbk_vector mask, int mask_mode) {
    const DevCSR& m = use_csr ? A->csr : A->csc;
    const int64_t n = A->n;
    hipStream_t st = stream();
    // frontier prefix sums
    void *pdeg, *poffs;
    GBK_TRY(scratch(2, sizeof(int32_t) * (x_nvals + 1), &pdeg));
    GBK_TRY(scratch(3, sizeof(int32_t) * (x_nvals + 1), &poffs));
    int32_t* deg = (int32_t*)pdeg;
    int32_t* offs = (int32_t*)poffs;
    int32_t nedges = 0;
    if (x_nvals > 0) {
        hipLaunchKernelGGL(k_frontier_degrees, grid_for(x_nvals), dim3(256), 0, st, m.row_ptr, x_ids, x_nvals, deg);
        GBK_LAUNCH_CHECK();
        GBK_TRY(exclusive_scan_i32(deg, offs, x_nvals, &nedges));
    }
    void* target = y;
    if (mask) target = A->ws.buf[1];
    return dispatch_type(A->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            GBK_TRY(fill_identity(A->type, ADD, target, n));
            if (nedges > 0) {
                hipLaunchKernelGGL((k_push<T, ADD, MUL>), grid_for(nedges), dim3(256), 0, st, m.row_ptr, m.col,
                                   (const T*)m.val, x_ids, (const T*)x_vals, x_nvals, offs, (int64_t)nedges, (T*)target);
                GBK_LAUNCH_CHECK();
            }
            if (!mask) return GBK_SUCCESS;
            T* yy = (T*)y;
            T* tt = (T*)target;
            const bool comp = (mask_mode == GBK_STR_COMP);
            if (mask->storage == GBK_STORAGE_DENSE) {
                hipLaunchKernelGGL(k_mask_dense<T>, grid_for(n), dim3(256), 0, st, yy, (const T*)tt, n,
                                   (const void*)mask->dense, mask->type, comp ? 0 : 1, accum);
                GBK_LAUNCH_CHECK();
            } else if (!comp) {
                if (mask->nvals > 0) {
                    hipLaunchKernelGGL(k_mask_listed<T>, grid_for(mask->nvals), dim3(256), 0, st, (const T*)yy, tt,
                                       mask->ids, mask->nvals, accum);
                    GBK_LAUNCH_CHECK();
                }
                if (n) GBK_HIP(hipMemcpyAsync(yy, tt, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
            } else {
                uint8_t* flag = (uint8_t*)A->ws.buf[2];
                GBK_TRY(build_flag_mask(mask, flag));
                hipLaunchKernelGGL(k_mask_dense<T>, grid_for(n), dim3(256), 0, st, yy, (const T*)tt, n,
                                   (const void*)flag, (int)GBK_BOOL, 0, accum);
                GBK_LAUNCH_CHECK();
            }
            return GBK_SUCCESS;
        });
    });
}

}  // namespace gbk
