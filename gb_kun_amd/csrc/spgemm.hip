// Sparse matrix-matrix products (masked and unmasked) and triangle counting.
//
// Reference: lablas::mxm (src/cpp_graphblas/operations.hpp:411-437),
// backend::mxm (src/backend/operations/operations.h:436-486) with a mask:
// SpMSpM_masked_ikj (src/backend/spmspm/spmspm_masked_hash_based_gustavson.h:
// 9-170), and LAGr_TriangleCount (algorithms/tc/tc.hpp:80-291).
//
// C<M> = A (+.*) B with a structural mask M: C has exactly the entries of M;
// C(i,j) = fold over k of A(i,k) (*) B(k,j), or 0 where no k contributes
// (the reference zero-fills its result values).  The reference writes each
// row's sums in hash-map iteration order next to the mask's column ids, so
// its values can sit at the wrong columns; here every sum is stored at its
// own column.  A complemented mask is an error, as in the reference.
//
// GPU form (the IKJ order of the reference): the mask's rows are sorted by
// column once (stable device sorts); a group of 16 lanes takes one entry
// A(i,k) and walks row k of B, each lane binary-searching its column j in the
// sorted mask row i and combining the product into that slot atomically (an
// add for plus, a compare-and-swap loop for the other monoids).  Integer
// results do not depend on the order; fp32 sums are exact up to rounding
// order.
//
// Triangle counting: the methods of LAGr_TriangleCount (Burkhardt, Cohen,
// Sandia, Sandia2, SandiaDot, SandiaDot2) as masked products with the
// plus.one semiring, summed; the count is accumulated directly (one atomic
// per workgroup) instead of materialising C.  The default method (0, which
// the reference maps to Burkhardt) is Sandia over the degree orientation,
// the cheapest of the formulations on power-law graphs.  The reference divides the
// Cohen sum by 6 (tc.hpp), which gives a third of the triangles; here Cohen
// divides by 2, so every method returns the same count.
#include <vector>

#include "gbk_objects.h"

namespace gbk {

namespace {

constexpr int MXM_G = 16;  // lanes per A entry

dim3 grid_m(int64_t n, int per = 256) {
    const int64_t b = (n + per - 1) / per;
    return dim3((unsigned)(b > 0 ? (b < 65535 * 16 ? b : 65535 * 16) : 1));
}

template <typename T, int ADD> __device__ __forceinline__ void combine_atomic(T* p, T v) {
    if constexpr (ADD == GBK_ADD_PLUS && !is_bool_v<T>) {
        if constexpr (sizeof(T) == 8 && std::is_integral<T>::value)
            atomicAdd((unsigned long long*)p, (unsigned long long)v);
        else
            atomicAdd(p, v);
    } else if constexpr (sizeof(T) == 4 || sizeof(T) == 8) {
        using U = typename std::conditional<sizeof(T) == 4, unsigned int, unsigned long long>::type;
        U* a = (U*)p;
        U old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            T o;
            __builtin_memcpy(&o, &old, sizeof(T));
            const T nv = add_op<ADD, T>(o, v);
            U nb;
            __builtin_memcpy(&nb, &nv, sizeof(T));
            if (nb == old) return;
            const U prev = atomicCAS(a, old, nb);
            if (prev == old) return;
            old = prev;
        }
    } else {
        // one byte inside its aligned word
        const uintptr_t addr = (uintptr_t)p;
        unsigned int* a = (unsigned int*)(addr & ~(uintptr_t)3);
        const int sh = (int)(addr & 3) * 8;
        unsigned int old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            const T o = (T)((old >> sh) & 0xFFu);
            const T nv = add_op<ADD, T>(o, v);
            const unsigned int nw = (old & ~(0xFFu << sh)) | ((unsigned int)(uint8_t)nv << sh);
            if (nw == old) return;
            const unsigned int prev = atomicCAS(a, old, nw);
            if (prev == old) return;
            old = prev;
        }
    }
}

// acc[p] (sorted mask order) = fold of products; touched[p] = 1 where any.
template <typename T, int ADD, int MUL>
__global__ __launch_bounds__(256) void k_mxm_masked(const int32_t* __restrict__ a_row, const int32_t* __restrict__ a_col,
                                                    const T* __restrict__ a_val, int64_t a_nnz,
                                                    const int32_t* __restrict__ b_rp, const int32_t* __restrict__ b_col,
                                                    const T* __restrict__ b_val, const int32_t* __restrict__ m_rp,
                                                    const int32_t* __restrict__ m_scol, T* __restrict__ acc,
                                                    uint8_t* __restrict__ touched) {
    const int lane = threadIdx.x % MXM_G;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x / MXM_G);
    for (int64_t e = (int64_t)blockIdx.x * (blockDim.x / MXM_G) + threadIdx.x / MXM_G; e < a_nnz; e += groups) {
        const int i = a_row[e], k = a_col[e];
        const int ms = m_rp[i], me = m_rp[i + 1];
        if (ms == me) continue;
        const T a = a_val[e];
        const int bs = b_rp[k], be = b_rp[k + 1];
        for (int q = bs + lane; q < be; q += MXM_G) {
            const int32_t j = b_col[q];
            // the mask row is sorted: skip columns outside its range cheaply
            if (j < m_scol[ms] || j > m_scol[me - 1]) continue;
            int lo = ms, hi = me - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (m_scol[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (m_scol[lo] != j) continue;
            combine_atomic<T, ADD>(acc + lo, mul_op<MUL, T>(a, b_val[q]));
            touched[lo] = 1;
        }
    }
}

// Counting form for plus.one: adds 1 per (i, k, j) with j in mask row i.
__global__ __launch_bounds__(256) void k_mxm_count(const int32_t* __restrict__ a_row, const int32_t* __restrict__ a_col,
                                                   int64_t a_nnz, const int32_t* __restrict__ b_rp,
                                                   const int32_t* __restrict__ b_col, const int32_t* __restrict__ m_rp,
                                                   const int32_t* __restrict__ m_scol,
                                                   unsigned long long* __restrict__ total) {
    __shared__ unsigned long long s_t;
    if (threadIdx.x == 0) s_t = 0;
    __syncthreads();
    const int lane = threadIdx.x % MXM_G;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x / MXM_G);
    unsigned long long c = 0;
    for (int64_t e = (int64_t)blockIdx.x * (blockDim.x / MXM_G) + threadIdx.x / MXM_G; e < a_nnz; e += groups) {
        const int i = a_row[e], k = a_col[e];
        const int ms = m_rp[i], me = m_rp[i + 1];
        if (ms == me) continue;
        const int32_t jlo = m_scol[ms], jhi = m_scol[me - 1];
        const int bs = b_rp[k], be = b_rp[k + 1];
        for (int q = bs + lane; q < be; q += MXM_G) {
            const int32_t j = b_col[q];
            if (j < jlo || j > jhi) continue;
            int lo = ms, hi = me - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (m_scol[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            c += (m_scol[lo] == j) ? 1ull : 0ull;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_t, c);
    __syncthreads();
    if (threadIdx.x == 0 && s_t) atomicAdd(total, s_t);
}

template <typename T, int ADD>
__global__ void k_fill_identity(T* __restrict__ acc, uint8_t* __restrict__ touched, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        acc[i] = add_identity<ADD, T>();
        touched[i] = 0;
    }
}

// C values in the mask's own entry order; 0 where nothing contributed.
template <typename T>
__global__ void k_unsort_vals(const T* __restrict__ acc, const uint8_t* __restrict__ touched,
                              const int32_t* __restrict__ perm, int64_t m, T* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) out[perm[p]] = touched[p] ? acc[p] : T(0);
}

// keep[e] = 1 for entries strictly below (mode 0) / above (mode 1) the
// diagonal, or (mode 2) from the lower- to the higher-ranked end, ranking by
// (degree, id)
__global__ void k_tri_flags(const int32_t* __restrict__ rowid, const int32_t* __restrict__ col, int64_t nnz, int mode,
                            const int32_t* __restrict__ rp, int32_t* __restrict__ keep) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int32_t r = rowid[e], c = col[e];
    if (mode == 0) keep[e] = c < r;
    else if (mode == 1) keep[e] = c > r;
    else {
        const int32_t dr = rp[r + 1] - rp[r], dc = rp[c + 1] - rp[c];
        keep[e] = (dr < dc) || (dr == dc && r < c);
    }
}

__global__ void k_tri_compact(const int32_t* __restrict__ keep, const int32_t* __restrict__ pos,
                              const int32_t* __restrict__ col, int64_t nnz, int32_t* __restrict__ out_col) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz && keep[e]) out_col[pos[e]] = col[e];
}

// new row pointers: rp'[r] = pos[rp[r]] (pos has nnz + 1 entries)
__global__ void k_tri_rowptr(const int32_t* __restrict__ rp, const int32_t* __restrict__ pos, int64_t n,
                             int32_t* __restrict__ out_rp) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= n) out_rp[r] = pos[rp[r]];
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { dev_free(p); }
    int alloc(size_t bytes) { return dev_alloc(&p, bytes > 0 ? bytes : 1); }
    template <typename V> V* as() const { return (V*)p; }
};

// Row-sorted copy of a CSR's column ids (scol) and, for each sorted
// position, the entry's original position (perm).
int sort_rows(const DevCSR& m, DevBuf& scol, DevBuf& perm, hipStream_t st) {
    const int64_t nnz = m.nnz, n = m.nrows;
    const int64_t z1 = nnz > 0 ? nnz : 1;
    GBK_TRY(scol.alloc(4 * z1));
    GBK_TRY(perm.alloc(4 * z1));
    if (nnz == 0) return GBK_SUCCESS;
    DevBuf iota, k1, p1, rowid, rkey, rkey_out;
    GBK_TRY(iota.alloc(4 * z1));
    GBK_TRY(k1.alloc(4 * z1));
    GBK_TRY(p1.alloc(4 * z1));
    GBK_TRY(rowid.alloc(4 * z1));
    GBK_TRY(rkey.alloc(4 * z1));
    GBK_TRY(rkey_out.alloc(4 * z1));
    GBK_TRY(dev_iota(iota.as<int32_t>(), nnz, st));
    // by column, then stably by row: (row, column) order
    GBK_TRY(dev_sort_pairs((const uint32_t*)m.col, k1.as<uint32_t>(), iota.as<int32_t>(), p1.as<int32_t>(), nnz,
                           bits_for(n > 0 ? n - 1 : 0), st));
    GBK_TRY(dev_row_ids(m.row_ptr, n, rowid.as<int32_t>(), st));
    GBK_TRY(dev_gather_i32(rowid.as<int32_t>(), p1.as<int32_t>(), nnz, rkey.as<int32_t>(), st));
    GBK_TRY(dev_sort_pairs(rkey.as<uint32_t>(), rkey_out.as<uint32_t>(), p1.as<int32_t>(), perm.as<int32_t>(), nnz,
                           bits_for(n > 0 ? n - 1 : 0), st));
    GBK_TRY(dev_gather_i32(m.col, perm.as<int32_t>(), nnz, scol.as<int32_t>(), st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

// mode 0: L = tril(A, -1), 1: U = triu(A, 1), 2: the entries oriented from
// the lower- to the higher-ranked vertex by (degree, id); an INT32 matrix of
// ones.
int tri_part(gbk_matrix A, int mode, gbk_matrix out) {
    hipStream_t st = stream();
    const DevCSR& m = A->csr;
    const int64_t n = A->n, nnz = A->nnz;
    const int64_t z1 = nnz > 0 ? nnz : 1;
    DevBuf rowid, keep, pos;
    GBK_TRY(rowid.alloc(4 * z1));
    GBK_TRY(keep.alloc(4 * (z1 + 1)));
    GBK_TRY(pos.alloc(4 * (z1 + 1)));
    GBK_TRY(dev_row_ids(m.row_ptr, n, rowid.as<int32_t>(), st));
    int32_t kept = 0;
    if (nnz > 0) {
        hipLaunchKernelGGL(k_tri_flags, grid_m(nnz), dim3(256), 0, st, rowid.as<int32_t>(), m.col, nnz, mode,
                           m.row_ptr, keep.as<int32_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(exclusive_scan_i32(keep.as<int32_t>(), pos.as<int32_t>(), nnz, &kept));  // pos[0..nnz]
    } else {
        GBK_HIP(hipMemsetAsync(pos.p, 0, 4, st));
    }
    int32_t *rp = nullptr, *col = nullptr;
    void* val = nullptr;
    GBK_TRY(dev_alloc((void**)&rp, 4 * (n + 1)));
    int rc = dev_alloc((void**)&col, 4 * (size_t)(kept > 0 ? kept : 1));
    if (rc == GBK_SUCCESS) rc = dev_alloc(&val, 4 * (size_t)(kept > 0 ? kept : 1));
    if (rc != GBK_SUCCESS) {
        dev_free(rp);
        dev_free(col);
        dev_free(val);
        return rc;
    }
    hipLaunchKernelGGL(k_tri_rowptr, grid_m(n + 1), dim3(256), 0, st, m.row_ptr, pos.as<int32_t>(), n, rp);
    if (nnz > 0)
        hipLaunchKernelGGL(k_tri_compact, grid_m(nnz), dim3(256), 0, st, keep.as<int32_t>(), pos.as<int32_t>(), m.col,
                           nnz, col);
    GBK_TRY(fill_value(GBK_INT32, val, kept, 1.0));
    GBK_LAUNCH_CHECK();
    GBK_HIP(hipStreamSynchronize(st));
    out->type = GBK_INT32;
    return build_from_device_csr(out, n, kept, rp, col, val);
}

// Sum over the entries of C<M> = A (plus.one) B, without forming C; every
// operand is given by its CSR (a transpose by the other orientation).
int masked_count(const DevCSR& M, const DevCSR& A, const DevCSR& B, uint64_t* out) {
    hipStream_t st = stream();
    DevBuf scol, perm, arow, tot;
    GBK_TRY(sort_rows(M, scol, perm, st));
    const int64_t a_nnz = A.nnz;
    GBK_TRY(arow.alloc(4 * (a_nnz > 0 ? a_nnz : 1)));
    GBK_TRY(dev_row_ids(A.row_ptr, A.nrows, arow.as<int32_t>(), st));
    GBK_TRY(tot.alloc(8));
    GBK_HIP(hipMemsetAsync(tot.p, 0, 8, st));
    if (a_nnz > 0) {
        const int64_t blocks = std::min<int64_t>((a_nnz + 15) / 16, (int64_t)device_cu_count() * 64);
        hipLaunchKernelGGL(k_mxm_count, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, st,
                           arow.as<int32_t>(), A.col, a_nnz, B.row_ptr, B.col, M.row_ptr, scol.as<int32_t>(),
                           tot.as<unsigned long long>());
        GBK_LAUNCH_CHECK();
    }
    unsigned long long h = 0;
    GBK_HIP(hipMemcpyAsync(&h, tot.p, 8, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    *out = h;
    return GBK_SUCCESS;
}

// ------------------------------------------------------- unmasked product --
// C = A (+.*) B without a mask (SpMSpM_unmasked_ikj,
// src/backend/spmspm/spmspm_hash_based_gustavson.h:9-153): expand, sort,
// compress.  Every product A(i,k) B(k,j) is written at a position in IKJ
// order (A's entries in row order, then row k of B in order) with the key
// i n + j; a stable radix sort of the keys keeps that order within a key, so
// one thread folding a key's products from the identity forms the
// reference's C(i,j) = add(... add(identity, p1) ..., pm) bit for bit;
// entries whose value is 0 are dropped, as the reference does.  Columns come
// out ascending (the reference writes them in hash-map order).

__global__ void k_prod_count(const int32_t* __restrict__ acol, int64_t annz, const int32_t* __restrict__ brp,
                             int64_t* __restrict__ cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < annz) cnt[e] = brp[acol[e] + 1] - brp[acol[e]];
}

// A group of MXM_G lanes per entry of A writes its row of B's products.
template <typename T, int MUL>
__global__ __launch_bounds__(256) void k_expand(const int32_t* __restrict__ arow, const int32_t* __restrict__ acol,
                                                const T* __restrict__ aval, int64_t annz,
                                                const int32_t* __restrict__ brp, const int32_t* __restrict__ bcol,
                                                const T* __restrict__ bval, const int64_t* __restrict__ offs,
                                                int64_t n, uint64_t* __restrict__ keys, T* __restrict__ prod) {
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / MXM_G;
    const int q = threadIdx.x % MXM_G;
    if (g >= annz) return;
    const uint64_t rowkey = (uint64_t)arow[g] * (uint64_t)n;
    const int k = acol[g];
    const T a = aval[g];
    const int b0 = brp[k], len = brp[k + 1] - b0;
    const int64_t o = offs[g];
    for (int t = q; t < len; t += MXM_G) {
        keys[o + t] = rowkey + (uint64_t)bcol[b0 + t];
        prod[o + t] = mul_op<MUL, T>(a, bval[b0 + t]);
    }
}

__global__ void k_key_starts(const uint64_t* __restrict__ skeys, int64_t F, int32_t* __restrict__ flag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < F) flag[p] = (p == 0 || skeys[p] != skeys[p - 1]) ? 1 : 0;
}

__global__ void k_run_starts(const int32_t* __restrict__ flag, const int32_t* __restrict__ uidx, int64_t F,
                             int64_t U, int32_t* __restrict__ start) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < F && flag[p]) start[uidx[p]] = (int32_t)p;
    if (p == 0) start[U] = (int32_t)F;
}

// One thread per distinct key: the left fold of its products in sorted
// (= IKJ) order; keep[u] = (sum != 0).
template <typename T, int ADD>
__global__ void k_run_fold(const uint64_t* __restrict__ skeys, const int32_t* __restrict__ perm,
                           const T* __restrict__ prod, const int32_t* __restrict__ start, int64_t U, int64_t n,
                           int32_t* __restrict__ urow, int32_t* __restrict__ ucol, T* __restrict__ uval,
                           int32_t* __restrict__ keep) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U) return;
    const int32_t p0 = start[u], p1 = start[u + 1];
    T acc = add_identity<ADD, T>();
    for (int32_t p = p0; p < p1; ++p) acc = add_op<ADD, T>(acc, prod[perm[p]]);
    const uint64_t key = skeys[p0];
    urow[u] = (int32_t)(key / (uint64_t)n);
    ucol[u] = (int32_t)(key % (uint64_t)n);
    uval[u] = acc;
    keep[u] = acc != T(0) ? 1 : 0;
}

template <typename T>
__global__ void k_keep_compact(const int32_t* __restrict__ keep, const int32_t* __restrict__ kidx, int64_t U,
                               const int32_t* __restrict__ urow, const int32_t* __restrict__ ucol,
                               const T* __restrict__ uval, int32_t* __restrict__ crow, int32_t* __restrict__ ccol,
                               T* __restrict__ cval) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U || !keep[u]) return;
    const int32_t o = kidx[u];
    crow[o] = urow[u];
    ccol[o] = ucol[u];
    cval[o] = uval[u];
}

// C's CSR arrays (allocated here, handed to the caller) of A (+.*) B.
int unmasked_product(const DevCSR& A, const DevCSR& B, int type, int add, int mul, int64_t n, int32_t** crp_out,
                     int32_t** ccol_out, void** cval_out, int64_t* nnz_out) {
    hipStream_t st = stream();
    const size_t vs = type_size(type);
    const int64_t annz = A.nnz;
    DevBuf cnt, offs, arow;
    GBK_TRY(cnt.alloc(8 * (annz + 1)));
    GBK_TRY(offs.alloc(8 * (annz + 1)));
    GBK_TRY(arow.alloc(4 * (annz > 0 ? annz : 1)));
    GBK_HIP(hipMemsetAsync(cnt.as<int64_t>() + annz, 0, 8, st));
    if (annz > 0) {
        hipLaunchKernelGGL(k_prod_count, grid_m(annz), dim3(256), 0, st, A.col, annz, B.row_ptr, cnt.as<int64_t>());
        GBK_LAUNCH_CHECK();
    }
    GBK_TRY(dev_exclusive_scan_i64(cnt.as<int64_t>(), offs.as<int64_t>(), annz + 1, st));
    int64_t F = 0;
    GBK_HIP(hipMemcpyAsync(&F, offs.as<int64_t>() + annz, 8, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    if (F >= (int64_t)INT32_MAX) return fail(GBK_INVALID_VALUE, "mxm: %lld products exceed 2^31", (long long)F);
    if (n > 0 && (uint64_t)n > (~0ull) / (uint64_t)n) return fail(GBK_INVALID_VALUE, "mxm: dimension too large");
    GBK_TRY(dev_row_ids(A.row_ptr, n, arow.as<int32_t>(), st));
    const int64_t F1 = F > 0 ? F : 1;
    DevBuf keys, skeys, prod, iota, perm, flag, uidx;
    GBK_TRY(keys.alloc(8 * F1));
    GBK_TRY(skeys.alloc(8 * F1));
    GBK_TRY(prod.alloc(vs * F1));
    GBK_TRY(iota.alloc(4 * F1));
    GBK_TRY(perm.alloc(4 * F1));
    GBK_TRY(flag.alloc(4 * F1));
    GBK_TRY(uidx.alloc(4 * (F1 + 1)));
    int rc = dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int MUL = decltype(b)::value;
            (void)a;
            if (annz > 0 && F > 0) {
                hipLaunchKernelGGL((k_expand<T, MUL>), grid_m(annz * MXM_G), dim3(256), 0, st, arow.as<int32_t>(),
                                   A.col, (const T*)A.val, annz, B.row_ptr, B.col, (const T*)B.val,
                                   offs.as<int64_t>(), n, keys.as<uint64_t>(), prod.as<T>());
                GBK_LAUNCH_CHECK();
            }
            return GBK_SUCCESS;
        });
    });
    GBK_TRY(rc);
    GBK_TRY(dev_iota(iota.as<int32_t>(), F, st));
    const uint64_t maxkey = n > 0 ? (uint64_t)n * (uint64_t)n - 1 : 0;
    int kb = 1;
    while (kb < 64 && (maxkey >> kb) != 0) ++kb;
    GBK_TRY(dev_sort_pairs64(keys.as<uint64_t>(), skeys.as<uint64_t>(), iota.as<int32_t>(), perm.as<int32_t>(), F,
                             kb, st));
    int32_t U = 0;
    if (F > 0) {
        hipLaunchKernelGGL(k_key_starts, grid_m(F), dim3(256), 0, st, skeys.as<uint64_t>(), F, flag.as<int32_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(exclusive_scan_i32(flag.as<int32_t>(), uidx.as<int32_t>(), F, &U));
    }
    const int64_t U1 = U > 0 ? U : 1;
    DevBuf start, urow, ucol, uval, keep, kidx, crow;
    GBK_TRY(start.alloc(4 * (U1 + 1)));
    GBK_TRY(urow.alloc(4 * U1));
    GBK_TRY(ucol.alloc(4 * U1));
    GBK_TRY(uval.alloc(vs * U1));
    GBK_TRY(keep.alloc(4 * U1));
    GBK_TRY(kidx.alloc(4 * (U1 + 1)));
    int32_t nz = 0;
    int32_t *crp = nullptr, *ccol = nullptr;
    void* cval = nullptr;
    rc = dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value;
            (void)b;
            if (U > 0) {
                hipLaunchKernelGGL(k_run_starts, grid_m(F), dim3(256), 0, st, flag.as<int32_t>(), uidx.as<int32_t>(),
                                   F, (int64_t)U, start.as<int32_t>());
                GBK_LAUNCH_CHECK();
                hipLaunchKernelGGL((k_run_fold<T, ADD>), grid_m(U), dim3(256), 0, st, skeys.as<uint64_t>(),
                                   perm.as<int32_t>(), prod.as<T>(), start.as<int32_t>(), (int64_t)U, n,
                                   urow.as<int32_t>(), ucol.as<int32_t>(), uval.as<T>(), keep.as<int32_t>());
                GBK_LAUNCH_CHECK();
                GBK_TRY(exclusive_scan_i32(keep.as<int32_t>(), kidx.as<int32_t>(), U, &nz));
            }
            const int64_t nz1 = nz > 0 ? nz : 1;
            GBK_TRY(crow.alloc(4 * nz1));
            GBK_TRY(dev_alloc((void**)&crp, 4 * (n + 1)));
            GBK_TRY(dev_alloc((void**)&ccol, 4 * nz1));
            GBK_TRY(dev_alloc(&cval, vs * nz1));
            if (nz > 0) {
                hipLaunchKernelGGL(k_keep_compact<T>, grid_m(U), dim3(256), 0, st, keep.as<int32_t>(),
                                   kidx.as<int32_t>(), (int64_t)U, urow.as<int32_t>(), ucol.as<int32_t>(),
                                   uval.as<T>(), crow.as<int32_t>(), ccol, (T*)cval);
                GBK_LAUNCH_CHECK();
            }
            return dev_bounds_sorted((const uint32_t*)crow.p, nz, n, crp, st);
        });
    });
    if (rc == GBK_SUCCESS && hipStreamSynchronize(st) != hipSuccess) rc = fail(GBK_PANIC, "sync failed");
    if (rc != GBK_SUCCESS) {
        dev_free(crp);
        dev_free(ccol);
        dev_free(cval);
        return rc;
    }
    *crp_out = crp;
    *ccol_out = ccol;
    *cval_out = cval;
    *nnz_out = nz;
    return GBK_SUCCESS;
}

}  // namespace

}  // namespace gbk

using namespace gbk;

static int check_built(gbk_matrix A) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    return ensure_init();
}

extern "C" int gbk_mxm(gbk_matrix C, gbk_matrix M, int accum, gbk_semiring op, gbk_matrix A, gbk_matrix B,
                       const gbk_desc* desc) {
    (void)accum;  // the reference rebuilds C from the product (operations.h:436-486)
    if (!C) return fail(GBK_UNINITIALIZED_OBJECT, "null result matrix");
    GBK_TRY(check_built(A));
    GBK_TRY(check_built(B));
    if (desc && (desc->mask == GBK_COMP || desc->mask == GBK_STR_COMP))
        return fail(GBK_INVALID_VALUE, "Error: complementary mask is not supported yet");
    if (!M) {
        if (A->n != B->n) return fail(GBK_DIMENSION_MISMATCH, "mxm operands differ in size");
        if (A->type != B->type) return fail(GBK_DOMAIN_MISMATCH, "mxm operands differ in type");
        if (C == A || C == B) return fail(GBK_INVALID_VALUE, "mxm result aliases an operand");
        int32_t *crp = nullptr, *ccol = nullptr;
        void* cval = nullptr;
        int64_t nz = 0;
        GBK_TRY(unmasked_product(A->csr, B->csr, A->type, op.add, op.mul, A->n, &crp, &ccol, &cval, &nz));
        C->type = A->type;
        return build_from_device_csr(C, A->n, nz, crp, ccol, cval);
    }
    GBK_TRY(check_built(M));
    if (A->n != B->n || A->n != M->n) return fail(GBK_DIMENSION_MISMATCH, "mxm operands differ in size");
    if (A->type != B->type) return fail(GBK_DOMAIN_MISMATCH, "mxm operands differ in type");
    if (C == A || C == B || C == M) return fail(GBK_INVALID_VALUE, "mxm result aliases an operand");
    hipStream_t st = stream();
    const int type = A->type;
    const size_t vs = type_size(type);
    const int64_t n = M->n, m = M->nnz;
    DevBuf scol, perm, arow, acc, touched;
    GBK_TRY(sort_rows(M->csr, scol, perm, st));
    GBK_TRY(arow.alloc(4 * (A->nnz > 0 ? A->nnz : 1)));
    GBK_TRY(dev_row_ids(A->csr.row_ptr, A->n, arow.as<int32_t>(), st));
    GBK_TRY(acc.alloc(vs * (m > 0 ? m : 1)));
    GBK_TRY(touched.alloc(m > 0 ? m : 1));
    // C's arrays: the mask's structure, values in its entry order
    int32_t *crp = nullptr, *ccol = nullptr;
    void* cval = nullptr;
    GBK_TRY(dev_alloc((void**)&crp, 4 * (n + 1)));
    int rc = dev_alloc((void**)&ccol, 4 * (size_t)(m > 0 ? m : 1));
    if (rc == GBK_SUCCESS) rc = dev_alloc(&cval, vs * (size_t)(m > 0 ? m : 1));
    if (rc == GBK_SUCCESS) {
        rc = dispatch_type(type, [&](auto tv) -> int {
            typedef decltype(tv) T;
            return dispatch_semiring(op.add, op.mul, [&](auto a, auto b) -> int {
                constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
                if (m > 0) {
                    hipLaunchKernelGGL((k_fill_identity<T, ADD>), grid_m(m), dim3(256), 0, st, acc.as<T>(),
                                       touched.as<uint8_t>(), m);
                    GBK_LAUNCH_CHECK();
                }
                if (A->nnz > 0 && m > 0) {
                    const int64_t blocks = std::min<int64_t>((A->nnz + 15) / 16, (int64_t)device_cu_count() * 64);
                    hipLaunchKernelGGL((k_mxm_masked<T, ADD, MUL>), dim3((unsigned)blocks), dim3(256), 0, st,
                                       arow.as<int32_t>(), A->csr.col, (const T*)A->csr.val, A->nnz, B->csr.row_ptr,
                                       B->csr.col, (const T*)B->csr.val, M->csr.row_ptr, scol.as<int32_t>(),
                                       acc.as<T>(), touched.as<uint8_t>());
                    GBK_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_unsort_vals<T>, grid_m(m), dim3(256), 0, st, acc.as<T>(),
                                       touched.as<uint8_t>(), perm.as<int32_t>(), m, (T*)cval);
                    GBK_LAUNCH_CHECK();
                }
                return GBK_SUCCESS;
            });
        });
    }
    if (rc == GBK_SUCCESS && hipMemcpyAsync(crp, M->csr.row_ptr, 4 * (n + 1), hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = fail(GBK_PANIC, "hipMemcpyAsync failed");
    if (rc == GBK_SUCCESS && m > 0 &&
        hipMemcpyAsync(ccol, M->csr.col, 4 * m, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = fail(GBK_PANIC, "hipMemcpyAsync failed");
    if (rc == GBK_SUCCESS && hipStreamSynchronize(st) != hipSuccess) rc = fail(GBK_PANIC, "sync failed");
    if (rc != GBK_SUCCESS) {
        dev_free(crp);
        dev_free(ccol);
        dev_free(cval);
        return rc;
    }
    C->type = type;
    return build_from_device_csr(C, n, m, crp, ccol, cval);
}

extern "C" int gbk_tricount(gbk_matrix A, int method, uint64_t* ntriangles) {
    GBK_TRY(check_built(A));
    if (!ntriangles) return fail(GBK_NULL_POINTER, "null result");
    if (method < 0 || method > 6) return fail(GBK_INVALID_VALUE, "unknown triangle count method %d", method);
    uint64_t s = 0;
    if (method == 0) {
        // default: Sandia over the degree orientation D (each edge kept from
        // its lower- to its higher-ranked end): sum((D D) .* D) counts every
        // triangle once, and no vertex keeps more out-edges than about
        // sqrt(2 nnz), so hub rows stay short
        gbk_matrix D = nullptr;
        int rc = gbk_matrix_new(&D, GBK_INT32);
        if (rc == GBK_SUCCESS) rc = tri_part(A, 2, D);
        if (rc == GBK_SUCCESS) rc = masked_count(D->csr, D->csr, D->csr, &s);
        gbk_matrix_free(D);
        GBK_TRY(rc);
        *ntriangles = s;
        return GBK_SUCCESS;
    }
    if (method == 1) {
        // sum((A^2) .* A) / 6
        GBK_TRY(masked_count(A->csr, A->csr, A->csr, &s));
        *ntriangles = s / 6;
        return GBK_SUCCESS;
    }
    gbk_matrix L = nullptr, U = nullptr;
    int rc = gbk_matrix_new(&L, GBK_INT32);
    if (rc == GBK_SUCCESS) rc = gbk_matrix_new(&U, GBK_INT32);
    if (rc == GBK_SUCCESS) rc = tri_part(A, 0, L);
    if (rc == GBK_SUCCESS) rc = tri_part(A, 1, U);
    if (rc == GBK_SUCCESS) {
        switch (method) {
            case 2: rc = masked_count(A->csr, L->csr, U->csr, &s); s /= 2; break;  // sum((L U) .* A) / 2
            case 3: rc = masked_count(L->csr, L->csr, L->csr, &s); break;          // sum((L L) .* L)
            case 4: rc = masked_count(U->csr, U->csr, U->csr, &s); break;          // sum((U U) .* U)
            case 5: rc = masked_count(L->csr, L->csr, U->csc, &s); break;          // sum((L U') .* L)
            default: rc = masked_count(U->csr, U->csr, L->csc, &s); break;         // sum((U L') .* U)
        }
    }
    gbk_matrix_free(L);
    gbk_matrix_free(U);
    GBK_TRY(rc);
    *ntriangles = s;
    return GBK_SUCCESS;
}
