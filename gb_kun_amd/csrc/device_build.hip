// Device-side construction steps shared by the matrix build and the format
// builders: stable key sorts (rocPRIM radix sort, a library routine of ROCm),
// row ids of CSR entries, byte gathers, and the CSR -> CSC transpose.
//
// The transpose replaces csr_to_csc / transpose_parallel
// (src/backend/matrix/transpose.hpp:69-206).  The reference builds the CSC
// with an atomic fetch-add placement, so its intra-column order is
// unspecified; here the entries of a column are in ascending row order (CSR
// order, a stable sort by column), which makes the CSC deterministic and
// identical to the host counting-sort transpose.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "gbk_objects.h"

namespace gbk {

namespace {

dim3 grid_n(int64_t n, int per = 256) {
    const int64_t b = (n + per - 1) / per;
    return dim3((unsigned)(b > 0 ? b : 1));
}

__global__ void k_iota(int32_t* __restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

// rowid[e] = r for rp[r] <= e < rp[r + 1].  A workgroup takes 256
// consecutive rows, whose entries are contiguous, and walks those entries
// together (each finds its row by a binary search over the row pointers in
// LDS), so a row of many entries is not left to one thread.
__global__ __launch_bounds__(256) void k_row_ids(const int32_t* __restrict__ rp, int64_t n,
                                                 int32_t* __restrict__ rowid) {
    __shared__ int32_t srp[257];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int rows = (int)min((int64_t)256, n - r0);
    for (int i = threadIdx.x; i <= rows; i += 256) srp[i] = rp[r0 + i];
    __syncthreads();
    const int32_t e1 = srp[rows];
    for (int32_t e = srp[0] + (int32_t)threadIdx.x; e < e1; e += 256) {
        int lo = 0, hi = rows - 1;  // the last row starting at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (srp[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        rowid[e] = (int32_t)(r0 + lo);
    }
}

// ptr[b] = first i with key[i] >= b (keys ascending), b in [0, nbins].
__global__ void k_lower_bounds(const uint32_t* __restrict__ key, int64_t m, int64_t nbins, int32_t* __restrict__ ptr) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbins) return;
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)key[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    ptr[b] = (int32_t)lo;
}

__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ idx, int64_t m,
                             int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = src[idx[i]];
}

template <typename V>
__global__ void k_gather_t(const V* __restrict__ src, const int32_t* __restrict__ idx, int64_t m, V* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = src[idx[i]];
}

__global__ void k_max_diff(const int32_t* __restrict__ p, int64_t n, int32_t* __restrict__ mx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t v = 0;
    if (i < n) v = p[i + 1] - p[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(mx, v);
}

}  // namespace

int bits_for(int64_t maxval) {
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= maxval) ++b;
    return b;
}

// Stable sort of (key, value) pairs by the low `bits` bits of the key.
int dev_sort_pairs(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t m, int bits,
                   hipStream_t st) {
    if (m <= 0) return GBK_SUCCESS;
    size_t tmp_bytes = 0;
    GBK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, (size_t)m, 0, bits, st));
    void* tmp = nullptr;
    GBK_TRY(dev_alloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1));
    const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (size_t)m, 0, bits, st);
    const hipError_t e2 = hipStreamSynchronize(st);
    dev_free(tmp);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(GBK_PANIC, "radix sort failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    return GBK_SUCCESS;
}

// The same with 64-bit keys (bits up to 64).
int dev_sort_pairs64(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t m, int bits,
                     hipStream_t st) {
    if (m <= 0) return GBK_SUCCESS;
    size_t tmp_bytes = 0;
    GBK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, (size_t)m, 0, bits, st));
    void* tmp = nullptr;
    GBK_TRY(dev_alloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1));
    const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (size_t)m, 0, bits, st);
    const hipError_t e2 = hipStreamSynchronize(st);
    dev_free(tmp);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(GBK_PANIC, "radix sort failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    return GBK_SUCCESS;
}

// out[i] = sum_{j<i} in[j] over m int64 elements (rocPRIM scan).
int dev_exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t m, hipStream_t st) {
    if (m <= 0) return GBK_SUCCESS;
    size_t tmp_bytes = 0;
    GBK_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, in, out, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(), st));
    void* tmp = nullptr;
    GBK_TRY(dev_alloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1));
    const hipError_t e =
        rocprim::exclusive_scan(tmp, tmp_bytes, in, out, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(), st);
    const hipError_t e2 = hipStreamSynchronize(st);
    dev_free(tmp);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(GBK_PANIC, "scan failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    return GBK_SUCCESS;
}

int dev_iota(int32_t* v, int64_t n, hipStream_t st) {
    if (n <= 0) return GBK_SUCCESS;
    hipLaunchKernelGGL(k_iota, grid_n(n), dim3(256), 0, st, v, n);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

int dev_row_ids(const int32_t* rp, int64_t n, int32_t* rowid, hipStream_t st) {
    if (n <= 0) return GBK_SUCCESS;
    hipLaunchKernelGGL(k_row_ids, grid_n(n), dim3(256), 0, st, rp, n, rowid);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

// Bucket pointers of ascending keys: ptr[b] = number of keys below b for b in
// [0, nbins] (one binary search per bucket: no atomics, so heavily repeated
// keys cost nothing extra).
int dev_bounds_sorted(const uint32_t* key, int64_t m, int64_t nbins, int32_t* ptr, hipStream_t st) {
    hipLaunchKernelGGL(k_lower_bounds, grid_n(nbins + 1), dim3(256), 0, st, key, m, nbins, ptr);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

int dev_gather_i32(const int32_t* src, const int32_t* idx, int64_t m, int32_t* out, hipStream_t st) {
    if (m <= 0) return GBK_SUCCESS;
    hipLaunchKernelGGL(k_gather_i32, grid_n(m), dim3(256), 0, st, src, idx, m, out);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

// out[i] = src[idx[i]] for elements of vs bytes (1, 4 or 8).
int dev_gather_bytes(const void* src, const int32_t* idx, int64_t m, size_t vs, void* out, hipStream_t st) {
    if (m <= 0) return GBK_SUCCESS;
    if (vs == 8)
        hipLaunchKernelGGL(k_gather_t<uint64_t>, grid_n(m), dim3(256), 0, st, (const uint64_t*)src, idx, m,
                           (uint64_t*)out);
    else if (vs == 4)
        hipLaunchKernelGGL(k_gather_t<uint32_t>, grid_n(m), dim3(256), 0, st, (const uint32_t*)src, idx, m,
                           (uint32_t*)out);
    else if (vs == 1)
        hipLaunchKernelGGL(k_gather_t<uint8_t>, grid_n(m), dim3(256), 0, st, (const uint8_t*)src, idx, m,
                           (uint8_t*)out);
    else
        return fail(GBK_INVALID_VALUE, "unsupported value size %zu", vs);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

// CSC of the n x n CSR `in` (device arrays of `out` are allocated here).
int device_transpose(const DevCSR& in, int type, int64_t n, DevCSR& out) {
    hipStream_t st = stream();
    const size_t vs = type_size(type);
    const int64_t nnz = in.nnz;
    out.nrows = out.ncols = n;
    out.nnz = nnz;
    GBK_TRY(dev_alloc((void**)&out.row_ptr, sizeof(int32_t) * (n + 1)));
    GBK_TRY(dev_alloc((void**)&out.col, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
    GBK_TRY(dev_alloc(&out.val, vs * (nnz > 0 ? nnz : 1)));
    // stable sort of entry ids by column: rows ascend inside every column;
    // the column pointers are the bucket bounds of the sorted columns
    int32_t *perm_in = nullptr, *perm = nullptr, *rowid = nullptr;
    uint32_t* keys_out = nullptr;
    const int64_t m1 = nnz > 0 ? nnz : 1;
    int rc = dev_alloc((void**)&perm_in, sizeof(int32_t) * m1);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&perm, sizeof(int32_t) * m1);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&keys_out, sizeof(uint32_t) * m1);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&rowid, sizeof(int32_t) * m1);
    if (rc == GBK_SUCCESS) rc = dev_iota(perm_in, nnz, st);
    if (rc == GBK_SUCCESS)
        rc = dev_sort_pairs((const uint32_t*)in.col, keys_out, perm_in, perm, nnz, bits_for(n > 0 ? n - 1 : 0), st);
    if (rc == GBK_SUCCESS) rc = dev_bounds_sorted(keys_out, nnz, n, out.row_ptr, st);
    if (rc == GBK_SUCCESS) rc = dev_row_ids(in.row_ptr, n, rowid, st);
    if (rc == GBK_SUCCESS) rc = dev_gather_i32(rowid, perm, nnz, out.col, st);
    if (rc == GBK_SUCCESS) rc = dev_gather_bytes(in.val, perm, nnz, vs, out.val, st);
    int32_t* mx = nullptr;
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&mx, sizeof(int32_t));
    if (rc == GBK_SUCCESS && hipMemsetAsync(mx, 0, sizeof(int32_t), st) != hipSuccess) rc = fail(GBK_PANIC, "memset");
    if (rc == GBK_SUCCESS && n > 0) {
        hipLaunchKernelGGL(k_max_diff, grid_n(n), dim3(256), 0, st, out.row_ptr, n, mx);
        if (hipGetLastError() != hipSuccess) rc = fail(GBK_PANIC, "kernel launch failed");
    }
    int32_t mxh = 0;
    if (rc == GBK_SUCCESS && hipMemcpyAsync(&mxh, mx, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(GBK_PANIC, "copy");
    if (rc == GBK_SUCCESS && hipStreamSynchronize(st) != hipSuccess) rc = fail(GBK_PANIC, "sync");
    dev_free(perm_in);
    dev_free(perm);
    dev_free(keys_out);
    dev_free(rowid);
    dev_free(mx);
    GBK_TRY(rc);
    out.max_row = mxh;
    return build_merge_tiles(out, type);
}

}  // namespace gbk
