// Device-resident matrix and vector objects and the internal entry points
// the C ABI is built from.
//
// Layout in HBM (per matrix, per orientation — CSR rows and CSC columns):
//   row_ptr  int32[n+1]        nnz < 2^31 is required
//   col      int32[nnz]
//   val      T[nnz]
//   tiles    int2[ntiles+1]    merge-path start coordinates (row, nz) of
//                              every 2048-item tile of the SpMV
//   carry    T[2*ntiles]       per-tile partial sums of rows that cross tiles
// plus the optional optimized representations (SELL-C-sigma, segmented CSR,
// LAV) and a workspace of n-sized scratch vectors.
#pragma once

#include "gbk_common.h"

namespace gbk {

struct SellC;
struct SegCSR;
struct Lav;

struct DevCSR {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int32_t* row_ptr = nullptr;
    int32_t* col = nullptr;
    void* val = nullptr;
    // merge-path tiling
    int64_t ntiles = 0;
    int2* tiles = nullptr;
    void* carry = nullptr;  // 2*ntiles elements of the value type (tail, head)
    // row-length statistics (for kernel selection)
    int32_t max_row = 0;
    // merge-path start coordinates of the SpMM lane-group tiles (built on
    // first multi-vector use)
    int64_t spmm_ntiles = 0;
    int2* spmm_tiles = nullptr;
};

struct Workspace {
    void* buf[4] = {nullptr, nullptr, nullptr, nullptr};  // n * 8 bytes each
    size_t bytes = 0;
};

}  // namespace gbk

struct gbk_matrix_s {
    int type = GBK_FP32;
    int64_t n = 0;    // square: nrows == ncols
    int64_t nnz = 0;
    bool built = false;
    int format = GBK_FMT_CSR;
    int sell_c = 64, sigma = 512;
    gbk::DevCSR csr, csc;
    gbk::SellC* sell[2] = {nullptr, nullptr};
    gbk::SegCSR* seg[2] = {nullptr, nullptr};
    gbk::Lav* lav[2] = {nullptr, nullptr};
    gbk::Workspace ws;
    // vertices with more in-edges than the BFS pull kernel gives one lane,
    // listed on the first BFS (bfs_nheavy < 0: not yet)
    int32_t* bfs_heavy = nullptr;
    int64_t bfs_nheavy = -1;
};

struct gbk_vector_s {
    int type = GBK_FP32;
    int64_t n = 0;
    int storage = GBK_STORAGE_SPARSE;
    void* dense = nullptr;     // T[n]
    int32_t* ids = nullptr;    // int32[n] (sparse ids, ascending after compaction)
    void* svals = nullptr;     // T[n]
    int64_t nvals = 0;         // entries of the sparse form
};

namespace gbk {

// Owner lookup for edge-balanced expansions over a queue: offs[0..nq] are the
// exclusive prefix sums of the queue entries' degrees (offs[nq] = total), and
// thread e of the grid handles edge e.  One thread binary-searches the global
// offsets for the block's first edge; the block stages the next EO_WIN + 1
// offsets in LDS and every thread searches there, falling back to the global
// array when the block's edges span more entries than the window.  Every
// thread of the block must call it (it synchronizes the block).
constexpr int EO_WIN = 256;

__device__ __forceinline__ int64_t edge_owner(const int32_t* __restrict__ offs, int64_t nq, int64_t e,
                                              int32_t* s_offs, int64_t* s_i0) {
    const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
    if (threadIdx.x == 0) {
        int64_t lo = 0, hi = nq - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= e0) lo = mid;
            else hi = mid - 1;
        }
        *s_i0 = lo;
    }
    __syncthreads();
    const int64_t i0 = *s_i0;
    const int64_t wn = (nq - i0 < EO_WIN) ? nq - i0 : EO_WIN;
    for (int64_t j = threadIdx.x; j <= wn; j += blockDim.x) s_offs[j] = offs[i0 + j];
    __syncthreads();
    int64_t lo, hi;
    if (e < s_offs[wn]) {
        lo = 0;
        hi = wn - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (s_offs[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        return i0 + lo;
    }
    lo = i0;
    hi = nq - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// runtime.hip
int ensure_init();
hipStream_t stream();
void set_thread_stream(hipStream_t s);
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
int device_cu_count();
int64_t device_lds_bytes();  // LDS a workgroup may allocate
// scratch buffers owned by the library (grown on demand, reused across calls)
int scratch(int slot, size_t bytes, void** p);

// scans / compaction (runtime.hip)
// out[i] = sum_{j<i} in[j] for i in [0, n]; returns total through *total
// (host) if total != nullptr.  in may be nullptr with deg given by row_ptr
// differences of the frontier (see spmspv).
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* total_host);
// Dense -> sparse: ids/vals of nonzero entries, ascending.  Returns count.
int compact_nonzero(int type, const void* dense, int64_t n, int32_t* ids, void* vals, int64_t* count);
int count_nonzero(int type, const void* dense, int64_t n, int64_t* count);
int fill_value(int type, void* dst, int64_t n, double value);
int fill_identity(int type, int add, void* dst, int64_t n);
int scatter_sparse(int type, const int32_t* ids, const void* vals, int64_t nvals, void* dense);

// vector state (vector.hip)
int vec_force_dense(gbk_vector v);
int vec_force_sparse(gbk_vector v);
int vec_convert_if_required(gbk_vector v);
int vec_nvals(gbk_vector v, int64_t* nvals);

// spmv.hip
int spmv_unmasked(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum);
int spmv_masked_dense(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum,
                      const void* mask, int mask_type, bool comp);
int spmv_masked_rows(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum,
                     const int32_t* rows, int64_t nrows_list);
int build_merge_tiles(DevCSR& m, int type);

// Device construction steps (device_build.hip).
int bits_for(int64_t maxval);
int dev_sort_pairs(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t m, int bits,
                   hipStream_t st);
int dev_sort_pairs64(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t m, int bits,
                     hipStream_t st);
int dev_exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t m, hipStream_t st);
int dev_iota(int32_t* v, int64_t n, hipStream_t st);
int dev_row_ids(const int32_t* rp, int64_t n, int32_t* rowid, hipStream_t st);
int dev_bounds_sorted(const uint32_t* key, int64_t m, int64_t nbins, int32_t* ptr, hipStream_t st);
int dev_gather_i32(const int32_t* src, const int32_t* idx, int64_t m, int32_t* out, hipStream_t st);
int dev_gather_bytes(const void* src, const int32_t* idx, int64_t m, size_t vs, void* out, hipStream_t st);
int device_transpose(const DevCSR& in, int type, int64_t n, DevCSR& out);
int build_from_device_csr(gbk_matrix A, int64_t n, int64_t nnz, int32_t* rp, int32_t* col, void* val);

// spmspv.hip
int spmspv(gbk_matrix A, bool use_csr, const int32_t* x_ids, const void* x_vals, int64_t x_nvals,
           void* y, int add, int mul, int accum, gbk_vector mask, int mask_mode);

// formats (formats.hip)
int format_build(gbk_matrix A);
void format_free(gbk_matrix A);
int format_spmv(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum, bool* handled);
int64_t format_bytes(gbk_matrix A);

}  // namespace gbk
