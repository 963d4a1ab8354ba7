// gbk_matrix: device-resident CSR + CSC (+ optional optimized format).
//
// Replaces lablas::Matrix<T> / backend::Matrix<T>
// (src/cpp_graphblas/matrix.hpp:11-144, src/backend/matrix/matrix.h:34-207).
// Like the reference, a matrix always keeps both orientations: mxv with a
// dense vector pulls over CSR rows, vxm over CSC rows, and SpMSpV pushes over
// the other one (spmspv.h:126-146).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>
#include <omp.h>

#include "gbk_objects.h"
#include "host_build.h"

namespace gbk {

static void free_csr(DevCSR& m) {
    dev_free(m.row_ptr);
    dev_free(m.col);
    dev_free(m.val);
    dev_free(m.tiles);
    dev_free(m.carry);
    dev_free(m.spmm_tiles);
    m = DevCSR();
}

static void free_matrix_storage(gbk_matrix A) {
    format_free(A);
    free_csr(A->csr);
    free_csr(A->csc);
    for (auto& b : A->ws.buf) { dev_free(b); b = nullptr; }
    A->ws.bytes = 0;
    dev_free(A->bfs_heavy);
    A->bfs_heavy = nullptr;
    A->bfs_nheavy = -1;
    A->built = false;
}

static int upload_csr(DevCSR& m, int type, int64_t n, const int32_t* rp, const int32_t* col, const void* val) {
    const size_t vs = type_size(type);
    const int64_t nnz = rp[n];
    m.nrows = m.ncols = n;
    m.nnz = nnz;
    GBK_TRY(dev_alloc((void**)&m.row_ptr, sizeof(int32_t) * (n + 1)));
    GBK_TRY(dev_alloc((void**)&m.col, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
    GBK_TRY(dev_alloc(&m.val, vs * (nnz > 0 ? nnz : 1)));
    hipStream_t st = stream();
    GBK_HIP(hipMemcpyAsync(m.row_ptr, rp, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    if (nnz > 0) {
        GBK_HIP(hipMemcpyAsync(m.col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
        GBK_HIP(hipMemcpyAsync(m.val, val, vs * nnz, hipMemcpyHostToDevice, st));
    }
    int32_t mx = 0;
#pragma omp parallel for reduction(max : mx) schedule(static)
    for (int64_t r = 0; r < n; ++r) mx = std::max(mx, rp[r + 1] - rp[r]);
    m.max_row = mx;
    GBK_HIP(hipStreamSynchronize(st));
    return build_merge_tiles(m, type);
}

// Builds the device matrix from a host CSR with int32 indices (n x n).
static int build_common(gbk_matrix A, int64_t n, const int32_t* rp, const int32_t* col, const void* val) {
    GBK_TRY(ensure_init());
    free_matrix_storage(A);
    const size_t vs = type_size(A->type);
    const int64_t nnz = rp[n];
    A->n = n;
    A->nnz = nnz;
    GBK_TRY(upload_csr(A->csr, A->type, n, rp, col, val));
    // CSC by a stable sort of the entries by column on the device
    GBK_TRY(device_transpose(A->csr, A->type, n, A->csc));
    (void)vs;
    for (auto& b : A->ws.buf) GBK_TRY(dev_alloc(&b, sizeof(double) * (n > 0 ? n : 1)));
    A->ws.bytes = 4 * sizeof(double) * (n > 0 ? n : 1);
    A->built = true;
    if (A->format != GBK_FMT_CSR) GBK_TRY(format_build(A));
    return GBK_SUCCESS;
}

__global__ void k_row_len_max(const int32_t* __restrict__ rp, int64_t n, int32_t* __restrict__ mx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t v = 0;
    if (i < n) v = rp[i + 1] - rp[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(mx, v);
}

// Builds A from an n x n CSR already on the device; A takes ownership of
// the three arrays (released by gbk_matrix_free or the next build).
int build_from_device_csr(gbk_matrix A, int64_t n, int64_t nnz, int32_t* rp, int32_t* col, void* val) {
    GBK_TRY(ensure_init());
    free_matrix_storage(A);
    A->n = n;
    A->nnz = nnz;
    DevCSR& m = A->csr;
    m.nrows = m.ncols = n;
    m.nnz = nnz;
    m.row_ptr = rp;
    m.col = col;
    m.val = val;
    hipStream_t st = stream();
    int32_t* mx = nullptr;
    GBK_TRY(dev_alloc((void**)&mx, sizeof(int32_t)));
    int32_t mxh = 0;
    int rc = GBK_SUCCESS;
    if (hipMemsetAsync(mx, 0, sizeof(int32_t), st) != hipSuccess) rc = fail(GBK_PANIC, "hipMemsetAsync failed");
    if (rc == GBK_SUCCESS && n > 0) {
        hipLaunchKernelGGL(k_row_len_max, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rp, n, mx);
        if (hipGetLastError() != hipSuccess) rc = fail(GBK_PANIC, "kernel launch failed");
    }
    if (rc == GBK_SUCCESS && hipMemcpyAsync(&mxh, mx, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(GBK_PANIC, "hipMemcpyAsync failed");
    if (rc == GBK_SUCCESS && hipStreamSynchronize(st) != hipSuccess) rc = fail(GBK_PANIC, "sync failed");
    dev_free(mx);
    GBK_TRY(rc);
    m.max_row = mxh;
    GBK_TRY(build_merge_tiles(m, A->type));
    GBK_TRY(device_transpose(A->csr, A->type, n, A->csc));
    for (auto& b : A->ws.buf) GBK_TRY(dev_alloc(&b, sizeof(double) * (n > 0 ? n : 1)));
    A->ws.bytes = 4 * sizeof(double) * (n > 0 ? n : 1);
    A->built = true;
    if (A->format != GBK_FMT_CSR) GBK_TRY(format_build(A));
    return GBK_SUCCESS;
}

static int check_nnz(int64_t nnz) {
    if (nnz < 0) return fail(GBK_INVALID_VALUE, "negative nnz");
    if (nnz >= (int64_t)INT32_MAX) return fail(GBK_INVALID_VALUE, "nnz = %lld does not fit the int32 index layout", (long long)nnz);
    return GBK_SUCCESS;
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_matrix_new(gbk_matrix* A, gbk_type type) {
    if (!A) return fail(GBK_NULL_POINTER, "null matrix handle pointer");
    if (!valid_type(type)) return fail(GBK_DOMAIN_MISMATCH, "unknown type %d", (int)type);
    *A = new gbk_matrix_s();
    (*A)->type = type;
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_free(gbk_matrix A) {
    if (!A) return GBK_SUCCESS;
    if (A->built) (void)hipStreamSynchronize(stream());
    free_matrix_storage(A);
    delete A;
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_set_format(gbk_matrix A, int format, int sell_c, int sigma) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (format != GBK_FMT_CSR && format != GBK_FMT_CSR_SEG && format != GBK_FMT_LAV && format != GBK_FMT_SELL_C)
        return fail(GBK_INVALID_VALUE, "unsupported matrix storage format %d", format);
    if (sell_c != 0 && sell_c != 32 && sell_c != 64) return fail(GBK_INVALID_VALUE, "SELL-C chunk height must be 32 or 64");
    if (sigma < 0) return fail(GBK_INVALID_VALUE, "sigma must be positive");
    A->format = format;
    A->sell_c = sell_c ? sell_c : 64;
    A->sigma = sigma ? sigma : 512;
    if (A->built) {
        format_free(A);
        if (format != GBK_FMT_CSR) return format_build(A);
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_get_format(gbk_matrix A, int* format) {
    if (!A || !format) return fail(GBK_NULL_POINTER, "null argument");
    *format = A->format;
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_build_csr32(gbk_matrix A, const int64_t* row_ptr, const int32_t* col_ids, const void* vals,
                                      int64_t nrows, int64_t nnz) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!row_ptr || !col_ids || !vals) return fail(GBK_NULL_POINTER, "null CSR array");
    GBK_TRY(check_nnz(nnz));
    if (nrows < 0) return fail(GBK_INVALID_VALUE, "negative nrows");
    if (row_ptr[0] != 0 || row_ptr[nrows] != nnz) return fail(GBK_INVALID_OBJECT, "row_ptr must start at 0 and end at nnz");
    int64_t maxc = -1;
    bool bad = false;
#pragma omp parallel for reduction(max : maxc) reduction(|| : bad) schedule(static)
    for (int64_t e = 0; e < nnz; ++e) {
        int64_t c = col_ids[e];
        if (c < 0) bad = true;
        maxc = std::max(maxc, c);
    }
    if (bad) return fail(GBK_INVALID_INDEX, "negative column id");
    for (int64_t r = 0; r < nrows; ++r)
        if (row_ptr[r + 1] < row_ptr[r]) return fail(GBK_INVALID_OBJECT, "row_ptr is not monotone at row %lld", (long long)r);
    const int64_t n = std::max(nrows, maxc + 1);
    if (n >= (int64_t)INT32_MAX) return fail(GBK_INVALID_VALUE, "matrix dimension does not fit int32");
    std::vector<int32_t> rp(n + 1);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r <= n; ++r) rp[r] = (int32_t)(r <= nrows ? row_ptr[r] : nnz);
    return build_common(A, n, rp.data(), col_ids, vals);
}

extern "C" int gbk_matrix_build_csr(gbk_matrix A, const int64_t* row_ptr, const int64_t* col_ids, const void* vals,
                                    int64_t nrows, int64_t nnz) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!row_ptr || !col_ids || !vals) return fail(GBK_NULL_POINTER, "null CSR array");
    GBK_TRY(check_nnz(nnz));
    std::vector<int32_t> col(nnz > 0 ? nnz : 1);
    bool bad = false;
#pragma omp parallel for reduction(|| : bad) schedule(static)
    for (int64_t e = 0; e < nnz; ++e) {
        int64_t c = col_ids[e];
        if (c < 0 || c >= (int64_t)INT32_MAX) bad = true;
        col[e] = (int32_t)c;
    }
    if (bad) return fail(GBK_INVALID_INDEX, "column id out of the int32 range");
    return gbk_matrix_build_csr32(A, row_ptr, col.data(), vals, nrows, nnz);
}

// COO -> CSR on the host for an n x n matrix: a stable counting sort by row
// keeps the list order inside each row.
static int build_from_coo(gbk_matrix A, int64_t n, const int64_t* rows, const int64_t* cols, const void* vals,
                          size_t val_stride, int64_t nvals) {
    if (n >= (int64_t)INT32_MAX) return fail(GBK_INVALID_VALUE, "matrix dimension does not fit int32");
    std::vector<int32_t> rp(n + 1, 0);
    for (int64_t e = 0; e < nvals; ++e) rp[rows[e] + 1]++;
    for (int64_t r = 0; r < n; ++r) rp[r + 1] += rp[r];
    std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
    std::vector<int32_t> col(nvals);
    const size_t vs = type_size(A->type);
    std::vector<char> v(vs * nvals);
    for (int64_t e = 0; e < nvals; ++e) {
        int64_t o = pos[rows[e]]++;
        col[o] = (int32_t)cols[e];
        std::memcpy(&v[o * vs], (const char*)vals + e * val_stride, vs);
    }
    return build_common(A, n, rp.data(), col.data(), v.data());
}

extern "C" int gbk_matrix_build_coo(gbk_matrix A, const int64_t* rows, const int64_t* cols, const void* vals,
                                    int64_t nvals) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!rows || !cols || !vals) return fail(GBK_NULL_POINTER, "null COO array");
    if (nvals <= 0) return fail(GBK_NO_VALUE, "empty edge list");
    GBK_TRY(check_nnz(nvals));
    int64_t maxid = -1;
    bool bad = false;
#pragma omp parallel for reduction(max : maxid) reduction(|| : bad) schedule(static)
    for (int64_t e = 0; e < nvals; ++e) {
        if (rows[e] < 0 || cols[e] < 0) bad = true;
        maxid = std::max(maxid, std::max(rows[e], cols[e]));
    }
    if (bad) return fail(GBK_INVALID_INDEX, "negative vertex id");
    return build_from_coo(A, maxid + 1, rows, cols, vals, type_size(A->type), nvals);
}

namespace gbk {
int mtx_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t** rows, int64_t** cols);
}

// Matrix::init_from_mtx (read_file.hpp:556-615): every entry gets the value
// 1 (EDGE_VAL); the matrix is square of size max(rows, cols).
extern "C" int gbk_matrix_read_mtx(gbk_matrix A, const char* path) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    int64_t nr = 0, nc = 0, nnz = 0;
    int64_t *rp = nullptr, *cp = nullptr;
    GBK_TRY(mtx_read(path, &nr, &nc, &nnz, &rp, &cp));
    std::unique_ptr<int64_t, decltype(&std::free)> rows(rp, &std::free), cols(cp, &std::free);
    GBK_TRY(check_nnz(nnz));
    const int64_t n = std::max(nr, nc);
    if (n == 0) return fail(GBK_NO_VALUE, "empty matrix file");
    alignas(8) unsigned char one[8] = {0};
    switch (A->type) {
        case GBK_FP32: { float v = 1.0f; std::memcpy(one, &v, 4); break; }
        case GBK_FP64: { double v = 1.0; std::memcpy(one, &v, 8); break; }
        case GBK_INT32: { int32_t v = 1; std::memcpy(one, &v, 4); break; }
        case GBK_INT64: { int64_t v = 1; std::memcpy(one, &v, 8); break; }
        default: one[0] = 1; break;
    }
    return build_from_coo(A, n, rows.get(), cols.get(), one, 0, nnz);
}

extern "C" int gbk_matrix_nrows(gbk_matrix A, int64_t* n) {
    if (!A || !n) return fail(GBK_NULL_POINTER, "null argument");
    *n = A->n;
    return GBK_SUCCESS;
}
extern "C" int gbk_matrix_ncols(gbk_matrix A, int64_t* n) { return gbk_matrix_nrows(A, n); }
extern "C" int gbk_matrix_nvals(gbk_matrix A, int64_t* nnz) {
    if (!A || !nnz) return fail(GBK_NULL_POINTER, "null argument");
    *nnz = A->nnz;
    return GBK_SUCCESS;
}
extern "C" int gbk_matrix_type(gbk_matrix A, int* type) {
    if (!A || !type) return fail(GBK_NULL_POINTER, "null argument");
    *type = A->type;
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_export(gbk_matrix A, int transposed, int64_t* row_ptr, int32_t* col_ids, void* vals) {
    if (!A) return fail(GBK_NULL_POINTER, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    const DevCSR& m = transposed ? A->csc : A->csr;
    hipStream_t st = stream();
    if (row_ptr) {
        std::vector<int32_t> rp(m.nrows + 1);
        GBK_HIP(hipMemcpyAsync(rp.data(), m.row_ptr, sizeof(int32_t) * (m.nrows + 1), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i <= m.nrows; ++i) row_ptr[i] = rp[i];
    }
    if (col_ids && m.nnz) GBK_HIP(hipMemcpyAsync(col_ids, m.col, sizeof(int32_t) * m.nnz, hipMemcpyDeviceToHost, st));
    if (vals && m.nnz) GBK_HIP(hipMemcpyAsync(vals, m.val, type_size(A->type) * m.nnz, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_degrees(gbk_matrix A, int transposed, int64_t* degrees) {
    if (!A || !degrees) return fail(GBK_NULL_POINTER, "null argument");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    const DevCSR& m = transposed ? A->csc : A->csr;
    std::vector<int32_t> rp(m.nrows + 1);
    GBK_HIP(hipMemcpyAsync(rp.data(), m.row_ptr, sizeof(int32_t) * (m.nrows + 1), hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    for (int64_t i = 0; i < m.nrows; ++i) degrees[i] = rp[i + 1] - rp[i];
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_device_csr(gbk_matrix A, int transposed, const int32_t** row_ptr, const int32_t** col_ids,
                                     const void** vals) {
    if (!A) return fail(GBK_NULL_POINTER, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    const DevCSR& m = transposed ? A->csc : A->csr;
    if (row_ptr) *row_ptr = m.row_ptr;
    if (col_ids) *col_ids = m.col;
    if (vals) *vals = m.val;
    return GBK_SUCCESS;
}

extern "C" int gbk_matrix_device_bytes(gbk_matrix A, int64_t* bytes) {
    if (!A || !bytes) return fail(GBK_NULL_POINTER, "null argument");
    const size_t vs = type_size(A->type);
    int64_t b = 0;
    for (const DevCSR* m : {&A->csr, &A->csc}) {
        b += sizeof(int32_t) * (m->nrows + 1) + (sizeof(int32_t) + vs) * m->nnz;
        b += sizeof(int2) * (m->ntiles + 1) + 2 * 8 * m->ntiles;
    }
    b += (int64_t)A->ws.bytes + format_bytes(A);
    *bytes = b;
    return GBK_SUCCESS;
}
