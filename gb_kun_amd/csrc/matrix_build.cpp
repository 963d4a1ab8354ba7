// Host-side construction of the CSR and CSC arrays (OpenMP), before upload.
//
// Follows backend::Matrix::build / build_from_csr_arrays
// (src/backend/matrix/build.hpp:123-210): the matrix is square with
// n = max(rows, max column + 1); the CSC is the transpose with entries of a
// column in ascending row order, which is the order MatrixCSR::build gives
// for an edge list sorted by row (csr/build.hpp:68-83,
// format_conversions.h:5-49).  The transpose is a stable parallel counting
// sort: per-thread column histograms over contiguous row blocks, prefixed in
// (column, thread) order, then scattered — deterministic for any thread count.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

#include "host_build.h"

namespace gbk_host {

void transpose_csr(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vsize,
                   int32_t* t_rp, int32_t* t_col, void* t_val) {
    const int64_t nnz = rp[n];
    const int nthreads = std::max(1, std::min(omp_get_max_threads(), 64));
    // contiguous row blocks with balanced nnz
    std::vector<int64_t> rb(nthreads + 1);
    for (int t = 0; t <= nthreads; ++t) {
        int64_t target = nnz * t / nthreads;
        rb[t] = std::lower_bound(rp, rp + n + 1, (int32_t)std::min<int64_t>(target, INT32_MAX)) - rp;
        if (rb[t] > n) rb[t] = n;
    }
    rb[0] = 0;
    rb[nthreads] = n;
    std::vector<int32_t> hist((size_t)nthreads * (size_t)n, 0);
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        int32_t* h = &hist[(size_t)t * n];
        for (int64_t r = rb[t]; r < rb[t + 1]; ++r)
            for (int64_t e = rp[r]; e < rp[r + 1]; ++e) h[col[e]]++;
    }
    // t_rp[c] = sum of all counts of columns < c; per-(thread, column) start offsets
    std::vector<int64_t> colsum(n + 1, 0);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t c = 0; c < n; ++c) {
        int64_t s = 0;
        for (int t = 0; t < nthreads; ++t) s += hist[(size_t)t * n + c];
        colsum[c + 1] = s;
    }
    for (int64_t c = 0; c < n; ++c) colsum[c + 1] += colsum[c];
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t c = 0; c < n; ++c) {
        int64_t off = colsum[c];
        for (int t = 0; t < nthreads; ++t) {
            int64_t cnt = hist[(size_t)t * n + c];
            hist[(size_t)t * n + c] = (int32_t)off;
            off += cnt;
        }
    }
    for (int64_t c = 0; c <= n; ++c) t_rp[c] = (int32_t)colsum[c];
    const char* v = (const char*)val;
    char* tv = (char*)t_val;
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        int32_t* h = &hist[(size_t)t * n];
        for (int64_t r = rb[t]; r < rb[t + 1]; ++r)
            for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
                int64_t o = h[col[e]]++;
                t_col[o] = (int32_t)r;
                if (!v) continue;
                switch (vsize) {
                    case 4: ((uint32_t*)tv)[o] = ((const uint32_t*)v)[e]; break;
                    case 8: ((uint64_t*)tv)[o] = ((const uint64_t*)v)[e]; break;
                    case 1: tv[o] = v[e]; break;
                    default: std::memcpy(tv + o * vsize, v + e * vsize, vsize);
                }
            }
    }
}

}  // namespace gbk_host
