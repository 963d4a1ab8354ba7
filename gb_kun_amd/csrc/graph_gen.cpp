// Seeded graph generation on the host (OpenMP).
//
// Replaces GraphGenerationAPI::RMAT / random_uniform
// (src/helpers/graph_generation/graph_generation.hpp:37-190), which seed from
// time() and cannot be reproduced.  Here every random draw is a pure function
// of (seed, stream, counter), so the output does not depend on the thread
// count.  The R-MAT recursion is the textbook one (Chakrabarti, Zhan,
// Faloutsos 2004) without noise: at each of `scale` levels a quadrant is
// drawn with probabilities (a, b, c, 1-a-b-c).  Like the reference, vertex
// ids are shuffled by a random permutation (graph_generation.hpp:10-33), here
// a seeded Fisher-Yates.  Unlike the reference, duplicates and self-loops are
// removed, so the result is a simple graph given as a CSR with ascending
// columns.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../../include/gbk.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    // SplitMix64 finalizer (Steele, Lea, Flood 2014).
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed * 0xD1B54A32D192ED03ull + stream) ^ (i * 0x9E3779B97F4A7C15ull));
}

inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Parallel LSD radix sort of 64-bit keys whose values are below 2^bits.
void radix_sort_u64(std::vector<uint64_t>& keys, int bits) {
    const size_t n = keys.size();
    if (n < 2) return;
    std::vector<uint64_t> tmp(n);
    const int D = 11, R = 1 << D;
    const int nthreads = omp_get_max_threads();
    std::vector<size_t> hist((size_t)nthreads * R);
    uint64_t* src = keys.data();
    uint64_t* dst = tmp.data();
    int passes = (bits + D - 1) / D;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * D;
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
            size_t* h = &hist[(size_t)t * R];
            for (size_t i = lo; i < hi; ++i) h[(src[i] >> shift) & (R - 1)]++;
#pragma omp barrier
#pragma omp single
            {
                size_t sum = 0;
                for (int d = 0; d < R; ++d)
                    for (int tt = 0; tt < nthreads; ++tt) {
                        size_t c = hist[(size_t)tt * R + d];
                        hist[(size_t)tt * R + d] = sum;
                        sum += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) {
                uint64_t k = src[i];
                dst[h[(k >> shift) & (R - 1)]++] = k;
            }
        }
        std::swap(src, dst);
    }
    if (src != keys.data()) std::memcpy(keys.data(), src, n * sizeof(uint64_t));
}

// Random permutation of [0, n) by a seeded Fisher-Yates shuffle.
std::vector<uint32_t> seeded_permutation(int64_t n, uint64_t seed) {
    std::vector<uint32_t> p(n);
    for (int64_t i = 0; i < n; ++i) p[i] = (uint32_t)i;
    for (int64_t i = n - 1; i > 0; --i) {
        uint64_t r = hash3(seed, 77, (uint64_t)i);
        int64_t j = (int64_t)(((unsigned __int128)r * (uint64_t)(i + 1)) >> 64);
        std::swap(p[i], p[j]);
    }
    return p;
}

// keys -> sorted, unique, no self loops -> CSR.
int keys_to_csr(std::vector<uint64_t>& keys, int scale, int64_t n, int64_t* n_out, int64_t* nnz_out,
                int64_t** row_ptr_out, int32_t** col_out) {
    radix_sort_u64(keys, 2 * scale);
    const uint64_t mask = (scale == 0) ? 0 : ((1ull << scale) - 1);
    // unique (parallel: mark first occurrences, then compact)
    const size_t m = keys.size();
    const int nthreads = omp_get_max_threads();
    std::vector<size_t> cnt(nthreads + 1, 0);
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        const size_t lo = m * t / nthreads, hi = m * (t + 1) / nthreads;
        size_t c = 0;
        for (size_t i = lo; i < hi; ++i) {
            uint64_t k = keys[i];
            bool self = ((k >> scale) == (k & mask));
            if (!self && (i == 0 || keys[i - 1] != k)) ++c;
        }
        cnt[t + 1] = c;
    }
    for (int t = 0; t < nthreads; ++t) cnt[t + 1] += cnt[t];
    const int64_t nnz = (int64_t)cnt[nthreads];
    int64_t* row_ptr = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
    int32_t* col = (int32_t*)std::malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    std::vector<int64_t> deg(n + 1, 0);
    if (!row_ptr || !col) { std::free(row_ptr); std::free(col); return GBK_OUT_OF_MEMORY; }
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        const size_t lo = m * t / nthreads, hi = m * (t + 1) / nthreads;
        size_t o = cnt[t];
        for (size_t i = lo; i < hi; ++i) {
            uint64_t k = keys[i];
            bool self = ((k >> scale) == (k & mask));
            if (!self && (i == 0 || keys[i - 1] != k)) {
                col[o++] = (int32_t)(k & mask);
            }
        }
    }
    // row pointers: count rows of the unique keys
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        const size_t lo = m * t / nthreads, hi = m * (t + 1) / nthreads;
        for (size_t i = lo; i < hi; ++i) {
            uint64_t k = keys[i];
            bool self = ((k >> scale) == (k & mask));
            if (!self && (i == 0 || keys[i - 1] != k)) {
#pragma omp atomic
                deg[(k >> scale) + 1]++;
            }
        }
    }
    row_ptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) row_ptr[r + 1] = row_ptr[r] + deg[r + 1];
    *n_out = n;
    *nnz_out = nnz;
    *row_ptr_out = row_ptr;
    *col_out = col;
    return GBK_SUCCESS;
}

}  // namespace

extern "C" int gbk_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                            int symmetric, int permute, int64_t* n_out, int64_t* nnz_out,
                            int64_t** row_ptr, int32_t** col_ids) {
    if (!n_out || !nnz_out || !row_ptr || !col_ids) return GBK_NULL_POINTER;
    if (scale < 1 || scale > 30 || edge_factor < 1) return GBK_INVALID_VALUE;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GBK_INVALID_VALUE;
    const int64_t n = 1ll << scale;
    const int64_t m = n * (int64_t)edge_factor;
    std::vector<uint32_t> perm;
    if (permute) perm = seeded_permutation(n, seed);
    std::vector<uint64_t> keys((size_t)(symmetric ? 2 * m : m));
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        uint64_t s = hash3(seed, 1, (uint64_t)e);
        uint64_t row = 0, col = 0;
        for (int l = 0; l < scale; ++l) {
            s += 0x9E3779B97F4A7C15ull;
            double u = u01(mix64(s));
            int q = (u < a) ? 0 : (u < ab) ? 1 : (u < abc) ? 2 : 3;
            row = (row << 1) | (uint64_t)(q >> 1);
            col = (col << 1) | (uint64_t)(q & 1);
        }
        if (permute) { row = perm[row]; col = perm[col]; }
        keys[e] = (row << scale) | col;
        if (symmetric) keys[m + e] = (col << scale) | row;
    }
    perm.clear();
    perm.shrink_to_fit();
    return keys_to_csr(keys, scale, n, n_out, nnz_out, row_ptr, col_ids);
}

extern "C" int gbk_gen_uniform(int scale, int edge_factor, uint64_t seed, int symmetric,
                               int64_t* n_out, int64_t* nnz_out, int64_t** row_ptr, int32_t** col_ids) {
    if (!n_out || !nnz_out || !row_ptr || !col_ids) return GBK_NULL_POINTER;
    if (scale < 1 || scale > 30 || edge_factor < 1) return GBK_INVALID_VALUE;
    const int64_t n = 1ll << scale;
    const int64_t m = n * (int64_t)edge_factor;
    const uint64_t mask = (uint64_t)n - 1;
    std::vector<uint64_t> keys((size_t)(symmetric ? 2 * m : m));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        uint64_t r = hash3(seed, 2, (uint64_t)e);
        uint64_t row = (r >> 32) & mask, col = r & mask;
        if (scale > 32) { row = hash3(seed, 3, (uint64_t)e) & mask; }
        keys[e] = (row << scale) | col;
        if (symmetric) keys[m + e] = (col << scale) | row;
    }
    return keys_to_csr(keys, scale, n, n_out, nnz_out, row_ptr, col_ids);
}

extern "C" void gbk_host_free(void* p) { std::free(p); }

extern "C" void gbk_fill_uniform_f32(uint64_t seed, int64_t count, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i)
        out[i] = (float)(hash3(seed, 5, (uint64_t)i) >> 40) * (1.0f / 16777216.0f);
}

extern "C" void gbk_fill_inv_coldeg_f32(int64_t n, int64_t nnz, const int32_t* col_ids, float* vals) {
    std::vector<int64_t> deg(n, 0);
    const int nthreads = omp_get_max_threads();
    std::vector<std::vector<int64_t>> local(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        std::vector<int64_t>& h = local[t];
        h.assign(n, 0);
        const int64_t lo = nnz * t / nthreads, hi = nnz * (t + 1) / nthreads;
        for (int64_t e = lo; e < hi; ++e) h[col_ids[e]]++;
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t v = 0; v < n; ++v) {
            int64_t s = 0;
            for (int tt = 0; tt < nthreads; ++tt) s += local[tt][v];
            deg[v] = s;
        }
#pragma omp for schedule(static)
        for (int64_t e = 0; e < nnz; ++e) vals[e] = 1.0f / (float)deg[col_ids[e]];
    }
}

extern "C" void gbk_fill_sym_weights_i32(int64_t n, const int64_t* row_ptr, const int32_t* col_ids,
                                         uint64_t seed, int32_t lo, int32_t hi, int32_t* vals) {
    const uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 0; r < n; ++r) {
        for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            uint64_t u = (uint64_t)r, v = (uint64_t)col_ids[e];
            uint64_t key = (u < v) ? ((u << 32) | v) : ((v << 32) | u);
            uint64_t h = hash3(seed, 9, key);
            vals[e] = (int32_t)(lo + (int64_t)(((unsigned __int128)h * span) >> 64));
        }
    }
}
