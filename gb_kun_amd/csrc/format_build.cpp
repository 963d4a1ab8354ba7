// Host-side (OpenMP) construction of the optimized formats from a CSR.
//
// SELL-C-sigma — replaces MatrixSellC::build / construct_sell_c_sigma
// (src/backend/matrix/containers/sell_c/build.hpp:28-295): rows are sorted by
// length, descending and stable, inside windows of sigma rows (a trailing
// window is sorted only when it holds more than C rows, build.hpp:92-97);
// chunks of C sorted rows are stored column-major, element j of lane r at
// chunk_ptr[c] + j*C + r, padded to the chunk's longest row.  A chunk whose
// padded size is at least twice its entries and which holds more than 1024
// entries is "problematic" (build.hpp:147-155): it is not stored in the
// SELL-C arrays, and its rows are listed for a row-split CSR kernel (the
// reference's fallback loop for them returns the identity, a defect not
// reproduced here).  Rows longer than long_thr leave their chunks the same
// way, so that no lane of the chunk kernel walks a power-law hub row alone.
//
// Segmented CSR — replaces MatrixSegmentedCSR::build
// (seg_csr/build.hpp:24-169): columns are cut into segments of seg_cols;
// each segment is a CSR over the rows that have entries in it.  Here the
// segments are concatenated into one "virtual" CSR (one virtual row per
// (segment, row) pair) whose per-pair sums are merged per row, in segment
// order, through row_pairs.
//
// LAV — replaces MatrixLAV::build (lav/build.hpp:103-212): columns are
// relabelled by descending in-degree (stable by id); the leading columns that
// cover 80% of the entries form dense segments of seg_cols columns, the rest
// one sparse segment; then it is a segmented CSR over the relabelled columns,
// with x permuted on each call.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>

#include "host_formats.h"

namespace gbk_host {

void build_sell_c(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vs, int C, int sigma,
                  int64_t long_thr, SellCHost& out) {
    out = SellCHost();
    out.C = C;
    out.sigma = sigma;
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    auto len = [&](int32_t r) { return rp[r + 1] - rp[r]; };
    const int64_t nwin = sigma > 0 ? n / sigma : 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t w = 0; w < nwin; ++w) {
        std::stable_sort(perm.begin() + w * sigma, perm.begin() + (w + 1) * sigma,
                         [&](int32_t a, int32_t b) { return len(a) > len(b); });
    }
    const int64_t rest = sigma > 0 ? n % sigma : 0;
    if (rest > C) {
        std::stable_sort(perm.begin() + nwin * sigma, perm.end(), [&](int32_t a, int32_t b) { return len(a) > len(b); });
    }
    const int64_t nchunks = (n + C - 1) / C;
    out.nchunks = nchunks;
    out.chunk_ptr.assign(nchunks + 1, 0);
    out.chunk_len.assign(nchunks, 0);
    out.row_perm.assign(nchunks * C, -1);
    out.row_len.assign(nchunks * C, 0);
    std::vector<char> problematic(nchunks, 0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunks; ++c) {
        int64_t mx = 0, tot = 0;
        for (int r = 0; r < C; ++r) {
            int64_t p = c * C + r;
            if (p < n) {
                int32_t row = perm[p];
                int64_t L = len(row);
                out.row_perm[p] = row;
                out.row_len[p] = (int32_t)L;
                if (L > long_thr) continue;  // a long row leaves the chunk (see below)
                mx = std::max(mx, L);
                tot += L;
            }
        }
        if (mx * C >= 2 * tot && tot > 1024) {
            problematic[c] = 1;
            out.chunk_len[c] = 0;
        } else {
            out.chunk_len[c] = (int32_t)mx;
        }
    }
    for (int64_t c = 0; c < nchunks; ++c) {
        out.chunk_ptr[c + 1] = out.chunk_ptr[c] + (int64_t)out.chunk_len[c] * C;
        if (!problematic[c]) {
            // long rows are served by the row-split kernel: their slots stay empty
            for (int r = 0; r < C; ++r) {
                int64_t p = c * C + r;
                if (p < n && out.row_len[p] > long_thr) {
                    out.long_rows.push_back(out.row_perm[p]);
                    out.row_len[p] = 0;
                    out.row_perm[p] = -1;
                }
            }
        }
        if (problematic[c]) {
            for (int r = 0; r < C; ++r) {
                int64_t p = c * C + r;
                if (p >= n) continue;
                // every row of the chunk (empty ones too) goes to a row kernel;
                // the chunk kernel skips positions whose row is -1
                if (out.row_len[p] > long_thr) out.long_rows.push_back(out.row_perm[p]);
                else out.problem_rows.push_back(out.row_perm[p]);
                out.row_len[p] = 0;
            }
        }
    }
    for (int64_t c = 0; c < nchunks; ++c) {
        if (problematic[c]) {
            for (int r = 0; r < C; ++r) {
                int64_t p = c * C + r;
                if (p < n) out.row_perm[p] = -1;
            }
        }
    }
    const int64_t total = out.chunk_ptr[nchunks];
    out.padded = total;
    out.col.assign(total, 0);
    out.val.assign(total * vs, 0);
    const char* v = (const char*)val;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t base = out.chunk_ptr[c];
        const int32_t L = out.chunk_len[c];
        for (int r = 0; r < C; ++r) {
            int64_t p = c * C + r;
            if (p >= n) continue;
            int32_t row = out.row_perm[p];
            int32_t rl = out.row_len[p];
            for (int32_t j = 0; j < L; ++j) {
                int64_t e = base + (int64_t)j * C + r;
                if (j < rl) {
                    int64_t k = rp[row] + j;
                    out.col[e] = col[k];
                    std::memcpy(&out.val[e * vs], v + k * vs, vs);
                } else {
                    out.col[e] = -1;  // padding
                }
            }
        }
    }
}

void build_segmented(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vs,
                     const std::vector<int64_t>& seg_bounds, const int32_t* col_relabel, SegHost& out) {
    out = SegHost();
    const int nseg = (int)seg_bounds.size() - 1;
    out.nseg = nseg;
    out.seg_bounds = seg_bounds;
    // per (row, segment) counts
    auto seg_of = [&](int32_t c) -> int {
        int64_t cc = col_relabel ? col_relabel[c] : c;
        return (int)(std::upper_bound(seg_bounds.begin(), seg_bounds.end(), cc) - seg_bounds.begin()) - 1;
    };
    // count entries per segment per row
    std::vector<std::vector<int32_t>> cnt(nseg, std::vector<int32_t>(n, 0));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) cnt[seg_of(col[k])][r]++;
    }
    // virtual rows: (segment, row) pairs with entries, segment-major
    std::vector<int64_t> seg_pairs(nseg + 1, 0), seg_nnz(nseg + 1, 0);
    for (int s = 0; s < nseg; ++s) {
        int64_t p = 0, z = 0;
#pragma omp parallel for reduction(+ : p, z)
        for (int64_t r = 0; r < n; ++r) {
            if (cnt[s][r]) { ++p; z += cnt[s][r]; }
        }
        seg_pairs[s + 1] = seg_pairs[s] + p;
        seg_nnz[s + 1] = seg_nnz[s] + z;
    }
    const int64_t P = seg_pairs[nseg], Z = seg_nnz[nseg];
    out.npairs = P;
    out.seg_pair_start = seg_pairs;
    out.vrp.assign(P + 1, 0);
    out.pair_row.assign(P, 0);
    out.col.assign(Z, 0);
    out.val.assign(Z * vs, 0);
    for (int s = 0; s < nseg; ++s) {
        int64_t p = seg_pairs[s];
        int64_t z = seg_nnz[s];
        for (int64_t r = 0; r < n; ++r) {
            if (!cnt[s][r]) continue;
            out.pair_row[p] = (int32_t)r;
            out.vrp[p] = (int32_t)z;
            z += cnt[s][r];
            ++p;
        }
    }
    out.vrp[P] = (int32_t)Z;
    // fill entries (CSR order inside each row and segment)
    std::vector<int64_t> row_pair_first(n, -1);
    // positions: for each segment, pair index of row r
    std::vector<std::vector<int64_t>> pos(nseg);
    for (int s = 0; s < nseg; ++s) {
        pos[s].assign(n, -1);
        for (int64_t p = seg_pairs[s]; p < seg_pairs[s + 1]; ++p) pos[s][out.pair_row[p]] = p;
    }
    const char* v = (const char*)val;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
        std::vector<int64_t> fill(nseg, 0);
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            int s = seg_of(col[k]);
            int64_t p = pos[s][r];
            int64_t e = out.vrp[p] + fill[s]++;
            out.col[e] = col_relabel ? col_relabel[col[k]] : col[k];
            std::memcpy(&out.val[e * vs], v + k * vs, vs);
        }
    }
    // row -> pairs (segment order)
    out.rpp.assign(n + 1, 0);
    for (int64_t r = 0; r < n; ++r) {
        int64_t c = 0;
        for (int s = 0; s < nseg; ++s) c += (pos[s][r] >= 0);
        out.rpp[r + 1] = out.rpp[r] + (int32_t)c;
    }
    out.row_pairs.assign(P, 0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        int64_t o = out.rpp[r];
        for (int s = 0; s < nseg; ++s)
            if (pos[s][r] >= 0) out.row_pairs[o++] = (int32_t)pos[s][r];
    }
}

void lav_relabel(int64_t n, const int32_t* rp, const int32_t* col, double dense_frac, int64_t seg_cols,
                 std::vector<int32_t>& old_to_new, std::vector<int32_t>& new_to_old, std::vector<int64_t>& bounds) {
    std::vector<int64_t> deg(n, 0);
    const int64_t nnz = rp[n];
    for (int64_t k = 0; k < nnz; ++k) deg[col[k]]++;
    new_to_old.resize(n);
    std::iota(new_to_old.begin(), new_to_old.end(), 0);
    std::stable_sort(new_to_old.begin(), new_to_old.end(), [&](int32_t a, int32_t b) { return deg[a] > deg[b]; });
    old_to_new.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) old_to_new[new_to_old[i]] = (int32_t)i;
    // dense columns: prefix covering dense_frac of the entries
    int64_t acc = 0, K = 0;
    while (K < n && (double)acc < dense_frac * (double)nnz) acc += deg[new_to_old[K++]];
    bounds.clear();
    bounds.push_back(0);
    for (int64_t b = seg_cols; b < K; b += seg_cols) bounds.push_back(b);
    if (K > 0) bounds.push_back(K);
    if (K < n) bounds.push_back(n);
    if (bounds.size() < 2) bounds.push_back(n);
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
}

}  // namespace gbk_host
