// Host-side images of the optimized formats, built on the CPU and uploaded.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace gbk_host {

struct SellCHost {
    int C = 64, sigma = 512;
    int64_t nchunks = 0, padded = 0;
    std::vector<int64_t> chunk_ptr;   // nchunks + 1
    std::vector<int32_t> chunk_len;   // nchunks
    std::vector<int32_t> row_perm;    // nchunks * C: sorted position -> row (-1 beyond n)
    std::vector<int32_t> row_len;     // nchunks * C: row length (0 for rows of problematic chunks)
    std::vector<int32_t> col;         // padded, -1 = padding
    std::vector<char> val;            // padded * vsize
    std::vector<int32_t> problem_rows;  // rows of problematic chunks with at most long_thr entries
    std::vector<int32_t> long_rows;     // rows with more than long_thr entries
};

struct SegHost {
    int nseg = 0;
    int64_t npairs = 0;
    std::vector<int64_t> seg_bounds;      // nseg + 1 column bounds
    std::vector<int64_t> seg_pair_start;  // nseg + 1
    std::vector<int32_t> vrp;             // npairs + 1: virtual CSR row pointers
    std::vector<int32_t> pair_row;        // npairs: row of each virtual row
    std::vector<int32_t> col;             // entries (relabelled columns for LAV)
    std::vector<char> val;
    std::vector<int32_t> rpp;             // n + 1: row -> its pairs
    std::vector<int32_t> row_pairs;       // npairs, segment order within a row
};

void build_sell_c(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vs, int C, int sigma,
                  int64_t long_thr, SellCHost& out);
void build_segmented(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vs,
                     const std::vector<int64_t>& seg_bounds, const int32_t* col_relabel, SegHost& out);
void lav_relabel(int64_t n, const int32_t* rp, const int32_t* col, double dense_frac, int64_t seg_cols,
                 std::vector<int32_t>& old_to_new, std::vector<int32_t>& new_to_old, std::vector<int64_t>& bounds);

}  // namespace gbk_host
