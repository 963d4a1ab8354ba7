// Host-side (OpenMP) helpers used while building device matrices.
#pragma once
#include <cstddef>
#include <cstdint>

namespace gbk_host {
// Stable transpose of an n x n CSR (int32 indices) into t_rp / t_col / t_val.
// val / t_val may be null; vsize is the element size in bytes.
void transpose_csr(int64_t n, const int32_t* rp, const int32_t* col, const void* val, size_t vsize,
                   int32_t* t_rp, int32_t* t_col, void* t_val);
}
