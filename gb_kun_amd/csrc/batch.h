// Internal interface of the batched kernels (batch.hip) and of the query
// server built on them (serve.cpp).  Not part of the C ABI.
#pragma once

#include <memory>
#include <vector>

#include "gbk_objects.h"

namespace gbk {

// merge-path items (rows + nonzeros) walked by one lane group of the SpMM
constexpr int SPMM_TILE = 256;

// scratch slots owned by the batched entry points
enum {
    SLOT_SPMM_CARRY = 8,
    SLOT_MSBFS = 9,
    SLOT_SOURCES = 10,
    SLOT_MSBFS_OUT = 11,
    SLOT_PPR = 12,
    SLOT_MSSSP = 13,
};

// Scan on an explicit stream with caller-provided partials
// (scan_partial_bytes(n) bytes).
size_t scan_partial_bytes(int64_t n);
int exclusive_scan_i32_on(const int32_t* in, int32_t* out, int64_t n, int32_t* total_host, hipStream_t st,
                          void* partial_ws, size_t ws_bytes);

// SpMM tiles are built on first use and kept with the matrix orientation.
int ensure_spmm_tiles(DevCSR& m, hipStream_t st);
size_t spmm_carry_bytes(const DevCSR& m, int type, int k);
// k columns of X and Y, rows ld apart (ld = 0: k).
int spmm_run(DevCSR& m, int type, const void* X, void* Y, int k, int add, int mul, int accum, void* carry,
             hipStream_t st, int ld = 0);

size_t merge_carry_bytes(const DevCSR& m);
int spmv_merge_ws(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum, void* carry);

size_t msbfs_ws_bytes(int64_t n);
int msbfs_run(gbk_matrix A, const int32_t* src_dev, int nsrc, void* ws, size_t ws_bytes, int32_t* levels_out,
              int64_t* stats, hipStream_t st);

size_t ppr_ws_bytes(gbk_matrix A, int k);
size_t mssp_ws_bytes(gbk_matrix A, int k);

// One (graph, query kind) batch of the query server: up to slots() queries in
// flight, advanced together one level / iteration / round per step() on the
// engine's stream.  A query admitted into a free slot joins at the next step;
// step() reports the slots whose query completed in it; extract() writes a
// slot's result (result_bytes()) to device memory; release() frees the slot.
struct Engine {
    virtual ~Engine() {}
    virtual int slots() const = 0;
    virtual size_t bytes() const = 0;
    virtual size_t result_bytes() const = 0;
    virtual int admit(int slot, int32_t source) = 0;
    virtual int step(std::vector<int>& finished) = 0;
    virtual int extract(int slot, void* out_dev) = 0;
    virtual void release(int slot) = 0;
};

int make_bfs_engine(gbk_matrix A, hipStream_t st, std::unique_ptr<Engine>* out);
int make_ppr_engine(gbk_matrix A, hipStream_t st, int k, double damping, double tol, int itermax,
                    std::unique_ptr<Engine>* out);
int make_sssp_engine(gbk_matrix A, hipStream_t st, int k, std::unique_ptr<Engine>* out);

}  // namespace gbk
