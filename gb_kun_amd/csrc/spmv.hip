// Pull SpMV kernels (dense x): y = accum(y, A (+.*) x) over CSR rows (mxv)
// or CSC rows (vxm).
//
// Reference semantics (src/backend/spmv/spmv_csr.h):
//   * all-active (341-384): every row r gets y[r] = accum(y[r], fold_r), the
//     fold starting from the monoid identity;
//   * dense mask (267-336): rows whose mask passes get accum(y, fold), rows
//     that fail get 0 ("if assign"); only GrB_STR_COMP complements;
//   * sparse mask (176-262): without STR_COMP only the listed rows are
//     computed and the others keep y; with STR_COMP listed rows get 0 and the
//     others the fold.
//
// Unmasked SpMV uses a merge-path decomposition (Merrill & Garland, SC'16)
// so that every 256-thread block consumes exactly 2048 items of the merge of
// row ends and nonzeros: power-law hub rows and runs of empty rows cost the
// same per item.  A block stages its row ends and its products
// A[k] (x) x[col[k]] in LDS (coalesced dword loads of col/val, gathers of x
// through L2/MALL), each thread walks 8 merge items, and a wave64 segmented
// scan (shuffles) plus a 4-wave LDS spine carries partial row sums across
// threads.  The partial sum of the row that crosses a tile boundary goes to a
// per-tile carry slot; a second small kernel adds the carries of those rows
// in tile order, so the result is deterministic.
//
// Masked SpMV and order-sensitive monoids use a row-group kernel: G lanes per
// row (G = 1 is the reference's sequential left fold), xor-shuffle tree.
#include "gbk_objects.h"

namespace gbk {

constexpr int MP_THREADS = 256;
constexpr int MP_ITEMS = 8;
constexpr int MP_TILE = MP_THREADS * MP_ITEMS;

template <typename T> __device__ __forceinline__ T shfl_up_t(T v, int off) {
    if constexpr (sizeof(T) >= 4) return __shfl_up(v, off, 64);
    else return (T)__shfl_up((int)v, off, 64);
}
template <typename T> __device__ __forceinline__ T shfl_xor_t(T v, int off, int width) {
    if constexpr (sizeof(T) >= 4) return __shfl_xor(v, off, width);
    else return (T)__shfl_xor((int)v, off, width);
}

__device__ __forceinline__ bool mask_test(const void* m, int mtype, int64_t i) {
    switch (mtype) {
        case GBK_FP32: return ((const float*)m)[i] != 0.0f;
        case GBK_FP64: return ((const double*)m)[i] != 0.0;
        case GBK_INT32: return ((const int32_t*)m)[i] != 0;
        case GBK_INT64: return ((const int64_t*)m)[i] != 0;
        default: return ((const uint8_t*)m)[i] != 0;
    }
}

// Merge-path start coordinate (row, nz) of every tile: diagonal t*MP_TILE.
__global__ void k_merge_tiles(const int32_t* __restrict__ rp, int64_t nrows, int64_t nnz, int64_t ntiles,
                              int2* __restrict__ tiles) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    int64_t diag = t * MP_TILE;
    if (diag > nrows + nnz) diag = nrows + nnz;
    int64_t lo = diag - nnz > 0 ? diag - nnz : 0;
    int64_t hi = diag < nrows ? diag : nrows;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)rp[mid + 1] <= diag - mid - 1) lo = mid + 1;
        else hi = mid;
    }
    tiles[t] = make_int2((int)lo, (int)(diag - lo));
}

int build_merge_tiles(DevCSR& m, int type) {
    dev_free(m.tiles);
    dev_free(m.carry);
    m.tiles = nullptr;
    m.carry = nullptr;
    m.ntiles = (m.nrows + m.nnz + MP_TILE - 1) / MP_TILE;
    GBK_TRY(dev_alloc((void**)&m.tiles, sizeof(int2) * (m.ntiles + 1)));
    GBK_TRY(dev_alloc(&m.carry, 2 * sizeof(double) * (m.ntiles > 0 ? m.ntiles : 1)));
    (void)type;
    hipLaunchKernelGGL(k_merge_tiles, dim3((unsigned)((m.ntiles + 1 + 255) / 256)), dim3(256), 0, stream(),
                       m.row_ptr, m.nrows, m.nnz, m.ntiles, m.tiles);
    GBK_LAUNCH_CHECK();
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}

template <typename T, int ADD, int MUL>
__global__ __launch_bounds__(MP_THREADS) void k_spmv_merge(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                           const T* __restrict__ val, const T* __restrict__ x,
                                                           T* __restrict__ y, int32_t nrows,
                                                           const int2* __restrict__ tiles, T* __restrict__ carry,
                                                           int accum, const int32_t* __restrict__ tile_map,
                                                           const int32_t* __restrict__ row_map) {
    __shared__ int32_t s_end[MP_TILE + 1];
    __shared__ T s_prod[MP_TILE];
    __shared__ T s_y[MP_TILE];
    __shared__ T s_wv[MP_THREADS / 64];
    __shared__ int s_wf[MP_THREADS / 64];

    // tile_map reorders tiles across the dispatch order (workgroups go to the
    // XCDs round-robin), e.g. to keep a column segment on one XCD's L2
    const int tile = tile_map ? tile_map[blockIdx.x] : (int)blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int2 c0 = tiles[tile], c1 = tiles[tile + 1];
    const int r0 = c0.x, k0 = c0.y;
    const int nr = c1.x - r0, nk = c1.y - k0;
    const T ident = add_identity<ADD, T>();

    // Stage row ends, relative to k0.
    for (int i = tid; i <= nr; i += MP_THREADS) {
        int r = r0 + 1 + i;
        s_end[i] = (r <= nrows) ? rp[r] - k0 : nk;
    }
    // Stage products: all loads of the tile are issued before the first use.
    int ci[MP_ITEMS];
    T av[MP_ITEMS], xv[MP_ITEMS];
#pragma unroll
    for (int it = 0; it < MP_ITEMS; ++it) {
        const int j = it * MP_THREADS + tid;
        ci[it] = 0;
        av[it] = T(0);
        if (j < nk) {
            ci[it] = __builtin_nontemporal_load(col + k0 + j);
            av[it] = __builtin_nontemporal_load(val + k0 + j);
        }
    }
#pragma unroll
    for (int it = 0; it < MP_ITEMS; ++it) {
        const int j = it * MP_THREADS + tid;
        xv[it] = T(0);
        if (j < nk) xv[it] = x[ci[it]];
    }
#pragma unroll
    for (int it = 0; it < MP_ITEMS; ++it) {
        const int j = it * MP_THREADS + tid;
        if (j < nk) s_prod[j] = mul_op<MUL, T>(av[it], xv[it]);
    }
    __syncthreads();

    // Thread-level merge over 8 items.
    const int total = nr + nk;
    const int d = min(tid * MP_ITEMS, total);
    int lo = max(d - nk, 0), hi = min(d, nr);
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s_end[mid] <= d - mid - 1) lo = mid + 1;
        else hi = mid;
    }
    int ri = lo, kj = d - lo;
    const int dend = min(d + MP_ITEMS, total);
    T acc = ident;
    int first = -1;
    for (int it = d; it < dend; ++it) {
        if (kj < s_end[ri]) {
            acc = add_op<ADD, T>(acc, s_prod[kj]);
            ++kj;
        } else {
            s_y[ri] = acc;
            if (first < 0) first = ri;
            acc = ident;
            ++ri;
        }
    }

    // Segmented scan of (trailing partial, has-row-end) across the block.
    T v = acc;
    int f = first >= 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T ov = shfl_up_t(v, off);
        int of = __shfl_up(f, off, 64);
        if (lane >= off) {
            if (!f) v = add_op<ADD, T>(ov, v);
            f |= of;
        }
    }
    if (lane == 63) {
        s_wv[wid] = v;
        s_wf[wid] = f;
    }
    T ev = shfl_up_t(v, 1);
    int ef = __shfl_up(f, 1, 64);
    if (lane == 0) {
        ev = ident;
        ef = 0;
    }
    __syncthreads();
    T pv = ident;
#pragma unroll
    for (int w = 0; w < MP_THREADS / 64; ++w) {
        if (w < wid) pv = s_wf[w] ? s_wv[w] : add_op<ADD, T>(pv, s_wv[w]);
    }
    const T carry_in = ef ? ev : add_op<ADD, T>(pv, ev);
    if (first >= 0) s_y[first] = add_op<ADD, T>(carry_in, s_y[first]);
    __syncthreads();

    // Row r0 continues a row begun in an earlier tile: its partial is left
    // for the fix-up kernel.
    const bool defer0 = (nr > 0) && (rp[r0] < k0);
    if (tid == 0) {
        T tv = ident;
#pragma unroll
        for (int w = 0; w < MP_THREADS / 64; ++w) tv = s_wf[w] ? s_wv[w] : add_op<ADD, T>(tv, s_wv[w]);
        carry[2 * tile] = tv;
        if (defer0) carry[2 * tile + 1] = s_y[0];
    }
    // row_map names the output row of each CSR row (a CSR over a subset of
    // the rows, e.g. the long rows of a SELL-C matrix)
    for (int i = tid; i < nr; i += MP_THREADS) {
        if (i == 0 && defer0) continue;
        const int r = row_map ? row_map[r0 + i] : r0 + i;
        const T res = s_y[i];
        y[r] = (accum == GBK_ACC_NONE) ? res : binop<T>(accum, y[r], res);
    }
}

template <typename T, int ADD>
__global__ void k_spmv_merge_fixup(const int32_t* __restrict__ rp, const int2* __restrict__ tiles, int64_t ntiles,
                                   const T* __restrict__ carry, T* __restrict__ y, int accum,
                                   const int32_t* __restrict__ row_map) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int2 c0 = tiles[t], c1 = tiles[t + 1];
    const int r0 = c0.x, k0 = c0.y;
    if (c1.x == r0) return;  // no row ends in this tile
    const int start = rp[r0];
    if (start >= k0) return;  // row r0 is complete inside the tile
    T total = carry[2 * t + 1];
    int64_t u = t - 1;
    while (u >= 0) {
        total = add_op<ADD, T>(carry[2 * u], total);
        if (tiles[u].y <= start) break;
        --u;
    }
    const int ro = row_map ? row_map[r0] : r0;
    y[ro] = (accum == GBK_ACC_NONE) ? total : binop<T>(accum, y[ro], total);
}

// Row-group kernel: G lanes per row.  rows == nullptr: all rows (count = n);
// else the listed rows.  mask == nullptr: no mask; else rows pass when
// (mask != 0) == pass_on_true, and failing rows get 0.
template <typename T, int ADD, int MUL, int G>
__global__ __launch_bounds__(256) void k_spmv_rows(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                   const T* __restrict__ val, const T* __restrict__ x,
                                                   T* __restrict__ y, int64_t count, const int32_t* __restrict__ rows,
                                                   const void* __restrict__ mask, int mtype, int pass_on_true,
                                                   int accum) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = gid / G;
    const int lane = (int)(gid % G);
    if (i >= count) return;
    const int64_t row = rows ? rows[i] : i;
    if (mask) {
        const bool m = mask_test(mask, mtype, row);
        if (m != (pass_on_true != 0)) {
            if (lane == 0) y[row] = T(0);
            return;
        }
    }
    T acc = add_identity<ADD, T>();
    const int e1 = rp[row + 1];
    for (int e = rp[row] + lane; e < e1; e += G) acc = add_op<ADD, T>(acc, mul_op<MUL, T>(val[e], x[col[e]]));
    if constexpr (G > 1) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc = add_op<ADD, T>(acc, shfl_xor_t(acc, off, G));
    }
    if (lane == 0) y[row] = (accum == GBK_ACC_NONE) ? acc : binop<T>(accum, y[row], acc);
}

static bool order_sensitive(int add) { return add == GBK_ADD_LESS || add == GBK_ADD_FIRST || add == GBK_ADD_SECOND; }

static int pick_group(const DevCSR& m, int add) {
    if (order_sensitive(add)) return 1;
    double avg = m.nrows > 0 ? (double)m.nnz / (double)m.nrows : 0.0;
    int g = 4;
    while (g < 64 && g < avg) g <<= 1;
    return g;
}

template <typename T, int ADD, int MUL>
static int launch_rows(const DevCSR& m, const T* x, T* y, int64_t count, const int32_t* rows, const void* mask,
                       int mtype, int pass_on_true, int accum, int G) {
    if (count <= 0) return GBK_SUCCESS;
    const int64_t threads = count * G;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    const T* val = (const T*)m.val;
    hipStream_t st = stream();
#define GBK_ROWS(GG)                                                                                      \
    case GG:                                                                                              \
        hipLaunchKernelGGL((k_spmv_rows<T, ADD, MUL, GG>), grid, block, 0, st, m.row_ptr, m.col, val, x, y, \
                           count, rows, mask, mtype, pass_on_true, accum);                                \
        break;
    switch (G) {
        GBK_ROWS(1)
        GBK_ROWS(4)
        GBK_ROWS(8)
        GBK_ROWS(16)
        GBK_ROWS(32)
        GBK_ROWS(64)
        default: return fail(GBK_PANIC, "bad row group %d", G);
    }
#undef GBK_ROWS
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

template <typename T, int ADD, int MUL>
static int launch_merge(const DevCSR& m, const T* x, T* y, int accum, const int32_t* tile_map = nullptr,
                        void* carry_ws = nullptr, const int32_t* row_map = nullptr) {
    if (m.ntiles == 0) return GBK_SUCCESS;
    hipStream_t st = stream();
    T* carry = (T*)(carry_ws ? carry_ws : m.carry);
    hipLaunchKernelGGL((k_spmv_merge<T, ADD, MUL>), dim3((unsigned)m.ntiles), dim3(MP_THREADS), 0, st, m.row_ptr, m.col,
                       (const T*)m.val, x, y, (int32_t)m.nrows, m.tiles, carry, accum, tile_map, row_map);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_spmv_merge_fixup<T, ADD>), dim3((unsigned)((m.ntiles + 255) / 256)), dim3(256), 0, st,
                       m.row_ptr, m.tiles, m.ntiles, (const T*)carry, y, accum, row_map);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

int spmv_csr_unmasked(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            if (order_sensitive(ADD))
                return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, m.nrows, nullptr, nullptr, 0, 1, accum, 1);
            return launch_merge<T, ADD, MUL>(m, (const T*)x, (T*)y, accum);
        });
    });
}

// Merge-path SpMV with a caller-owned carry buffer (merge_carry_bytes(m)
// bytes), so that SpMVs on one matrix may run concurrently on several streams.
size_t merge_carry_bytes(const DevCSR& m) { return 2 * sizeof(double) * (size_t)(m.ntiles > 0 ? m.ntiles : 1); }

int spmv_merge_ws(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum, void* carry) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            if (order_sensitive(ADD))
                return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, m.nrows, nullptr, nullptr, 0, 1, accum, 1);
            return launch_merge<T, ADD, MUL>(m, (const T*)x, (T*)y, accum, nullptr, carry);
        });
    });
}

// Merge-path SpMV of a CSR with an optional tile order (used by the segmented formats).
int spmv_merge_mapped(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum,
                      const int32_t* tile_map, const int32_t* row_map) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            if (order_sensitive(ADD))
                return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, m.nrows, nullptr, nullptr, 0, 1, accum, 1);
            if (row_map && order_sensitive(ADD))
                return fail(GBK_INVALID_VALUE, "order-sensitive monoid on a row-mapped CSR");
            return launch_merge<T, ADD, MUL>(m, (const T*)x, (T*)y, accum, tile_map, nullptr, row_map);
        });
    });
}

// Row-group SpMV over a list of rows of a CSR: `group` lanes per row (64 when
// group is 0; always 1, the sequential fold, for order-sensitive monoids).
int spmv_rows_list(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum,
                   const int32_t* rows, int64_t count, int group) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            const int G = order_sensitive(ADD) ? 1 : (group > 0 ? group : 64);
            return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, count, rows, nullptr, 0, 1, accum, G);
        });
    });
}

int spmv_unmasked(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum) {
    bool handled = false;
    if (A->format != GBK_FMT_CSR) {
        GBK_TRY(format_spmv(A, transposed, x, y, add, mul, accum, &handled));
        if (handled) return GBK_SUCCESS;
    }
    return spmv_csr_unmasked(transposed ? A->csc : A->csr, A->type, x, y, add, mul, accum);
}

int spmv_masked_dense(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum,
                      const void* mask, int mask_type, bool comp) {
    const DevCSR& m = transposed ? A->csc : A->csr;
    return dispatch_type(A->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, m.nrows, nullptr, mask, mask_type, comp ? 0 : 1,
                                            accum, pick_group(m, ADD));
        });
    });
}

int spmv_masked_rows(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum,
                     const int32_t* rows, int64_t nrows_list) {
    const DevCSR& m = transposed ? A->csc : A->csr;
    return dispatch_type(A->type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            return launch_rows<T, ADD, MUL>(m, (const T*)x, (T*)y, nrows_list, rows, nullptr, 0, 1, accum,
                                            pick_group(m, ADD));
        });
    });
}

template <typename T>
__global__ void k_accum_into(T* __restrict__ y, const T* __restrict__ t, int64_t n, int accum) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = binop<T>(accum, y[i], t[i]);
}

int accum_into(int type, void* y, const void* t, int64_t n, int accum) {
    if (n <= 0) return GBK_SUCCESS;
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        hipLaunchKernelGGL(k_accum_into<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), (T*)y,
                           (const T*)t, n, accum);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    });
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_spmv_device(gbk_matrix A, int transposed, const void* x, void* y, gbk_semiring op, int accum) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!x || !y) return fail(GBK_NULL_POINTER, "null vector pointer");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (accum < GBK_ACC_NONE || accum > GBK_ACC_NE) return fail(GBK_INVALID_VALUE, "unknown accumulator %d", accum);
    GBK_TRY(ensure_init());
    if (x == y) {
        // same-vectors path (spmv_csr.h:433-486): fold into a workspace, then accumulate
        void* t = A->ws.buf[0];
        GBK_TRY(spmv_unmasked(A, transposed, x, t, op.add, op.mul, GBK_ACC_NONE));
        return accum_into(A->type, y, t, A->n, accum);
    }
    return spmv_unmasked(A, transposed, x, y, op.add, op.mul, accum);
}
