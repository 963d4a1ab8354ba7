// Optimized matrix representations (SELL-C-sigma, segmented CSR, LAV) on the
// device, built on the device from the CSR / CSC, one object per
// orientation as the reference keeps data / transposed_data
// (src/backend/matrix/build.hpp:6-118).
//
// SELL-C-sigma (spmv_sell_c.h:15-170): one lane per row, C = 64 (one wave per
// chunk) or C = 32 (two chunks per wave); at step j the lanes of a chunk read
// consecutive addresses chunk_ptr + j*C + r, so col/val stream coalesced.  A
// lane folds its row in CSR order, the reference's left fold, so with FMA
// contraction off the fp32 result equals the reference's CSR bits.  Rows of
// "problematic" chunks run on the row-split CSR kernel.
//
// Segmented CSR (spmv_seg.h:9-308) and LAV (spmv_lav.h:8-156): per-segment
// sums over a virtual CSR of (segment, row) pairs, each segment's work on one
// XCD so its slice of x stays in that XCD's L2; then a per-row merge of the
// pair sums in segment order (deterministic, no atomics).  The segmented
// format runs k_seg_flag (pair ends flagged in the column words, the
// segment's most referenced x values in LDS); LAV runs the merge-path kernel
// over XCD-ordered tiles after gathering x into its degree-sorted order.
#include <algorithm>
#include <mutex>
#include <vector>

#include "gbk_objects.h"

namespace gbk {

int spmv_merge_mapped(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum,
                      const int32_t* tile_map, const int32_t* row_map = nullptr);
int spmv_rows_list(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum,
                   const int32_t* rows, int64_t count, int group = 0);

struct SellC {
    int C = 64, sigma = 512;
    int64_t nchunks = 0, padded = 0, nproblem = 0;
    int64_t* chunk_ptr = nullptr;
    int32_t* chunk_len = nullptr;
    int32_t* row_perm = nullptr;
    int32_t* row_len = nullptr;
    int32_t* col = nullptr;
    void* val = nullptr;
    int32_t* problem_rows = nullptr;  // short rows of problematic chunks (sequential fold)
    int64_t nlong = 0;
    int32_t* long_rows = nullptr;     // rows longer than 4C entries, ascending
    DevCSR lng;                       // CSR of the long rows (row i = long_rows[i])
};

// Hot columns of a segment in LDS: the persistent kernel's blocks on XCD x
// serve the segments s with s % 8 == x; each block first copies the hot x
// values of its segment into LDS, so entries of hot columns read LDS and the
// rest read the segment's x slice through that XCD's L2.
struct SegHot {
    int H = 0;                    // LDS slots per segment
    int32_t* hcol = nullptr;      // during the build: entries, hot columns as -1 - slot
    int32_t* hot_ids = nullptr;   // nseg * H original columns (-1: unused slot)
    void* xh = nullptr;           // nseg * H gathered x values, per call
    int32_t* seg_pair = nullptr;  // nseg + 1: first pair of each segment
    // flagged-entry layout (k_seg_flag): fcol[e] = (e ends its pair) << 31 |
    // (hot ? 1 << 30 | slot : column), padded with zeros to whole
    // FLAG_CHUNK-entry chunks; a tile is one chunk's entries of one segment, ftiles[t] =
    // {first entry, first pair | (that pair began in an earlier tile) << 31}.
    int32_t* fcol = nullptr;
    int64_t nftiles = 0;
    int2* ftiles = nullptr;        // nftiles + 1 (the last = {nnz, 0})
    int64_t* fseg_tile = nullptr;  // nseg + 1
    void* fcarry = nullptr;        // 2 * nftiles values
    int32_t* fback = nullptr;      // nftiles: tiles back to the start of the pair that began
                                   // earlier and ends in tile t (-1: none)
};

struct SegCSR {
    int nseg = 0;
    int64_t npairs = 0;
    DevCSR v;  // virtual CSR: one row per (segment, row) pair
    int32_t* tile_map = nullptr;
    int32_t* rpp = nullptr;        // row -> its pairs (rpp, row_pairs), for the plain kernels
    int32_t* row_pairs = nullptr;
    int32_t* pair_row = nullptr;   // pair -> row (while building)
    uint16_t* pair_off = nullptr;  // pair -> row % MERGE_ROWS, and for every block of MERGE_ROWS
    int32_t* block_pair = nullptr; // rows and segment its first pair (the per-row merge of CSR_SEG)
    void* partial = nullptr;
    int64_t n = 0;
    SegHot hot;  // built for CSR_SEG (hot.fcol == nullptr otherwise)
};

struct Lav {
    SegCSR seg;
    int32_t* new_to_old = nullptr;
    void* xperm = nullptr;
    int nseg_dense = 0;
};

int build_merge_tiles(DevCSR& m, int type);

template <typename V> static int upload(V** dst, const std::vector<V>& src) {
    GBK_TRY(dev_alloc((void**)dst, sizeof(V) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) GBK_HIP(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
    return GBK_SUCCESS;
}

// ------------------------------------------------------------ SELL-C-sigma --
static void free_sell(SellC* s) {
    if (!s) return;
    dev_free(s->chunk_ptr);
    dev_free(s->chunk_len);
    dev_free(s->row_perm);
    dev_free(s->row_len);
    dev_free(s->col);
    dev_free(s->val);
    dev_free(s->problem_rows);
    dev_free(s->long_rows);
    dev_free(s->lng.row_ptr);
    dev_free(s->lng.col);
    dev_free(s->lng.val);
    dev_free(s->lng.tiles);
    dev_free(s->lng.carry);
    delete s;
}

// Device -> host copy ordered after the work queued on st (the library
// stream does not synchronize with the null stream).
static int d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
    GBK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

// Device scratch of one build step, released when it leaves scope.
struct TmpBuf {
    void* p = nullptr;
    ~TmpBuf() { dev_free(p); }
    int alloc(size_t bytes) { return dev_alloc(&p, bytes > 0 ? bytes : 1); }
    template <typename V> V* as() const { return (V*)p; }
};

// ------------------------------------------------- SELL-C-sigma, device build --
// Replaces MatrixSellC::build / construct_sell_c_sigma
// (src/backend/matrix/containers/sell_c/build.hpp:28-295), which runs on one
// host thread.  Rows are sorted by length, descending and stable, inside
// windows of sigma rows (a trailing window only when it holds more than C
// rows, build.hpp:92-97) by one radix sort of (window, ~length) keys; chunks
// of C sorted rows are stored column-major, element j of lane r at
// chunk_ptr[c] + j*C + r, padded with column -1 to the chunk's longest row.
// A chunk whose padded size is at least twice its entries and which holds
// more than 1024 entries is "problematic" (build.hpp:147-155): it is not
// stored, and its rows are listed for a row kernel (the reference's fallback
// loop returns the identity for them, a defect not reproduced).  Rows longer
// than long_thr leave their chunks the same way (listed ascending, with a CSR
// of their own for the merge-path kernel), so no lane walks a hub row alone.

__global__ void k_sell_key(const int32_t* __restrict__ rp, int64_t n, int64_t sigma, int64_t nwin, int sort_rest,
                           uint64_t* __restrict__ key) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t len = (uint32_t)(rp[r + 1] - rp[r]);
    const int64_t w = sigma > 0 ? min(r / sigma, nwin) : 0;
    const bool sorted = sigma > 0 && (w < nwin || sort_rest);
    key[r] = ((uint64_t)w << 32) | (sorted ? (uint64_t)(0xFFFFFFFFu - len) : 0);
}

// Sorted position p -> row and its length (-1 / 0 past n).
__global__ void k_sell_place(const int32_t* __restrict__ perm, const int32_t* __restrict__ rp, int64_t n, int64_t np,
                             int32_t* __restrict__ row_perm, int32_t* __restrict__ row_len) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    if (p < n) {
        const int32_t row = perm[p];
        row_perm[p] = row;
        row_len[p] = rp[row + 1] - rp[row];
    } else {
        row_perm[p] = -1;
        row_len[p] = 0;
    }
}

// One thread per chunk: its padded length over the rows it keeps, and
// whether it is problematic.  csize[c] = stored slots (chunk_len * C).
__global__ void k_sell_chunks(const int32_t* __restrict__ row_len, int64_t n, int64_t nchunks, int C,
                              int64_t long_thr, int32_t* __restrict__ chunk_len, int64_t* __restrict__ csize,
                              uint8_t* __restrict__ prob) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    int64_t mx = 0, tot = 0;
    for (int r = 0; r < C; ++r) {
        const int64_t p = c * C + r;
        if (p >= n) break;
        const int64_t L = row_len[p];
        if (L > long_thr) continue;
        mx = max(mx, L);
        tot += L;
    }
    const bool pr = mx * C >= 2 * tot && tot > 1024;
    prob[c] = pr ? 1 : 0;
    chunk_len[c] = pr ? 0 : (int32_t)mx;
    csize[c] = pr ? 0 : mx * C;
}

// pflag[p] = 1 for a short row of a problematic chunk; lflag[row] = 1 for a long row.
__global__ void k_sell_flags(const int32_t* __restrict__ row_perm, const int32_t* __restrict__ row_len,
                             const uint8_t* __restrict__ prob, int64_t n, int C, int64_t long_thr,
                             int32_t* __restrict__ pflag, int32_t* __restrict__ lflag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const bool lng = row_len[p] > long_thr;
    pflag[p] = (!lng && prob[p / C]) ? 1 : 0;
    if (lng) lflag[row_perm[p]] = 1;
}

// Lists the problematic rows (position order) and takes listed rows out of
// the chunks.
__global__ void k_sell_unlink(int32_t* __restrict__ row_perm, int32_t* __restrict__ row_len,
                              const int32_t* __restrict__ pflag, const int32_t* __restrict__ pidx, int64_t n,
                              int64_t long_thr, int32_t* __restrict__ problem_rows) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (pflag[p]) problem_rows[pidx[p]] = row_perm[p];
    if (pflag[p] || row_len[p] > long_thr) {
        row_perm[p] = -1;
        row_len[p] = 0;
    }
}

__global__ void k_list_flagged(const int32_t* __restrict__ flag, const int32_t* __restrict__ idx, int64_t n,
                               int32_t* __restrict__ list) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) list[idx[i]] = (int32_t)i;
}

// One thread per slot position: writes the chunk column of its row, padding
// (column -1, value 0) past the row's end.
template <typename V>
__global__ void k_sell_fill(const int64_t* __restrict__ chunk_ptr, const int32_t* __restrict__ chunk_len,
                            const int32_t* __restrict__ row_perm, const int32_t* __restrict__ row_len,
                            const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const V* __restrict__ val, int64_t np, int C, int32_t* __restrict__ scol,
                            V* __restrict__ sval) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t c = p / C;
    const int r = (int)(p % C);
    const int32_t L = chunk_len[c];
    const int32_t row = row_perm[p];
    const int32_t rl = row >= 0 ? row_len[p] : 0;
    const int64_t start = row >= 0 ? rp[row] : 0;
    const int64_t base = chunk_ptr[c] + r;
    for (int32_t j = 0; j < L; ++j) {
        const int64_t e = base + (int64_t)j * C;
        if (j < rl) {
            scol[e] = col[start + j];
            sval[e] = val[start + j];
        } else {
            scol[e] = -1;
            sval[e] = V(0);
        }
    }
}

__global__ void k_list_len(const int32_t* __restrict__ rows, const int32_t* __restrict__ rp, int64_t m,
                           int32_t* __restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) len[i] = rp[rows[i] + 1] - rp[rows[i]];
}

// One wave per listed row: copies its entries into the sub-CSR.
template <typename V>
__global__ void k_rows_copy(const int32_t* __restrict__ rows, const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ col, const V* __restrict__ val, int64_t m,
                            const int32_t* __restrict__ lrp, int32_t* __restrict__ lcol, V* __restrict__ lval) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= m) return;
    const int32_t r = rows[i];
    const int64_t s = rp[r], L = rp[r + 1] - s, d = lrp[i];
    for (int64_t j = lane; j < L; j += 64) {
        lcol[d + j] = col[s + j];
        lval[d + j] = val[s + j];
    }
}

static dim3 grid_b(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return dim3((unsigned)(b > 0 ? b : 1));
}

template <typename F> static int by_value_size(size_t vs, F&& f) {
    if (vs == 8) return f(uint64_t());
    if (vs == 4) return f(uint32_t());
    if (vs == 1) return f(uint8_t());
    return fail(GBK_INVALID_VALUE, "unsupported value size %zu", vs);
}

static int build_sell_device(const DevCSR& m, int type, int C, int sigma, int64_t long_thr, SellC& s) {
    hipStream_t st = stream();
    const size_t vs = type_size(type);
    const int64_t n = m.nrows;
    const int64_t nchunks = (n + C - 1) / C;
    const int64_t np = nchunks * C;
    s.C = C;
    s.sigma = sigma;
    s.nchunks = nchunks;
    const int64_t n1 = n > 0 ? n : 1, np1 = np > 0 ? np : 1;
    GBK_TRY(dev_alloc((void**)&s.chunk_ptr, 8 * (nchunks + 1)));
    GBK_TRY(dev_alloc((void**)&s.chunk_len, 4 * (nchunks > 0 ? nchunks : 1)));
    GBK_TRY(dev_alloc((void**)&s.row_perm, 4 * np1));
    GBK_TRY(dev_alloc((void**)&s.row_len, 4 * np1));
    if (n == 0) {
        GBK_HIP(hipMemsetAsync(s.chunk_ptr, 0, 8, st));
        GBK_TRY(dev_alloc((void**)&s.col, 4));
        GBK_TRY(dev_alloc(&s.val, vs));
        GBK_HIP(hipStreamSynchronize(st));
        return GBK_SUCCESS;
    }
    // sigma-window sort of the rows
    {
        const int64_t nwin = sigma > 0 ? n / sigma : 0;
        const int64_t rest = sigma > 0 ? n % sigma : 0;
        TmpBuf key, key_out, ids, perm;
        GBK_TRY(key.alloc(8 * n1));
        GBK_TRY(key_out.alloc(8 * n1));
        GBK_TRY(ids.alloc(4 * n1));
        GBK_TRY(perm.alloc(4 * n1));
        hipLaunchKernelGGL(k_sell_key, grid_b(n), dim3(256), 0, st, m.row_ptr, n, (int64_t)sigma, nwin,
                           rest > C ? 1 : 0, key.as<uint64_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(dev_iota(ids.as<int32_t>(), n, st));
        GBK_TRY(dev_sort_pairs64(key.as<uint64_t>(), key_out.as<uint64_t>(), ids.as<int32_t>(), perm.as<int32_t>(),
                                 n, 32 + bits_for(nwin), st));
        hipLaunchKernelGGL(k_sell_place, grid_b(np), dim3(256), 0, st, perm.as<int32_t>(), m.row_ptr, n, np,
                           s.row_perm, s.row_len);
        GBK_LAUNCH_CHECK();
    }
    // chunk lengths, problematic chunks, chunk pointers
    TmpBuf csize, prob, pflag, pidx, lflag, lidx;
    GBK_TRY(csize.alloc(8 * (nchunks + 1)));
    GBK_TRY(prob.alloc(nchunks));
    hipLaunchKernelGGL(k_sell_chunks, grid_b(nchunks), dim3(256), 0, st, s.row_len, n, nchunks, C, long_thr,
                       s.chunk_len, csize.as<int64_t>(), prob.as<uint8_t>());
    GBK_LAUNCH_CHECK();
    GBK_HIP(hipMemsetAsync(csize.as<int64_t>() + nchunks, 0, 8, st));
    GBK_TRY(dev_exclusive_scan_i64(csize.as<int64_t>(), s.chunk_ptr, nchunks + 1, st));
    GBK_TRY(d2h(&s.padded, s.chunk_ptr + nchunks, 8, st));
    // listed rows: short rows of problematic chunks, long rows
    GBK_TRY(pflag.alloc(4 * n1));
    GBK_TRY(pidx.alloc(4 * (n1 + 1)));
    GBK_TRY(lflag.alloc(4 * (n1 + 1)));
    GBK_TRY(lidx.alloc(4 * (n1 + 1)));
    GBK_HIP(hipMemsetAsync(lflag.p, 0, 4 * (n1 + 1), st));
    hipLaunchKernelGGL(k_sell_flags, grid_b(n), dim3(256), 0, st, s.row_perm, s.row_len, prob.as<uint8_t>(), n, C,
                       long_thr, pflag.as<int32_t>(), lflag.as<int32_t>());
    GBK_LAUNCH_CHECK();
    int32_t npr = 0, nlg = 0;
    GBK_TRY(exclusive_scan_i32(pflag.as<int32_t>(), pidx.as<int32_t>(), n, &npr));
    GBK_TRY(exclusive_scan_i32(lflag.as<int32_t>(), lidx.as<int32_t>(), n, &nlg));
    s.nproblem = npr;
    s.nlong = nlg;
    GBK_TRY(dev_alloc((void**)&s.problem_rows, 4 * (npr > 0 ? npr : 1)));
    GBK_TRY(dev_alloc((void**)&s.long_rows, 4 * (nlg > 0 ? nlg : 1)));
    hipLaunchKernelGGL(k_sell_unlink, grid_b(n), dim3(256), 0, st, s.row_perm, s.row_len, pflag.as<int32_t>(),
                       pidx.as<int32_t>(), n, long_thr, s.problem_rows);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_list_flagged, grid_b(n), dim3(256), 0, st, lflag.as<int32_t>(), lidx.as<int32_t>(), n,
                       s.long_rows);
    GBK_LAUNCH_CHECK();
    // the chunk slots
    const int64_t pad1 = s.padded > 0 ? s.padded : 1;
    GBK_TRY(dev_alloc((void**)&s.col, 4 * pad1));
    GBK_TRY(dev_alloc(&s.val, vs * pad1));
    GBK_TRY(by_value_size(vs, [&](auto tv) -> int {
        typedef decltype(tv) V;
        hipLaunchKernelGGL(k_sell_fill<V>, grid_b(np), dim3(256), 0, st, s.chunk_ptr, s.chunk_len, s.row_perm,
                           s.row_len, m.row_ptr, m.col, (const V*)m.val, np, C, s.col, (V*)s.val);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    }));
    // CSR of the long rows, ascending
    if (nlg > 0) {
        TmpBuf llen;
        GBK_TRY(llen.alloc(4 * nlg));
        DevCSR& L = s.lng;
        GBK_TRY(dev_alloc((void**)&L.row_ptr, 4 * (nlg + 1)));
        hipLaunchKernelGGL(k_list_len, grid_b(nlg), dim3(256), 0, st, s.long_rows, m.row_ptr, (int64_t)nlg,
                           llen.as<int32_t>());
        GBK_LAUNCH_CHECK();
        int32_t lnnz = 0;
        GBK_TRY(exclusive_scan_i32(llen.as<int32_t>(), L.row_ptr, nlg, &lnnz));
        L.nrows = nlg;
        L.ncols = m.ncols;
        L.nnz = lnnz;
        GBK_TRY(dev_alloc((void**)&L.col, 4 * (lnnz > 0 ? lnnz : 1)));
        GBK_TRY(dev_alloc(&L.val, vs * (lnnz > 0 ? lnnz : 1)));
        GBK_TRY(by_value_size(vs, [&](auto tv) -> int {
            typedef decltype(tv) V;
            hipLaunchKernelGGL(k_rows_copy<V>, grid_b((int64_t)nlg * 64), dim3(256), 0, st, s.long_rows, m.row_ptr,
                               m.col, (const V*)m.val, (int64_t)nlg, L.row_ptr, L.col, (V*)L.val);
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        }));
        GBK_TRY(build_merge_tiles(L, type));
    }
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

static int build_sell(const DevCSR& m, int type, int C, int sigma, SellC** out) {
    if (C <= 0 || C > 64) return fail(GBK_INVALID_VALUE, "SELL-C chunk size %d out of range (1..64)", C);
    SellC* s = new SellC();
    // rows longer than 4 C entries go to the merge-path kernel
    const int rc = build_sell_device(m, type, C, sigma, 4 * (int64_t)C, *s);
    if (rc != GBK_SUCCESS) {
        free_sell(s);
        return rc;
    }
    *out = s;
    return GBK_SUCCESS;
}

template <typename T, int ADD, int MUL>
__global__ __launch_bounds__(256) void k_sell_c(const int64_t* __restrict__ chunk_ptr, const int32_t* __restrict__ row_perm,
                                                const int32_t* __restrict__ row_len, const int32_t* __restrict__ col,
                                                const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
                                                int64_t nchunks, int C, int accum) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = gid / C;
    const int r = (int)(gid % C);
    if (c >= nchunks) return;
    const int64_t p = c * C + r;
    const int32_t row = row_perm[p];
    if (row < 0) return;
    const int32_t L = row_len[p];
    const int64_t base = chunk_ptr[c] + r;
    T acc = add_identity<ADD, T>();
    int32_t j = 0;
    // four slices per step, all loads issued before the fold (the fold
    // itself stays in storage order: the reference's left fold)
    for (; j + 4 <= L; j += 4) {
        int32_t cc[4];
        T vv[4], xx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t e = base + (int64_t)(j + u) * C;
            cc[u] = __builtin_nontemporal_load(col + e);
            vv[u] = __builtin_nontemporal_load(val + e);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) xx[u] = x[cc[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = add_op<ADD, T>(acc, mul_op<MUL, T>(vv[u], xx[u]));
    }
    for (; j < L; ++j) {
        const int64_t e = base + (int64_t)j * C;
        acc = add_op<ADD, T>(acc, mul_op<MUL, T>(__builtin_nontemporal_load(val + e), x[__builtin_nontemporal_load(col + e)]));
    }
    y[row] = (accum == GBK_ACC_NONE) ? acc : binop<T>(accum, y[row], acc);
}

static bool order_sensitive_monoid(int add) {
    return add == GBK_ADD_LESS || add == GBK_ADD_FIRST || add == GBK_ADD_SECOND;
}

static int sell_spmv(const SellC* s, const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum) {
    GBK_TRY(dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            const int64_t threads = s->nchunks * s->C;
            if (threads > 0) {
                hipLaunchKernelGGL((k_sell_c<T, ADD, MUL>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                                   stream(), s->chunk_ptr, s->row_perm, s->row_len, s->col, (const T*)s->val,
                                   (const T*)x, (T*)y, s->nchunks, s->C, accum);
                GBK_LAUNCH_CHECK();
            }
            return GBK_SUCCESS;
        });
    }));
    // short rows of problematic chunks: one lane per row, the left fold
    if (s->nproblem > 0) GBK_TRY(spmv_rows_list(m, type, x, y, add, mul, accum, s->problem_rows, s->nproblem, 1));
    // long rows: merge path over their own CSR (order-sensitive monoids fold sequentially)
    if (s->nlong > 0) {
        if (order_sensitive_monoid(add))
            GBK_TRY(spmv_rows_list(m, type, x, y, add, mul, accum, s->long_rows, s->nlong, 1));
        else
            GBK_TRY(spmv_merge_mapped(s->lng, type, x, y, add, mul, accum, nullptr, s->long_rows));
    }
    return GBK_SUCCESS;
}

// --------------------------------------------------------- segmented CSR --
static void free_seg_fields(SegCSR& s) {
    dev_free(s.v.row_ptr);
    dev_free(s.v.col);
    dev_free(s.v.val);
    dev_free(s.v.tiles);
    dev_free(s.v.carry);
    dev_free(s.tile_map);
    dev_free(s.rpp);
    dev_free(s.row_pairs);
    dev_free(s.pair_row);
    dev_free(s.pair_off);
    dev_free(s.block_pair);
    dev_free(s.partial);
    dev_free(s.hot.hcol);
    dev_free(s.hot.hot_ids);
    dev_free(s.hot.xh);
    dev_free(s.hot.seg_pair);
    dev_free(s.hot.fcol);
    dev_free(s.hot.ftiles);
    dev_free(s.hot.fseg_tile);
    dev_free(s.hot.fcarry);
    dev_free(s.hot.fback);
    s = SegCSR();
}

// Tile order for the XCD-aware launch: block b runs on XCD b % 8 and gets a
// tile of a segment assigned to that XCD (segment s -> XCD s % 8) while any
// remain, else a tile of the group with the most left.
static std::vector<int32_t> xcd_tile_map(const std::vector<int2>& tiles, const std::vector<int64_t>& seg_pair_start,
                                         int nseg) {
    const int64_t nt = (int64_t)tiles.size() - 1;
    const int G = 8;
    std::vector<std::vector<int32_t>> grp(G);
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t row = tiles[t].x;
        int s = (int)(std::upper_bound(seg_pair_start.begin(), seg_pair_start.end(), row) - seg_pair_start.begin()) - 1;
        s = std::max(0, std::min(s, nseg - 1));
        grp[s % G].push_back((int32_t)t);
    }
    std::vector<size_t> at(G, 0);
    std::vector<int32_t> map;
    map.reserve(nt);
    for (int64_t b = 0; b < nt; ++b) {
        int g = (int)(b % G);
        if (at[g] >= grp[g].size()) {
            size_t best = 0;
            for (int h = 0; h < G; ++h)
                if (grp[h].size() - at[h] > best) {
                    best = grp[h].size() - at[h];
                    g = h;
                }
        }
        map.push_back(grp[g][at[g]++]);
    }
    return map;
}

static int64_t seg_cols_default(int64_t n) {
    // 2 MiB of fp32 x per segment: half of one XCD's 4 MiB L2
    (void)n;
    return 512 * 1024;
}

// Persistent hot-column kernel shape: 16 waves per block, the block's LDS
// holding the hot x values of its segment.
constexpr int HOT_WAVES = 16;
// 16-byte pieces of hot x values each thread loads at once while its block
// fills the LDS (160 KiB = 10 per thread of a 1024-thread block).
constexpr int HOT_FILL = 10;
// Flagged tiles: FLAG_EPL consecutive entries per lane of a wave.
constexpr int FLAG_EPL = 4;
constexpr int FLAG_CHUNK = 64 * FLAG_EPL;
constexpr int FLAG_CNT_BITS = 3;  // bits of a lane's flag count (0..FLAG_EPL)
static_assert(FLAG_EPL <= 16 && (1 << FLAG_CNT_BITS) > FLAG_EPL, "flag counts fit the ballots");

// LDS slots per segment for values of vs bytes: what one workgroup's LDS
// holds, a multiple of 64.
static int hot_slots(size_t vs) {
    const int64_t room = device_lds_bytes();
    if (room <= 0) return 0;
    int64_t h = (room / (int64_t)vs) & ~(int64_t)63;
    return h < 1024 ? 0 : (int)std::min<int64_t>(h, 1 << 20);
}

// ------------------------------------------- segmented CSR, device build --
// The (segment, row) pairs, the hot column slots and the wave tiles are built
// on the device from the CSR already there: entries are stably sorted by
// segment (CSR order is kept inside a segment), pairs start where the
// (segment, row) key changes, pairs are stably sorted by row for the merge,
// columns are sorted by (segment, descending in-degree, id) for the slots.
// Replaces MatrixSegmentedCSR::build (seg_csr/build.hpp:24-169), which runs
// single-threaded on the host.

__global__ void k_seg_key(const int32_t* __restrict__ col, int64_t nnz, int64_t w, uint32_t* __restrict__ key) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) key[e] = (uint32_t)(col[e] / w);
}

// flag[i] = 1 where sorted entry i opens a new (segment, row) pair.
__global__ void k_pair_flags(const uint32_t* __restrict__ skey, const int32_t* __restrict__ perm,
                             const int32_t* __restrict__ rowid, int64_t nnz, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    flag[i] = (i == 0 || skey[i] != skey[i - 1] || rowid[perm[i]] != rowid[perm[i - 1]]) ? 1 : 0;
}

__global__ void k_pair_fill(const int32_t* __restrict__ flag, const int32_t* __restrict__ pidx,
                            const uint32_t* __restrict__ skey, const int32_t* __restrict__ perm,
                            const int32_t* __restrict__ rowid, int64_t nnz, int32_t* __restrict__ vrp,
                            int32_t* __restrict__ pair_row, int32_t* __restrict__ pair_seg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz || !flag[i]) return;
    const int32_t p = pidx[i];
    vrp[p] = (int32_t)i;
    pair_row[p] = rowid[perm[i]];
    pair_seg[p] = (int32_t)skey[i];
}

// deg[c] = entries of column c = row length of c in the other orientation.
__global__ void k_deg_from_ptr(const int32_t* __restrict__ optr, int64_t n, int32_t* __restrict__ deg) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) deg[c] = optr[c + 1] - optr[c];
}

__global__ void k_set_i32(int32_t* __restrict__ p, int64_t i, int32_t v) {
    if (threadIdx.x == 0) p[i] = v;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Column sort key: segment in the high bits, then descending in-degree
// (saturated), so a stable sort of column ids puts each segment's columns
// together, most referenced first, ties by id.
__global__ void k_hot_key(const int32_t* __restrict__ deg, int64_t n, int64_t w, int db, uint32_t* __restrict__ key) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const uint32_t dmax = (1u << db) - 1u;
    const uint32_t d = (uint32_t)min((int64_t)deg[c], (int64_t)dmax);
    key[c] = ((uint32_t)(c / w) << db) | (dmax - d);
}

// Sorted position i of segment s holds its (i - s w)-th most referenced column.
__global__ void k_hot_assign(const int32_t* __restrict__ order, const int32_t* __restrict__ deg, int64_t n, int64_t w,
                             int H, int32_t* __restrict__ slot, int32_t* __restrict__ hot_ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = order[i];
    const int64_t sg = c / w;
    const int64_t rank = i - sg * w;
    if (rank < H && deg[c] >= 2) {
        slot[c] = (int32_t)rank;
        hot_ids[sg * H + rank] = c;
    }
}

// The hot slots renumbered in column order inside each segment (the set of
// hot columns is unchanged): pos is the exclusive prefix count of hot
// columns, so the per-SpMV gather of their x values reads x ascending.
__global__ void k_hot_flag(const int32_t* __restrict__ slot, int64_t n, int32_t* __restrict__ flag) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) flag[c] = slot[c] >= 0 ? 1 : 0;
}
__global__ void k_hot_reslot(const int32_t* __restrict__ pos, int64_t n, int64_t w, int H, int32_t* __restrict__ slot,
                             int32_t* __restrict__ hot_ids) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || slot[c] < 0) return;
    const int64_t sg = c / w;
    const int32_t ns = pos[c] - pos[sg * w];
    slot[c] = ns;
    hot_ids[sg * H + ns] = (int32_t)c;
}

__global__ void k_hot_encode(const int32_t* __restrict__ vcol, const int32_t* __restrict__ slot, int64_t nnz,
                             int32_t* __restrict__ hcol) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int32_t c = vcol[i];
    const int32_t sl = slot[c];
    hcol[i] = sl >= 0 ? -1 - sl : c;
}

static dim3 grid_e(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// v.col[i] = relabel[col[perm[i]]] (col[perm[i]] without a relabelling).
__global__ void k_gather_cols(const int32_t* __restrict__ col, const int32_t* __restrict__ perm,
                              const int32_t* __restrict__ relabel, int64_t nnz, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int32_t c = col[perm[i]];
    out[i] = relabel ? relabel[c] : c;
}

// The (segment, row) pairs of a segmented CSR whose entry e lies in segment
// ekey[e]: the virtual CSR (columns relabelled when `relabel` is given), the
// row -> pairs map and the pair sums' buffer.  seg_pair receives the first
// pair of every segment.
static int seg_pairs_device(const DevCSR& m, int type, int64_t nseg, const uint32_t* ekey, const int32_t* relabel,
                            SegCSR& s, std::vector<int32_t>& seg_pair) {
    hipStream_t st = stream();
    const size_t vs = type_size(type);
    const int64_t n = m.nrows, nnz = m.nnz;
    const int64_t z1 = nnz > 0 ? nnz : 1;
    s.nseg = (int)nseg;
    s.n = n;
    // entries stably sorted by segment
    TmpBuf skey, iota, perm, rowid, flag, pidx;
    GBK_TRY(skey.alloc(4 * z1));
    GBK_TRY(iota.alloc(4 * z1));
    GBK_TRY(perm.alloc(4 * z1));
    GBK_TRY(rowid.alloc(4 * z1));
    GBK_TRY(flag.alloc(4 * z1));
    GBK_TRY(pidx.alloc(4 * (z1 + 1)));  // the scan also writes the total at [nnz]
    GBK_TRY(dev_iota(iota.as<int32_t>(), nnz, st));
    GBK_TRY(dev_sort_pairs(ekey, skey.as<uint32_t>(), iota.as<int32_t>(), perm.as<int32_t>(), nnz,
                           bits_for(nseg - 1), st));
    GBK_TRY(dev_row_ids(m.row_ptr, n, rowid.as<int32_t>(), st));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_pair_flags, grid_e(nnz), dim3(256), 0, st, skey.as<uint32_t>(), perm.as<int32_t>(),
                           rowid.as<int32_t>(), nnz, flag.as<int32_t>());
        GBK_LAUNCH_CHECK();
    }
    int32_t np32 = 0;
    if (nnz > 0) GBK_TRY(exclusive_scan_i32(flag.as<int32_t>(), pidx.as<int32_t>(), nnz, &np32));
    const int64_t npairs = np32;
    s.npairs = npairs;
    // the virtual CSR: one row per pair
    TmpBuf pair_seg;
    int32_t* pair_row = nullptr;
    GBK_TRY(pair_seg.alloc(4 * (npairs > 0 ? npairs : 1)));
    GBK_TRY(dev_alloc((void**)&pair_row, 4 * (npairs > 0 ? npairs : 1)));
    s.pair_row = pair_row;  // owned by the format from here on
    s.v.nrows = s.v.ncols = npairs;
    s.v.nnz = nnz;
    GBK_TRY(dev_alloc((void**)&s.v.row_ptr, 4 * (npairs + 1)));
    GBK_TRY(dev_alloc((void**)&s.v.col, 4 * z1));
    // values padded with zeros to whole chunks (the flagged kernel loads
    // whole chunks)
    const int64_t zpad = (nnz + FLAG_CHUNK - 1) / FLAG_CHUNK * FLAG_CHUNK + FLAG_CHUNK;
    GBK_TRY(dev_alloc(&s.v.val, vs * zpad));
    GBK_HIP(hipMemsetAsync((char*)s.v.val + vs * nnz, 0, vs * (zpad - nnz), st));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_pair_fill, grid_e(nnz), dim3(256), 0, st, flag.as<int32_t>(), pidx.as<int32_t>(),
                           skey.as<uint32_t>(), perm.as<int32_t>(), rowid.as<int32_t>(), nnz, s.v.row_ptr, pair_row,
                           pair_seg.as<int32_t>());
        GBK_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(64), 0, st, s.v.row_ptr, npairs, (int32_t)nnz);
    GBK_LAUNCH_CHECK();
    if (nnz > 0) {
        hipLaunchKernelGGL(k_gather_cols, grid_e(nnz), dim3(256), 0, st, m.col, perm.as<int32_t>(), relabel, nnz,
                           s.v.col);
        GBK_LAUNCH_CHECK();
    }
    GBK_TRY(dev_gather_bytes(m.val, perm.as<int32_t>(), nnz, vs, s.v.val, st));
    // first pair of every segment (pairs are in segment order)
    TmpBuf segptr;
    GBK_TRY(segptr.alloc(4 * (nseg + 1)));
    GBK_TRY(dev_bounds_sorted((const uint32_t*)pair_seg.p, npairs, nseg, segptr.as<int32_t>(), st));
    seg_pair.assign(nseg + 1, 0);
    GBK_TRY(d2h(seg_pair.data(), segptr.p, 4 * (nseg + 1), st));
    // rows -> their pairs in segment order: pairs stably sorted by row
    TmpBuf piota, pkey_out;
    GBK_TRY(piota.alloc(4 * (npairs > 0 ? npairs : 1)));
    GBK_TRY(pkey_out.alloc(4 * (npairs > 0 ? npairs : 1)));
    GBK_TRY(dev_alloc((void**)&s.row_pairs, 4 * (npairs > 0 ? npairs : 1)));
    GBK_TRY(dev_alloc((void**)&s.rpp, 4 * (n + 1)));
    GBK_TRY(dev_iota(piota.as<int32_t>(), npairs, st));
    GBK_TRY(dev_sort_pairs((const uint32_t*)pair_row, pkey_out.as<uint32_t>(), piota.as<int32_t>(), s.row_pairs,
                           npairs, bits_for(n > 0 ? n - 1 : 0), st));
    GBK_TRY(dev_bounds_sorted(pkey_out.as<uint32_t>(), npairs, n, s.rpp, st));  // rpp[0..n]
    GBK_TRY(dev_alloc(&s.partial, vs * (npairs > 0 ? npairs : 1)));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

// Merge-path tiles of the virtual CSR, ordered so that a segment's tiles run
// on one XCD (the kernel of the formats without hot columns).
static int seg_plain_tiles(SegCSR& s, int type, const std::vector<int32_t>& seg_pair) {
    GBK_TRY(build_merge_tiles(s.v, type));
    std::vector<int2> tiles(s.v.ntiles + 1);
    GBK_TRY(d2h(tiles.data(), s.v.tiles, sizeof(int2) * (s.v.ntiles + 1), stream()));
    std::vector<int64_t> sp(seg_pair.begin(), seg_pair.end());
    return upload(&s.tile_map, xcd_tile_map(tiles, sp, s.nseg));
}

constexpr int MERGE_THREADS = 256;
constexpr int MERGE_RPT = 4;                               // rows per thread
constexpr int MERGE_ROWS = MERGE_THREADS * MERGE_RPT;      // rows per workgroup of the per-row merge

// block_pair[b * nseg + g] = first pair of segment g whose row is at least
// b * MERGE_ROWS (pairs are in row order inside a segment), b in [0, nblocks].
__global__ void k_block_pairs(const int32_t* __restrict__ pair_row, const int32_t* __restrict__ seg_pair, int nseg,
                              int64_t nblocks, int32_t* __restrict__ block_pair) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nblocks + 1) * nseg) return;
    const int64_t b = t / nseg;
    const int g = (int)(t % nseg);
    const int64_t row = b * MERGE_ROWS;
    int64_t lo = seg_pair[g], hi = seg_pair[g + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)pair_row[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    block_pair[t] = (int32_t)lo;
}

static_assert((MERGE_ROWS & (MERGE_ROWS - 1)) == 0 && MERGE_ROWS <= 65536, "row offsets fit 16 bits");

__global__ void k_pair_off(const int32_t* __restrict__ pair_row, int64_t npairs, uint16_t* __restrict__ pair_off) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npairs) pair_off[p] = (uint16_t)(pair_row[p] & (MERGE_ROWS - 1));
}

// The per-row merge's tables: block_pair, and each pair's row as an offset
// in its block of MERGE_ROWS rows (2 bytes a pair instead of 4).
static int build_block_pairs(SegCSR& s) {
    const int64_t nblocks = (s.n + MERGE_ROWS - 1) / MERGE_ROWS;
    const int64_t m = (nblocks + 1) * s.nseg;
    GBK_TRY(dev_alloc((void**)&s.block_pair, 4 * (m > 0 ? m : 1)));
    if (m > 0) {
        hipLaunchKernelGGL(k_block_pairs, grid_e(m), dim3(256), 0, stream(), s.pair_row, s.hot.seg_pair, s.nseg,
                           nblocks, s.block_pair);
        GBK_LAUNCH_CHECK();
    }
    GBK_TRY(dev_alloc((void**)&s.pair_off, 2 * (s.npairs > 0 ? s.npairs : 1)));
    if (s.npairs > 0) {
        hipLaunchKernelGGL(k_pair_off, grid_e(s.npairs), dim3(256), 0, stream(), s.pair_row, s.npairs, s.pair_off);
        GBK_LAUNCH_CHECK();
    }
    return GBK_SUCCESS;
}


__global__ void k_flag_encode(const int32_t* __restrict__ hcol, int64_t nnz, int64_t padded,
                              int32_t* __restrict__ fcol) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= padded) return;
    if (e >= nnz) {
        fcol[e] = 0;  // padding: column 0, never counted
        return;
    }
    const int32_t c = hcol[e];
    fcol[e] = c < 0 ? (int32_t)(0x40000000u | (uint32_t)(-1 - c)) : c;
}

__global__ void k_flag_ends(const int32_t* __restrict__ vrp, int64_t npairs, int32_t* __restrict__ fcol) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npairs) fcol[vrp[p + 1] - 1] |= (int32_t)0x80000000u;
}

// Tile t: the part of chunk c in segment s, with its first pair (the pair
// holding its first entry) and whether that pair began before the tile.
__global__ void k_flag_tiles(const int32_t* __restrict__ vrp, const int64_t* __restrict__ fseg_tile,
                             const int32_t* __restrict__ seg_pair, int nseg, int64_t ntiles, int64_t nnz,
                             int2* __restrict__ ftiles) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        ftiles[t] = make_int2((int)nnz, 0);
        return;
    }
    int lo_s = 0, hi_s = nseg - 1;  // last segment with fseg_tile[s] <= t
    while (lo_s < hi_s) {
        const int mid = (lo_s + hi_s + 1) >> 1;
        if (fseg_tile[mid] <= t) lo_s = mid;
        else hi_s = mid - 1;
    }
    const int64_t p0 = seg_pair[lo_s], p1 = seg_pair[lo_s + 1];
    const int64_t sk = vrp[p0];
    const int64_t k0 = max(sk, (sk / FLAG_CHUNK + (t - fseg_tile[lo_s])) * FLAG_CHUNK);
    int64_t lo = p0, hi = p1 - 1;  // last pair of the segment starting at or before k0
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (vrp[mid] <= k0) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t defer = vrp[lo] < k0 ? 0x80000000u : 0u;
    ftiles[t] = make_int2((int)k0, (int)((uint32_t)lo | defer));
}

// The flagged layout from the hot encoding, the virtual row pointers and the
// segments' pair / entry bounds.
// fback[t]: for the pair that began in an earlier tile and ends in tile t,
// the number of tiles t - 1, t - 2, ... down to the one holding its first
// entry (the tiles whose trailing sums k_flag_fix adds); -1 when no such
// pair ends in t.
__global__ void k_flag_back(const int32_t* __restrict__ vrp, const int2* __restrict__ ftiles, int64_t ntiles,
                            int32_t* __restrict__ fback) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int32_t b = -1;
    const int2 ti = ftiles[t];
    if (ti.y < 0) {
        const int p = ti.y & 0x7fffffff;
        if (vrp[p + 1] <= ftiles[t + 1].x) {
            const int start = vrp[p];
            int64_t u = t - 1;
            if (u >= 0) {
                while (u > 0 && ftiles[u].x > start) --u;
                b = (int32_t)(t - u);
            } else {
                b = 0;
            }
        }
    }
    fback[t] = b;
}

static int build_flagged(SegCSR& s, size_t vs, const std::vector<int32_t>& seg_k) {
    hipStream_t st = stream();
    SegHot& o = s.hot;
    const int64_t nnz = s.v.nnz, nseg = s.nseg;
    const int64_t padded = (nnz + FLAG_CHUNK - 1) / FLAG_CHUNK * FLAG_CHUNK + FLAG_CHUNK;
    GBK_TRY(dev_alloc((void**)&o.fcol, 4 * padded));
    hipLaunchKernelGGL(k_flag_encode, grid_e(padded), dim3(256), 0, st, o.hcol, nnz, padded, o.fcol);
    GBK_LAUNCH_CHECK();
    if (s.npairs > 0) {
        hipLaunchKernelGGL(k_flag_ends, grid_e(s.npairs), dim3(256), 0, st, s.v.row_ptr, s.npairs, o.fcol);
        GBK_LAUNCH_CHECK();
    }
    std::vector<int64_t> fseg_tile(nseg + 1, 0);
    for (int64_t g = 0; g < nseg; ++g) {
        const int64_t a = seg_k[g], b = seg_k[g + 1];
        fseg_tile[g + 1] = fseg_tile[g] + (b > a ? (b - 1) / FLAG_CHUNK - a / FLAG_CHUNK + 1 : 0);
    }
    o.nftiles = fseg_tile[nseg];
    GBK_TRY(upload(&o.fseg_tile, fseg_tile));
    GBK_TRY(dev_alloc((void**)&o.ftiles, sizeof(int2) * (o.nftiles + 1)));
    hipLaunchKernelGGL(k_flag_tiles, grid_e(o.nftiles + 1), dim3(256), 0, st, s.v.row_ptr, o.fseg_tile, o.seg_pair,
                       (int)nseg, o.nftiles, nnz, o.ftiles);
    GBK_LAUNCH_CHECK();
    GBK_TRY(dev_alloc(&o.fcarry, 2 * vs * (size_t)(o.nftiles > 0 ? o.nftiles : 1)));
    GBK_TRY(dev_alloc((void**)&o.fback, 4 * (size_t)(o.nftiles > 0 ? o.nftiles : 1)));
    if (o.nftiles > 0) {
        hipLaunchKernelGGL(k_flag_back, grid_e(o.nftiles), dim3(256), 0, st, s.v.row_ptr, o.ftiles, o.nftiles,
                           o.fback);
        GBK_LAUNCH_CHECK();
    }
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

static int build_seg_device(const DevCSR& m, const int32_t* optr, int type, int64_t nseg, int64_t w, SegCSR& s) {
    hipStream_t st = stream();
    const size_t vs = type_size(type);
    const int64_t n = m.nrows, nnz = m.nnz;
    const int64_t z1 = nnz > 0 ? nnz : 1;
    std::vector<int32_t> seg_pair;
    {
        TmpBuf key;
        GBK_TRY(key.alloc(4 * z1));
        if (nnz > 0) {
            hipLaunchKernelGGL(k_seg_key, grid_e(nnz), dim3(256), 0, st, m.col, nnz, w, key.as<uint32_t>());
            GBK_LAUNCH_CHECK();
        }
        GBK_TRY(seg_pairs_device(m, type, nseg, key.as<uint32_t>(), nullptr, s, seg_pair));
    }
    // hot slots: columns by (segment, descending in-degree, id)
    SegHot& o = s.hot;
    const int H = hot_slots(vs);
    o.H = H;
    {
        TmpBuf deg, ckey, ckey_out, corder, ciota, slot;
        GBK_TRY(deg.alloc(4 * (n + 1)));
        GBK_TRY(ckey.alloc(4 * (n > 0 ? n : 1)));
        GBK_TRY(ckey_out.alloc(4 * (n > 0 ? n : 1)));
        GBK_TRY(corder.alloc(4 * (n > 0 ? n : 1)));
        GBK_TRY(ciota.alloc(4 * (n > 0 ? n : 1)));
        GBK_TRY(slot.alloc(4 * (n > 0 ? n : 1)));
        if (n > 0) {
            hipLaunchKernelGGL(k_deg_from_ptr, grid_e(n), dim3(256), 0, st, optr, n, deg.as<int32_t>());
            GBK_LAUNCH_CHECK();
        }
        const int segbits = bits_for(nseg - 1);
        const int db = 32 - segbits;
        const size_t nh = (size_t)nseg * (H > 0 ? H : 1);
        GBK_TRY(dev_alloc((void**)&o.hot_ids, 4 * nh));
        GBK_TRY(dev_alloc(&o.xh, vs * nh));
        hipLaunchKernelGGL(k_fill_i32, grid_e((int64_t)nh), dim3(256), 0, st, o.hot_ids, (int64_t)nh, -1);
        GBK_LAUNCH_CHECK();
        if (n > 0) {
            hipLaunchKernelGGL(k_fill_i32, grid_e(n), dim3(256), 0, st, slot.as<int32_t>(), n, -1);
            GBK_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_hot_key, grid_e(n), dim3(256), 0, st, deg.as<int32_t>(), n, w, db,
                               ckey.as<uint32_t>());
            GBK_LAUNCH_CHECK();
            GBK_TRY(dev_iota(ciota.as<int32_t>(), n, st));
            GBK_TRY(dev_sort_pairs(ckey.as<uint32_t>(), ckey_out.as<uint32_t>(), ciota.as<int32_t>(),
                                   corder.as<int32_t>(), n, 32, st));
            if (H > 0) {
                hipLaunchKernelGGL(k_hot_assign, grid_e(n), dim3(256), 0, st, corder.as<int32_t>(),
                                   deg.as<int32_t>(), n, w, H, slot.as<int32_t>(), o.hot_ids);
                GBK_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_hot_flag, grid_e(n), dim3(256), 0, st, slot.as<int32_t>(), n,
                                   ckey_out.as<int32_t>());
                GBK_LAUNCH_CHECK();
                // the scan writes n + 1 entries (the total last): deg, no
                // longer needed, has room for them
                GBK_TRY(exclusive_scan_i32(ckey_out.as<int32_t>(), deg.as<int32_t>(), n, nullptr));
                hipLaunchKernelGGL(k_hot_reslot, grid_e(n), dim3(256), 0, st, deg.as<int32_t>(), n, w, H,
                                   slot.as<int32_t>(), o.hot_ids);
                GBK_LAUNCH_CHECK();
            }
        }
        GBK_TRY(dev_alloc((void**)&o.hcol, 4 * z1));
        if (nnz > 0) {
            hipLaunchKernelGGL(k_hot_encode, grid_e(nnz), dim3(256), 0, st, s.v.col, slot.as<int32_t>(), nnz,
                               o.hcol);
            GBK_LAUNCH_CHECK();
        }
        GBK_HIP(hipStreamSynchronize(st));
    }
    // first entry of every segment
    std::vector<int32_t> seg_k(nseg + 1);
    {
        TmpBuf sp_d, sk_d;
        GBK_TRY(sp_d.alloc(4 * (nseg + 1)));
        GBK_TRY(sk_d.alloc(4 * (nseg + 1)));
        GBK_HIP(hipMemcpyAsync(sp_d.p, seg_pair.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, st));
        GBK_TRY(dev_gather_i32(s.v.row_ptr, sp_d.as<int32_t>(), nseg + 1, sk_d.as<int32_t>(), st));
        GBK_TRY(d2h(seg_k.data(), sk_d.p, 4 * (nseg + 1), st));
    }
    GBK_TRY(upload(&o.seg_pair, seg_pair));
    if (n <= (int64_t)0x40000000) {
        // the flagged layout (columns below 2^30) replaces the column arrays
        GBK_TRY(build_flagged(s, vs, seg_k));
        GBK_TRY(build_block_pairs(s));
        GBK_HIP(hipStreamSynchronize(st));
        dev_free(o.hcol);
        o.hcol = nullptr;
        dev_free(s.v.col);
        s.v.col = nullptr;
        dev_free(s.rpp);
        s.rpp = nullptr;
        dev_free(s.row_pairs);
        s.row_pairs = nullptr;
        dev_free(s.pair_row);
        s.pair_row = nullptr;
    } else {
        dev_free(o.hcol);
        o.hcol = nullptr;
        GBK_TRY(seg_plain_tiles(s, type, seg_pair));
        GBK_HIP(hipStreamSynchronize(st));
        dev_free(s.pair_row);
        s.pair_row = nullptr;
    }
    return GBK_SUCCESS;
}

static int build_seg(const DevCSR& m, const DevCSR& other, int type, SegCSR** out) {
    const int64_t n = m.nrows;
    if (m.nnz >= (int64_t)INT32_MAX / 2) return fail(GBK_INVALID_VALUE, "too many entries for the segmented format");
    // a multiple of 8 segments (one group per XCD), each at most 512 Ki
    // columns wide
    const int64_t sc = seg_cols_default(n);
    const int64_t nseg = 8 * std::max<int64_t>(1, (n + 8 * sc - 1) / (8 * sc));
    const int64_t w = std::max<int64_t>(1, (n + nseg - 1) / nseg);
    SegCSR* s = new SegCSR();
    const int rc = build_seg_device(m, other.row_ptr, type, nseg, w, *s);
    if (rc != GBK_SUCCESS) {
        free_seg_fields(*s);
        delete s;
        return rc;
    }
    *out = s;
    return GBK_SUCCESS;
}

template <typename T, int ADD>
__global__ void k_seg_merge(const int32_t* __restrict__ rpp, const int32_t* __restrict__ row_pairs,
                            const T* __restrict__ partial, T* __restrict__ y, int64_t n, int accum) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    T acc = add_identity<ADD, T>();
    for (int p = rpp[r]; p < rpp[r + 1]; ++p) acc = add_op<ADD, T>(acc, partial[row_pairs[p]]);
    y[r] = (accum == GBK_ACC_NONE) ? acc : binop<T>(accum, y[r], acc);
}

// Per-row merge by blocks of MERGE_ROWS rows.  A segment has at most one
// pair per row, so a block's range of a segment holds at most MERGE_ROWS
// pairs: each thread loads its MERGE_RPT pairs of the range of each of (up
// to) 8 segments at once (coalesced, all in flight together), then the block
// adds them into the LDS slots of their rows segment by segment, in segment
// order: the same fold as k_seg_merge, without the row -> pair indirection
// and its dependent loads.
template <typename T, int ADD>
__global__ __launch_bounds__(MERGE_THREADS) void k_seg_merge_rows(const uint16_t* __restrict__ pair_off,
                                                                  const T* __restrict__ partial,
                                                                  const int32_t* __restrict__ block_pair, int nseg,
                                                                  int64_t n, T* __restrict__ y, int accum) {
    __shared__ T acc[MERGE_ROWS];
    const int64_t r0 = (int64_t)blockIdx.x * MERGE_ROWS;
    const T ident = add_identity<ADD, T>();
#pragma unroll
    for (int u = 0; u < MERGE_RPT; ++u) acc[threadIdx.x + u * MERGE_THREADS] = ident;
    const int32_t* bp = block_pair + (int64_t)blockIdx.x * nseg;
    for (int g0 = 0; g0 < nseg; g0 += 8) {
        int rr[8][MERGE_RPT];
        T pv[8][MERGE_RPT];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int g = g0 + q;
            const int lo = g < nseg ? bp[g] : 0, hi = g < nseg ? bp[nseg + g] : 0;
#pragma unroll
            for (int u = 0; u < MERGE_RPT; ++u) {
                const int p = lo + (int)threadIdx.x + u * MERGE_THREADS;
                const bool ok = p < hi;
                const int pc = ok ? p : 0;  // an in-range index for the unused lanes
                const int pr = pair_off[pc];
                rr[q][u] = ok ? pr : -1;
                pv[q][u] = partial[pc];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
#pragma unroll
            for (int u = 0; u < MERGE_RPT; ++u)
                if (rr[q][u] >= 0) acc[rr[q][u]] = add_op<ADD, T>(acc[rr[q][u]], pv[q][u]);
            __syncthreads();
        }
    }
#pragma unroll
    for (int u = 0; u < MERGE_RPT; ++u) {
        const int i = threadIdx.x + u * MERGE_THREADS;
        const int64_t r = r0 + i;
        if (r < n) y[r] = (accum == GBK_ACC_NONE) ? acc[i] : binop<T>(accum, y[r], acc[i]);
    }
}

template <typename T> __device__ __forceinline__ T hot_shfl_up(T v, int off) {
    if constexpr (sizeof(T) >= 4) return __shfl_up(v, off, 64);
    else return (T)__shfl_up((int)v, off, 64);
}
template <typename T> __device__ __forceinline__ T hot_shfl(T v, int lane) {
    if constexpr (sizeof(T) >= 4) return __shfl(v, lane, 64);
    else return (T)__shfl((int)v, lane, 64);
}

// xh[s * H + i] = x[hot_ids[s * H + i]] (unused slots: the identity, never read).
template <typename T, int ADD>
__global__ void k_hot_x(const int32_t* __restrict__ ids, const T* __restrict__ x, T* __restrict__ xh, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const int32_t c = ids[i];
        xh[i] = c >= 0 ? x[c] : add_identity<ADD, T>();
    }
}

// FLAG_EPL consecutive elements from an address aligned to their size.
typedef unsigned int gbk_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int gbk_u32x2 __attribute__((ext_vector_type(2)));
template <typename V> __device__ __forceinline__ void loadn_nt(const V* __restrict__ p, V (&out)[FLAG_EPL]) {
    constexpr int B = FLAG_EPL * (int)sizeof(V);
    if constexpr (B % 4 != 0) {
        static_assert(B % 2 == 0, "whole 2-byte pieces");
        unsigned short u[B / 2];
#pragma unroll
        for (int i = 0; i < B / 2; ++i) u[i] = __builtin_nontemporal_load((const unsigned short*)p + i);
        __builtin_memcpy(out, u, B);
    } else if constexpr (B % 8 != 0) {
        unsigned int u[B / 4];
#pragma unroll
        for (int i = 0; i < B / 4; ++i) u[i] = __builtin_nontemporal_load((const unsigned int*)p + i);
        __builtin_memcpy(out, u, B);
    } else if constexpr (B % 16 != 0) {
        gbk_u32x2 u[B / 8];
#pragma unroll
        for (int i = 0; i < B / 8; ++i) u[i] = __builtin_nontemporal_load((const gbk_u32x2*)p + i);
        __builtin_memcpy(out, u, B);
    } else {
        gbk_u32x4 u[B / 16];
#pragma unroll
        for (int i = 0; i < B / 16; ++i) u[i] = __builtin_nontemporal_load((const gbk_u32x4*)p + i);
        __builtin_memcpy(out, u, B);
    }
}

// Persistent SpMV over the flagged layout, hot x of the block's segment in
// LDS.  Grid: a multiple of 8 workgroups; workgroup b runs on XCD b % 8 and
// serves segments b % 8, b % 8 + 8, ...; within a segment, wave w of block j
// takes tiles fseg_tile[s] + j * HOT_WAVES + w in steps of (grid / 8) *
// HOT_WAVES.  Every (segment, row) pair has at least one entry, so a pair
// ends exactly at an entry whose flag is set and tiles need no row items.
// Lane L of a wave folds the FLAG_EPL entries chunk + FLAG_EPL L ... in
// order; its n-th flag ends pair first_pair + (flags of lanes below L) + n,
// the prefix taken from one ballot per bit of the per-lane counts.  A lane's sum up to its first flag is
// joined with the lanes below by a wave64 segmented scan; the tile's trailing
// sum and a deferred first pair go to carry (k_flag_fix).
template <typename T, int ADD, int MUL>
__global__ __launch_bounds__(HOT_WAVES * 64) void k_seg_flag(
    const int32_t* __restrict__ fcol, const T* __restrict__ vval, const T* __restrict__ x, const T* __restrict__ xh,
    int H, int nseg, T* __restrict__ partial, const int2* __restrict__ ftiles, const int64_t* __restrict__ fseg_tile,
    T* __restrict__ carry) {
    extern __shared__ __align__(16) unsigned char hot_lds[];
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T* s_x = (T*)hot_lds;
    const int j = blockIdx.x >> 3, nbx = gridDim.x >> 3;
    const T ident = add_identity<ADD, T>();
    const uint64_t below = (1ull << lane) - 1ull;
    for (int seg = blockIdx.x & 7; seg < nseg; seg += 8) {
        {
            // the segment's hot x values into LDS in 16-byte pieces (H is a
            // multiple of 64 slots), all of a thread's loads issued before
            // its first LDS write
            const gbk_u32x4* src = (const gbk_u32x4*)(xh + (int64_t)seg * H);
            gbk_u32x4* dst = (gbk_u32x4*)s_x;
            constexpr int ROUND = HOT_FILL * HOT_WAVES * 64;
            const int nq = (int)((int64_t)H * (int64_t)sizeof(T) / 16);
            const int full = nq / ROUND * ROUND;
            for (int q0 = threadIdx.x; q0 < full; q0 += ROUND) {
                gbk_u32x4 v[HOT_FILL];
#pragma unroll
                for (int u = 0; u < HOT_FILL; ++u) v[u] = src[q0 + u * HOT_WAVES * 64];
#pragma unroll
                for (int u = 0; u < HOT_FILL; ++u) dst[q0 + u * HOT_WAVES * 64] = v[u];
            }
            for (int q = full + (int)threadIdx.x; q < nq; q += HOT_WAVES * 64) dst[q] = src[q];
        }
        __syncthreads();
        const int64_t t_lo = fseg_tile[seg], t_hi = fseg_tile[seg + 1];
        const int64_t step = (int64_t)nbx * HOT_WAVES;
        int64_t tile = t_lo + (int64_t)j * HOT_WAVES + wid;
        if (tile < t_hi) {
            // tile t of the segment covers (part of) chunk chunk0 + t - t_lo
            const int chunk0 = ftiles[t_lo].x / FLAG_CHUNK;
            int2 ti = ftiles[tile];
            int k1 = ftiles[tile + 1].x;
            int c[FLAG_EPL];
            T a[FLAG_EPL];
            {
                const int base = (chunk0 + (int)(tile - t_lo)) * FLAG_CHUNK + lane * FLAG_EPL;
                loadn_nt(fcol + base, c);
                loadn_nt(vval + base, a);
            }
            while (true) {
                // the next tile (clamped to the last one): its descriptor
                // and, after this tile's gathers, its entries are loaded
                // while this tile is reduced
                const int64_t nt = min(tile + step, t_hi - 1);
                const int2 ti_n = ftiles[nt];
                const int k1_n = ftiles[nt + 1].x;
                T xv[FLAG_EPL];
#pragma unroll
                for (int m = 0; m < FLAG_EPL; ++m) {
                    const int idx = c[m] & 0x3fffffff;
                    if (c[m] & 0x40000000) xv[m] = s_x[idx];
                    else xv[m] = x[(uint32_t)idx];
                }
                int cn[FLAG_EPL];
                T an[FLAG_EPL];
                {
                    const int base_n = (chunk0 + (int)(nt - t_lo)) * FLAG_CHUNK + lane * FLAG_EPL;
                    loadn_nt(fcol + base_n, cn);
                    loadn_nt(vval + base_n, an);
                }
                const int k0 = ti.x;
                const int pfirst = ti.y & 0x7fffffff;
                const bool defer0 = ti.y < 0;
                const int base = (k0 & ~(FLAG_CHUNK - 1)) + lane * FLAG_EPL;
                // entries of this tile: [k0, k1) of the chunk
                const int mlo = min(max(k0 - base, 0), FLAG_EPL), mhi = min(max(k1 - base, 0), FLAG_EPL);
                const uint32_t valid = ((1u << mhi) - 1u) & ~((1u << mlo) - 1u);
                uint32_t fl = 0;
#pragma unroll
                for (int m = 0; m < FLAG_EPL; ++m) fl |= ((uint32_t)c[m] >> 31) << m;
                fl &= valid;
                const int cnt = __builtin_popcount(fl);
                int before = 0;
#pragma unroll
                for (int q = 0; q < FLAG_CNT_BITS; ++q) before += __builtin_popcountll(__ballot((cnt >> q) & 1) & below) << q;
                T acc = ident, firstsum = ident;
                int nf = 0;
                const int pbase = pfirst + before;
#pragma unroll
                for (int m = 0; m < FLAG_EPL; ++m) {
                    const T pm = ((valid >> m) & 1u) ? mul_op<MUL, T>(a[m], xv[m]) : ident;
                    acc = add_op<ADD, T>(acc, pm);
                    if ((fl >> m) & 1u) {
                        if (nf == 0) firstsum = acc;
                        else partial[pbase + nf] = acc;
                        ++nf;
                        acc = ident;
                    }
                }
                // segmented scan of (trailing sum, has-flag) over the wave
                T v = acc;
                int f = cnt > 0;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const T ov = hot_shfl_up(v, off);
                    const int of = __shfl_up(f, off, 64);
                    if (lane >= off) {
                        if (!f) v = add_op<ADD, T>(ov, v);
                        f |= of;
                    }
                }
                T ev = hot_shfl_up(v, 1);
                if (lane == 0) ev = ident;
                const T tail = hot_shfl(v, 63);
                if (cnt > 0) {
                    const T tot = add_op<ADD, T>(ev, firstsum);
                    if (before == 0 && defer0) carry[2 * tile + 1] = tot;
                    else partial[pbase] = tot;
                }
                if (lane == 0) carry[2 * tile] = tail;
                tile += step;
                if (tile >= t_hi) break;
                ti = ti_n;
                k1 = k1_n;
#pragma unroll
                for (int m = 0; m < FLAG_EPL; ++m) {
                    c[m] = cn[m];
                    a[m] = an[m];
                }
            }
        }
        __syncthreads();
    }
}

// Pairs that began in an earlier tile and end in tile t: the deferred part
// plus the trailing sums of the tiles back to the one holding the pair start,
// added from the nearest tile back; how many tiles that is was stored per
// tile at build time (fback, k_flag_back).  A thread walks up to FIX_STEPS tiles
// back itself (most pairs span one or two tiles); the walks that go further
// (pairs of hub rows) are finished by the whole wave, one at a time, 64
// earlier tiles per round (lane l: tile u - l), each round's sums combined by
// a fixed butterfly, so the result is the same on every run.
constexpr int FIX_STEPS = 8;

template <typename T, int ADD>
__global__ __launch_bounds__(256) void k_flag_fix(const int32_t* __restrict__ fback, const int2* __restrict__ ftiles,
                                                  int64_t ntiles, const T* __restrict__ carry,
                                                  T* __restrict__ partial) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const T ident = add_identity<ADD, T>();
    bool open = false;  // the walk continues from tile `next` down to tile `lo`
    int p = 0;
    int64_t next = 0, lo = 0;
    T total = ident;
    if (t < ntiles) {
        const int b = fback[t];
        if (b >= 0) {  // a pair that began earlier ends in this tile
            p = ftiles[t].y & 0x7fffffff;
            total = carry[2 * t + 1];
#pragma unroll
            for (int k = 0; k < FIX_STEPS; ++k)
                if (k < b) total = add_op<ADD, T>(carry[2 * (t - 1 - k)], total);
            if (b <= FIX_STEPS) partial[p] = total;
            else {
                open = true;
                next = t - 1 - FIX_STEPS;
                lo = t - b;
            }
        }
    }
    uint64_t todo = __ballot(open);
    while (todo) {
        const int src = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const int64_t hi0 = (int64_t)__shfl((int)next, src, 64);
        const int64_t lo0 = (int64_t)__shfl((int)lo, src, 64);
        T acc = hot_shfl(total, src);
        for (int64_t hi = hi0; hi >= lo0; hi -= 64) {
            const int64_t u = hi - lane;
            T v = u >= lo0 ? carry[2 * u] : ident;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v = add_op<ADD, T>(v, hot_shfl(v, lane ^ off));
            acc = add_op<ADD, T>(v, acc);
        }
        if (lane == src) partial[p] = acc;
    }
}

template <typename T, int ADD, int MUL> static int hot_prepare() {
    // the dynamic LDS of the persistent kernel exceeds the 64 KiB default
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        err = hipFuncSetAttribute((const void*)k_seg_flag<T, ADD, MUL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)device_lds_bytes());
    });
    if (err != hipSuccess) return fail(GBK_PANIC, "hipFuncSetAttribute: %s", hipGetErrorString(err));
    return GBK_SUCCESS;
}

template <typename T, int ADD, int MUL>
static int seg_hot_spmv(const SegCSR* s, const T* x, T* y, int accum) {
    const SegHot& h = s->hot;
    hipStream_t st = stream();
    const int64_t nh = (int64_t)s->nseg * h.H;
    if (nh > 0) {
        hipLaunchKernelGGL((k_hot_x<T, ADD>), dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, h.hot_ids, x,
                           (T*)h.xh, nh);
        GBK_LAUNCH_CHECK();
    }
    if (h.nftiles > 0) {
        GBK_TRY((hot_prepare<T, ADD, MUL>()));
        const size_t lds = (size_t)h.H * sizeof(T);
        const int grid = 8 * std::max(1, device_cu_count() / 8);
        hipLaunchKernelGGL((k_seg_flag<T, ADD, MUL>), dim3(grid), dim3(HOT_WAVES * 64), lds, st, h.fcol,
                           (const T*)s->v.val, x, (const T*)h.xh, h.H, s->nseg, (T*)s->partial, h.ftiles,
                           h.fseg_tile, (T*)h.fcarry);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_flag_fix<T, ADD>), dim3((unsigned)((h.nftiles + 255) / 256)), dim3(256), 0, st,
                           h.fback, h.ftiles, h.nftiles, (const T*)h.fcarry, (T*)s->partial);
        GBK_LAUNCH_CHECK();
    }
    if (s->n > 0) {
        hipLaunchKernelGGL((k_seg_merge_rows<T, ADD>), dim3((unsigned)((s->n + MERGE_ROWS - 1) / MERGE_ROWS)),
                           dim3(MERGE_THREADS), 0, st, s->pair_off, (const T*)s->partial, s->block_pair, s->nseg, s->n,
                           y, accum);
        GBK_LAUNCH_CHECK();
    }
    return GBK_SUCCESS;
}

static int seg_spmv(const SegCSR* s, int type, const void* x, void* y, int add, int mul, int accum) {
    if (s->hot.fcol && !order_sensitive_monoid(add)) {
        return dispatch_type(type, [&](auto tv) -> int {
            typedef decltype(tv) T;
            return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
                constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
                return seg_hot_spmv<T, ADD, MUL>(s, (const T*)x, (T*)y, accum);
            });
        });
    }
    GBK_TRY(spmv_merge_mapped(s->v, type, x, s->partial, add, mul, GBK_ACC_NONE, s->tile_map));
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value;
            (void)b;
            if (s->n > 0) {
                hipLaunchKernelGGL((k_seg_merge<T, ADD>), dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0,
                                   stream(), s->rpp, s->row_pairs, (const T*)s->partial, (T*)y, s->n, accum);
                GBK_LAUNCH_CHECK();
            }
            return GBK_SUCCESS;
        });
    });
}

// ------------------------------------------------------------------- LAV --
// LAV, device build.  Replaces MatrixLAV::build (lav/build.hpp:103-212):
// columns are relabelled by descending in-degree (stable by id, one radix
// sort); the leading columns that cover 80% of the entries form dense
// segments of 512 Ki columns, the rest one sparse segment; then it is the
// segmented CSR over the relabelled columns, with x permuted on each call.

__global__ void k_desc_key(const int32_t* __restrict__ deg, int64_t n, uint32_t* __restrict__ key) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) key[c] = 0xFFFFFFFFu - (uint32_t)deg[c];
}

__global__ void k_invert(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (int32_t)i;
}

// K = number of leading relabelled columns whose preceding entries number
// fewer than `target`: the shortest prefix that covers it.
__global__ void k_count_below(const int32_t* __restrict__ excl, int64_t n, double target, int32_t* __restrict__ K) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool below = i < n && (double)excl[i] < target;
    const uint64_t b = __ballot(below);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(K, (int32_t)__popcll(b));
}

// Segment of entry e: relabelled column c' < K lies in dense segment
// c' / w, the rest in the sparse segment ndense.
__global__ void k_lav_key(const int32_t* __restrict__ col, const int32_t* __restrict__ o2n, int64_t nnz, int64_t K,
                          int64_t w, int64_t ndense, uint32_t* __restrict__ key) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int64_t c = o2n[col[e]];
    key[e] = (uint32_t)(c < K ? c / w : ndense);
}

static int build_lav_device(const DevCSR& m, const int32_t* optr, int type, double dense_frac, int64_t w, Lav& l) {
    hipStream_t st = stream();
    const int64_t n = m.nrows, nnz = m.nnz;
    const int64_t n1 = n > 0 ? n : 1, z1 = nnz > 0 ? nnz : 1;
    GBK_TRY(dev_alloc((void**)&l.new_to_old, 4 * n1));
    GBK_TRY(dev_alloc(&l.xperm, type_size(type) * n1));
    TmpBuf deg, key, key_out, ids, o2n, sdeg, excl, Kd;
    GBK_TRY(deg.alloc(4 * (n1 + 1)));
    GBK_TRY(key.alloc(4 * n1));
    GBK_TRY(key_out.alloc(4 * n1));
    GBK_TRY(ids.alloc(4 * n1));
    GBK_TRY(o2n.alloc(4 * n1));
    GBK_TRY(sdeg.alloc(4 * n1));
    GBK_TRY(excl.alloc(4 * (n1 + 1)));
    GBK_TRY(Kd.alloc(4));
    GBK_HIP(hipMemsetAsync(Kd.p, 0, 4, st));
    int32_t K = 0;
    if (n > 0) {
        hipLaunchKernelGGL(k_deg_from_ptr, grid_b(n), dim3(256), 0, st, optr, n, deg.as<int32_t>());
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_desc_key, grid_b(n), dim3(256), 0, st, deg.as<int32_t>(), n, key.as<uint32_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(dev_iota(ids.as<int32_t>(), n, st));
        GBK_TRY(dev_sort_pairs(key.as<uint32_t>(), key_out.as<uint32_t>(), ids.as<int32_t>(), l.new_to_old, n, 32,
                               st));
        hipLaunchKernelGGL(k_invert, grid_b(n), dim3(256), 0, st, l.new_to_old, n, o2n.as<int32_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(dev_gather_i32(deg.as<int32_t>(), l.new_to_old, n, sdeg.as<int32_t>(), st));
        int32_t tot = 0;
        GBK_TRY(exclusive_scan_i32(sdeg.as<int32_t>(), excl.as<int32_t>(), n, &tot));
        hipLaunchKernelGGL(k_count_below, grid_b(n), dim3(256), 0, st, excl.as<int32_t>(), n,
                           dense_frac * (double)nnz, Kd.as<int32_t>());
        GBK_LAUNCH_CHECK();
        GBK_TRY(d2h(&K, Kd.p, 4, st));
    }
    const int64_t ndense = K > 0 ? ((int64_t)K + w - 1) / w : 0;
    const int64_t nseg = ndense + ((int64_t)K < n || n == 0 ? 1 : 0);
    l.nseg_dense = (int)ndense;
    std::vector<int32_t> seg_pair;
    {
        TmpBuf ekey;
        GBK_TRY(ekey.alloc(4 * z1));
        if (nnz > 0) {
            hipLaunchKernelGGL(k_lav_key, grid_b(nnz), dim3(256), 0, st, m.col, o2n.as<int32_t>(), nnz, (int64_t)K,
                               w, ndense, ekey.as<uint32_t>());
            GBK_LAUNCH_CHECK();
        }
        GBK_TRY(seg_pairs_device(m, type, nseg, ekey.as<uint32_t>(), o2n.as<int32_t>(), l.seg, seg_pair));
    }
    dev_free(l.seg.pair_row);
    l.seg.pair_row = nullptr;
    return seg_plain_tiles(l.seg, type, seg_pair);
}

static int build_lav(const DevCSR& m, const DevCSR& other, int type, Lav** out) {
    Lav* l = new Lav();
    const int rc = build_lav_device(m, other.row_ptr, type, 0.8, seg_cols_default(m.nrows), *l);
    if (rc != GBK_SUCCESS) {
        free_seg_fields(l->seg);
        dev_free(l->new_to_old);
        dev_free(l->xperm);
        delete l;
        return rc;
    }
    *out = l;
    return GBK_SUCCESS;
}

template <typename T>
__global__ void k_gather(const int32_t* __restrict__ idx, const T* __restrict__ x, T* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[idx[i]];
}

static int lav_spmv(const Lav* l, int type, const void* x, void* y, int add, int mul, int accum) {
    const int64_t n = l->seg.n;
    if (n > 0) {
        GBK_TRY(dispatch_type(type, [&](auto tv) -> int {
            typedef decltype(tv) T;
            hipLaunchKernelGGL(k_gather<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), l->new_to_old,
                               (const T*)x, (T*)l->xperm, n);
            GBK_LAUNCH_CHECK();
            return GBK_SUCCESS;
        }));
    }
    return seg_spmv(&l->seg, type, l->xperm, y, add, mul, accum);
}

// ------------------------------------------------------------ interface --
void format_free(gbk_matrix A) {
    for (int t = 0; t < 2; ++t) {
        free_sell(A->sell[t]);
        A->sell[t] = nullptr;
        if (A->seg[t]) {
            free_seg_fields(*A->seg[t]);
            delete A->seg[t];
            A->seg[t] = nullptr;
        }
        if (A->lav[t]) {
            free_seg_fields(A->lav[t]->seg);
            dev_free(A->lav[t]->new_to_old);
            dev_free(A->lav[t]->xperm);
            delete A->lav[t];
            A->lav[t] = nullptr;
        }
    }
}

int format_build(gbk_matrix A) {
    GBK_TRY(ensure_init());
    format_free(A);
    for (int t = 0; t < 2; ++t) {
        const DevCSR& m = t ? A->csc : A->csr;
        const DevCSR& other = t ? A->csr : A->csc;  // its row lengths are m's column counts
        int rc = GBK_SUCCESS;
        if (A->format == GBK_FMT_SELL_C) rc = build_sell(m, A->type, A->sell_c, A->sigma, &A->sell[t]);
        else if (A->format == GBK_FMT_CSR_SEG) rc = build_seg(m, other, A->type, &A->seg[t]);
        else if (A->format == GBK_FMT_LAV) rc = build_lav(m, other, A->type, &A->lav[t]);
        if (rc != GBK_SUCCESS) {
            format_free(A);
            return rc;
        }
    }
    return GBK_SUCCESS;
}

int format_spmv(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum, bool* handled) {
    *handled = false;
    const int t = transposed ? 1 : 0;
    const DevCSR& m = t ? A->csc : A->csr;
    if (A->format == GBK_FMT_SELL_C && A->sell[t]) {
        *handled = true;
        return sell_spmv(A->sell[t], m, A->type, x, y, add, mul, accum);
    }
    // Order-sensitive monoids (less, first, second) are not associative: a
    // fold split by column segments would depend on the segmentation, so they
    // take the CSR left fold (*handled stays false).
    if (A->format == GBK_FMT_CSR_SEG && A->seg[t]) {
        if (order_sensitive_monoid(add)) return GBK_SUCCESS;
        *handled = true;
        return seg_spmv(A->seg[t], A->type, x, y, add, mul, accum);
    }
    if (A->format == GBK_FMT_LAV && A->lav[t]) {
        if (transposed) return fail(GBK_INVALID_VALUE, "unsupported matrix storage format in VSpM");
        if (order_sensitive_monoid(add)) return GBK_SUCCESS;
        *handled = true;
        return lav_spmv(A->lav[t], A->type, x, y, add, mul, accum);
    }
    return GBK_SUCCESS;
}

int64_t format_bytes(gbk_matrix A) {
    const size_t vs = type_size(A->type);
    int64_t b = 0;
    for (int t = 0; t < 2; ++t) {
        if (const SellC* s = A->sell[t])
            b += (int64_t)(s->padded * (4 + vs) + s->nchunks * (8 + 4 + 8 * s->C) + (s->nproblem + s->nlong) * 4 +
                           s->lng.nnz * (4 + vs) + s->lng.nrows * 4);
        const SegCSR* sg = A->seg[t] ? A->seg[t] : (A->lav[t] ? &A->lav[t]->seg : nullptr);
        if (sg) {
            b += (int64_t)(sg->npairs * (4 + vs) + sg->v.nnz * vs + sg->v.ntiles * 28);
            if (sg->v.col) b += sg->v.nnz * 4;
            if (sg->rpp) b += (sg->n + 1) * 4;
            if (sg->row_pairs) b += sg->npairs * 4;
            if (sg->pair_row) b += sg->npairs * 4;
            if (sg->pair_off) b += sg->npairs * 2;
            if (sg->block_pair) b += ((sg->n + MERGE_ROWS - 1) / MERGE_ROWS + 1) * sg->nseg * 4;
        }
        if (sg && sg->hot.fcol)
            b += (int64_t)(sg->v.nnz * 4 + (int64_t)sg->nseg * sg->hot.H * (4 + vs) + sg->hot.nftiles * (12 + 2 * vs) +
                           (sg->nseg + 1) * 16);
        if (A->lav[t]) b += A->lav[t]->seg.n * (4 + vs);
    }
    return b;
}

}  // namespace gbk

using namespace gbk;

// Per-matrix format autotuner: builds each candidate representation, times
// `trials` plus.times SpMVs (mxv over rows) with HIP events on the library
// stream, keeps the fastest.  Candidates: CSR, segmented CSR, LAV,
// SELL-C-sigma with C = 64 and C = 32 (sigma = 512).
extern "C" int gbk_matrix_autotune(gbk_matrix A, int trials, int* chosen_format, int* chosen_c, double* times_us) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (trials < 1) trials = 3;
    GBK_TRY(ensure_init());
    struct Cand { int fmt, C; };
    const Cand cands[5] = {{GBK_FMT_CSR, 0}, {GBK_FMT_CSR_SEG, 0}, {GBK_FMT_LAV, 0}, {GBK_FMT_SELL_C, 64},
                           {GBK_FMT_SELL_C, 32}};
    const int64_t n = A->n;
    const size_t vs = type_size(A->type);
    void *x = nullptr, *y = nullptr;
    GBK_TRY(dev_alloc(&x, vs * (n > 0 ? n : 1)));
    int rc = dev_alloc(&y, vs * (n > 0 ? n : 1));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == GBK_SUCCESS && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(GBK_PANIC, "hipEventCreate failed");
    if (rc == GBK_SUCCESS) rc = fill_value(A->type, x, n, 1.0);
    const int add = A->type == GBK_BOOL ? GBK_ADD_LOR : GBK_ADD_PLUS;
    const int mul = A->type == GBK_BOOL ? GBK_MUL_LAND : GBK_MUL_TIMES;
    double best = -1.0;
    int best_i = 0;
    for (int i = 0; i < 5 && rc == GBK_SUCCESS; ++i) {
        A->format = cands[i].fmt;
        if (cands[i].C) {
            A->sell_c = cands[i].C;
            A->sigma = 512;
        }
        rc = format_build(A);
        if (rc != GBK_SUCCESS) break;
        rc = spmv_unmasked(A, 0, x, y, add, mul, GBK_ACC_NONE);  // warm-up
        if (rc != GBK_SUCCESS) break;
        (void)hipEventRecord(e0, stream());
        for (int t = 0; t < trials && rc == GBK_SUCCESS; ++t) rc = spmv_unmasked(A, 0, x, y, add, mul, GBK_ACC_NONE);
        (void)hipEventRecord(e1, stream());
        if (rc != GBK_SUCCESS) break;
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(GBK_PANIC, "autotune timing failed");
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double us = 1000.0 * ms / trials;
        if (times_us) times_us[i] = us;
        if (best < 0 || us < best) {
            best = us;
            best_i = i;
        }
    }
    if (rc == GBK_SUCCESS) {
        A->format = cands[best_i].fmt;
        if (cands[best_i].C) {
            A->sell_c = cands[best_i].C;
            A->sigma = 512;
        }
        rc = format_build(A);
        if (chosen_format) *chosen_format = cands[best_i].fmt;
        if (chosen_c) *chosen_c = cands[best_i].C;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream());
    dev_free(x);
    dev_free(y);
    return rc;
}

// Format statistics for the autotuner and reports: [0] padded SELL-C entries,
// [1] problematic SELL-C rows, [2] segments, [3] (segment, row) pairs.
extern "C" int gbk_matrix_format_stats(gbk_matrix A, int transposed, int64_t* stats) {
    if (!A || !stats) return fail(GBK_NULL_POINTER, "null argument");
    const int t = transposed ? 1 : 0;
    for (int i = 0; i < 4; ++i) stats[i] = 0;
    if (A->sell[t]) {
        stats[0] = A->sell[t]->padded;
        stats[1] = A->sell[t]->nproblem + A->sell[t]->nlong;
    }
    const SegCSR* sg = A->seg[t] ? A->seg[t] : (A->lav[t] ? &A->lav[t]->seg : nullptr);
    if (sg) {
        stats[2] = sg->nseg;
        stats[3] = sg->npairs;
    }
    return GBK_SUCCESS;
}
