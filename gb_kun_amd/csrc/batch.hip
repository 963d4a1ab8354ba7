// Batched (multi-vector) kernels: many right-hand sides in one launch.
//
//   SpMM   Y[n][k] = accum(Y, A (+.*) X[n][k]) for k <= 64 right-hand sides
//          stored row-major (the k values of one vertex are contiguous, so
//          one gather of a neighbour fetches all k of them in one coalesced
//          k*sizeof(T) access).  It is the multi-vector form of the pull SpMV
//          (src/backend/spmv/spmv_csr.h:341-384): one pass over col/val
//          serves k vectors.  Work is split by merge path (rows + nonzeros)
//          into tiles of SPMM_TILE items; one group of LW lanes (LW = k
//          rounded up to a power of two, at least 8) walks one tile, lane q
//          owning column q.  A group loads 8 col/val pairs with 8 lanes,
//          broadcasts them by shuffle and issues the 8 gathers of X before it
//          folds them, so 8 neighbour rows are in flight per group.  Rows
//          that cross a tile boundary leave their partial sums in a carry
//          buffer and a fix-up kernel adds them in tile order, so results do
//          not depend on scheduling.
//
//   MS-BFS 64 breadth-first searches at once (multi-source BFS) as a boolean
//          SpMM over 64-bit frontier words: bit q of F[v] says v is in the
//          frontier of search q.  Pull steps fold F over the in-edges of each
//          vertex that is still unvisited in some search (lor.land over 64
//          searches in one 64-bit OR) and stop scanning a vertex once every
//          missing bit has been found; push steps expand the vertices with a
//          non-empty frontier word, edge-balanced, with 64-bit atomicOr.  The
//          step direction follows the frontier's edge count as in the single
//          BFS.  Levels are 1-based with 0 = unreached, exactly as the
//          single-source driver (algorithms/bfs/bfs.hpp:94-142) reports them.
#include <algorithm>
#include <memory>
#include <vector>

#include "gbk_objects.h"
#include "batch.h"

namespace gbk {


static dim3 grid_of(int64_t n, int per = 256) {
    int64_t b = (n + per - 1) / per;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// ------------------------------------------------------------------ tiles --
__global__ void k_merge_tiles_sized(const int32_t* __restrict__ rp, int64_t nrows, int64_t nnz, int64_t ntiles,
                                    int64_t tile, int2* __restrict__ tiles) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    int64_t diag = t * tile;
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

int ensure_spmm_tiles(DevCSR& m, hipStream_t st) {
    if (m.spmm_tiles) return GBK_SUCCESS;
    const int64_t nt = (m.nrows + m.nnz + SPMM_TILE - 1) / SPMM_TILE;
    GBK_TRY(dev_alloc((void**)&m.spmm_tiles, sizeof(int2) * (nt + 1)));
    hipLaunchKernelGGL(k_merge_tiles_sized, grid_of(nt + 1), dim3(256), 0, st, m.row_ptr, m.nrows, m.nnz, nt,
                       (int64_t)SPMM_TILE, m.spmm_tiles);
    GBK_LAUNCH_CHECK();
    m.spmm_ntiles = nt;
    return GBK_SUCCESS;
}

// ------------------------------------------------------------------ SpMM --
template <typename T> __device__ __forceinline__ T shfl_t(T v, int src, int width) {
    if constexpr (sizeof(T) >= 4) return __shfl(v, src, width);
    else return (T)__shfl((int)v, src, width);
}

template <typename T, int ADD, int MUL, int LW>
__global__ __launch_bounds__(256) void k_spmm(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                              const T* __restrict__ val, const T* __restrict__ X, T* __restrict__ Y,
                                              int32_t nrows, int K, int ldx, const int2* __restrict__ tiles,
                                              int64_t ntiles, T* __restrict__ carry_head, T* __restrict__ carry_tail,
                                              int accum) {
    constexpr int GPB = 256 / LW;
    // Row ends of the block's rows, staged in LDS by the whole block (eight
    // loads in flight per thread), so that the walk below reads them there
    // rather than one dependent global load per row.
    __shared__ int32_t s_rend[GPB * SPMM_TILE + 1];
    const int64_t tb0 = (int64_t)blockIdx.x * GPB;
    const int64_t tb1 = min(tb0 + GPB, ntiles);
    const int R0 = tiles[tb0].x;
    const int nrend = tiles[tb1].x - R0 + 1;
    for (int base = 0; base < nrend; base += 256 * 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = rp[(uint32_t)min(R0 + 1 + base + u * 256 + (int)threadIdx.x, nrows)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = base + u * 256 + (int)threadIdx.x;
            if (i < nrend) s_rend[i] = v[u];
        }
    }
    __syncthreads();
    const int64_t t = tb0 + threadIdx.x / LW;
    if (t >= ntiles) return;  // whole group leaves together
    const int lane = threadIdx.x % LW;
    const bool act = lane < K;
    const int2 c0 = tiles[t], c1 = tiles[t + 1];
    const T ident = add_identity<ADD, T>();
    int row = c0.x;
    const int row_lim = c1.x;
    int kk = c0.y;
    const int k_end = c1.y;
    const bool head_open = (row < row_lim) && (rp[row] < kk);
    int row_end = s_rend[row - R0];
    T acc = ident;

    auto flush = [&](int r, T v) {
        if (!act) return;
        if (r == c0.x && head_open) {
            carry_head[t * K + lane] = v;
        } else {
            T* py = Y + (int64_t)r * ldx + lane;
            *py = (accum == GBK_ACC_NONE) ? v : binop<T>(accum, *py, v);
        }
    };

    while (kk < k_end) {
        const int nb = min(8, k_end - kk);
        int cc = 0;
        T vv = T(0);
        if (lane < nb) {
            cc = __builtin_nontemporal_load(col + kk + lane);
            vv = __builtin_nontemporal_load(val + kk + lane);
        }
        T xv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int ci = __shfl(cc, i, LW);
            xv[i] = (i < nb && act) ? X[(int64_t)ci * ldx + lane] : T(0);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const T a = shfl_t(vv, i, LW);
            if (i < nb) {
                while (row_end <= kk + i) {
                    flush(row, acc);
                    acc = ident;
                    ++row;
                    row_end = s_rend[row - R0];
                }
                acc = add_op<ADD, T>(acc, mul_op<MUL, T>(a, xv[i]));
            }
        }
        kk += nb;
    }
    while (row < row_lim) {
        flush(row, acc);
        acc = ident;
        ++row;
    }
    if (act) carry_tail[t * K + lane] = acc;
}

// Adds the carried partials of every row that begins before its ending tile.
template <typename T, int ADD, int LW>
__global__ __launch_bounds__(256) void k_spmm_fixup(const int32_t* __restrict__ rp, const int2* __restrict__ tiles,
                                                    int64_t ntiles, int K, int ldx, const T* __restrict__ carry_head,
                                                    const T* __restrict__ carry_tail, T* __restrict__ Y, int accum) {
    constexpr int GPB = 256 / LW;
    const int64_t t = (int64_t)blockIdx.x * GPB + threadIdx.x / LW;
    const int lane = threadIdx.x % LW;
    if (t >= ntiles || lane >= K) return;
    const int2 c0 = tiles[t], c1 = tiles[t + 1];
    const int r0 = c0.x;
    if (c1.x == r0) return;
    const int start = rp[r0];
    if (start >= c0.y) return;
    T total = carry_head[t * K + lane];
    for (int64_t u = t - 1; u >= 0; --u) {
        total = add_op<ADD, T>(carry_tail[u * K + lane], total);
        if (tiles[u].y <= start) break;
    }
    T* py = Y + (int64_t)r0 * ldx + lane;
    *py = (accum == GBK_ACC_NONE) ? total : binop<T>(accum, *py, total);
}

static bool order_sensitive_add(int add) {
    return add == GBK_ADD_LESS || add == GBK_ADD_FIRST || add == GBK_ADD_SECOND;
}

size_t spmm_carry_bytes(const DevCSR& m, int type, int k) {
    const int64_t nt = (m.nrows + m.nnz + SPMM_TILE - 1) / SPMM_TILE;
    return 2 * type_size(type) * (size_t)(nt > 0 ? nt : 1) * (size_t)k;
}

int spmm_run(DevCSR& m, int type, const void* X, void* Y, int k, int add, int mul, int accum, void* carry,
             hipStream_t st, int ld) {
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "SpMM needs 1 <= k <= 64 right-hand sides (got %d)", k);
    if (ld == 0) ld = k;
    if (ld < k) return fail(GBK_INVALID_VALUE, "SpMM row stride %d below its %d columns", ld, k);
    if (order_sensitive_add(add))
        return fail(GBK_INVALID_VALUE, "SpMM needs a commutative monoid (got add=%d)", add);
    GBK_TRY(ensure_spmm_tiles(m, st));
    const int64_t nt = m.spmm_ntiles;
    if (nt == 0) return GBK_SUCCESS;
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        T* head = (T*)carry;
        T* tail = head + nt * k;
        return dispatch_semiring(add, mul, [&](auto a, auto b) -> int {
            constexpr int ADD = decltype(a)::value, MUL = decltype(b)::value;
            if constexpr (ADD == GBK_ADD_LESS || ADD == GBK_ADD_FIRST || ADD == GBK_ADD_SECOND) {
                return GBK_INVALID_VALUE;
            } else {
#define GBK_SPMM(LWV)                                                                                           \
    {                                                                                                           \
        constexpr int GPB = 256 / LWV;                                                                          \
        hipLaunchKernelGGL((k_spmm<T, ADD, MUL, LWV>), grid_of(nt, GPB), dim3(256), 0, st, m.row_ptr, m.col,     \
                           (const T*)m.val, (const T*)X, (T*)Y, (int32_t)m.nrows, k, ld, m.spmm_tiles, nt,      \
                           head, tail, accum);                                                                  \
        GBK_LAUNCH_CHECK();                                                                                     \
        hipLaunchKernelGGL((k_spmm_fixup<T, ADD, LWV>), grid_of(nt, GPB), dim3(256), 0, st, m.row_ptr,           \
                           m.spmm_tiles, nt, k, ld, (const T*)head, (const T*)tail, (T*)Y, accum);              \
        GBK_LAUNCH_CHECK();                                                                                     \
        return GBK_SUCCESS;                                                                                     \
    }
                if (k <= 8) GBK_SPMM(8)
                if (k <= 16) GBK_SPMM(16)
                if (k <= 32) GBK_SPMM(32)
                GBK_SPMM(64)
#undef GBK_SPMM
            }
        });
    });
}

// ---------------------------------------------------------------- MS-BFS --
__device__ __forceinline__ unsigned long long lanemask_lt_b() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ void k_msbfs_init(uint64_t* __restrict__ F, uint64_t* __restrict__ V, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        F[i] = 0;
        V[i] = 0;
    }
}

// One thread per search: sets the source bit; levels[v*64 + q] = 1.
__global__ void k_msbfs_sources(const int32_t* __restrict__ src, int nsrc, uint64_t* __restrict__ F,
                                uint64_t* __restrict__ V, int32_t* __restrict__ lev) {
    const int q = threadIdx.x;
    if (q < nsrc) {
        const int s = src[q];
        atomicOr((unsigned long long*)(F + s), 1ull << q);
        atomicOr((unsigned long long*)(V + s), 1ull << q);
        lev[(int64_t)s * 64 + q] = 1;
    }
}

// Pull step: G lanes per vertex.  A vertex still missing some search bits
// ORs the frontier words of its in-neighbours, stopping once every missing
// bit is found.  Writes the new frontier word of every vertex (0 if none).
// Block-level sums: `c` added and `m` ORed over the block, then one global
// atomic each per block (kernels run on a capped grid, so a step costs a few
// thousand global atomics, not one per wave).
__device__ __forceinline__ void block_flush(unsigned long long c, unsigned long long m,
                                            unsigned long long* __restrict__ counters,
                                            unsigned long long* __restrict__ active) {
    __shared__ unsigned long long s_c, s_m;
    if (threadIdx.x == 0) {
        s_c = 0;
        s_m = 0;
    }
    __syncthreads();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        m |= __shfl_xor(m, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c) atomicAdd(&s_c, c);
        if (m) atomicOr(&s_m, m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_c && counters) atomicAdd(counters, s_c);
        if (s_m && active) atomicOr(active, s_m);
    }
}

constexpr int64_t MSBFS_GRID = 8192;

static dim3 capped_grid(int64_t threads) {
    int64_t b = (threads + 255) / 256;
    if (b > MSBFS_GRID) b = MSBFS_GRID;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// Pull step: G lanes per vertex.  A vertex still missing some search bits
// ORs the frontier words of its in-neighbours, stopping once every missing
// bit is found.  Writes the new frontier word of every vertex (0 if none).
// Grid-stride over vertex groups; the stride is a multiple of G, so a group
// keeps one vertex per pass.
// Lanes per vertex of the MS-BFS pull step for an average in-degree.  A
// vertex stops only once every missing search bit is found, so it scans far
// more of its list than a single search does; wide groups win (directed
// R-MAT 22, 64 searches: G = 1: 16.2, 4: 8.5, 16: 5.9 ms).
static int msbfs_group(double avg) {
    int G = 4;
    while (G < 32 && G < avg) G <<= 1;
    return G;
}

template <int G>
__global__ __launch_bounds__(256) void k_msbfs_pull(const int32_t* __restrict__ crp, const int32_t* __restrict__ ccol,
                                                    int64_t n, uint64_t all, const uint64_t* __restrict__ F,
                                                    uint64_t* __restrict__ V, uint64_t* __restrict__ N,
                                                    int32_t* __restrict__ lev, int32_t next_level,
                                                    unsigned long long* __restrict__ counters,
                                                    unsigned long long* __restrict__ active) {
    unsigned long long c = 0, m = 0;
    const int64_t total = n * G;
    for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
         gid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = gid / G;
        const int lane = (int)(gid % G);
        uint64_t found = 0;
        const uint64_t vis = V[v];
        const uint64_t want = all & ~vis;
        if (want) {
            // group-uniform trip count; the group stops once its lanes together hold every missing bit
            const int e0 = crp[v], e1 = crp[v + 1];
            for (int base = e0; base < e1; base += 4 * G) {
                uint64_t f[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int e = base + k * G + lane;
                    f[k] = e < e1 ? F[ccol[e]] : 0ull;
                }
                found |= (f[0] | f[1] | f[2] | f[3]) & want;
                uint64_t gf = found;
                if constexpr (G > 1) {
#pragma unroll
                    for (int off = G / 2; off > 0; off >>= 1)
                        gf |= (uint64_t)__shfl_xor((unsigned long long)gf, off, G);
                }
                if (gf == want) break;
            }
            if constexpr (G > 1) {
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1)
                    found |= (uint64_t)__shfl_xor((unsigned long long)found, off, G);
            }
        }
        if (lane == 0) {
            N[v] = found;
            if (found) {
                V[v] = vis | found;
                c += (unsigned long long)__popcll(found);
                m |= found;
            }
        }
        // level writes: lanes of the group split the 64 search bits
        if (found) {
            for (int q = lane; q < 64; q += G)
                if ((found >> q) & 1ull) lev[v * 64 + q] = next_level;
        }
    }
    block_flush(c, m, counters, active);
}

// Frontier statistics: vertices with a non-empty frontier word and the sum
// of their out-degrees (the push / pull decision).
__global__ __launch_bounds__(256) void k_msbfs_fstats(const int32_t* __restrict__ rp, const uint64_t* __restrict__ F,
                                                      int64_t n, unsigned long long* __restrict__ nq,
                                                      unsigned long long* __restrict__ edges) {
    unsigned long long c = 0, d = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        if (F[v]) {
            ++c;
            d += (unsigned long long)(rp[v + 1] - rp[v]);
        }
    }
    __shared__ unsigned long long s_c, s_d;
    if (threadIdx.x == 0) {
        s_c = 0;
        s_d = 0;
    }
    __syncthreads();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c) atomicAdd(&s_c, c);
        if (d) atomicAdd(&s_d, d);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_c) atomicAdd(nq, s_c);
        if (s_d) atomicAdd(edges, s_d);
    }
}

// Queue of the vertices with a non-empty frontier word, in vertex order:
// per-chunk counts, a scan, then a ballot / prefix-sum write.
constexpr int FQ_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_fq_count(const uint64_t* __restrict__ F, int64_t n, int32_t* __restrict__ counts) {
    __shared__ int s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * FQ_CHUNK;
    int c = 0;
    for (int i = threadIdx.x; i < FQ_CHUNK; i += 256) {
        const int64_t v = base + i;
        if (v < n && F[v]) ++c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_c, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_c;
}

__global__ __launch_bounds__(256) void k_fq_compact(const uint64_t* __restrict__ F, int64_t n,
                                                    const int32_t* __restrict__ offs, int32_t* __restrict__ q) {
    __shared__ int s_w[4];
    const int64_t base = (int64_t)blockIdx.x * FQ_CHUNK;
    int o = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int r = 0; r < FQ_CHUNK; r += 256) {
        const int64_t v = base + r + threadIdx.x;
        const bool p = v < n && F[v] != 0;
        const unsigned long long b = __ballot(p);
        if (lane == 0) s_w[wid] = __popcll(b);
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            if (w < wid) pre += s_w[w];
            tot += s_w[w];
        }
        if (p) q[o + pre + __popcll(b & lanemask_lt_b())] = (int32_t)v;
        o += tot;
        __syncthreads();
    }
}

// Frontier queue (vertices with a non-empty frontier word, ascending) built
// without a shared counter; `counts` and `offs` need n / FQ_CHUNK + 1 entries.
static int build_frontier_queue(const uint64_t* F, int64_t n, int32_t* q, int32_t* counts, int32_t* offs,
                                void* partial, size_t part_bytes, hipStream_t st, int64_t* nq) {
    const int64_t nb = (n + FQ_CHUNK - 1) / FQ_CHUNK;
    hipLaunchKernelGGL(k_fq_count, dim3((unsigned)(nb > 0 ? nb : 1)), dim3(256), 0, st, F, n, counts);
    GBK_LAUNCH_CHECK();
    int32_t total = 0;
    GBK_TRY(exclusive_scan_i32_on(counts, offs, nb, &total, st, partial, part_bytes));
    hipLaunchKernelGGL(k_fq_compact, dim3((unsigned)(nb > 0 ? nb : 1)), dim3(256), 0, st, F, n, offs, q);
    GBK_LAUNCH_CHECK();
    *nq = total;
    return GBK_SUCCESS;
}

__global__ void k_queue_degrees(const int32_t* __restrict__ rp, const int32_t* __restrict__ q, int64_t nq,
                                int32_t* __restrict__ deg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        const int u = q[i];
        deg[i] = rp[u + 1] - rp[u];
    }
}

// Push step: thread e owns edge e of the queue's concatenated adjacency.
__global__ __launch_bounds__(256) void k_msbfs_push(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ q, int64_t nq,
                                                    const int32_t* __restrict__ offs, int64_t nedges,
                                                    const uint64_t* __restrict__ F, const uint64_t* __restrict__ V,
                                                    uint64_t* __restrict__ N) {
    __shared__ int32_t s_offs[EO_WIN + 1];
    __shared__ int64_t s_i0;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = edge_owner(offs, nq, e, s_offs, &s_i0);
    if (e >= nedges) return;
    const int u = q[i];
    const int v = col[rp[u] + (e - offs[i])];
    const uint64_t bits = F[u] & ~V[v];
    if (bits) atomicOr((unsigned long long*)(N + v), (unsigned long long)bits);
}

// After a push step: new = N & ~V per vertex; marks visited and levels.
__global__ __launch_bounds__(256) void k_msbfs_settle(int64_t n, uint64_t* __restrict__ V, uint64_t* __restrict__ N,
                                                      int32_t* __restrict__ lev, int32_t next_level,
                                                      unsigned long long* __restrict__ counters,
                                                      unsigned long long* __restrict__ active) {
    unsigned long long c = 0, m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t vis = V[v];
        const uint64_t nw = N[v] & ~vis;
        N[v] = nw;
        if (nw) {
            V[v] = vis | nw;
            uint64_t b = nw;
            while (b) {
                const int qb = __ffsll((long long)b) - 1;
                lev[v * 64 + qb] = next_level;
                b &= b - 1;
            }
            c += (unsigned long long)__popcll(nw);
            m |= nw;
        }
    }
    block_flush(c, m, counters, active);
}

__global__ void k_zero_u64(uint64_t* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// lev[v*64 + q] -> out[q*n + v] for q < nsrc.
__global__ void k_msbfs_transpose(const int32_t* __restrict__ lev, int64_t n, int nsrc, int32_t* __restrict__ out) {
    __shared__ int32_t tile[64][65];
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 256 threads: 4 rows per pass
    for (int r = ty; r < 64; r += 4) {
        const int64_t v = v0 + r;
        tile[r][tx] = (v < n) ? lev[v * 64 + tx] : 0;
    }
    __syncthreads();
    for (int q = ty; q < nsrc; q += 4) {
        const int64_t v = v0 + tx;
        if (v < n) out[(int64_t)q * n + v] = tile[tx][q];
    }
}

// Workspace of one multi-source BFS over n vertices, carved in this order
// (each piece 256-byte aligned): F, V, N (u64[n]); levels int32[n*64];
// queue, degrees, offsets int32[n+1]; scan partials; 3 counters.
struct MsbfsLayout {
    uint64_t *F, *V, *N;
    int32_t *lev, *q, *deg, *offs;
    void* partial;
    size_t part_bytes;
    unsigned long long* ctr;
    size_t total;
};

static size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

static MsbfsLayout msbfs_layout(int64_t n, char* base) {
    MsbfsLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* r = base ? base + off : nullptr;
        off = up256(off + bytes);
        return r;
    };
    L.F = (uint64_t*)take(sizeof(uint64_t) * n);
    L.V = (uint64_t*)take(sizeof(uint64_t) * n);
    L.N = (uint64_t*)take(sizeof(uint64_t) * n);
    L.lev = (int32_t*)take(sizeof(int32_t) * 64 * (size_t)n);
    L.q = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.deg = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.offs = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.part_bytes = scan_partial_bytes(n);
    L.partial = take(L.part_bytes);
    L.ctr = (unsigned long long*)take(4 * sizeof(unsigned long long));
    L.total = off;
    return L;
}

size_t msbfs_ws_bytes(int64_t n) { return msbfs_layout(n, nullptr).total + 256; }

int msbfs_run(gbk_matrix A, const int32_t* src_dev, int nsrc, void* ws, size_t ws_bytes, int32_t* levels_out,
              int64_t* stats, hipStream_t st) {
    const int64_t n = A->n;
    if (nsrc < 1 || nsrc > 64) return fail(GBK_INVALID_VALUE, "multi-source BFS takes 1..64 sources (got %d)", nsrc);
    if (ws_bytes < msbfs_ws_bytes(n)) return fail(GBK_INSUFFICIENT_SPACE, "multi-source BFS workspace too small");
    const DevCSR& out = A->csr;
    const DevCSR& in = A->csc;
    char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    MsbfsLayout L = msbfs_layout(n, base);
    uint64_t* F = L.F;
    uint64_t* V = L.V;
    uint64_t* N = L.N;
    int32_t* lev = L.lev;
    int32_t* q = L.q;
    int32_t* deg = L.deg;
    int32_t* offs = L.offs;
    void* partial = L.partial;
    const size_t part_bytes = L.part_bytes;
    unsigned long long* ctr = L.ctr;  // [0] new bits, [1] frontier edges, [2] queue length
    const uint64_t all = nsrc == 64 ? ~0ull : ((1ull << nsrc) - 1);

    hipLaunchKernelGGL(k_msbfs_init, grid_of(n), dim3(256), 0, st, F, V, n);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zero_u64, grid_of(32 * n), dim3(256), 0, st, (uint64_t*)lev, 32 * n);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_msbfs_sources, dim3(1), dim3(64), 0, st, src_dev, nsrc, F, V, lev);
    GBK_LAUNCH_CHECK();

    const int G = msbfs_group(n ? (double)in.nnz / (double)n : 0);
    int64_t levels = 1, pushes = 0, pulls = 0, reached = nsrc, edges_scanned = 0;
    int32_t cur = 1;
    unsigned long long h[3];
    while (true) {
        // frontier size and out-edge count decide the direction
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, (uint64_t*)ctr, (int64_t)3);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_msbfs_fstats, capped_grid(n), dim3(256), 0, st, out.row_ptr, F, n, ctr + 2, ctr + 1);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        int64_t nq = (int64_t)h[2];
        const unsigned long long m_f = h[1];
        if (nq == 0) break;
        edges_scanned += (int64_t)m_f;
        const bool push = (double)m_f * 14.0 < (double)out.nnz;
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, (uint64_t*)ctr, (int64_t)1);
        GBK_LAUNCH_CHECK();
        if (push) {
            ++pushes;
            GBK_TRY(build_frontier_queue(F, n, q, deg, offs, partial, part_bytes, st, &nq));
            hipLaunchKernelGGL(k_zero_u64, grid_of(n), dim3(256), 0, st, N, n);
            GBK_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_queue_degrees, grid_of(nq), dim3(256), 0, st, out.row_ptr, q, nq, deg);
            GBK_LAUNCH_CHECK();
            int32_t ne = 0;
            GBK_TRY(exclusive_scan_i32_on(deg, offs, nq, &ne, st, partial, part_bytes));
            if (ne > 0) {
                hipLaunchKernelGGL(k_msbfs_push, grid_of(ne), dim3(256), 0, st, out.row_ptr, out.col, q, nq, offs,
                                   (int64_t)ne, F, V, N);
                GBK_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_msbfs_settle, capped_grid(n), dim3(256), 0, st, n, V, N, lev, cur + 1, ctr,
                               (unsigned long long*)nullptr);
            GBK_LAUNCH_CHECK();
        } else {
            ++pulls;
#define GBK_MSPULL(GG)                                                                                         \
    case GG:                                                                                                   \
        hipLaunchKernelGGL(k_msbfs_pull<GG>, capped_grid(n * GG), dim3(256), 0, st, in.row_ptr, in.col, n, all, F, \
                           V, N, lev, cur + 1, ctr, (unsigned long long*)nullptr);                             \
        break;
            switch (G) {
                GBK_MSPULL(4)
                GBK_MSPULL(8)
                GBK_MSPULL(16)
                GBK_MSPULL(32)
            }
#undef GBK_MSPULL
            GBK_LAUNCH_CHECK();
        }
        unsigned long long newbits = 0;
        GBK_HIP(hipMemcpyAsync(&newbits, ctr, sizeof(newbits), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        if (newbits == 0) break;
        reached += (int64_t)newbits;
        std::swap(F, N);
        ++cur;
        ++levels;
    }
    hipLaunchKernelGGL(k_msbfs_transpose, grid_of(n, 64), dim3(256), 0, st, lev, n, nsrc, levels_out);
    GBK_LAUNCH_CHECK();
    if (stats) {
        stats[0] = levels;
        stats[1] = pushes;
        stats[2] = pulls;
        stats[3] = edges_scanned;
        stats[4] = reached;
    }
    return GBK_SUCCESS;
}


// ------------------------------------------------------ BFS serving engine --
// Multi-source BFS as a continuously refilled batch: 64 slots, one search
// per slot.  A slot that finishes (its frontier is empty) is retired and can
// take a new search at the next step; levels are stored as global step
// numbers and rebased per slot on extraction.

// Clears the search bits of `mask` in every vertex (a slot being reused).
__global__ void k_bfs_slot_reset(int64_t n, uint64_t mask, uint64_t* __restrict__ F, uint64_t* __restrict__ V,
                                 uint64_t* __restrict__ N, int32_t* __restrict__ lev) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    F[v] &= ~mask;
    N[v] &= ~mask;
    const uint64_t had = V[v] & mask;
    if (had) {
        V[v] &= ~mask;
        uint64_t b = had;
        while (b) {
            const int q = __ffsll((long long)b) - 1;
            lev[v * 64 + q] = 0;
            b &= b - 1;
        }
    }
}

// Seeds search `slot[i]` at vertex `src[i]` (one thread per seed).
__global__ void k_bfs_seed(const int32_t* __restrict__ src, const int32_t* __restrict__ slot, int count,
                           uint64_t* __restrict__ F, uint64_t* __restrict__ V, int32_t* __restrict__ lev, int32_t cur) {
    const int i = threadIdx.x;
    if (i < count) {
        const int64_t s = src[i];
        const int q = slot[i];
        atomicOr((unsigned long long*)(F + s), 1ull << q);
        atomicOr((unsigned long long*)(V + s), 1ull << q);
        lev[s * 64 + q] = cur;
    }
}

// out[v] = level of v in search q (1-based, 0 = unreached).
__global__ void k_bfs_extract(const int32_t* __restrict__ lev, int64_t n, int q, int32_t base, int32_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        const int32_t l = lev[v * 64 + q];
        out[v] = l ? l - base + 1 : 0;
    }
}

namespace {

class BfsEngine : public Engine {
  public:
    BfsEngine(gbk_matrix A, hipStream_t st) : A_(A), st_(st), n_(A->n) {}
    ~BfsEngine() override {
        dev_free(ws_);
        dev_free(seeds_);
        if (hctr_) (void)hipHostFree(hctr_);
    }
    int init() {
        ws_bytes_ = msbfs_ws_bytes(n_) + 512;
        GBK_TRY(dev_alloc(&ws_, ws_bytes_));
        char* base = (char*)(((uintptr_t)ws_ + 255) & ~(uintptr_t)255);
        L_ = msbfs_layout(n_, base);
        GBK_TRY(dev_alloc((void**)&seeds_, sizeof(int32_t) * 128));
        GBK_HIP(hipHostMalloc((void**)&hctr_, 4 * sizeof(unsigned long long), hipHostMallocDefault));
        hipLaunchKernelGGL(k_msbfs_init, grid_of(n_), dim3(256), 0, st_, L_.F, L_.V, n_);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_msbfs_init, grid_of(n_), dim3(256), 0, st_, L_.N, L_.N, n_);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_u64, grid_of(32 * n_), dim3(256), 0, st_, (uint64_t*)L_.lev, 32 * n_);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipStreamSynchronize(st_));
        G_ = msbfs_group(n_ ? (double)A_->csc.nnz / (double)n_ : 0);
        return GBK_SUCCESS;
    }
    int slots() const override { return 64; }
    size_t bytes() const override { return ws_bytes_; }
    size_t result_bytes() const override { return sizeof(int32_t) * (size_t)n_; }
    int admit(int slot, int32_t source) override {
        if (source < 0 || source >= n_) return fail(GBK_INVALID_INDEX, "source out of range");
        pend_src_.push_back(source);
        pend_slot_.push_back(slot);
        return GBK_SUCCESS;
    }
    int step(std::vector<int>& finished) override {
        finished.clear();
        const DevCSR& out = A_->csr;
        const DevCSR& in = A_->csc;
        unsigned long long* ctr = L_.ctr;
        if (!pend_slot_.empty()) {
            uint64_t mask = 0;
            for (int q : pend_slot_) mask |= 1ull << q;
            hipLaunchKernelGGL(k_bfs_slot_reset, grid_of(n_), dim3(256), 0, st_, n_, mask, L_.F, L_.V, L_.N, L_.lev);
            GBK_LAUNCH_CHECK();
            std::vector<int32_t> buf(pend_src_);
            buf.insert(buf.end(), pend_slot_.begin(), pend_slot_.end());
            GBK_HIP(hipMemcpyAsync(seeds_, buf.data(), sizeof(int32_t) * buf.size(), hipMemcpyHostToDevice, st_));
            const int cnt = (int)pend_slot_.size();
            hipLaunchKernelGGL(k_bfs_seed, dim3(1), dim3(64), 0, st_, seeds_, seeds_ + cnt, cnt, L_.F, L_.V, L_.lev, cur_);
            GBK_LAUNCH_CHECK();
            for (int q : pend_slot_) {
                base_[q] = cur_;
                occupied_ |= 1ull << q;
            }
            GBK_HIP(hipStreamSynchronize(st_));  // buf leaves scope
            pend_src_.clear();
            pend_slot_.clear();
        }
        if (!occupied_) return GBK_SUCCESS;
        // frontier size and out-edge count decide the direction
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st_, (uint64_t*)ctr, (int64_t)4);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_msbfs_fstats, capped_grid(n_), dim3(256), 0, st_, out.row_ptr, L_.F, n_, ctr + 2, ctr + 1);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipMemcpyAsync(hctr_, ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
        GBK_HIP(hipStreamSynchronize(st_));
        int64_t nq = (int64_t)hctr_[2];
        const bool push = (double)hctr_[1] * 14.0 < (double)out.nnz;
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st_, (uint64_t*)ctr, (int64_t)4);
        GBK_LAUNCH_CHECK();
        if (nq > 0) {
            if (push) {
                GBK_TRY(build_frontier_queue(L_.F, n_, L_.q, L_.deg, L_.offs, L_.partial, L_.part_bytes, st_, &nq));
                hipLaunchKernelGGL(k_zero_u64, grid_of(n_), dim3(256), 0, st_, L_.N, n_);
                GBK_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_queue_degrees, grid_of(nq), dim3(256), 0, st_, out.row_ptr, L_.q, nq, L_.deg);
                GBK_LAUNCH_CHECK();
                int32_t ne = 0;
                GBK_TRY(exclusive_scan_i32_on(L_.deg, L_.offs, nq, &ne, st_, L_.partial, L_.part_bytes));
                if (ne > 0) {
                    hipLaunchKernelGGL(k_msbfs_push, grid_of(ne), dim3(256), 0, st_, out.row_ptr, out.col, L_.q, nq,
                                       L_.offs, (int64_t)ne, L_.F, L_.V, L_.N);
                    GBK_LAUNCH_CHECK();
                }
                hipLaunchKernelGGL(k_msbfs_settle, capped_grid(n_), dim3(256), 0, st_, n_, L_.V, L_.N, L_.lev, cur_ + 1,
                                   ctr, ctr + 3);
                GBK_LAUNCH_CHECK();
            } else {
#define GBK_SPULL(GG)                                                                                       \
    case GG:                                                                                                \
        hipLaunchKernelGGL(k_msbfs_pull<GG>, capped_grid(n_ * GG), dim3(256), 0, st_, in.row_ptr, in.col, n_, \
                           (uint64_t)occupied_, L_.F, L_.V, L_.N, L_.lev, cur_ + 1, ctr, ctr + 3);          \
        break;
                switch (G_) {
                    GBK_SPULL(4)
                    GBK_SPULL(8)
                    GBK_SPULL(16)
                    GBK_SPULL(32)
                }
#undef GBK_SPULL
                GBK_LAUNCH_CHECK();
            }
            GBK_HIP(hipMemcpyAsync(hctr_, ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
            GBK_HIP(hipStreamSynchronize(st_));
        } else {
            hctr_[3] = 0;
        }
        const uint64_t active = hctr_[3] & occupied_;
        for (int q = 0; q < 64; ++q)
            if (((occupied_ >> q) & 1ull) && !((active >> q) & 1ull)) finished.push_back(q);
        std::swap(L_.F, L_.N);
        ++cur_;
        return GBK_SUCCESS;
    }
    int extract(int slot, void* out_dev) override {
        hipLaunchKernelGGL(k_bfs_extract, grid_of(n_), dim3(256), 0, st_, L_.lev, n_, slot, base_[slot],
                           (int32_t*)out_dev);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    }
    void release(int slot) override { occupied_ &= ~(1ull << slot); }

  private:
    gbk_matrix A_;
    hipStream_t st_;
    int64_t n_;
    void* ws_ = nullptr;
    int32_t* seeds_ = nullptr;  // source, slot pairs of the next seeding (<= 64 each)
    size_t ws_bytes_ = 0;
    MsbfsLayout L_{};
    unsigned long long* hctr_ = nullptr;
    uint64_t occupied_ = 0;
    int32_t cur_ = 1;
    int32_t base_[64] = {0};
    int G_ = 4;
    std::vector<int32_t> pend_src_, pend_slot_;
};

}  // namespace

int make_bfs_engine(gbk_matrix A, hipStream_t st, std::unique_ptr<Engine>* out) {
    auto e = std::make_unique<BfsEngine>(A, st);
    GBK_TRY(e->init());
    *out = std::move(e);
    return GBK_SUCCESS;
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_spmm_device(gbk_matrix A, int transposed, const void* X, void* Y, int k, gbk_semiring op,
                               int accum) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!X || !Y) return fail(GBK_NULL_POINTER, "null dense block pointer");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (X == Y) return fail(GBK_INVALID_VALUE, "SpMM input and output blocks must not alias");
    if (accum < GBK_ACC_NONE || accum > GBK_ACC_NE) return fail(GBK_INVALID_VALUE, "unknown accumulator %d", accum);
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "SpMM needs 1 <= k <= 64 right-hand sides (got %d)", k);
    GBK_TRY(ensure_init());
    DevCSR& m = transposed ? A->csc : A->csr;
    void* carry;
    GBK_TRY(scratch(SLOT_SPMM_CARRY, spmm_carry_bytes(m, A->type, k), &carry));
    return spmm_run(m, A->type, X, Y, k, op.add, op.mul, accum, carry, stream());
}

extern "C" int gbk_msbfs_device(gbk_matrix A, const int64_t* sources, int nsrc, int32_t* levels_dev, int64_t* stats) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!sources || !levels_dev) return fail(GBK_NULL_POINTER, "null pointer");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (nsrc < 1 || nsrc > 64) return fail(GBK_INVALID_VALUE, "multi-source BFS takes 1..64 sources (got %d)", nsrc);
    std::vector<int32_t> s32(nsrc);
    for (int i = 0; i < nsrc; ++i) {
        if (sources[i] < 0 || sources[i] >= A->n)
            return fail(GBK_INVALID_INDEX, "source %lld out of range [0, %lld)", (long long)sources[i],
                        (long long)A->n);
        s32[i] = (int32_t)sources[i];
    }
    GBK_TRY(ensure_init());
    hipStream_t st = stream();
    void *ws, *src_dev;
    GBK_TRY(scratch(SLOT_MSBFS, msbfs_ws_bytes(A->n), &ws));
    GBK_TRY(scratch(SLOT_SOURCES, sizeof(int32_t) * 64, &src_dev));
    GBK_HIP(hipMemcpyAsync(src_dev, s32.data(), sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
    return msbfs_run(A, (const int32_t*)src_dev, nsrc, ws, msbfs_ws_bytes(A->n), levels_dev, stats, st);
}

extern "C" int gbk_msbfs(gbk_matrix A, const int64_t* sources, int nsrc, int32_t* levels, int64_t* stats) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!levels) return fail(GBK_NULL_POINTER, "null levels");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (nsrc < 1 || nsrc > 64) return fail(GBK_INVALID_VALUE, "multi-source BFS takes 1..64 sources (got %d)", nsrc);
    GBK_TRY(ensure_init());
    void* out;
    GBK_TRY(scratch(SLOT_MSBFS_OUT, sizeof(int32_t) * (size_t)nsrc * (size_t)A->n, &out));
    GBK_TRY(gbk_msbfs_device(A, sources, nsrc, (int32_t*)out, stats));
    GBK_HIP(hipMemcpyAsync(levels, out, sizeof(int32_t) * (size_t)nsrc * (size_t)A->n, hipMemcpyDeviceToHost,
                           stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}
