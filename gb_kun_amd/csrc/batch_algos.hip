// Multi-query graph drivers on the batched kernels: personalized PageRank
// for k sources at once (one SpMM per iteration instead of k SpMVs) and
// min.plus SSSP for k sources at once.
//
// Personalized PageRank (the per-source form of LAGraph_page_rank_sinks,
// algorithms/pr/pr.hpp:20-153): on a matrix whose entry (v, u) is the
// transition probability 1/outdeg(u) of edge u -> v,
//     r_q' = d * A r_q + ((1 - d) + d * sink_q) e_{s_q},
//     sink_q = sum of r_q over vertices without out-edges,
// iterated per column until the L1 change of that column drops below tol
// (columns that converge are frozen while the others continue).  With a
// uniform start and teleport vector instead of e_{s_q} this is the
// reference's PageRank with sinks.
//
// Multi-source SSSP (sssp_bellman_ford_blast, algorithms/sssp/sssp_blast.hpp:8-72,
// for k sources): dense Bellman-Ford rounds, each one min.plus SpMM over the
// k distance columns (int32, saturating at INT32_MAX), until no column
// changes: the reference's own formulation (a dense vxm with MinimumPlus per
// iteration), batched over sources.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "batch.h"
#include "gbk_objects.h"

namespace gbk {

static dim3 grid_b(int64_t n, int per = 256) {
    int64_t b = (n + per - 1) / per;
    return dim3((unsigned)(b > 0 ? b : 1));
}

static size_t up256b(size_t v) { return (v + 255) & ~(size_t)255; }

// ------------------------------------------------------------------- PPR --
__global__ void k_ppr_init(float* __restrict__ R, int64_t n, int k, const int32_t* __restrict__ src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * k) {
        const int q = (int)(i % k);
        const int64_t v = i / k;
        R[i] = (src[q] == v) ? 1.0f : 0.0f;
    }
}

// out-degree = entries in the column = row length of the CSC
__global__ void k_sink_flags(const int32_t* __restrict__ crp, int64_t n, int32_t* __restrict__ flag) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = (crp[v + 1] == crp[v]) ? 1 : 0;
}

__global__ void k_compact_flags(const int32_t* __restrict__ flag, const int32_t* __restrict__ offs, int64_t n,
                                int32_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && flag[v]) out[offs[v]] = (int32_t)v;
}

// sink[q] = sum over sinks of R[v][q] (fp64 accumulation, one atomic per
// wave-column pair).  Block = 256 threads = 256/k sinks x k columns.
__global__ __launch_bounds__(256) void k_ppr_sink(const float* __restrict__ R, int k, int ld,
                                                  const int32_t* __restrict__ sinks, int64_t nsinks,
                                                  double* __restrict__ sink) {
    __shared__ double s[256];
    const int per = 256 / k;
    const int q = threadIdx.x % k;
    const int j = threadIdx.x / k;
    double acc = 0.0;
    if (j < per) {
        for (int64_t i = (int64_t)blockIdx.x * per + j; i < nsinks; i += (int64_t)gridDim.x * per)
            acc += (double)R[(int64_t)sinks[i] * ld + q];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = 0.0;
        for (int jj = 0; jj < per; ++jj) t += s[jj * k + threadIdx.x];
        atomicAdd(sink + threadIdx.x, t);
    }
}

// Rn = d * Y + teleport on the source rows; frozen columns keep R.  Adds the
// L1 change of every column to delta[q] and, when sink_flag is given, the new
// ranks of the sinks to sink_next[q] (the next iteration's sink sums, so that
// no separate pass gathers the sink rows).
// Rn may be Y (each element is read, then written, by the same thread).
__global__ __launch_bounds__(256) void k_ppr_update(const float* __restrict__ R, const float* Y, float* Rn, int64_t n,
                                                    int k, int ld,
                                                    const int32_t* __restrict__ src, const double* __restrict__ sink,
                                                    const int32_t* __restrict__ done, float damping,
                                                    double* __restrict__ delta, const int32_t* __restrict__ sink_flag,
                                                    double* __restrict__ sink_next) {
    __shared__ double s[256], s2[256];
    const int per = 256 / k;
    const int q = threadIdx.x % k;
    const int j = threadIdx.x / k;
    double acc = 0.0, sacc = 0.0;
    if (j < per) {
        const float tele = (float)((1.0 - (double)damping) + (double)damping * sink[q]);
        const int sq = src[q];
        const bool frozen = done[q] != 0;
        for (int64_t v = (int64_t)blockIdx.x * per + j; v < n; v += (int64_t)gridDim.x * per) {
            const int64_t i = v * ld + q;
            const float old = R[i];
            float nv;
            if (frozen) {
                nv = old;
            } else {
                nv = damping * Y[i];
                if (v == sq) nv += tele;
            }
            Rn[i] = nv;
            acc += fabs((double)nv - (double)old);
            if (sink_flag && sink_flag[v]) sacc += (double)nv;
        }
    }
    s[threadIdx.x] = acc;
    s2[threadIdx.x] = sacc;
    __syncthreads();
    if (threadIdx.x < k) {
        double t = 0.0, t2 = 0.0;
        for (int jj = 0; jj < per; ++jj) {
            t += s[jj * k + threadIdx.x];
            t2 += s2[jj * k + threadIdx.x];
        }
        atomicAdd(delta + threadIdx.x, t);
        if (sink_flag) atomicAdd(sink_next + threadIdx.x, t2);
    }
}

__global__ void k_zero_f64(double* __restrict__ p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// R[v][q] -> out[q][v]
template <typename T>
__global__ void k_transpose_nk(const T* __restrict__ R, int64_t n, int k, T* __restrict__ out) {
    __shared__ T tile[64][65];
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int64_t v = v0 + r;
        tile[r][tx] = (v < n && tx < k) ? R[v * k + tx] : T(0);
    }
    __syncthreads();
    for (int q = ty; q < k; q += 4) {
        const int64_t v = v0 + tx;
        if (v < n) out[(int64_t)q * n + v] = tile[tx][q];
    }
}

struct PprLayout {
    float *R, *Y;
    void* carry;
    int32_t *flag, *offs, *sinks;
    void* partial;
    size_t part_bytes;
    double *sink, *sink2, *delta;
    int32_t* done;
    int32_t* src;
    size_t total;
};

static PprLayout ppr_layout(const DevCSR& m, int64_t n, int k, char* base) {
    PprLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* r = base ? base + off : nullptr;
        off = up256b(off + bytes);
        return r;
    };
    L.R = (float*)take(sizeof(float) * (size_t)n * k);
    L.Y = (float*)take(sizeof(float) * (size_t)n * k);
    L.carry = take(k == 1 ? merge_carry_bytes(m) : spmm_carry_bytes(m, GBK_FP32, k));
    L.flag = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.offs = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.sinks = (int32_t*)take(sizeof(int32_t) * (n + 1));
    L.part_bytes = scan_partial_bytes(n);
    L.partial = take(L.part_bytes);
    L.sink = (double*)take(sizeof(double) * 64);
    L.sink2 = (double*)take(sizeof(double) * 64);
    L.delta = (double*)take(sizeof(double) * 64);
    L.done = (int32_t*)take(sizeof(int32_t) * 64);
    L.src = (int32_t*)take(sizeof(int32_t) * 64);
    L.total = off;
    return L;
}

size_t ppr_ws_bytes(gbk_matrix A, int k) { return ppr_layout(A->csr, A->n, k, nullptr).total + 256; }

// ranks_out: k*n floats, source-major, on the device.  iters[q]: iterations
// column q ran (itermax if it did not converge).
int ppr_run(gbk_matrix A, const int32_t* src_host, int k, void* ws, size_t ws_bytes, float* ranks_out, int itermax,
            double damping, double tol, int* iters, hipStream_t st) {
    const int64_t n = A->n;
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "personalized PageRank takes 1..64 sources (got %d)", k);
    if (A->type != GBK_FP32) return fail(GBK_DOMAIN_MISMATCH, "personalized PageRank needs an FP32 matrix");
    if (ws_bytes < ppr_ws_bytes(A, k)) return fail(GBK_INSUFFICIENT_SPACE, "PageRank workspace too small");
    DevCSR& m = A->csr;
    char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    PprLayout L = ppr_layout(m, n, k, base);
    GBK_HIP(hipMemcpyAsync(L.src, src_host, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ppr_init, grid_b(n * k), dim3(256), 0, st, L.R, n, k, L.src);
    GBK_LAUNCH_CHECK();
    // sink list
    hipLaunchKernelGGL(k_sink_flags, grid_b(n), dim3(256), 0, st, A->csc.row_ptr, n, L.flag);
    GBK_LAUNCH_CHECK();
    int32_t nsinks = 0;
    GBK_TRY(exclusive_scan_i32_on(L.flag, L.offs, n, &nsinks, st, L.partial, L.part_bytes));
    hipLaunchKernelGGL(k_compact_flags, grid_b(n), dim3(256), 0, st, L.flag, L.offs, n, L.sinks);
    GBK_LAUNCH_CHECK();
    std::vector<int32_t> done(k, 0);
    for (int q = 0; q < k; ++q) iters[q] = itermax;
    GBK_HIP(hipMemcpyAsync(L.done, done.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    const int per = 256 / k;
    // the update streams n k values: a few workgroups per CU suffice, and
    // fewer of them add fewer per-column partial sums atomically
    const int64_t ub = std::min<int64_t>((n + per - 1) / per, (int64_t)device_cu_count() * 4);
    const int64_t sb = std::min<int64_t>(((int64_t)nsinks + per - 1) / per, (int64_t)device_cu_count() * 8);
    std::vector<double> delta(k);
    float* R = L.R;
    float* Y = L.Y;
    double* sink_cur = L.sink;
    double* sink_next = L.sink2;
    int remaining = k;
    hipStream_t prev = stream();
    set_thread_stream(st);
    // the sink sums of the starting ranks; each update forms the next ones
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, sink_cur, 64);
    if (nsinks > 0 && sb > 0)
        hipLaunchKernelGGL(k_ppr_sink, dim3((unsigned)sb), dim3(256), 0, st, R, k, k, L.sinks, (int64_t)nsinks,
                           sink_cur);
    int rc = GBK_SUCCESS;
    for (int it = 1; it <= itermax && remaining > 0; ++it) {
        if (k == 1) rc = spmv_merge_ws(m, GBK_FP32, R, Y, GBK_ADD_PLUS, GBK_MUL_TIMES, GBK_ACC_NONE, L.carry);
        else rc = spmm_run(m, GBK_FP32, R, Y, k, GBK_ADD_PLUS, GBK_MUL_TIMES, GBK_ACC_NONE, L.carry, st);
        if (rc != GBK_SUCCESS) break;
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, sink_next, 64);
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, L.delta, 64);
        // Y is overwritten with the new ranks, then R and Y swap roles
        hipLaunchKernelGGL(k_ppr_update, dim3((unsigned)std::max<int64_t>(ub, 1)), dim3(256), 0, st, R, Y, Y, n, k, k,
                           L.src, sink_cur, L.done, (float)damping, L.delta, L.flag, sink_next);
        if (hipGetLastError() != hipSuccess) {
            rc = fail(GBK_PANIC, "PageRank kernel launch failed");
            break;
        }
        if (hipMemcpyAsync(delta.data(), L.delta, sizeof(double) * k, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(GBK_PANIC, "PageRank delta copy failed");
            break;
        }
        std::swap(R, Y);
        std::swap(sink_cur, sink_next);
        bool changed = false;
        for (int q = 0; q < k; ++q) {
            if (!done[q] && tol > 0.0 && delta[q] < tol) {
                done[q] = 1;
                iters[q] = it;
                --remaining;
                changed = true;
            }
        }
        if (changed && remaining > 0)
            if (hipMemcpyAsync(L.done, done.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, st) != hipSuccess) {
                rc = fail(GBK_PANIC, "PageRank flag copy failed");
                break;
            }
    }
    set_thread_stream(prev);
    GBK_TRY(rc);
    if (k == 1) {
        GBK_HIP(hipMemcpyAsync(ranks_out, R, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    } else {
        hipLaunchKernelGGL(k_transpose_nk<float>, grid_b(n, 64), dim3(256), 0, st, R, n, k, ranks_out);
        GBK_LAUNCH_CHECK();
    }
    return GBK_SUCCESS;
}

// ------------------------------------------------------------ multi-SSSP --
// Dense Bellman-Ford rounds: each round is one min.plus SpMM over the k
// distance columns (Y = A (min.+) D, saturating int32) and an element-wise
// D = min(D, Y) that records which columns changed.  A column that did not
// change in a round is at its fixed point, the shortest distances.  The
// rounds run on the merge-path SpMM, with no atomics on distances.

// D[v][q] = INT32_MAX for every q in `cols` (a bit mask), and 0 at the
// column's source.
__global__ void k_sssp_reset_cols(int32_t* __restrict__ D, int64_t n, int k, uint64_t cols,
                                  const int32_t* __restrict__ src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * k) return;
    const int q = (int)(i % k);
    if (!((cols >> q) & 1ull)) return;
    const int64_t v = i / k;
    D[i] = (src[q] == v) ? 0 : INT32_MAX;
}

// D = min(D, Y) over the first kw of the ld columns of n rows; bit q of
// *changed is set if column q improved anywhere.
__global__ __launch_bounds__(256) void k_sssp_min_update(int32_t* __restrict__ D, const int32_t* __restrict__ Y,
                                                         int64_t n, int kw, int ld,
                                                         unsigned long long* __restrict__ changed) {
    __shared__ unsigned long long s_m;
    if (threadIdx.x == 0) s_m = 0;
    __syncthreads();
    unsigned long long m = 0;
    const int64_t total = n * kw;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(i % kw);
        const int64_t e = (i / kw) * ld + q;
        const int32_t y = Y[e];
        if (y < D[e]) {
            D[e] = y;
            m |= 1ull << q;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off, 64);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(&s_m, m);
    __syncthreads();
    if (threadIdx.x == 0 && s_m) atomicOr(changed, s_m);
}

__global__ void k_zero_mask(unsigned long long* m) {
    if (threadIdx.x == 0) *m = 0;
}

struct MsspLayout {
    int32_t *D, *Y, *src;
    void* carry;
    unsigned long long* changed;
    size_t total;
};

static MsspLayout mssp_layout(const DevCSR& m, int64_t n, int k, char* base) {
    MsspLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* r = base ? base + off : nullptr;
        off = up256b(off + bytes);
        return r;
    };
    L.D = (int32_t*)take(sizeof(int32_t) * (size_t)n * k);
    L.Y = (int32_t*)take(sizeof(int32_t) * (size_t)n * k);
    L.carry = take(spmm_carry_bytes(m, GBK_INT32, k));
    L.src = (int32_t*)take(sizeof(int32_t) * 64);
    L.changed = (unsigned long long*)take(sizeof(unsigned long long) * 2);
    L.total = off;
    return L;
}

size_t mssp_ws_bytes(gbk_matrix A, int k) { return mssp_layout(A->csr, A->n, k, nullptr).total + 256; }

static dim3 capped_b(int64_t threads) {
    int64_t b = (threads + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b > 0 ? b : 1));
}

// One round: Y = A (min.+) D, D = min(D, Y); *changed_out receives the mask of
// columns that changed.
static int mssp_round(gbk_matrix A, const MsspLayout& L, int kw, int ld, hipStream_t st, unsigned long long* hmask,
                      uint64_t* changed_out) {
    const int64_t n = A->n;
    GBK_TRY(spmm_run(A->csr, GBK_INT32, L.D, L.Y, kw, GBK_ADD_MIN, GBK_MUL_PLUS, GBK_ACC_NONE, L.carry, st, ld));
    hipLaunchKernelGGL(k_zero_mask, dim3(1), dim3(64), 0, st, L.changed);
    hipLaunchKernelGGL(k_sssp_min_update, capped_b(n * kw), dim3(256), 0, st, L.D, L.Y, n, kw, ld, L.changed);
    GBK_LAUNCH_CHECK();
    GBK_HIP(hipMemcpyAsync(hmask, L.changed, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    *changed_out = *hmask;
    return GBK_SUCCESS;
}

// dist_out: k*n int32, source-major, on the device.
int mssp_run(gbk_matrix A, const int32_t* src_host, int k, void* ws, size_t ws_bytes, int32_t* dist_out,
             int64_t* stats, hipStream_t st) {
    const int64_t n = A->n;
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "multi-source SSSP takes 1..64 sources (got %d)", k);
    if (A->type != GBK_INT32) return fail(GBK_DOMAIN_MISMATCH, "multi-source SSSP needs an INT32 weight matrix");
    if (ws_bytes < mssp_ws_bytes(A, k)) return fail(GBK_INSUFFICIENT_SPACE, "SSSP workspace too small");
    GBK_TRY(ensure_spmm_tiles(A->csr, st));
    char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    MsspLayout L = mssp_layout(A->csr, n, k, base);
    GBK_HIP(hipMemcpyAsync(L.src, src_host, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    const uint64_t all = k == 64 ? ~0ull : ((1ull << k) - 1);
    hipLaunchKernelGGL(k_sssp_reset_cols, grid_b(n * k), dim3(256), 0, st, L.D, n, k, all, L.src);
    GBK_LAUNCH_CHECK();
    unsigned long long hmask = 0;
    int64_t rounds = 0;
    uint64_t changed = all;
    while (changed) {
        GBK_TRY(mssp_round(A, L, k, k, st, &hmask, &changed));
        ++rounds;
    }
    hipLaunchKernelGGL(k_transpose_nk<int32_t>, grid_b(n, 64), dim3(256), 0, st, L.D, n, k, dist_out);
    GBK_LAUNCH_CHECK();
    if (stats) {
        stats[0] = rounds;
        stats[1] = rounds * A->csr.nnz;
    }
    return GBK_SUCCESS;
}

// --------------------------------------------------- PPR / SSSP engines --
// Column q of the dense block belongs to slot q.  Seeding a slot rewrites its
// column; the other columns are untouched, so a query computes the same
// numbers alone and in a mix (the SpMM folds every column in the same order
// whatever the other columns hold).

// R[v][slot[i]] = (v == src[i]) for every seeded slot i.
__global__ void k_ppr_seed(float* __restrict__ R, int64_t n, int k, const int32_t* __restrict__ src,
                           const int32_t* __restrict__ slot, int count) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    for (int i = 0; i < count; ++i) R[v * k + slot[i]] = (src[i] == v) ? 1.0f : 0.0f;
}

// A freshly seeded column is one-hot: its sink sum is 1 when its source is a
// sink, else 0.
__global__ void k_ppr_seed_sink(double* __restrict__ sink, const int32_t* __restrict__ sink_flag,
                                const int32_t* __restrict__ src, const int32_t* __restrict__ slot, int count) {
    const int i = threadIdx.x;
    if (i < count) sink[slot[i]] = sink_flag[src[i]] ? 1.0 : 0.0;
}

template <typename T>
__global__ void k_extract_col(const T* __restrict__ M, int64_t n, int k, int q, T* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = M[v * k + q];
}

namespace {

class PprEngine : public Engine {
  public:
    PprEngine(gbk_matrix A, hipStream_t st, int k, double damping, double tol, int itermax)
        : A_(A), st_(st), n_(A->n), k_(k), damping_(damping), tol_(tol), itermax_(itermax) {}
    ~PprEngine() override {
        dev_free(ws_);
        dev_free(seeds_);
        if (hdelta_) (void)hipHostFree(hdelta_);
    }
    int init() {
        if (A_->type != GBK_FP32) return fail(GBK_DOMAIN_MISMATCH, "personalized PageRank needs an FP32 matrix");
        GBK_TRY(ensure_spmm_tiles(A_->csr, st_));
        ws_bytes_ = ppr_ws_bytes(A_, k_) + 512;
        GBK_TRY(dev_alloc(&ws_, ws_bytes_));
        GBK_TRY(dev_alloc((void**)&seeds_, sizeof(int32_t) * 128));
        GBK_HIP(hipHostMalloc((void**)&hdelta_, sizeof(double) * 64, hipHostMallocDefault));
        char* base = (char*)(((uintptr_t)ws_ + 255) & ~(uintptr_t)255);
        L_ = ppr_layout(A_->csr, n_, k_, base);
        // sink flags (the update pass forms the sink sums from them)
        hipLaunchKernelGGL(k_sink_flags, grid_b(n_), dim3(256), 0, st_, A_->csc.row_ptr, n_, L_.flag);
        GBK_LAUNCH_CHECK();
        // every column starts frozen at zero
        std::vector<int32_t> ones(64, 1), zeros(64, 0);
        GBK_HIP(hipMemcpyAsync(L_.done, ones.data(), sizeof(int32_t) * 64, hipMemcpyHostToDevice, st_));
        GBK_HIP(hipMemcpyAsync(L_.src, zeros.data(), sizeof(int32_t) * 64, hipMemcpyHostToDevice, st_));
        hipLaunchKernelGGL(k_ppr_init, grid_b(n_ * k_), dim3(256), 0, st_, L_.R, n_, k_, L_.src);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st_, L_.sink, 64);
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st_, L_.sink2, 64);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipStreamSynchronize(st_));
        for (int q = 0; q < 64; ++q) done_[q] = 1;
        return GBK_SUCCESS;
    }
    int slots() const override { return k_; }
    size_t bytes() const override { return ws_bytes_; }
    size_t result_bytes() const override { return sizeof(float) * (size_t)n_; }
    int admit(int slot, int32_t source) override {
        if (source < 0 || source >= n_) return fail(GBK_INVALID_INDEX, "source out of range");
        pend_src_.push_back(source);
        pend_slot_.push_back(slot);
        return GBK_SUCCESS;
    }
    int step(std::vector<int>& finished) override {
        finished.clear();
        bool flags_dirty = false;
        if (!pend_slot_.empty()) {
            std::vector<int32_t> buf(pend_src_);
            buf.insert(buf.end(), pend_slot_.begin(), pend_slot_.end());
            GBK_HIP(hipMemcpyAsync(seeds_, buf.data(), sizeof(int32_t) * buf.size(), hipMemcpyHostToDevice, st_));
            const int cnt = (int)pend_slot_.size();
            hipLaunchKernelGGL(k_ppr_seed, grid_b(n_), dim3(256), 0, st_, L_.R, n_, k_, seeds_, seeds_ + cnt, cnt);
            GBK_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_ppr_seed_sink, dim3(1), dim3(64), 0, st_, L_.sink, L_.flag, seeds_, seeds_ + cnt, cnt);
            GBK_LAUNCH_CHECK();
            for (int i = 0; i < cnt; ++i) {
                const int q = pend_slot_[i];
                src_[q] = pend_src_[i];
                done_[q] = 0;
                iters_[q] = 0;
                occupied_ |= 1ull << q;
            }
            flags_dirty = true;
            GBK_HIP(hipStreamSynchronize(st_));
            pend_src_.clear();
            pend_slot_.clear();
        }
        if (!occupied_) return GBK_SUCCESS;
        if (flags_dirty) {
            GBK_HIP(hipMemcpyAsync(L_.done, done_, sizeof(int32_t) * k_, hipMemcpyHostToDevice, st_));
            GBK_HIP(hipMemcpyAsync(L_.src, src_, sizeof(int32_t) * k_, hipMemcpyHostToDevice, st_));
        }
        // Only the columns up to the highest occupied slot are advanced (the
        // server fills the lowest free slot first): a lightly loaded engine
        // runs a narrow SpMM.  Columns beyond it are reseeded in full when a
        // query is admitted there.
        const int kw = 64 - __builtin_clzll(occupied_);
        hipStream_t prev = stream();
        set_thread_stream(st_);
        int rc;
        if (k_ == 1) rc = spmv_merge_ws(A_->csr, GBK_FP32, L_.R, L_.Y, GBK_ADD_PLUS, GBK_MUL_TIMES, GBK_ACC_NONE, L_.carry);
        else rc = spmm_run(A_->csr, GBK_FP32, L_.R, L_.Y, kw, GBK_ADD_PLUS, GBK_MUL_TIMES, GBK_ACC_NONE, L_.carry, st_, k_);
        set_thread_stream(prev);
        GBK_TRY(rc);
        const int per = 256 / kw;
        const int64_t ub = std::max<int64_t>(std::min<int64_t>((n_ + per - 1) / per, (int64_t)device_cu_count() * 4), 1);
        // the sink sums of the current ranks are L_.sink (formed by the last
        // update, or set when a column was seeded); the update forms the next
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st_, L_.sink2, 64);
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st_, L_.delta, 64);
        hipLaunchKernelGGL(k_ppr_update, dim3((unsigned)ub), dim3(256), 0, st_, L_.R, L_.Y, L_.Y, n_, kw, k_, L_.src,
                           L_.sink, L_.done, (float)damping_, L_.delta, L_.flag, L_.sink2);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipMemcpyAsync(hdelta_, L_.delta, sizeof(double) * kw, hipMemcpyDeviceToHost, st_));
        GBK_HIP(hipStreamSynchronize(st_));
        std::swap(L_.R, L_.Y);
        std::swap(L_.sink, L_.sink2);
        bool changed = false;
        for (int q = 0; q < kw; ++q) {
            if (!((occupied_ >> q) & 1ull) || done_[q]) continue;
            ++iters_[q];
            if ((tol_ > 0.0 && hdelta_[q] < tol_) || iters_[q] >= itermax_) {
                done_[q] = 1;
                changed = true;
                finished.push_back(q);
            }
        }
        if (changed) GBK_HIP(hipMemcpyAsync(L_.done, done_, sizeof(int32_t) * k_, hipMemcpyHostToDevice, st_));
        return GBK_SUCCESS;
    }
    int extract(int slot, void* out_dev) override {
        hipLaunchKernelGGL(k_extract_col<float>, grid_b(n_), dim3(256), 0, st_, L_.R, n_, k_, slot, (float*)out_dev);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    }
    void release(int slot) override { occupied_ &= ~(1ull << slot); }
    int iterations(int slot) const { return iters_[slot]; }

  private:
    gbk_matrix A_;
    hipStream_t st_;
    int64_t n_;
    int k_;
    double damping_, tol_;
    int itermax_;
    void* ws_ = nullptr;
    int32_t* seeds_ = nullptr;
    size_t ws_bytes_ = 0;
    PprLayout L_{};
    double* hdelta_ = nullptr;
    uint64_t occupied_ = 0;
    int32_t done_[64];
    int32_t src_[64] = {0};
    int iters_[64] = {0};
    std::vector<int32_t> pend_src_, pend_slot_;
};

class SsspEngine : public Engine {
  public:
    SsspEngine(gbk_matrix A, hipStream_t st, int k) : A_(A), st_(st), n_(A->n), k_(k) {}
    ~SsspEngine() override {
        dev_free(ws_);
        if (hmask_) (void)hipHostFree(hmask_);
    }
    int init() {
        if (A_->type != GBK_INT32) return fail(GBK_DOMAIN_MISMATCH, "SSSP needs an INT32 weight matrix");
        GBK_TRY(ensure_spmm_tiles(A_->csr, st_));
        ws_bytes_ = mssp_ws_bytes(A_, k_) + 512;
        GBK_TRY(dev_alloc(&ws_, ws_bytes_));
        GBK_HIP(hipHostMalloc((void**)&hmask_, sizeof(unsigned long long), hipHostMallocDefault));
        char* base = (char*)(((uintptr_t)ws_ + 255) & ~(uintptr_t)255);
        L_ = mssp_layout(A_->csr, n_, k_, base);
        // every column starts as "source 0 at distance 0", unused until seeded
        std::vector<int32_t> zeros(64, 0);
        GBK_HIP(hipMemcpyAsync(L_.src, zeros.data(), sizeof(int32_t) * 64, hipMemcpyHostToDevice, st_));
        const uint64_t all = k_ == 64 ? ~0ull : ((1ull << k_) - 1);
        hipLaunchKernelGGL(k_sssp_reset_cols, grid_b(n_ * k_), dim3(256), 0, st_, L_.D, n_, k_, all, L_.src);
        GBK_LAUNCH_CHECK();
        GBK_HIP(hipStreamSynchronize(st_));
        return GBK_SUCCESS;
    }
    int slots() const override { return k_; }
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
        if (!pend_slot_.empty()) {
            uint64_t cols = 0;
            for (size_t i = 0; i < pend_slot_.size(); ++i) {
                src_[pend_slot_[i]] = pend_src_[i];
                cols |= 1ull << pend_slot_[i];
            }
            GBK_HIP(hipMemcpyAsync(L_.src, src_, sizeof(int32_t) * k_, hipMemcpyHostToDevice, st_));
            hipLaunchKernelGGL(k_sssp_reset_cols, grid_b(n_ * k_), dim3(256), 0, st_, L_.D, n_, k_, cols, L_.src);
            GBK_LAUNCH_CHECK();
            occupied_ |= cols;
            GBK_HIP(hipStreamSynchronize(st_));  // src_ may change before the next copy
            pend_src_.clear();
            pend_slot_.clear();
        }
        if (!occupied_) return GBK_SUCCESS;
        // only the columns up to the highest occupied slot take part
        const int kw = 64 - __builtin_clzll(occupied_);
        uint64_t changed = 0;
        GBK_TRY(mssp_round(A_, L_, kw, k_, st_, hmask_, &changed));
        for (int q = 0; q < kw; ++q)
            if (((occupied_ >> q) & 1ull) && !((changed >> q) & 1ull)) finished.push_back(q);
        return GBK_SUCCESS;
    }
    int extract(int slot, void* out_dev) override {
        hipLaunchKernelGGL(k_extract_col<int32_t>, grid_b(n_), dim3(256), 0, st_, L_.D, n_, k_, slot, (int32_t*)out_dev);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    }
    void release(int slot) override { occupied_ &= ~(1ull << slot); }

  private:
    gbk_matrix A_;
    hipStream_t st_;
    int64_t n_;
    int k_;
    void* ws_ = nullptr;
    unsigned long long* hmask_ = nullptr;
    size_t ws_bytes_ = 0;
    MsspLayout L_{};
    uint64_t occupied_ = 0;
    int32_t src_[64] = {0};
    std::vector<int32_t> pend_src_, pend_slot_;
};

}  // namespace

int make_ppr_engine(gbk_matrix A, hipStream_t st, int k, double damping, double tol, int itermax,
                    std::unique_ptr<Engine>* out) {
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "PageRank engine takes 1..64 slots");
    auto e = std::make_unique<PprEngine>(A, st, k, damping, tol, itermax);
    GBK_TRY(e->init());
    *out = std::move(e);
    return GBK_SUCCESS;
}

int make_sssp_engine(gbk_matrix A, hipStream_t st, int k, std::unique_ptr<Engine>* out) {
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "SSSP engine takes 1..64 slots");
    auto e = std::make_unique<SsspEngine>(A, st, k);
    GBK_TRY(e->init());
    *out = std::move(e);
    return GBK_SUCCESS;
}

}  // namespace gbk

using namespace gbk;

static int check_sources(gbk_matrix A, const int64_t* sources, int k, std::vector<int32_t>& s32) {
    if (k < 1 || k > 64) return fail(GBK_INVALID_VALUE, "a batch takes 1..64 sources (got %d)", k);
    if (!sources) return fail(GBK_NULL_POINTER, "null sources");
    s32.resize(k);
    for (int i = 0; i < k; ++i) {
        if (sources[i] < 0 || sources[i] >= A->n)
            return fail(GBK_INVALID_INDEX, "source %lld out of range [0, %lld)", (long long)sources[i],
                        (long long)A->n);
        s32[i] = (int32_t)sources[i];
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_ppr_multi(gbk_matrix A, const int64_t* sources, int k, float* ranks, int itermax, double damping,
                             double tol, int* iters) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (!ranks || !iters) return fail(GBK_NULL_POINTER, "null output");
    if (itermax < 1) return fail(GBK_INVALID_VALUE, "itermax must be positive");
    std::vector<int32_t> s32;
    GBK_TRY(check_sources(A, sources, k, s32));
    GBK_TRY(ensure_init());
    hipStream_t st = stream();
    void *ws, *out;
    const size_t wb = ppr_ws_bytes(A, k);
    GBK_TRY(scratch(SLOT_PPR, wb, &ws));
    GBK_TRY(scratch(SLOT_MSBFS_OUT, sizeof(float) * (size_t)k * A->n, &out));
    GBK_TRY(ppr_run(A, s32.data(), k, ws, wb, (float*)out, itermax, damping, tol, iters, st));
    GBK_HIP(hipMemcpyAsync(ranks, out, sizeof(float) * (size_t)k * A->n, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

extern "C" int gbk_ppr_multi_device(gbk_matrix A, const int64_t* sources, int k, float* ranks_dev, int itermax,
                                    double damping, double tol, int* iters) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (!ranks_dev || !iters) return fail(GBK_NULL_POINTER, "null output");
    if (itermax < 1) return fail(GBK_INVALID_VALUE, "itermax must be positive");
    std::vector<int32_t> s32;
    GBK_TRY(check_sources(A, sources, k, s32));
    GBK_TRY(ensure_init());
    hipStream_t st = stream();
    void* ws;
    const size_t wb = ppr_ws_bytes(A, k);
    GBK_TRY(scratch(SLOT_PPR, wb, &ws));
    GBK_TRY(ppr_run(A, s32.data(), k, ws, wb, ranks_dev, itermax, damping, tol, iters, st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}

extern "C" int gbk_sssp_multi(gbk_matrix A, const int64_t* sources, int k, int32_t* dist, int64_t* stats) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    if (!dist) return fail(GBK_NULL_POINTER, "null output");
    std::vector<int32_t> s32;
    GBK_TRY(check_sources(A, sources, k, s32));
    GBK_TRY(ensure_init());
    hipStream_t st = stream();
    void *ws, *out;
    const size_t wb = mssp_ws_bytes(A, k);
    GBK_TRY(scratch(SLOT_MSSSP, wb, &ws));
    GBK_TRY(scratch(SLOT_MSBFS_OUT, sizeof(int32_t) * (size_t)k * A->n, &out));
    GBK_TRY(mssp_run(A, s32.data(), k, ws, wb, (int32_t*)out, stats, st));
    GBK_HIP(hipMemcpyAsync(dist, out, sizeof(int32_t) * (size_t)k * A->n, hipMemcpyDeviceToHost, st));
    GBK_HIP(hipStreamSynchronize(st));
    return GBK_SUCCESS;
}
