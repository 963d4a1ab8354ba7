// Many small graphs in one launch: a BFS from one source in each of up to
// thousands of independent graphs, run by one persistent kernel that is
// captured once in a HIP graph and replayed.
//
// The reference runs one driver call per graph (bfs_blast,
// algorithms/bfs/bfs.hpp:94-142, levels 1-based, 0 = unreached).  Here the
// graphs are packed back to back in HBM (CSR and CSC, local vertex ids) and
// a persistent grid of workgroups takes graphs from a work queue, largest
// first (one atomic per graph).  A workgroup runs the whole BFS of its graph
// with __syncthreads only — no inter-workgroup waits, so every workgroup
// exits when the queue is empty.  The frontier, next-frontier and visited
// sets are bitmaps in LDS (graphs of up to MAX_V vertices); each level is a
// push step (waves walk the out-edges of frontier vertices, atomicOr into the
// LDS bitmap) while the frontier is small and a pull step (one thread per
// unvisited vertex scans its in-edges and stops at the first frontier
// vertex) once it is large, then a settle pass writes the new levels.
#include <algorithm>
#include <numeric>
#include <vector>

#include "gbk_objects.h"

namespace gbk {

constexpr int GB_THREADS = 1024;
constexpr int GB_MAX_V = 1 << 17;  // vertices per graph (3 bitmaps of 16 KiB in LDS)
constexpr int GB_WORDS = GB_MAX_V / 32;

struct GraphBatchDev {
    int ngraphs = 0;
    int64_t total_v = 0, total_e = 0;
    int64_t* voff = nullptr;   // ngraphs + 1 vertex offsets
    int64_t* eoff = nullptr;   // ngraphs + 1 edge offsets
    int32_t* rp = nullptr;     // total_v + ngraphs: per-graph local row pointers (n_g + 1 each)
    int32_t* col = nullptr;    // total_e local ids
    int32_t* crp = nullptr;    // CSC, same layout
    int32_t* ccol = nullptr;
    int32_t* order = nullptr;  // graphs by descending edge count
    int32_t* src = nullptr;    // ngraphs local sources
    int32_t* levels = nullptr; // total_v
    int32_t* counter = nullptr;
};

__global__ void k_gb_reset(int32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

__device__ __forceinline__ bool bm_test(const uint32_t* bm, int v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

__global__ __launch_bounds__(GB_THREADS) void k_gb_bfs(const int64_t* __restrict__ voff, const int64_t* __restrict__ eoff,
                                                       const int32_t* __restrict__ rp_all,
                                                       const int32_t* __restrict__ col_all,
                                                       const int32_t* __restrict__ crp_all,
                                                       const int32_t* __restrict__ ccol_all,
                                                       const int32_t* __restrict__ order, const int32_t* __restrict__ src,
                                                       int ngraphs, int32_t* __restrict__ levels,
                                                       int32_t* __restrict__ counter) {
    __shared__ uint32_t F[GB_WORDS], N[GB_WORDS], V[GB_WORDS];
    __shared__ int s_g;
    __shared__ unsigned long long s_nf, s_mf;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    constexpr int NWAVES = GB_THREADS / 64;
    while (true) {
        if (tid == 0) s_g = atomicAdd(counter, 1);
        __syncthreads();
        const int gi = s_g;
        __syncthreads();
        if (gi >= ngraphs) break;
        const int g = order[gi];
        const int64_t vb = voff[g];
        const int n = (int)(voff[g + 1] - vb);
        const int32_t* rp = rp_all + vb + g;  // n + 1 entries
        const int32_t* col = col_all + eoff[g];
        const int32_t* crp = crp_all + vb + g;
        const int32_t* ccol = ccol_all + eoff[g];
        int32_t* lev = levels + vb;
        const int words = (n + 31) >> 5;
        for (int w = tid; w < words; w += GB_THREADS) {
            F[w] = 0;
            N[w] = 0;
            V[w] = 0;
        }
        for (int v = tid; v < n; v += GB_THREADS) lev[v] = 0;
        __syncthreads();
        const int s = src[g];
        if (tid == 0) {
            F[s >> 5] = 1u << (s & 31);
            V[s >> 5] = 1u << (s & 31);
            lev[s] = 1;
            s_nf = 1;
            s_mf = (unsigned long long)(rp[s + 1] - rp[s]);
        }
        __syncthreads();
        const int64_t m = rp[n];
        int32_t cur = 1;
        while (true) {
            const unsigned long long nf = s_nf, mf = s_mf;
            __syncthreads();
            if (nf == 0) break;
            if (tid == 0) {
                s_nf = 0;
                s_mf = 0;
            }
            if ((double)mf * 14.0 < (double)m) {
                // push: each wave takes frontier words; all lanes walk one vertex's out-edges
                for (int w = wid; w < words; w += NWAVES) {
                    uint32_t bits = F[w];
                    while (bits) {
                        const int b = __ffs(bits) - 1;
                        bits &= bits - 1;
                        const int v = w * 32 + b;
                        const int e1 = rp[v + 1];
                        for (int e = rp[v] + lane; e < e1; e += 64) {
                            const int u = col[e];
                            if (!bm_test(V, u)) atomicOr(&N[u >> 5], 1u << (u & 31));
                        }
                    }
                }
            } else {
                // pull: one thread per unvisited vertex, first frontier in-neighbour wins
                for (int v = tid; v < n; v += GB_THREADS) {
                    if (bm_test(V, v)) continue;
                    const int e1 = crp[v + 1];
                    for (int e = crp[v]; e < e1; ++e) {
                        if (bm_test(F, ccol[e])) {
                            atomicOr(&N[v >> 5], 1u << (v & 31));
                            break;
                        }
                    }
                }
            }
            __syncthreads();
            // settle: new = N & ~V; frontier = new; levels; counts for the next direction choice
            unsigned long long cnt = 0, deg = 0;
            for (int w = tid; w < words; w += GB_THREADS) {
                const uint32_t nw = N[w] & ~V[w];
                N[w] = 0;
                F[w] = nw;
                if (nw) {
                    V[w] |= nw;
                    uint32_t b = nw;
                    while (b) {
                        const int i = __ffs(b) - 1;
                        b &= b - 1;
                        const int v = w * 32 + i;
                        lev[v] = cur + 1;
                        deg += (unsigned long long)(rp[v + 1] - rp[v]);
                        ++cnt;
                    }
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                cnt += __shfl_xor(cnt, off, 64);
                deg += __shfl_xor(deg, off, 64);
            }
            if (lane == 0 && cnt) {
                atomicAdd(&s_nf, cnt);
                atomicAdd(&s_mf, deg);
            }
            __syncthreads();
            ++cur;
        }
    }
}

}  // namespace gbk

struct gbk_graph_batch_s {
    gbk::GraphBatchDev d;
    std::vector<int64_t> voff;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int grid = 0;
};

using namespace gbk;

static void gb_free_dev(GraphBatchDev& d) {
    dev_free(d.voff);
    dev_free(d.eoff);
    dev_free(d.rp);
    dev_free(d.col);
    dev_free(d.crp);
    dev_free(d.ccol);
    dev_free(d.order);
    dev_free(d.src);
    dev_free(d.levels);
    dev_free(d.counter);
    d = GraphBatchDev();
}

template <typename V> static int gb_upload(V** dst, const std::vector<V>& src) {
    GBK_TRY(dev_alloc((void**)dst, sizeof(V) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) GBK_HIP(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
    return GBK_SUCCESS;
}

extern "C" int gbk_graph_batch_free(gbk_graph_batch b) {
    if (!b) return GBK_SUCCESS;
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->st) (void)hipStreamDestroy(b->st);
    gb_free_dev(b->d);
    delete b;
    return GBK_SUCCESS;
}

extern "C" int gbk_graph_batch_new(gbk_graph_batch* out, int ngraphs, const int64_t* nverts,
                                   const int64_t* const* row_ptrs, const int32_t* const* col_ids) {
    if (!out || !nverts || !row_ptrs || !col_ids) return fail(GBK_NULL_POINTER, "null argument");
    if (ngraphs < 1) return fail(GBK_INVALID_VALUE, "a batch needs at least one graph");
    std::vector<int64_t> voff(ngraphs + 1, 0), eoff(ngraphs + 1, 0);
    for (int g = 0; g < ngraphs; ++g) {
        const int64_t n = nverts[g];
        if (n < 1 || n > GB_MAX_V)
            return fail(GBK_INVALID_VALUE, "graph %d has %lld vertices; a batched graph holds 1..%d", g, (long long)n,
                        GB_MAX_V);
        if (!row_ptrs[g] || (!col_ids[g] && row_ptrs[g][n] > 0)) return fail(GBK_NULL_POINTER, "graph %d arrays", g);
        const int64_t m = row_ptrs[g][n];
        if (m < 0 || m > INT32_MAX) return fail(GBK_INVALID_VALUE, "graph %d has too many edges", g);
        for (int64_t e = 0; e < m; ++e)
            if (col_ids[g][e] < 0 || col_ids[g][e] >= n)
                return fail(GBK_INDEX_OUT_OF_BOUNDS, "graph %d: column id %d out of range", g, col_ids[g][e]);
        voff[g + 1] = voff[g] + n;
        eoff[g + 1] = eoff[g] + m;
    }
    const int64_t tv = voff[ngraphs], te = eoff[ngraphs];
    if (te > INT32_MAX || tv + ngraphs > INT32_MAX) return fail(GBK_INVALID_VALUE, "batch too large");
    std::vector<int32_t> rp(tv + ngraphs), col(te), crp(tv + ngraphs), ccol(te);
#pragma omp parallel for schedule(dynamic, 4)
    for (int g = 0; g < ngraphs; ++g) {
        const int64_t n = nverts[g], vb = voff[g] + g, eb = eoff[g];
        const int64_t* r = row_ptrs[g];
        for (int64_t v = 0; v <= n; ++v) rp[vb + v] = (int32_t)r[v];
        for (int64_t e = 0; e < r[n]; ++e) col[eb + e] = col_ids[g][e];
        // CSC by counting sort (rows ascending inside a column)
        std::vector<int32_t> cnt(n + 1, 0);
        for (int64_t e = 0; e < r[n]; ++e) cnt[col_ids[g][e] + 1]++;
        for (int64_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
        for (int64_t v = 0; v <= n; ++v) crp[vb + v] = cnt[v];
        for (int64_t u = 0; u < n; ++u)
            for (int64_t e = r[u]; e < r[u + 1]; ++e) ccol[eb + cnt[col_ids[g][e]]++] = (int32_t)u;
    }
    std::vector<int32_t> order(ngraphs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return eoff[a + 1] - eoff[a] + voff[a + 1] - voff[a] > eoff[b + 1] - eoff[b] + voff[b + 1] - voff[b];
    });
    GBK_TRY(ensure_init());
    gbk_graph_batch b = new gbk_graph_batch_s();
    b->d.ngraphs = ngraphs;
    b->d.total_v = tv;
    b->d.total_e = te;
    b->voff = voff;
    int rc = gb_upload(&b->d.voff, voff);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.eoff, eoff);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.rp, rp);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.col, col);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.crp, crp);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.ccol, ccol);
    if (rc == GBK_SUCCESS) rc = gb_upload(&b->d.order, order);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&b->d.src, sizeof(int32_t) * ngraphs);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&b->d.levels, sizeof(int32_t) * (tv > 0 ? tv : 1));
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&b->d.counter, 256);
    if (rc == GBK_SUCCESS && hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
        rc = fail(GBK_PANIC, "hipStreamCreate failed");
    if (rc == GBK_SUCCESS && (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess))
        rc = fail(GBK_PANIC, "hipEventCreate failed");
    if (rc != GBK_SUCCESS) {
        gbk_graph_batch_free(b);
        return rc;
    }
    // two persistent workgroups per CU (the LDS bitmaps take 48 KiB each)
    b->grid = std::max(1, std::min(ngraphs, 2 * std::max(device_cu_count(), 1)));
    // capture: reset the work queue (a kernel, not a memset node), then the persistent BFS
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        gbk_graph_batch_free(b);
        return fail(GBK_PANIC, "hipStreamBeginCapture failed");
    }
    hipLaunchKernelGGL(k_gb_reset, dim3(1), dim3(64), 0, b->st, b->d.counter);
    hipLaunchKernelGGL(k_gb_bfs, dim3((unsigned)b->grid), dim3(GB_THREADS), 0, b->st, b->d.voff, b->d.eoff, b->d.rp,
                       b->d.col, b->d.crp, b->d.ccol, b->d.order, b->d.src, ngraphs, b->d.levels, b->d.counter);
    hipError_t ce = hipStreamEndCapture(b->st, &b->graph);
    if (ce != hipSuccess || hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        gbk_graph_batch_free(b);
        return fail(GBK_PANIC, "HIP graph capture failed: %s", hipGetErrorString(ce));
    }
    *out = b;
    return GBK_SUCCESS;
}

extern "C" int gbk_graph_batch_bfs(gbk_graph_batch b, const int64_t* sources, int32_t* levels, double* device_ms) {
    if (!b || !sources) return fail(GBK_NULL_POINTER, "null argument");
    const int G = b->d.ngraphs;
    std::vector<int32_t> s(G);
    for (int g = 0; g < G; ++g) {
        const int64_t n = b->voff[g + 1] - b->voff[g];
        if (sources[g] < 0 || sources[g] >= n)
            return fail(GBK_INVALID_INDEX, "source %lld of graph %d out of range", (long long)sources[g], g);
        s[g] = (int32_t)sources[g];
    }
    GBK_TRY(ensure_init());
    GBK_HIP(hipMemcpyAsync(b->d.src, s.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, b->st));
    GBK_HIP(hipEventRecord(b->e0, b->st));
    GBK_HIP(hipGraphLaunch(b->exec, b->st));
    GBK_HIP(hipEventRecord(b->e1, b->st));
    if (levels)
        GBK_HIP(hipMemcpyAsync(levels, b->d.levels, sizeof(int32_t) * b->d.total_v, hipMemcpyDeviceToHost, b->st));
    GBK_HIP(hipStreamSynchronize(b->st));
    if (device_ms) {
        float ms = 0.f;
        GBK_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
        *device_ms = ms;
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_graph_batch_info(gbk_graph_batch b, int64_t* info) {
    if (!b || !info) return fail(GBK_NULL_POINTER, "null argument");
    info[0] = b->d.ngraphs;
    info[1] = b->d.total_v;
    info[2] = b->d.total_e;
    info[3] = b->grid;
    return GBK_SUCCESS;
}
