// Graph drivers on top of the SpMV / SpMSpV path, with the element-wise
// passes of each iteration fused into the traversal kernels.
//
//   BFS  — replaces bfs_blast (algorithms/bfs/bfs.hpp:94-142).  The
//          reference alternates push (SpMSpV over the CSR + complement mask)
//          and pull (masked SpMV over the CSC) by the 30% density of the
//          frontier vector.  Here the switch is Beamer's edge heuristic:
//          push while the frontier's out-edges m_f < m_u / 14 (m_u = edges
//          of unvisited vertices), pull while the frontier has more than
//          n / 24 vertices.  Push: edge-balanced expansion over the frontier
//          queue, atomicCAS claims a vertex, wave-aggregated append to the
//          next queue.  Pull: one lane group per unvisited vertex scans its
//          in-edges and stops at the first parent in the frontier.  Levels
//          are 1-based and 0 means unreached, as in the reference; they do
//          not depend on the direction chosen.
//   SSSP — replaces sssp_bellman_ford_blast (sssp_blast.hpp:8-72): the same
//          fixed point (min.plus relaxation until no distance improves),
//          computed over the set of vertices improved in the last round
//          instead of all vertices.  int32 distances saturate at INT32_MAX.
//   CC   — replaces FastSV cc (cc.hpp:12-119): min-label hooking over the
//          edges plus pointer jumping, which converges to the same labels
//          (the minimum vertex id of each component).
//   PR   — replaces LAGraph_page_rank_sinks (pr.hpp:20-153): r = teleport +
//          A (+.second) (t ./ d), d = max(coldeg / damping, 1 / damping),
//          teleport = (1 - damping) / n + damping / n * sum of the sinks'
//          ranks, with the SpMV on the merge-path kernel.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "gbk_objects.h"

namespace gbk {

int spmv_csr_unmasked(const DevCSR& m, int type, const void* x, void* y, int add, int mul, int accum);

static dim3 grid_n(int64_t n, int per = 256) {
    int64_t b = (n + per - 1) / per;
    return dim3((unsigned)(b > 0 ? b : 1));
}

__device__ __forceinline__ unsigned long long lanemask_lt64() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Appends v to queue (wave-aggregated: one atomic per wave).
__device__ __forceinline__ void wave_append(bool take, int32_t v, int32_t* __restrict__ queue, int32_t* __restrict__ count) {
    unsigned long long b = __ballot(take);
    if (b == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)b) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, (int)__popcll(b));
    base = __shfl(base, leader, 64);
    if (take) queue[base + __popcll(b & lanemask_lt64())] = v;
}

// ------------------------------------------------------------------- BFS --
__global__ void k_frontier_deg(const int32_t* __restrict__ rp, const int32_t* __restrict__ q, int64_t nq,
                               int32_t* __restrict__ deg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        int u = q[i];
        deg[i] = rp[u + 1] - rp[u];
    }
}

// Push: thread e owns edge e of the frontier's concatenated adjacency lists.
__global__ __launch_bounds__(256) void k_bfs_push(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ q, int64_t nq,
                                                  const int32_t* __restrict__ offs, int64_t nedges,
                                                  int32_t* __restrict__ level, int32_t next_level,
                                                  int32_t* __restrict__ nq_out, int32_t* __restrict__ nq_count) {
    __shared__ int32_t s_offs[EO_WIN + 1];
    __shared__ int64_t s_i0;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = edge_owner(offs, nq, e, s_offs, &s_i0);
    bool take = false;
    int32_t v = 0;
    if (e < nedges) {
        const int u = q[i];
        v = col[rp[u] + (e - offs[i])];
        if (level[v] == 0) take = (atomicCAS(&level[v], 0, next_level) == 0);
    }
    wave_append(take, v, nq_out, nq_count);
}

// Pull: G lanes per vertex; an unvisited vertex joins the next frontier if
// any in-neighbour is in the current one.  The frontier is read from a
// bitmap (n/8 bytes, L2-resident) rather than the level array; each lane
// tests four in-edges per step with all loads issued first, and the group
// stops as soon as one lane has found a parent.  Vertices with more than
// BFS_HEAVY in-edges are left to k_bfs_pull_heavy.
constexpr int BFS_HEAVY = 2048;

__device__ __forceinline__ bool fb_test(const uint32_t* __restrict__ fb, int u) {
    return (fb[u >> 5] >> (u & 31)) & 1u;
}

template <int G>
__global__ __launch_bounds__(256) void k_bfs_pull(const int32_t* __restrict__ crp, const int32_t* __restrict__ ccol,
                                                  int64_t n, int32_t* __restrict__ level,
                                                  const uint32_t* __restrict__ fb, int32_t cur, int heavy) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t v = gid / G;
    const int lane = (int)(gid % G);
    bool found = false;
    if (v < n && level[v] == 0) {
        const int e0 = crp[v], e1 = crp[v + 1];
        if (e1 - e0 <= heavy) {
            const int gbase = (threadIdx.x & 63) & ~(G - 1);
            const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << gbase);
            // group-uniform trip count, so every lane of the group takes part in each ballot
            for (int base = e0; base < e1; base += 4 * G) {
                int u[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int e = base + k * G + lane;
                    u[k] = e < e1 ? ccol[e] : -1;
                }
                bool hit = false;
#pragma unroll
                for (int k = 0; k < 4; ++k) hit |= (u[k] >= 0) && fb_test(fb, u[k]);
                found |= hit;
                // the whole group stops once any of its lanes has found a parent
                if (__ballot(found) & gmask) break;
            }
            found = (__ballot(found) & gmask) != 0;
        }
    }
    // the next frontier is compacted from the levels afterwards (no shared counter)
    if (found && lane == 0) level[v] = cur + 1;
}

// Unvisited vertices with more than BFS_HEAVY in-edges: one workgroup each.
__global__ __launch_bounds__(256) void k_bfs_pull_heavy(const int32_t* __restrict__ crp, const int32_t* __restrict__ ccol,
                                                        const int32_t* __restrict__ heavy, int64_t nheavy,
                                                        int32_t* __restrict__ level, const uint32_t* __restrict__ fb,
                                                        int32_t cur) {
    __shared__ int s_found;
    const int64_t h = blockIdx.x;
    if (h >= nheavy) return;
    const int v = heavy[h];
    if (level[v] != 0) return;  // uniform across the block
    if (threadIdx.x == 0) s_found = 0;
    __syncthreads();
    const int e0 = crp[v], e1 = crp[v + 1];
    for (int base = e0; base < e1; base += 4 * 256) {
        int u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = base + k * 256 + (int)threadIdx.x;
            u[k] = e < e1 ? ccol[e] : -1;
        }
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) hit |= (u[k] >= 0) && fb_test(fb, u[k]);
        if (hit) s_found = 1;
        __syncthreads();
        const int f = s_found;
        __syncthreads();
        if (f) break;
    }
    if (threadIdx.x == 0 && s_found) level[v] = cur + 1;
}

// Compaction of the vertices whose level is `target` into a queue, in vertex
// order: per-chunk counts, an exclusive scan of the counts, then a ballot /
// prefix-sum write (no shared counter).
constexpr int LC_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_lvl_count(const int32_t* __restrict__ level, int64_t n, int32_t target,
                                                   int32_t* __restrict__ counts) {
    __shared__ int s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * LC_CHUNK;
    int c = 0;
    for (int i = threadIdx.x; i < LC_CHUNK; i += 256) {
        const int64_t v = base + i;
        if (v < n && level[v] == target) ++c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_c, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_c;
}

__global__ __launch_bounds__(256) void k_lvl_compact(const int32_t* __restrict__ level, int64_t n, int32_t target,
                                                     const int32_t* __restrict__ offs, int32_t* __restrict__ out) {
    __shared__ int s_w[4];
    const int64_t base = (int64_t)blockIdx.x * LC_CHUNK;
    int o = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int r = 0; r < LC_CHUNK; r += 256) {
        const int64_t v = base + r + threadIdx.x;
        const bool p = v < n && level[v] == target;
        const unsigned long long b = __ballot(p);
        if (lane == 0) s_w[wid] = __popcll(b);
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            if (w < wid) pre += s_w[w];
            tot += s_w[w];
        }
        if (p) out[o + pre + __popcll(b & lanemask_lt64())] = (int32_t)v;
        o += tot;
        __syncthreads();
    }
}

// The frontier bitmap: bit v set when level[v] == cur (the vertices of the
// current queue).  One ballot per 64 vertices writes two whole words (no
// atomics; setting bits from the queue made consecutive threads contend for
// the same word).
__global__ void k_fb_from_levels(const int32_t* __restrict__ level, int64_t n, int32_t cur,
                                 uint32_t* __restrict__ fb, int64_t words) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(v < n && level[v] == cur);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w = v >> 5;
        if (w < words) fb[w] = (uint32_t)m;
        if (w + 1 < words) fb[w + 1] = (uint32_t)(m >> 32);
    }
}
// Vertices with more than BFS_HEAVY in-edges.
__global__ void k_heavy_list(const int32_t* __restrict__ crp, int64_t n, int heavy, int32_t* __restrict__ out,
                             int32_t* __restrict__ count) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = v < n && (crp[v + 1] - crp[v]) > heavy;
    wave_append(take, (int32_t)v, out, count);
}

__global__ void k_set_level(int32_t* level, int64_t s, int32_t* q, int32_t* count) {
    level[s] = 1;
    q[0] = (int32_t)s;
    *count = 1;
}

struct BfsScratch {
    int32_t* q[2];
    int32_t* cnt;  // [2] queue counters, [2] heavy count
    int32_t* deg;
    int32_t* offs;
    uint32_t* fb;     // frontier bitmap
    int32_t* heavy;   // vertices with more than BFS_HEAVY in-edges
};

static int bfs_run(gbk_matrix A, int64_t source, int32_t* level, int64_t* stats) {
    const int64_t n = A->n;
    const DevCSR& out = A->csr;  // frontier advances along rows of A (vxm)
    const DevCSR& in = A->csc;
    hipStream_t st = stream();
    void* p;
    // slot 5: two queues + degrees + offsets + heavy list (n+1 each) + counters + frontier bitmap
    const size_t qbytes = sizeof(int32_t) * (size_t)(n + 1);
    const int64_t fwords = (n + 31) / 32;
    GBK_TRY(scratch(5, 5 * qbytes + 64 + sizeof(uint32_t) * (size_t)fwords + 64, &p));
    BfsScratch s;
    s.q[0] = (int32_t*)p;
    s.q[1] = s.q[0] + (n + 1);
    s.deg = s.q[1] + (n + 1);
    s.offs = s.deg + (n + 1);
    s.heavy = s.offs + (n + 1);
    s.cnt = s.heavy + (n + 1);
    s.fb = (uint32_t*)(((uintptr_t)(s.cnt + 4) + 15) & ~(uintptr_t)15);
    GBK_TRY(fill_value(GBK_INT32, level, n, 0.0));
    hipLaunchKernelGGL(k_set_level, dim3(1), dim3(1), 0, st, level, source, s.q[0], s.cnt);
    GBK_LAUNCH_CHECK();
    int64_t nq = 1, m_u = A->nnz, levels = 0, pushes = 0, pulls = 0, traversed = 0;
    int32_t cur = 1;
    bool pull = false;
    int qi = 0;
    // Beamer's switch: pull once the frontier's out-edges exceed 1/alpha of
    // the unexplored edges, push again once the frontier is below n/beta.
    // The pull step here is cheap, so it pays earlier than Beamer's alpha =
    // 14 (symmetric R-MAT 23: alpha 14: 85.8, 50: 93.1, 200: 91.7 GTEPS).
    const double alpha = 50.0, beta = 24.0;
    // Lanes per vertex in the pull step.  Most unvisited vertices meet a
    // frontier parent within their first few in-edges, so one lane each
    // (four edges per load round) beat wider groups on symmetric R-MAT 23
    // (G = 1: 71.7, 2: 68.3, 4: 64.1, 8: 55.9, 16: 45.1, 32: 35.3 GTEPS);
    // vertices with more than BFS_HEAVY in-edges take a workgroup each.
    const double avg_in = n ? (double)in.nnz / (double)n : 0;
    const int G = avg_in < 64 ? 1 : 4;
    const int heavy = BFS_HEAVY;
    while (nq > 0) {
        // out-degrees of the frontier and their prefix sums: the total is the
        // frontier's out-edge count m_f (Beamer's switch), the offsets are
        // the push step's edge-balanced expansion
        hipLaunchKernelGGL(k_frontier_deg, grid_n(nq), dim3(256), 0, st, out.row_ptr, s.q[qi], nq, s.deg);
        GBK_LAUNCH_CHECK();
        int32_t ne = 0;
        GBK_TRY(exclusive_scan_i32(s.deg, s.offs, nq, &ne));
        const int64_t m_f = ne;
        traversed += m_f;
        if (!pull && (double)m_f > (double)m_u / alpha) pull = true;
        else if (pull && (double)nq < (double)n / beta) pull = false;
        m_u -= m_f;
        int32_t* nxt = s.q[qi ^ 1];
        int32_t* ncnt = s.cnt + (qi ^ 1);
        int64_t next_nq = 0;
        if (pull) {
            ++pulls;
            if (A->bfs_nheavy < 0) {
                // the heavy vertices depend on the graph only: listed once per matrix
                GBK_TRY(fill_value(GBK_INT32, s.cnt + 2, 1, 0.0));
                hipLaunchKernelGGL(k_heavy_list, grid_n(n), dim3(256), 0, st, in.row_ptr, n, heavy, s.heavy, s.cnt + 2);
                GBK_LAUNCH_CHECK();
                int32_t hc = 0;
                GBK_HIP(hipMemcpyAsync(&hc, s.cnt + 2, sizeof(int32_t), hipMemcpyDeviceToHost, st));
                GBK_HIP(hipStreamSynchronize(st));
                GBK_TRY(dev_alloc((void**)&A->bfs_heavy, sizeof(int32_t) * (hc > 0 ? hc : 1)));
                if (hc > 0)
                    GBK_HIP(hipMemcpyAsync(A->bfs_heavy, s.heavy, sizeof(int32_t) * hc, hipMemcpyDeviceToDevice, st));
                A->bfs_nheavy = hc;
            }
            const int64_t nheavy = A->bfs_nheavy;
            hipLaunchKernelGGL(k_fb_from_levels, grid_n(fwords * 32), dim3(256), 0, st, level, n, cur, s.fb, fwords);
            GBK_LAUNCH_CHECK();
            if (nheavy > 0) {
                hipLaunchKernelGGL(k_bfs_pull_heavy, dim3((unsigned)nheavy), dim3(256), 0, st, in.row_ptr, in.col,
                                   A->bfs_heavy, nheavy, level, s.fb, cur);
                GBK_LAUNCH_CHECK();
            }
#define GBK_PULL(GG)                                                                                  \
    case GG:                                                                                          \
        hipLaunchKernelGGL(k_bfs_pull<GG>, grid_n(n * GG), dim3(256), 0, st, in.row_ptr, in.col, n, level, s.fb, \
                           cur, heavy);                                                               \
        break;
            switch (G) {
                GBK_PULL(1)
                GBK_PULL(4)
            }
#undef GBK_PULL
            GBK_LAUNCH_CHECK();
            // next frontier = vertices whose level became cur + 1, in vertex order
            const int64_t nb = (n + LC_CHUNK - 1) / LC_CHUNK;
            hipLaunchKernelGGL(k_lvl_count, dim3((unsigned)nb), dim3(256), 0, st, level, n, cur + 1, s.deg);
            GBK_LAUNCH_CHECK();
            int32_t total = 0;
            GBK_TRY(exclusive_scan_i32(s.deg, s.offs, nb, &total));
            hipLaunchKernelGGL(k_lvl_compact, dim3((unsigned)nb), dim3(256), 0, st, level, n, cur + 1, s.offs, nxt);
            GBK_LAUNCH_CHECK();
            next_nq = total;
        } else {
            ++pushes;
            GBK_TRY(fill_value(GBK_INT32, ncnt, 1, 0.0));
            if (ne > 0) {
                hipLaunchKernelGGL(k_bfs_push, grid_n(ne), dim3(256), 0, st, out.row_ptr, out.col, s.q[qi], nq,
                                   s.offs, (int64_t)ne, level, cur + 1, nxt, ncnt);
                GBK_LAUNCH_CHECK();
            }
            int32_t c = 0;
            GBK_HIP(hipMemcpyAsync(&c, ncnt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            GBK_HIP(hipStreamSynchronize(st));
            next_nq = c;
        }
        nq = next_nq;
        qi ^= 1;
        ++cur;
        ++levels;
    }
    if (stats) {
        stats[0] = levels;
        stats[1] = pushes;
        stats[2] = pulls;
        stats[3] = traversed;
    }
    return GBK_SUCCESS;
}

// ------------------------------------------------------------------ SSSP --
template <typename T> __device__ __forceinline__ bool atomic_min_improve(T* p, T v) {
    if constexpr (std::is_same<T, int32_t>::value) {
        return atomicMin(p, v) > v;
    } else {
        unsigned int* a = (unsigned int*)p;
        unsigned int old = __atomic_load_n(a, __ATOMIC_RELAXED);
        while (true) {
            float o = __uint_as_float(old);
            if (!(v < o)) return false;
            unsigned int prev = atomicCAS(a, old, __float_as_uint(v));
            if (prev == old) return true;
            old = prev;
        }
    }
}

template <typename T> __device__ __forceinline__ T dist_add(T d, T w) {
    if constexpr (std::is_same<T, int32_t>::value) return sat_add<int32_t>(d, w);
    else return d + w;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sssp_relax(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    const T* __restrict__ w, const int32_t* __restrict__ q, int64_t nq,
                                                    const int32_t* __restrict__ offs, int64_t nedges,
                                                    T* __restrict__ dist, int32_t* __restrict__ inq,
                                                    int32_t round) {
    __shared__ int32_t s_offs[EO_WIN + 1];
    __shared__ int64_t s_i0;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = edge_owner(offs, nq, e, s_offs, &s_i0);
    if (e < nedges) {
        const int u = q[i];
        const int64_t k = rp[u] + (e - offs[i]);
        const int v = col[k];
        const T nd = dist_add(dist[u], w[k]);
        // an improved vertex is marked with the round; the next queue is
        // compacted from the marks (no shared counter)
        if (nd < dist[v] && atomic_min_improve(dist + v, nd)) inq[v] = round;
    }
}

template <typename T> __global__ void k_sssp_init(T* dist, int64_t n, int64_t s, T inf, int32_t* q, int32_t* count,
                                                  int32_t* inq) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dist[i] = (i == s) ? T(0) : inf;
        inq[i] = -1;
    }
    if (i == 0) {
        q[0] = (int32_t)s;
        *count = 1;
    }
}

template <typename T> static int sssp_run(gbk_matrix A, int64_t source, T* dist_dev, int64_t* stats) {
    const int64_t n = A->n;
    const DevCSR& m = A->csr;
    hipStream_t st = stream();
    void* p;
    const size_t qbytes = sizeof(int32_t) * (size_t)(n + 1);
    GBK_TRY(scratch(5, 5 * qbytes + 64, &p));
    int32_t* q[2] = {(int32_t*)p, (int32_t*)p + (n + 1)};
    int32_t* deg = q[1] + (n + 1);
    int32_t* offs = deg + (n + 1);
    int32_t* inq = offs + (n + 1);
    int32_t* cnt = inq + (n + 1);
    const T inf = std::is_same<T, int32_t>::value ? (T)INT32_MAX : (T)FLT_MAX;
    hipLaunchKernelGGL(k_sssp_init<T>, grid_n(n), dim3(256), 0, st, dist_dev, n, source, inf, q[0], cnt, inq);
    GBK_LAUNCH_CHECK();
    int64_t nq = 1, rounds = 0, relaxed = 0;
    int qi = 0;
    while (nq > 0) {
        hipLaunchKernelGGL(k_frontier_deg, grid_n(nq), dim3(256), 0, st, m.row_ptr, q[qi], nq, deg);
        GBK_LAUNCH_CHECK();
        int32_t ne = 0;
        GBK_TRY(exclusive_scan_i32(deg, offs, nq, &ne));
        GBK_TRY(fill_value(GBK_INT32, cnt + (qi ^ 1), 1, 0.0));
        int32_t c = 0;
        if (ne > 0) {
            hipLaunchKernelGGL((k_sssp_relax<T>), grid_n(ne), dim3(256), 0, st, m.row_ptr, m.col, (const T*)m.val,
                               q[qi], nq, offs, (int64_t)ne, dist_dev, inq, (int32_t)rounds);
            GBK_LAUNCH_CHECK();
            // next queue: the vertices marked in this round, in vertex order
            const int64_t nb = (n + LC_CHUNK - 1) / LC_CHUNK;
            hipLaunchKernelGGL(k_lvl_count, dim3((unsigned)nb), dim3(256), 0, st, inq, n, (int32_t)rounds, deg);
            GBK_LAUNCH_CHECK();
            GBK_TRY(exclusive_scan_i32(deg, offs, nb, &c));
            hipLaunchKernelGGL(k_lvl_compact, dim3((unsigned)nb), dim3(256), 0, st, inq, n, (int32_t)rounds, offs,
                               q[qi ^ 1]);
            GBK_LAUNCH_CHECK();
        }
        relaxed += ne;
        nq = c;
        qi ^= 1;
        ++rounds;
    }
    if (stats) {
        stats[0] = rounds;
        stats[1] = relaxed;
    }
    return GBK_SUCCESS;
}

// -------------------------------------------------------------------- CC --
__global__ void k_cc_init(int32_t* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// Hook: for every edge (u, v) with different parents, the larger parent is
// lowered to the smaller one.
__global__ __launch_bounds__(256) void k_cc_hook(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                 int64_t n, int32_t* __restrict__ parent, int32_t* __restrict__ changed) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t u = gid >> 2;
    const int lane = (int)(gid & 3);
    if (u >= n) return;
    bool ch = false;
    const int e1 = rp[u + 1];
    for (int e = rp[u] + lane; e < e1; e += 4) {
        const int v = col[e];
        int pu = parent[u], pv = parent[v];
        if (pu == pv) continue;
        int hi = pu > pv ? pu : pv, lo = pu > pv ? pv : pu;
        if (atomicMin(parent + hi, lo) > lo) ch = true;
    }
    if (ch) *changed = 1;
}

__global__ void k_cc_jump(int32_t* __restrict__ parent, int64_t n, int32_t* __restrict__ changed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t p = parent[i];
    int32_t pp = parent[p];
    if (pp != p) {
        while (true) {
            int32_t q = parent[pp];
            if (q == pp) break;
            pp = q;
        }
        parent[i] = pp;
        *changed = 1;
    }
}

static int cc_run(gbk_matrix A, int32_t* parent, int64_t* stats) {
    const int64_t n = A->n;
    const DevCSR& m = A->csr;
    hipStream_t st = stream();
    void* p;
    GBK_TRY(scratch(6, 64, &p));
    int32_t* changed = (int32_t*)p;
    hipLaunchKernelGGL(k_cc_init, grid_n(n), dim3(256), 0, st, parent, n);
    GBK_LAUNCH_CHECK();
    int64_t iters = 0;
    while (true) {
        GBK_TRY(fill_value(GBK_INT32, changed, 1, 0.0));
        hipLaunchKernelGGL(k_cc_hook, grid_n(n * 4), dim3(256), 0, st, m.row_ptr, m.col, n, parent, changed);
        GBK_LAUNCH_CHECK();
        // pointer jumping to a fixed point
        hipLaunchKernelGGL(k_cc_jump, grid_n(n), dim3(256), 0, st, parent, n, changed);
        GBK_LAUNCH_CHECK();
        int32_t c = 0;
        GBK_HIP(hipMemcpyAsync(&c, changed, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        ++iters;
        if (!c) break;
        if (iters > 4 * 64 + n) return fail(GBK_PANIC, "connected components did not converge");
    }
    if (stats) stats[0] = iters;
    return GBK_SUCCESS;
}

// -------------------------------------------------------------------- PR --
// PageRank with sinks (LAGraph_page_rank_sinks, algorithms/pr/pr.hpp:20-153).
// Per iteration: one SpMV r' = A (+.second) w written without accumulation,
// then k_pr_post adds the teleport term (1 - damping)/n + damping/n * (sum of
// the sinks' ranks) to every row, accumulates the L1 change, and forms the
// next iteration's w = r' ./ d and sink sum.  The reference's t = r copy,
// w = t ./ d pass, teleport assign, accumulate and diff reduction are this one
// pass; the ranks alternate between two buffers.  The teleport term is added
// to the finished row sum (float addition commutes), so the bits equal
// "r = teleport; r += A w".  d[i] = max(deg / damping, 1 / damping) is kept
// negated for sinks (deg == 0).

// The per-vertex sums (sink ranks, L1 change) are reduced per block into
// partials and then by one block in a fixed order: no atomics on one address
// (65 K waves adding into one double serialize for most of a millisecond at
// n = 4 M) and the same sums on every run.
constexpr int PR_THREADS = 256;
constexpr int PR_MAX_BLOCKS = 2048;

static int pr_blocks(int64_t n) {
    const int64_t b = (n + PR_THREADS - 1) / PR_THREADS;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, PR_MAX_BLOCKS));
}

// Block sum of (a, b) into out[0], out[1] (thread 0 writes).
__device__ __forceinline__ void pr_block_sum2(double a, double b, double* __restrict__ out) {
    __shared__ double sa[PR_THREADS / 64], sb[PR_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[wv] = a;
        sb[wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        for (int i = 0; i < PR_THREADS / 64; ++i) {
            ta += sa[i];
            tb += sb[i];
        }
        out[0] = ta;
        out[1] = tb;
    }
}

__global__ __launch_bounds__(PR_THREADS) void k_pr_init(const int32_t* __restrict__ crp, int64_t n, float damping,
                                                        float r0, float* __restrict__ r, float* __restrict__ w,
                                                        float* __restrict__ d, double* __restrict__ part) {
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int deg = crp[i + 1] - crp[i];
        float di = (float)deg / damping;
        const float dmin = 1.0f / damping;
        if (di < dmin) di = dmin;
        d[i] = deg == 0 ? -di : di;
        r[i] = r0;
        w[i] = r0 / di;
        if (deg == 0) s += r0;
    }
    pr_block_sum2(s, 0.0, part + 2 * blockIdx.x);
}

__global__ __launch_bounds__(PR_THREADS) void k_pr_post(const float* __restrict__ rold, float* __restrict__ rnew,
                                                        const float* __restrict__ d, float* __restrict__ w, int64_t n,
                                                        const double* __restrict__ sink_in, double base, double coef,
                                                        double* __restrict__ part) {
    const float tele = (float)(base + coef * *sink_in);
    double s = 0.0, df = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float r = tele + rnew[i];
        rnew[i] = r;
        df += fabs((double)rold[i] - (double)r);
        const float di = d[i];
        w[i] = r / fabsf(di);
        if (di < 0.0f) s += r;
    }
    pr_block_sum2(s, df, part + 2 * blockIdx.x);
}

// out[0], out[1] = sums of the nb block partials (one block, fixed order).
__global__ __launch_bounds__(PR_THREADS) void k_pr_final(const double* __restrict__ part, int nb,
                                                         double* __restrict__ out) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += PR_THREADS) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    pr_block_sum2(a, b, out);
}

// A small pinned host buffer per thread for the L1 changes, so that reading
// one does not wait for work queued after it.
static double* pinned_diff() {
    thread_local double* p = nullptr;
    if (!p && hipHostMalloc((void**)&p, 64, hipHostMallocDefault) != hipSuccess) p = nullptr;
    return p;
}

struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~EventPair() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// With tol > 0 the host reads iteration k's L1 change while iteration k + 1
// runs; when it is below tol the result of iteration k + 1 is dropped, so the
// ranks and the iteration count equal those of a loop that waits each time.
static int pr_run(gbk_matrix A, float* ranks_dev, int itermax, double damping, double tol, int* iters_out) {
    if (A->type != GBK_FP32) return fail(GBK_DOMAIN_MISMATCH, "PageRank needs an FP32 matrix");
    const int64_t n = A->n;
    hipStream_t st = stream();
    float* R[2] = {ranks_dev, (float*)A->ws.buf[0]};
    float* w = (float*)A->ws.buf[1];
    float* d = w + n;  // ws buffers hold n * 8 bytes
    const int nb = pr_blocks(n);
    void* p;
    GBK_TRY(scratch(6, 64 + 16 * (size_t)nb, &p));
    double* acc = (double*)p;  // slot k & 1: [sink sum, L1 change] of iteration k
    double* part = acc + 8;    // block partials
    double* hd = nullptr;
    EventPair ev;
    if (tol > 0) {
        hd = pinned_diff();
        if (!hd) return fail(GBK_OUT_OF_MEMORY, "pinned host buffer");
        for (auto& e : ev.e) GBK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    hipLaunchKernelGGL(k_pr_init, dim3(nb), dim3(PR_THREADS), 0, st, A->csc.row_ptr, n, (float)damping,
                       (float)(1.0 / (double)n), R[0], w, d, part);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pr_final, dim3(1), dim3(PR_THREADS), 0, st, part, nb, acc);
    GBK_LAUNCH_CHECK();
    const double base = (1.0 - damping) / (double)n, coef = damping / (double)n;
    int res = itermax;
    for (int k = 1; k <= itermax; ++k) {
        double* out = acc + 2 * (k & 1);
        GBK_TRY(spmv_unmasked(A, 0, w, R[k & 1], GBK_ADD_PLUS, GBK_MUL_SECOND, GBK_ACC_NONE));
        hipLaunchKernelGGL(k_pr_post, dim3(nb), dim3(PR_THREADS), 0, st, R[(k - 1) & 1], R[k & 1], d, w, n,
                           acc + 2 * ((k - 1) & 1), base, coef, part);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_pr_final, dim3(1), dim3(PR_THREADS), 0, st, part, nb, out);
        GBK_LAUNCH_CHECK();
        if (tol > 0) {
            GBK_HIP(hipMemcpyAsync(hd + (k & 1), out + 1, sizeof(double), hipMemcpyDeviceToHost, st));
            GBK_HIP(hipEventRecord(ev.e[k & 1], st));
            if (k >= 2) {
                GBK_HIP(hipEventSynchronize(ev.e[(k - 1) & 1]));
                if (hd[(k - 1) & 1] < tol) {
                    res = k - 1;
                    break;
                }
            }
        }
    }
    GBK_HIP(hipStreamSynchronize(st));
    if (res & 1) GBK_HIP(hipMemcpyAsync(R[0], R[1], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    GBK_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = res;
    return GBK_SUCCESS;
}

}  // namespace gbk

using namespace gbk;

static int check_graph(gbk_matrix A) {
    if (!A) return fail(GBK_UNINITIALIZED_OBJECT, "null matrix");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    return ensure_init();
}

extern "C" int gbk_bfs_device(gbk_matrix A, int64_t source, int32_t* levels_dev, int64_t* stats) {
    GBK_TRY(check_graph(A));
    if (!levels_dev) return fail(GBK_NULL_POINTER, "null levels");
    if (source < 0 || source >= A->n) return fail(GBK_INVALID_INDEX, "source %lld out of range", (long long)source);
    return bfs_run(A, source, levels_dev, stats);
}

extern "C" int gbk_bfs(gbk_matrix A, int64_t source, int32_t* levels, int64_t* stats) {
    GBK_TRY(check_graph(A));
    if (!levels) return fail(GBK_NULL_POINTER, "null levels");
    if (source < 0 || source >= A->n) return fail(GBK_INVALID_INDEX, "source %lld out of range", (long long)source);
    int32_t* lv = (int32_t*)A->ws.buf[3];
    GBK_TRY(bfs_run(A, source, lv, stats));
    GBK_HIP(hipMemcpyAsync(levels, lv, sizeof(int32_t) * A->n, hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}

static int sssp_into(gbk_matrix A, int64_t source, void* d, int64_t* stats) {
    GBK_TRY(check_graph(A));
    if (!d) return fail(GBK_NULL_POINTER, "null distances");
    if (source < 0 || source >= A->n) return fail(GBK_INVALID_INDEX, "source %lld out of range", (long long)source);
    if (A->type == GBK_INT32) return sssp_run<int32_t>(A, source, (int32_t*)d, stats);
    if (A->type == GBK_FP32) return sssp_run<float>(A, source, (float*)d, stats);
    return fail(GBK_DOMAIN_MISMATCH, "SSSP needs INT32 or FP32 weights");
}

extern "C" int gbk_sssp_device(gbk_matrix A, int64_t source, void* dist_dev, int64_t* stats) {
    return sssp_into(A, source, dist_dev, stats);
}

extern "C" int gbk_sssp(gbk_matrix A, int64_t source, void* dist, int64_t* stats) {
    if (!dist) return fail(GBK_NULL_POINTER, "null distances");
    GBK_TRY(check_graph(A));
    void* d = A->ws.buf[3];
    GBK_TRY(sssp_into(A, source, d, stats));
    GBK_HIP(hipMemcpyAsync(dist, d, 4 * A->n, hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}

extern "C" int gbk_cc(gbk_matrix A, int32_t* labels, int64_t* stats) {
    GBK_TRY(check_graph(A));
    if (!labels) return fail(GBK_NULL_POINTER, "null labels");
    int32_t* p = (int32_t*)A->ws.buf[3];
    GBK_TRY(cc_run(A, p, stats));
    GBK_HIP(hipMemcpyAsync(labels, p, sizeof(int32_t) * A->n, hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}

extern "C" int gbk_pagerank_device(gbk_matrix A, float* ranks_dev, int itermax, double damping, double tol,
                                   int* iters) {
    GBK_TRY(check_graph(A));
    if (!ranks_dev) return fail(GBK_NULL_POINTER, "null ranks");
    if (itermax < 0 || damping <= 0 || damping >= 1) return fail(GBK_INVALID_VALUE, "bad PageRank parameters");
    if (ranks_dev == (float*)A->ws.buf[0] || ranks_dev == (float*)A->ws.buf[1])
        return fail(GBK_INVALID_VALUE, "ranks may not live in the matrix workspace");
    return pr_run(A, ranks_dev, itermax, damping, tol, iters);
}

extern "C" int gbk_pagerank(gbk_matrix A, float* ranks, int itermax, double damping, double tol, int* iters) {
    GBK_TRY(check_graph(A));
    if (!ranks) return fail(GBK_NULL_POINTER, "null ranks");
    if (itermax < 0 || damping <= 0 || damping >= 1) return fail(GBK_INVALID_VALUE, "bad PageRank parameters");
    float* r = (float*)A->ws.buf[2];
    GBK_TRY(pr_run(A, r, itermax, damping, tol, iters));
    GBK_HIP(hipMemcpyAsync(ranks, r, sizeof(float) * A->n, hipMemcpyDeviceToHost, stream()));
    GBK_HIP(hipStreamSynchronize(stream()));
    return GBK_SUCCESS;
}
