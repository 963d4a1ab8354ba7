// Query server: continuous batching of BFS / personalized-PageRank / SSSP
// queries on device-resident graphs.
//
// The reference has no serving layer; each query it answers is one call of
// its drivers (bfs_blast, algorithms/bfs/bfs.hpp:94-142; LAGraph_page_rank_sinks,
// algorithms/pr/pr.hpp:20-153; sssp_bellman_ford_blast,
// algorithms/sssp/sssp_blast.hpp:8-72).  Here every (graph, query kind) pair
// gets an engine (batch.h) that advances up to `slots` queries together,
// one level / iteration / relaxation round per step, in multi-vector
// launches.  Each engine runs on its own worker thread and HIP stream, so
// the steps of different graphs and kinds overlap on the device.
//
// Admission: queries carry an arrival time (seconds after gbk_server_run
// starts).  A worker admits an arrived query when its engine has a free
// slot and the server's HBM budget holds the query's footprint (its result
// vector); the engine's workspace is reserved from the same budget when the
// engine is created and returned when the engine has no more queries.  A
// query admitted while others run joins the batch at the next step
// (continuous batching): slots refill as queries finish.
//
// Results: at completion the worker writes the query's result vector
// (int32 levels, float ranks or int32 distances) to a per-slot device
// buffer, reduces it to a checksum (sum and count of reached / nonzero
// entries), and, when the server keeps results, copies it to host memory.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "gbk_objects.h"

namespace gbk {

namespace {

// sum (as double) and count of the entries that are reached: levels != 0,
// ranks != 0, distances != INT32_MAX.
template <typename T>
__global__ void k_checksum(const T* __restrict__ v, int64_t n, double* __restrict__ sum,
                           unsigned long long* __restrict__ cnt) {
    double s = 0.0;
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T x = v[i];
        bool reached;
        if constexpr (std::is_same<T, float>::value) reached = x != 0.0f;
        else reached = (x != 0) && (x != INT32_MAX);
        if (reached) {
            s += (double)x;
            ++c;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && (c || s != 0.0)) {
        atomicAdd(sum, s);
        atomicAdd(cnt, c);
    }
}

__global__ void k_zero_sum(double* sum, unsigned long long* cnt) {
    *sum = 0.0;
    *cnt = 0;
}

using clk = std::chrono::steady_clock;

struct Query {
    int graph = 0, kind = 0;
    int32_t source = 0;
    double arrival = 0.0, admit = -1.0, done = -1.0;
    int slot = -1, occupancy = 0;
    double sum = 0.0;
    int64_t count = 0;
    int status = GBK_SUCCESS;
    std::vector<char> result;  // kept results only
};

struct Graph {
    gbk_matrix m[3] = {nullptr, nullptr, nullptr};  // BFS, PPR, SSSP
};

}  // namespace
}  // namespace gbk

struct gbk_server_s {
    int64_t budget = 0;
    int ppr_slots = 32, sssp_slots = 32;
    double damping = 0.85, tol = 1e-6;
    int itermax = 100;
    bool keep = false;
    std::vector<gbk::Graph> graphs;
    std::vector<gbk::Query> queries;
    // run state
    std::mutex mu;
    int64_t used = 0, peak = 0;
    int engines_made = 0;
    std::atomic<int> waiters{0};  // workers waiting for budget to create an engine
    int64_t largest_engine = 0;   // largest engine reservation (workspace estimate + one result)
    int max_engines_live = 0, engines_live = 0;
    double wall = 0.0;
    int first_error = GBK_SUCCESS;
    char first_error_msg[256] = "";
};

namespace gbk {
namespace {

bool try_reserve(gbk_server s, int64_t bytes) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->used + bytes > s->budget) return false;
    s->used += bytes;
    s->peak = std::max(s->peak, s->used);
    return true;
}

void give_back(gbk_server s, int64_t bytes) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= bytes;
}

int64_t engine_estimate(gbk_matrix A, int kind, int k) {
    // workspace estimate used for admission before the engine exists
    const int64_t n = A->n;
    if (kind == GBK_QUERY_BFS) return (int64_t)msbfs_ws_bytes(n) + 1024;
    // the engines' own workspace sizes, plus the seed and host buffers
    if (kind == GBK_QUERY_PPR) return (int64_t)ppr_ws_bytes(A, k) + 512 + 8192;
    return (int64_t)mssp_ws_bytes(A, k) + 512 + 8192;
}

// One worker: the queries of one (graph, kind), in arrival order.
void worker(gbk_server s, int gi, int kind, std::vector<int64_t> qids, clk::time_point t0) {
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    auto record_error = [&](int rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->first_error == GBK_SUCCESS) {
            s->first_error = rc;
            snprintf(s->first_error_msg, sizeof(s->first_error_msg), "%s", gbk_last_error());
        }
    };
    if (ensure_init() != GBK_SUCCESS) {
        record_error(GBK_PANIC);
        return;
    }
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        record_error(fail(GBK_PANIC, "hipStreamCreate failed"));
        return;
    }
    set_thread_stream(st);
    gbk_matrix A = s->graphs[gi].m[kind];
    const int k = kind == GBK_QUERY_BFS ? 64 : kind == GBK_QUERY_PPR ? s->ppr_slots : s->sssp_slots;
    const int64_t n = A->n;
    const size_t rbytes = 4 * (size_t)n;
    std::unique_ptr<Engine> eng;
    int64_t eng_bytes = 0;
    std::vector<void*> rbuf(k, nullptr);
    std::vector<int64_t> slot_q(k, -1);
    int occupied = 0;
    int held = 0;  // result footprints reserved from the budget (one stays with the engine)
    bool waiting = false;
    size_t next = 0;
    double* dsum = nullptr;
    unsigned long long* dcnt = nullptr;
    int rc = dev_alloc((void**)&dsum, 256);
    if (rc == GBK_SUCCESS) rc = dev_alloc((void**)&dcnt, 256);
    std::vector<int> finished;
    while (rc == GBK_SUCCESS && (next < qids.size() || occupied > 0)) {
        const double t = now();
        if (!eng && next < qids.size() && s->queries[qids[next]].arrival <= t) {
            const int64_t est = engine_estimate(A, kind, k);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->largest_engine = std::max(s->largest_engine, est + (int64_t)rbytes);
            }
            if (est + (int64_t)rbytes > s->budget) {
                rc = fail(GBK_OUT_OF_MEMORY, "a query on graph %d needs %lld bytes of HBM, over the server budget", gi,
                          (long long)(est + (int64_t)rbytes));
                break;
            }
            // an engine is created with room for one result, so that every
            // engine that holds memory can always make progress
            if (!try_reserve(s, est + (int64_t)rbytes)) {
                if (!waiting) {
                    waiting = true;
                    ++s->waiters;
                }
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                continue;
            }
            if (waiting) {
                waiting = false;
                --s->waiters;
            }
            if (kind == GBK_QUERY_BFS) rc = make_bfs_engine(A, st, &eng);
            else if (kind == GBK_QUERY_PPR) rc = make_ppr_engine(A, st, k, s->damping, s->tol, s->itermax, &eng);
            else rc = make_sssp_engine(A, st, k, &eng);
            if (rc != GBK_SUCCESS) {
                give_back(s, est + (int64_t)rbytes);
                break;
            }
            eng_bytes = (int64_t)eng->bytes();
            give_back(s, est);
            held = 1;
            {
                // the real workspace is held whatever the estimate was
                std::lock_guard<std::mutex> lk(s->mu);
                s->used += eng_bytes;
                s->peak = std::max(s->peak, s->used);
                ++s->engines_made;
                ++s->engines_live;
                s->max_engines_live = std::max(s->max_engines_live, s->engines_live);
            }
        }
        // admit arrived queries into free slots while the budget holds them
        while (eng && next < qids.size() && occupied < k && s->queries[qids[next]].arrival <= now()) {
            if (occupied >= held) {
                if (!try_reserve(s, (int64_t)rbytes)) break;
                ++held;
            }
            int slot = 0;
            while (slot_q[slot] >= 0) ++slot;
            Query& q = s->queries[qids[next]];
            rc = eng->admit(slot, q.source);
            if (rc != GBK_SUCCESS) break;
            if (!rbuf[slot]) rc = dev_alloc(&rbuf[slot], rbytes);
            if (rc != GBK_SUCCESS) break;
            slot_q[slot] = qids[next];
            q.slot = slot;
            q.admit = now();
            q.occupancy = occupied + 1;
            ++occupied;
            ++next;
        }
        if (rc != GBK_SUCCESS) break;
        if (occupied == 0) {
            if (eng && next >= qids.size()) break;
            if (eng && s->waiters.load() > 0) {
                // idle while another worker waits for memory: hand the engine's memory back
                (void)hipStreamSynchronize(st);
                eng.reset();
                give_back(s, eng_bytes + (int64_t)held * (int64_t)rbytes);
                eng_bytes = 0;
                held = 0;
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    --s->engines_live;
                }
                continue;
            }
            // wait for the next arrival (or for budget)
            double wait = next < qids.size() ? s->queries[qids[next]].arrival - now() : 0.0;
            wait = std::min(std::max(wait, 0.0), 0.0002);
            std::this_thread::sleep_for(std::chrono::duration<double>(wait > 0 ? wait : 0.00005));
            continue;
        }
        rc = eng->step(finished);
        if (rc != GBK_SUCCESS) break;
        for (int slot : finished) {
            const int64_t qi = slot_q[slot];
            Query& q = s->queries[qi];
            rc = eng->extract(slot, rbuf[slot]);
            if (rc != GBK_SUCCESS) break;
            hipLaunchKernelGGL(k_zero_sum, dim3(1), dim3(1), 0, st, dsum, dcnt);
            const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256 > 0 ? (n + 255) / 256 : 1, 1024);
            if (kind == GBK_QUERY_PPR)
                hipLaunchKernelGGL(k_checksum<float>, dim3(blocks), dim3(256), 0, st, (const float*)rbuf[slot], n, dsum,
                                   dcnt);
            else
                hipLaunchKernelGGL(k_checksum<int32_t>, dim3(blocks), dim3(256), 0, st, (const int32_t*)rbuf[slot], n,
                                   dsum, dcnt);
            double hs = 0.0;
            unsigned long long hc = 0;
            if (hipMemcpyAsync(&hs, dsum, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&hc, dcnt, sizeof(hc), hipMemcpyDeviceToHost, st) != hipSuccess) {
                rc = fail(GBK_PANIC, "checksum copy failed");
                break;
            }
            if (s->keep) {
                q.result.resize(rbytes);
                if (hipMemcpyAsync(q.result.data(), rbuf[slot], rbytes, hipMemcpyDeviceToHost, st) != hipSuccess) {
                    rc = fail(GBK_PANIC, "result copy failed");
                    break;
                }
            }
            if (hipStreamSynchronize(st) != hipSuccess) {
                rc = fail(GBK_PANIC, "stream synchronize failed");
                break;
            }
            q.sum = hs;
            q.count = (int64_t)hc;
            q.done = now();
            eng->release(slot);
            slot_q[slot] = -1;
            --occupied;
            if (held > 1) {
                give_back(s, (int64_t)rbytes);
                --held;
            }
        }
    }
    if (waiting) --s->waiters;
    if (rc != GBK_SUCCESS) {
        record_error(rc);
        for (int slot = 0; slot < k; ++slot)
            if (slot_q[slot] >= 0) s->queries[slot_q[slot]].status = rc;
    }
    (void)hipStreamSynchronize(st);
    eng.reset();
    for (void* p : rbuf) dev_free(p);
    dev_free(dsum);
    dev_free(dcnt);
    if (eng_bytes) {
        give_back(s, eng_bytes + (int64_t)held * (int64_t)rbytes);
        std::lock_guard<std::mutex> lk(s->mu);
        --s->engines_live;
    }
    set_thread_stream(nullptr);
    (void)hipStreamDestroy(st);
}

}  // namespace
}  // namespace gbk

using namespace gbk;

extern "C" int gbk_server_new(gbk_server* out, int64_t hbm_budget_bytes, int ppr_slots, int sssp_slots,
                              double ppr_damping, double ppr_tol, int ppr_itermax, int keep_results) {
    if (!out) return fail(GBK_NULL_POINTER, "null server pointer");
    if (hbm_budget_bytes <= 0) return fail(GBK_INVALID_VALUE, "the HBM budget must be positive");
    if (ppr_slots < 1 || ppr_slots > 64 || sssp_slots < 1 || sssp_slots > 64)
        return fail(GBK_INVALID_VALUE, "engines take 1..64 slots");
    if (ppr_itermax < 1) return fail(GBK_INVALID_VALUE, "ppr_itermax must be positive");
    GBK_TRY(ensure_init());
    gbk_server s = new gbk_server_s();
    s->budget = hbm_budget_bytes;
    s->ppr_slots = ppr_slots;
    s->sssp_slots = sssp_slots;
    s->damping = ppr_damping;
    s->tol = ppr_tol;
    s->itermax = ppr_itermax;
    s->keep = keep_results != 0;
    *out = s;
    return GBK_SUCCESS;
}

extern "C" int gbk_server_free(gbk_server s) {
    delete s;
    return GBK_SUCCESS;
}

extern "C" int gbk_server_add_graph(gbk_server s, gbk_matrix bfs_A, gbk_matrix ppr_A, gbk_matrix sssp_A,
                                    int* graph_id) {
    if (!s || !graph_id) return fail(GBK_NULL_POINTER, "null argument");
    gbk_matrix ms[3] = {bfs_A, ppr_A, sssp_A};
    int64_t n = -1;
    for (int i = 0; i < 3; ++i) {
        if (!ms[i]) continue;
        if (!ms[i]->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
        if (n >= 0 && ms[i]->n != n) return fail(GBK_DIMENSION_MISMATCH, "the matrices of one graph differ in size");
        n = ms[i]->n;
    }
    if (n < 0) return fail(GBK_NULL_POINTER, "a graph needs at least one matrix");
    if (ppr_A && ppr_A->type != GBK_FP32) return fail(GBK_DOMAIN_MISMATCH, "the PageRank matrix must be FP32");
    if (sssp_A && sssp_A->type != GBK_INT32) return fail(GBK_DOMAIN_MISMATCH, "the SSSP matrix must be INT32");
    Graph g;
    for (int i = 0; i < 3; ++i) g.m[i] = ms[i];
    s->graphs.push_back(g);
    *graph_id = (int)s->graphs.size() - 1;
    return GBK_SUCCESS;
}

extern "C" int gbk_server_submit(gbk_server s, int graph_id, int kind, int64_t source, double arrival_s,
                                 int64_t* query_id) {
    if (!s) return fail(GBK_NULL_POINTER, "null server");
    if (graph_id < 0 || graph_id >= (int)s->graphs.size()) return fail(GBK_INVALID_INDEX, "unknown graph %d", graph_id);
    if (kind < GBK_QUERY_BFS || kind > GBK_QUERY_SSSP) return fail(GBK_INVALID_VALUE, "unknown query kind %d", kind);
    gbk_matrix A = s->graphs[graph_id].m[kind];
    if (!A) return fail(GBK_INVALID_VALUE, "graph %d has no matrix for query kind %d", graph_id, kind);
    if (source < 0 || source >= A->n) return fail(GBK_INVALID_INDEX, "source %lld out of range", (long long)source);
    Query q;
    q.graph = graph_id;
    q.kind = kind;
    q.source = (int32_t)source;
    q.arrival = arrival_s;
    s->queries.push_back(std::move(q));
    if (query_id) *query_id = (int64_t)s->queries.size() - 1;
    return GBK_SUCCESS;
}

extern "C" int gbk_server_run(gbk_server s) {
    if (!s) return fail(GBK_NULL_POINTER, "null server");
    GBK_TRY(ensure_init());
    // one worker per (graph, kind) that has queries, each with its queries in arrival order
    std::map<std::pair<int, int>, std::vector<int64_t>> work;
    for (int64_t i = 0; i < (int64_t)s->queries.size(); ++i) {
        Query& q = s->queries[i];
        if (q.done >= 0) continue;
        work[{q.graph, q.kind}].push_back(i);
    }
    for (auto& kv : work)
        std::stable_sort(kv.second.begin(), kv.second.end(),
                         [&](int64_t a, int64_t b) { return s->queries[a].arrival < s->queries[b].arrival; });
    s->first_error = GBK_SUCCESS;
    s->first_error_msg[0] = 0;
    s->waiters = 0;
    s->used = 0;
    s->peak = 0;
    const auto t0 = clk::now();
    std::vector<std::thread> threads;
    for (auto& kv : work) threads.emplace_back(worker, s, kv.first.first, kv.first.second, kv.second, t0);
    for (auto& t : threads) t.join();
    s->wall = std::chrono::duration<double>(clk::now() - t0).count();
    if (s->first_error != GBK_SUCCESS) return fail(s->first_error, "%s", s->first_error_msg);
    return GBK_SUCCESS;
}

extern "C" int gbk_server_query_info(gbk_server s, int64_t query_id, double* times, int64_t* info) {
    if (!s) return fail(GBK_NULL_POINTER, "null server");
    if (query_id < 0 || query_id >= (int64_t)s->queries.size()) return fail(GBK_INVALID_INDEX, "unknown query");
    const Query& q = s->queries[query_id];
    if (times) {
        times[0] = q.arrival;
        times[1] = q.admit;
        times[2] = q.done;
        times[3] = q.sum;
    }
    if (info) {
        info[0] = q.count;
        info[1] = q.occupancy;
        info[2] = q.slot;
        info[3] = q.status;
    }
    return GBK_SUCCESS;
}

extern "C" int gbk_server_result(gbk_server s, int64_t query_id, void* host_out) {
    if (!s || !host_out) return fail(GBK_NULL_POINTER, "null argument");
    if (query_id < 0 || query_id >= (int64_t)s->queries.size()) return fail(GBK_INVALID_INDEX, "unknown query");
    const Query& q = s->queries[query_id];
    if (q.result.empty()) return fail(GBK_NO_VALUE, "no kept result for query %lld", (long long)query_id);
    std::memcpy(host_out, q.result.data(), q.result.size());
    return GBK_SUCCESS;
}

extern "C" int gbk_server_stats(gbk_server s, double* stats) {
    if (!s || !stats) return fail(GBK_NULL_POINTER, "null argument");
    stats[0] = s->wall;
    stats[1] = (double)s->queries.size();
    stats[2] = (double)s->peak;
    stats[3] = (double)s->engines_made;
    stats[4] = (double)s->max_engines_live;
    stats[5] = (double)s->largest_engine;
    return GBK_SUCCESS;
}
