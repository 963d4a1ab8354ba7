// Library runtime: device selection, stream, errors, device memory, and the
// data-parallel primitives (scan, stream compaction, fills, scatter) that the
// vector state machine and SpMSpV are built from.
//
// The compaction is the wave-ballot / prefix-sum form: in each 256-element
// round a wave takes __ballot of its predicate and every lane counts the set
// bits below it; the four wave counts are prefixed through LDS.  Rounds keep
// index order, so ids come out ascending — the order the reference's
// SparseVector::convert produces (sparse_vector.hpp:36-92).
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "gbk_objects.h"
#include "batch.h"

namespace gbk {

// ----------------------------------------------------------------- errors --
static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// ---------------------------------------------------------------- runtime --
static std::mutex g_mu;
static bool g_init = false;
static int g_device = 0;
static hipStream_t g_own_stream = nullptr;
static hipStream_t g_stream = nullptr;
static int g_cus = 0;
static int64_t g_lds = 0;

struct ScratchSlot { void* p = nullptr; size_t bytes = 0; };
static ScratchSlot g_scratch[16];

int init_device(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) return GBK_SUCCESS;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(GBK_PANIC, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count) return fail(GBK_INVALID_VALUE, "device %d out of range", device);
    GBK_HIP(hipSetDevice(device));
    GBK_HIP(hipStreamCreateWithFlags(&g_own_stream, hipStreamNonBlocking));
    g_stream = g_own_stream;
    hipDeviceProp_t prop;
    GBK_HIP(hipGetDeviceProperties(&prop, device));
    g_cus = prop.multiProcessorCount;
    g_lds = (int64_t)prop.sharedMemPerBlock;
    g_device = device;
    g_init = true;
    return GBK_SUCCESS;
}

int ensure_init() {
    if (g_init) {
        // keep the library device current on this thread
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != g_device) GBK_HIP(hipSetDevice(g_device));
        return GBK_SUCCESS;
    }
    return init_device(0);
}

// A thread may route the library's launches to a stream of its own (the
// query server runs one stream per worker thread).
static thread_local hipStream_t t_stream = nullptr;
hipStream_t stream() { return t_stream ? t_stream : g_stream; }
void set_thread_stream(hipStream_t s) { t_stream = s; }
int device_cu_count() { return g_cus; }
int64_t device_lds_bytes() { return g_lds; }

int dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(GBK_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return GBK_SUCCESS;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

int scratch(int slot, size_t bytes, void** p) {
    ScratchSlot& s = g_scratch[slot];
    if (s.bytes < bytes) {
        if (s.p) {
            GBK_HIP(hipStreamSynchronize(stream()));
            dev_free(s.p);
            s.p = nullptr;
            s.bytes = 0;
        }
        size_t b = bytes + bytes / 4 + 4096;
        GBK_TRY(dev_alloc(&s.p, b));
        s.bytes = b;
    }
    *p = s.p;
    return GBK_SUCCESS;
}

// ------------------------------------------------------------------ scan --
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Inclusive wave scan of an int64 value (wave64).
__device__ __forceinline__ long long wave_incl_scan(long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        long long o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Exclusive block scan (256 threads); returns the block total through *total.
__device__ __forceinline__ long long block_excl_scan(long long v, long long* total) {
    __shared__ long long s_w[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    long long incl = wave_incl_scan(v);
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    long long pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        long long s = s_w[w];
        if (w < wid) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + incl - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(const int32_t* __restrict__ in, int64_t n,
                                                                 long long* __restrict__ partial) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t idx = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    long long tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single-block exclusive scan of the per-tile partials (in place).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_spine(long long* __restrict__ partial, int64_t nb) {
    long long carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN_THREADS) {
        int64_t idx = base + threadIdx.x;
        long long v = (idx < nb) ? partial[idx] : 0;
        long long tot;
        long long ex = block_excl_scan(v, &tot);
        if (idx < nb) partial[idx] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                              int64_t n, const long long* __restrict__ partial) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t idx = base + i;
        v[i] = (idx < n) ? in[idx] : 0;
        s += v[i];
    }
    long long tot;
    long long run = partial[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = (int32_t)run;
        run += v[i];
    }
    // the last element of the last tile writes out[n]
    if (base <= n - 1 && n - 1 < base + SCAN_ITEMS) out[n] = (int32_t)run;
}

size_t scan_partial_bytes(int64_t n) { return sizeof(long long) * (size_t)((n + SCAN_TILE - 1) / SCAN_TILE + 1); }

int exclusive_scan_i32_on(const int32_t* in, int32_t* out, int64_t n, int32_t* total_host, hipStream_t st,
                          void* partial_ws, size_t ws_bytes) {
    if (n <= 0) {
        int32_t z = 0;
        GBK_HIP(hipMemcpyAsync(out, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
        GBK_HIP(hipStreamSynchronize(st));
        if (total_host) *total_host = 0;
        return GBK_SUCCESS;
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ws_bytes < sizeof(long long) * (size_t)nb) return fail(GBK_INSUFFICIENT_SPACE, "scan workspace too small");
    long long* partial = (long long*)partial_ws;
    // The per-tile partials must use the same element-to-tile map as the
    // apply pass: tile b owns [b*TILE, (b+1)*TILE).
    hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(SCAN_THREADS), 0, st, in, n, partial);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(SCAN_THREADS), 0, st, partial, nb);
    GBK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, st, in, out, n, partial);
    GBK_LAUNCH_CHECK();
    if (total_host) {
        GBK_HIP(hipMemcpyAsync(total_host, out + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
    }
    return GBK_SUCCESS;
}

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* total_host) {
    void* p = nullptr;
    const size_t b = scan_partial_bytes(n);
    GBK_TRY(scratch(0, b, &p));
    return exclusive_scan_i32_on(in, out, n, total_host, stream(), p, b);
}

template <typename T> int fill_typed(T* dst, int64_t n, T v);

// ------------------------------------------------------------ compaction --
constexpr int CMP_THREADS = 256;
constexpr int CMP_ROUNDS = 16;
constexpr int CMP_TILE = CMP_THREADS * CMP_ROUNDS;

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__global__ __launch_bounds__(CMP_THREADS) void k_count_nz(const T* __restrict__ v, int64_t n,
                                                          long long* __restrict__ partial) {
    const int64_t base = (int64_t)blockIdx.x * CMP_TILE;
    long long c = 0;
#pragma unroll 4
    for (int r = 0; r < CMP_ROUNDS; ++r) {
        int64_t idx = base + (int64_t)r * CMP_THREADS + threadIdx.x;
        if (idx < n && truth(v[idx])) ++c;
    }
    long long tot;
    (void)block_excl_scan(c, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(CMP_THREADS) void k_compact_nz(const T* __restrict__ v, int64_t n,
                                                            const long long* __restrict__ offs,
                                                            int32_t* __restrict__ ids, T* __restrict__ vals) {
    __shared__ int s_cnt[CMP_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * CMP_TILE;
    long long out = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int r = 0; r < CMP_ROUNDS; ++r) {
        int64_t idx = base + (int64_t)r * CMP_THREADS + threadIdx.x;
        T x = (idx < n) ? v[idx] : T(0);
        bool p = (idx < n) && truth(x);
        unsigned long long b = __ballot(p);
        int below = __popcll(b & lanemask_lt());
        if (lane == 0) s_cnt[wid] = __popcll(b);
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < CMP_THREADS / 64; ++w) {
            int c = s_cnt[w];
            if (w < wid) pre += c;
            tot += c;
        }
        if (p) {
            long long o = out + pre + below;
            ids[o] = (int32_t)idx;
            if (vals) vals[o] = x;
        }
        out += tot;
        __syncthreads();
    }
}

int count_nonzero(int type, const void* dense, int64_t n, int64_t* count) {
    if (n <= 0) { *count = 0; return GBK_SUCCESS; }
    hipStream_t st = stream();
    const int64_t nb = (n + CMP_TILE - 1) / CMP_TILE;
    void* p;
    GBK_TRY(scratch(1, sizeof(long long) * (nb + 1), &p));
    long long* partial = (long long*)p;
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        GBK_TRY(fill_typed<long long>(partial + nb, 1, 0));
        hipLaunchKernelGGL(k_count_nz<T>, dim3(nb), dim3(CMP_THREADS), 0, st, (const T*)dense, n, partial);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(SCAN_THREADS), 0, st, partial, nb + 1);
        GBK_LAUNCH_CHECK();
        long long tot = 0;
        // partial[nb] was 0 before the exclusive scan, so it now holds the total
        GBK_HIP(hipMemcpyAsync(&tot, partial + nb, sizeof(long long), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        *count = tot;
        return GBK_SUCCESS;
    });
}

int compact_nonzero(int type, const void* dense, int64_t n, int32_t* ids, void* vals, int64_t* count) {
    if (n <= 0) { *count = 0; return GBK_SUCCESS; }
    hipStream_t st = stream();
    const int64_t nb = (n + CMP_TILE - 1) / CMP_TILE;
    void* p;
    GBK_TRY(scratch(1, sizeof(long long) * (nb + 1), &p));
    long long* partial = (long long*)p;
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        GBK_TRY(fill_typed<long long>(partial + nb, 1, 0));
        hipLaunchKernelGGL(k_count_nz<T>, dim3(nb), dim3(CMP_THREADS), 0, st, (const T*)dense, n, partial);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(SCAN_THREADS), 0, st, partial, nb + 1);
        GBK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_compact_nz<T>, dim3(nb), dim3(CMP_THREADS), 0, st, (const T*)dense, n,
                           (const long long*)partial, ids, (T*)vals);
        GBK_LAUNCH_CHECK();
        long long tot = 0;
        GBK_HIP(hipMemcpyAsync(&tot, partial + nb, sizeof(long long), hipMemcpyDeviceToHost, st));
        GBK_HIP(hipStreamSynchronize(st));
        *count = tot;
        return GBK_SUCCESS;
    });
}

// ----------------------------------------------------------------- fills --
template <typename T>
__global__ void k_fill(T* __restrict__ d, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = v;
}

template <typename T>
__global__ void k_scatter(const int32_t* __restrict__ ids, const T* __restrict__ vals, int64_t nv, T* __restrict__ d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) d[ids[i]] = vals[i];
}

static dim3 fill_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    int64_t cap = (int64_t)device_cu_count() * 16;
    if (cap < 1) cap = 1024;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

template <typename T> int fill_typed(T* dst, int64_t n, T v) {
    if (n <= 0) return GBK_SUCCESS;
    hipLaunchKernelGGL(k_fill<T>, fill_grid(n), dim3(256), 0, stream(), dst, n, v);
    GBK_LAUNCH_CHECK();
    return GBK_SUCCESS;
}

int fill_value(int type, void* dst, int64_t n, double value) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return fill_typed<T>((T*)dst, n, from_double<T>(value));
    });
}

int fill_identity(int type, int add, void* dst, int64_t n) {
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        return fill_typed<T>((T*)dst, n, identity_rt<T>(add));
    });
}

int scatter_sparse(int type, const int32_t* ids, const void* vals, int64_t nvals, void* dense) {
    if (nvals <= 0) return GBK_SUCCESS;
    return dispatch_type(type, [&](auto tv) -> int {
        typedef decltype(tv) T;
        hipLaunchKernelGGL(k_scatter<T>, dim3((unsigned)((nvals + 255) / 256)), dim3(256), 0, stream(), ids,
                           (const T*)vals, nvals, (T*)dense);
        GBK_LAUNCH_CHECK();
        return GBK_SUCCESS;
    });
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_init(int device) { return init_device(device); }

extern "C" int gbk_set_stream(void* s) {
    GBK_TRY(ensure_init());
    g_stream = s ? (hipStream_t)s : g_own_stream;
    return GBK_SUCCESS;
}

extern "C" void* gbk_get_stream(void) {
    if (ensure_init() != GBK_SUCCESS) return nullptr;
    return (void*)g_stream;
}

extern "C" int gbk_synchronize(void) {
    GBK_TRY(ensure_init());
    GBK_HIP(hipStreamSynchronize(g_stream));
    return GBK_SUCCESS;
}

extern "C" const char* gbk_last_error(void) { return g_err; }

extern "C" int gbk_device_info(int* cu_count, char* name, int name_len) {
    GBK_TRY(ensure_init());
    hipDeviceProp_t prop;
    GBK_HIP(hipGetDeviceProperties(&prop, g_device));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
    }
    return GBK_SUCCESS;
}
