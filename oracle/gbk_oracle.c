/*
 * gbk_oracle.c — CPU restatement of the reference's semiring SpMV / SpMSpV
 * path, used ONLY as a test checker and as the timed CPU baseline.
 *
 * This file is test infrastructure.  Nothing in the product (libgbk.so,
 * include/, gb_kun_amd/) links, loads or calls it; it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg only.
 *
 * Every function names the reference file:line it restates
 * (afanasyev-ilya/GB_Kun).  Each is pinned by tests/test_oracle.py before it
 * is trusted as a checker:
 *   orc_spmv            golden vectors of the compiled reference (fp32 bits,
 *                       int32, masks, accumulators, the trailing-row defect),
 *                       dense brute force, int32 min.plus saturation
 *   orc_spmv_f64_of_f32 dense FP64 product on exactly representable data
 *   orc_balance_rows +
 *   orc_spmv_plus_times_f32_omp   reference fp32 bits on the golden graph
 *   orc_spmspv          golden SpMSpV vectors of the compiled reference
 *   orc_bfs             golden BFS levels; |level(u) - level(v)| <= 1 on edges
 *   orc_dijkstra_i32    golden SSSP distances; triangle inequality / tree
 *   orc_dijkstra_f32    equals orc_dijkstra_i32 on integer weights
 *   orc_cc              golden CC labels; constant on edges, component minimum
 *   orc_pagerank_f64    golden PageRank of the compiled reference; fixed point
 *   orc_ppr_f64         closed-form linear solve; uniform teleport reduces it
 *                       to orc_pagerank_f64
 *   orc_mxm_unmasked    scipy sparse products (int64 plus.times), A * I = A,
 *                       brute-force min.plus on small matrices
 *   orc_mxm_masked      dense products on small matrices (plus.times,
 *                       min.plus), restricted to the mask's pattern
 *   orc_masked_count    trace(A^3) / 6 of symmetric graphs (scipy)
 *
 * Conventions: CSR with int64 row_ptr[n+1] and int32 col[nnz]; values and
 * vectors are arrays of one element type given by `type`
 * (0 f32, 1 f64, 2 i32, 3 i64, 4 bool/u8).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

enum { T_F32 = 0, T_F64 = 1, T_I32 = 2, T_I64 = 3, T_BOOL = 4 };
enum { A_PLUS = 0, A_LOR, A_MIN, A_MAX, A_FIRST, A_SECOND, A_LESS, A_FIRSTMIN };
enum { M_TIMES = 0, M_LAND, M_FIRST, M_SECOND, M_PLUS, M_ONEB, M_NE, M_MINUS, M_DIV };
enum { B_NONE = 0, B_FIRST, B_PLUS, B_MINUS, B_TIMES, B_DIV, B_MIN, B_MAX, B_LOR, B_LAND, B_LESS, B_NE };
enum { D_SCMP = 0, D_STR_COMP = 5, D_DEFAULT = 6 };

/* ------------------------------------------------------------------------ */
/* Scalar semantics of the functors, src/cpp_graphblas/types.hpp:105-232.   */
/* Each type gets its own set through a macro; integer plus used as the     */
/* multiplier saturates (the GPU path's documented deviation; the reference */
/* wraps, types.hpp:107-109 — tests only feed it operands that do not       */
/* overflow unless they test the saturation itself).                        */
/* ------------------------------------------------------------------------ */

#define ORC_DEFINE_OPS(SUF, T, TMAX, TLOW, ISINT, ISBOOL)                                    \
    static T add_id_##SUF(int a) {                                                          \
        if (a == A_MIN || a == A_LESS) return TMAX;                                          \
        if (a == A_MAX) return TLOW;                                                        \
        return (T)0;                                                                        \
    }                                                                                        \
    static T wadd_##SUF(T l, T r) {                                                         \
        if (ISBOOL) return (T)((l != 0) || (r != 0));                                       \
        if (ISINT) return (T)((uint64_t)(int64_t)l + (uint64_t)(int64_t)r);                \
        return l + r;                                                                       \
    }                                                                                        \
    static T wsub_##SUF(T l, T r) {                                                         \
        if (ISBOOL) return (T)((l != 0) != (r != 0));                                       \
        if (ISINT) return (T)((uint64_t)(int64_t)l - (uint64_t)(int64_t)r);                \
        return l - r;                                                                       \
    }                                                                                        \
    static T wmul_##SUF(T l, T r) {                                                         \
        if (ISBOOL) return (T)((l != 0) && (r != 0));                                       \
        if (ISINT) return (T)((uint64_t)(int64_t)l * (uint64_t)(int64_t)r);                \
        return l * r;                                                                       \
    }                                                                                        \
    static T sadd_##SUF(T l, T r) {                                                         \
        if (ISINT) {                                                                        \
            if (sizeof(T) == 4) {                                                           \
                int64_t s = (int64_t)l + (int64_t)r;                                        \
                if (s > INT32_MAX) return (T)INT32_MAX;                                     \
                if (s < INT32_MIN) return (T)INT32_MIN;                                     \
                return (T)s;                                                                \
            } else {                                                                        \
                int64_t res;                                                                \
                if (__builtin_add_overflow((int64_t)l, (int64_t)r, &res))                   \
                    return (T)((int64_t)l > 0 ? INT64_MAX : INT64_MIN);                     \
                return (T)res;                                                              \
            }                                                                               \
        }                                                                                   \
        return wadd_##SUF(l, r);                                                            \
    }                                                                                        \
    static T sdiv_##SUF(T l, T r) {                                                         \
        if (ISINT || ISBOOL) {                                                              \
            if (r == 0) return (T)0;                                                        \
            if (ISINT && r == (T)-1) return wsub_##SUF((T)0, l);                            \
        }                                                                                   \
        return l / r;                                                                       \
    }                                                                                        \
    static T add_##SUF(int a, T l, T r) {                                                   \
        switch (a) {                                                                        \
            case A_PLUS: return wadd_##SUF(l, r);                                           \
            case A_LOR: return (T)((l != 0) || (r != 0));                                   \
            case A_MIN: case A_FIRSTMIN: return (r < l) ? r : l;                            \
            case A_MAX: return (r > l) ? r : l;                                             \
            case A_FIRST: return l;                                                         \
            case A_SECOND: return r;                                                        \
            default: return (T)(l < r);                                                     \
        }                                                                                   \
    }                                                                                        \
    static T mul_##SUF(int m, T a, T x) {                                                   \
        switch (m) {                                                                        \
            case M_TIMES: return wmul_##SUF(a, x);                                          \
            case M_LAND: return (T)((a != 0) && (x != 0));                                  \
            case M_FIRST: return a;                                                         \
            case M_SECOND: return x;                                                        \
            case M_PLUS: return sadd_##SUF(a, x);                                           \
            case M_ONEB: return (T)1;                                                       \
            case M_NE: return (T)(a != x);                                                  \
            case M_MINUS: return wsub_##SUF(a, x);                                          \
            default: return sdiv_##SUF(a, x);                                               \
        }                                                                                   \
    }                                                                                        \
    static T bin_##SUF(int op, T l, T r) {                                                  \
        switch (op) {                                                                       \
            case B_NONE: return r;                                                          \
            case B_FIRST: return l;                                                         \
            case B_PLUS: return wadd_##SUF(l, r);                                           \
            case B_MINUS: return wsub_##SUF(l, r);                                          \
            case B_TIMES: return wmul_##SUF(l, r);                                          \
            case B_DIV: return sdiv_##SUF(l, r);                                            \
            case B_MIN: return (r < l) ? r : l;                                             \
            case B_MAX: return (r > l) ? r : l;                                             \
            case B_LOR: return (T)((l != 0) || (r != 0));                                   \
            case B_LAND: return (T)((l != 0) && (r != 0));                                  \
            case B_LESS: return (T)(l < r);                                                 \
            default: return (T)(l != r);                                                    \
        }                                                                                   \
    }

ORC_DEFINE_OPS(f32, float, FLT_MAX, -FLT_MAX, 0, 0)
ORC_DEFINE_OPS(f64, double, DBL_MAX, -DBL_MAX, 0, 0)
ORC_DEFINE_OPS(i32, int32_t, INT32_MAX, INT32_MIN, 1, 0)
ORC_DEFINE_OPS(i64, int64_t, INT64_MAX, INT64_MIN, 1, 0)
ORC_DEFINE_OPS(u8, uint8_t, 1, 0, 0, 1)

static int mask_true(const void* m, int mt, int64_t i) {
    switch (mt) {
        case T_F32: return ((const float*)m)[i] != 0.0f;
        case T_F64: return ((const double*)m)[i] != 0.0;
        case T_I32: return ((const int32_t*)m)[i] != 0;
        case T_I64: return ((const int64_t*)m)[i] != 0;
        default: return ((const uint8_t*)m)[i] != 0;
    }
}

/* ------------------------------------------------------------------------ */
/* balance_matrix_rows, src/helpers/balancing/balancing.h:5-33: thread t    */
/* gets rows [lower_bound(rp, t*q), min(n, lower_bound(rp, (t+1)*q))) with  */
/* q = (nnz-1)/T + 1.                                                       */
/* ------------------------------------------------------------------------ */
static int64_t lower_bound_i64(const int64_t* a, int64_t len, int64_t v) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

void orc_balance_rows(int64_t n, const int64_t* rp, int nthreads, int64_t* first, int64_t* last) {
    int64_t nnz = rp[n];
    int64_t q = (nnz - 1) / nthreads + 1;
    for (int t = 0; t < nthreads; ++t) {
        int64_t lo = lower_bound_i64(rp, n + 1, q * t);
        int64_t hi = lower_bound_i64(rp, n + 1, q * (t + 1));
        if (hi > n) hi = n;
        first[t] = lo;
        last[t] = hi;
    }
}

/* ------------------------------------------------------------------------ */
/* SpMV restatements.                                                       */
/*  mask_kind 0: none  — SpMV_all_active_diff_vectors (spmv_csr.h:341-384): */
/*               y[r] = accum(y[r], left fold from identity in CSR order).   */
/*               ref_partition_threads > 0 reproduces the reference's      */
/*               row partition, INCLUDING its defect that rows beyond the   */
/*               last thread's range are not visited (balancing.h:23-24);   */
/*               0 visits every row (the corrected semantics).              */
/*  mask_kind 1: dense mask — SpMV_dense (spmv_csr.h:267-336).              */
/*  mask_kind 2: sparse mask ids — SpMV_sparse (spmv_csr.h:176-262).        */
/* ------------------------------------------------------------------------ */
#define ORC_SPMV(SUF, T)                                                                                   \
    static void spmv_##SUF(int64_t n, const int64_t* rp, const int32_t* col, const T* val, const T* x, T* y, \
                           int add, int mul, int accum, int mask_kind, const void* mask, int mask_type,       \
                           const int32_t* mask_ids, int64_t mask_nvals, int desc_mask, int ref_threads) {   \
        const T id = add_id_##SUF(add);                                                                   \
        if (mask_kind == 0) {                                                                             \
            int64_t r_first = 0, r_last = n;                                                             \
            int64_t *fs = NULL, *ls = NULL;                                                               \
            int nt = ref_threads > 0 ? ref_threads : 1;                                                  \
            if (ref_threads > 0) {                                                                       \
                fs = (int64_t*)malloc(sizeof(int64_t) * nt);                                             \
                ls = (int64_t*)malloc(sizeof(int64_t) * nt);                                             \
                orc_balance_rows(n, rp, nt, fs, ls);                                                      \
            }                                                                                             \
            for (int t = 0; t < nt; ++t) {                                                               \
                if (ref_threads > 0) { r_first = fs[t]; r_last = ls[t]; }                                \
                for (int64_t r = r_first; r < r_last; ++r) {                                             \
                    T res = id;                                                                           \
                    for (int64_t j = rp[r]; j < rp[r + 1]; ++j) res = add_##SUF(add, res, mul_##SUF(mul, val[j], x[col[j]])); \
                    y[r] = bin_##SUF(accum, y[r], res);                                                  \
                }                                                                                         \
            }                                                                                             \
            free(fs);                                                                                     \
            free(ls);                                                                                     \
            return;                                                                                       \
        }                                                                                                 \
        int comp = (desc_mask == D_STR_COMP);                                                             \
        if (mask_kind == 1) {                                                                             \
            for (int64_t r = 0; r < n; ++r) {                                                            \
                int m = mask_true(mask, mask_type, r);                                                    \
                if ((!comp && m) || (comp && !m)) {                                                       \
                    T res = id;                                                                           \
                    for (int64_t j = rp[r]; j < rp[r + 1]; ++j) res = add_##SUF(add, res, mul_##SUF(mul, val[j], x[col[j]])); \
                    y[r] = bin_##SUF(accum, y[r], res);                                                  \
                } else {                                                                                  \
                    y[r] = (T)0;                                                                          \
                }                                                                                         \
            }                                                                                             \
            return;                                                                                       \
        }                                                                                                 \
        if (!comp) {                                                                                      \
            for (int64_t i = 0; i < mask_nvals; ++i) {                                                   \
                int64_t r = mask_ids[i];                                                                  \
                T res = id;                                                                               \
                for (int64_t j = rp[r]; j < rp[r + 1]; ++j) res = add_##SUF(add, res, mul_##SUF(mul, val[j], x[col[j]])); \
                y[r] = bin_##SUF(accum, y[r], res);                                                      \
            }                                                                                             \
        } else {                                                                                          \
            uint8_t* dm = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);                                   \
            for (int64_t i = 0; i < mask_nvals; ++i) dm[mask_ids[i]] = 1;                                 \
            for (int64_t r = 0; r < n; ++r) {                                                            \
                if (!dm[r]) {                                                                             \
                    T res = id;                                                                           \
                    for (int64_t j = rp[r]; j < rp[r + 1]; ++j) res = add_##SUF(add, res, mul_##SUF(mul, val[j], x[col[j]])); \
                    y[r] = bin_##SUF(accum, y[r], res);                                                  \
                } else {                                                                                  \
                    y[r] = (T)0;                                                                          \
                }                                                                                         \
            }                                                                                             \
            free(dm);                                                                                     \
        }                                                                                                 \
    }

ORC_SPMV(f32, float)
ORC_SPMV(f64, double)
ORC_SPMV(i32, int32_t)
ORC_SPMV(i64, int64_t)
ORC_SPMV(u8, uint8_t)

int orc_spmv(int type, int64_t n, const int64_t* rp, const int32_t* col, const void* val, const void* x, void* y,
             int add, int mul, int accum, int mask_kind, const void* mask, int mask_type, const int32_t* mask_ids,
             int64_t mask_nvals, int desc_mask, int ref_threads) {
    switch (type) {
        case T_F32: spmv_f32(n, rp, col, val, x, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask, ref_threads); return 0;
        case T_F64: spmv_f64(n, rp, col, val, x, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask, ref_threads); return 0;
        case T_I32: spmv_i32(n, rp, col, val, x, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask, ref_threads); return 0;
        case T_I64: spmv_i64(n, rp, col, val, x, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask, ref_threads); return 0;
        case T_BOOL: spmv_u8(n, rp, col, val, x, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask, ref_threads); return 0;
    }
    return -1;
}

/* FP64 plus.times of an fp32 matrix and vector: the closed-form reference  */
/* for fp32 plus.times error bounds.                                        */
void orc_spmv_f64_of_f32(int64_t n, const int64_t* rp, const int32_t* col, const float* val, const float* x,
                         double* y, double* abs_sum) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
        double s = 0.0, a = 0.0;
        for (int64_t j = rp[r]; j < rp[r + 1]; ++j) {
            double p = (double)val[j] * (double)x[col[j]];
            s += p;
            a += fabs(p);
        }
        y[r] = s;
        if (abs_sum) abs_sum[r] = a;
    }
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: SpMV_all_active_diff_vectors with plus.times fp32 and the  */
/* reference's OpenMP row partition (balancing.h:5-33) — trailing rows not  */
/* covered by the partition are visited too (defect corrected).             */
/* ------------------------------------------------------------------------ */
void orc_spmv_plus_times_f32_omp(int64_t n, const int64_t* rp, const int32_t* col, const float* val, const float* x,
                                 float* y, int nthreads, const int64_t* first, const int64_t* last) {
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        int64_t r0 = first[t], r1 = last[t];
        if (t == nthreads - 1) r1 = n;
        for (int64_t r = r0; r < r1; ++r) {
            float res = 0.0f;
            for (int64_t j = rp[r]; j < rp[r + 1]; ++j) res = res + val[j] * x[col[j]];
            y[r] = res;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SpMSpV + apply_mask, src/backend/spmspv/spmspv.h:17-158 and              */
/* spmspv_atomics.h:5-170, run sequentially.  The kernel family follows     */
/* add(3, 5): 8 -> y += mul(a, x); 1 -> y = 1; else y = add(y, mul(a, x)).   */
/* rp/col/val are the rows pushed along (CSC for mxv, CSR for vxm).         */
/* mask_kind 0 none, 1 dense (mask, mask_type), 2 sparse (mask_ids).        */
/* ------------------------------------------------------------------------ */
#define ORC_SPMSPV(SUF, T)                                                                                    \
    static void spmspv_##SUF(int64_t n, const int64_t* rp, const int32_t* col, const T* val, const int32_t* xi, \
                             const T* xv, int64_t xn, T* y, int add, int mul, int accum, int mask_kind,          \
                             const void* mask, int mask_type, const int32_t* mask_ids, int64_t mask_nvals,       \
                             int desc_mask) {                                                                    \
        T* old = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                                               \
        memcpy(old, y, sizeof(T) * (size_t)n);                                                                  \
        T probe = add_##SUF(add, (T)3, (T)5);                                                                   \
        int kind = ((int)probe == 8) ? 0 : ((int)probe == 1) ? 1 : 2;                                          \
        T id = add_id_##SUF(add);                                                                               \
        for (int64_t i = 0; i < n; ++i) y[i] = id;                                                             \
        for (int64_t i = 0; i < xn; ++i) {                                                                     \
            int64_t u = xi[i];                                                                                  \
            for (int64_t j = rp[u]; j < rp[u + 1]; ++j) {                                                      \
                int64_t d = col[j];                                                                             \
                if (kind == 0) y[d] = y[d] + mul_##SUF(mul, val[j], xv[i]);                                    \
                else if (kind == 1) y[d] = (T)1;                                                                \
                else y[d] = add_##SUF(add, y[d], mul_##SUF(mul, val[j], xv[i]));                               \
            }                                                                                                   \
        }                                                                                                       \
        if (mask_kind != 0) {                                                                                   \
            int comp = (desc_mask == D_STR_COMP);                                                               \
            if (mask_kind == 1) {                                                                               \
                for (int64_t i = 0; i < n; ++i) {                                                              \
                    int m = mask_true(mask, mask_type, i);                                                      \
                    y[i] = (comp ? !m : m) ? bin_##SUF(accum, old[i], y[i]) : (T)0;                           \
                }                                                                                               \
            } else if (!comp) {                                                                                 \
                for (int64_t k = 0; k < mask_nvals; ++k) {                                                     \
                    int64_t i = mask_ids[k];                                                                    \
                    y[i] = bin_##SUF(accum, old[i], y[i]);                                                     \
                }                                                                                               \
            } else {                                                                                            \
                uint8_t* dm = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));                                       \
                memset(dm, 1, (size_t)n);                                                                       \
                for (int64_t k = 0; k < mask_nvals; ++k) dm[mask_ids[k]] = 0;                                  \
                for (int64_t i = 0; i < n; ++i) y[i] = dm[i] ? bin_##SUF(accum, old[i], y[i]) : (T)0;         \
                free(dm);                                                                                       \
            }                                                                                                   \
        }                                                                                                       \
        free(old);                                                                                              \
    }

ORC_SPMSPV(f32, float)
ORC_SPMSPV(f64, double)
ORC_SPMSPV(i32, int32_t)
ORC_SPMSPV(i64, int64_t)
ORC_SPMSPV(u8, uint8_t)

int orc_spmspv(int type, int64_t n, const int64_t* rp, const int32_t* col, const void* val, const int32_t* xi,
               const void* xv, int64_t xn, void* y, int add, int mul, int accum, int mask_kind, const void* mask,
               int mask_type, const int32_t* mask_ids, int64_t mask_nvals, int desc_mask) {
    switch (type) {
        case T_F32: spmspv_f32(n, rp, col, val, xi, xv, xn, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask); return 0;
        case T_F64: spmspv_f64(n, rp, col, val, xi, xv, xn, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask); return 0;
        case T_I32: spmspv_i32(n, rp, col, val, xi, xv, xn, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask); return 0;
        case T_I64: spmspv_i64(n, rp, col, val, xi, xv, xn, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask); return 0;
        case T_BOOL: spmspv_u8(n, rp, col, val, xi, xv, xn, y, add, mul, accum, mask_kind, mask, mask_type, mask_ids, mask_nvals, desc_mask); return 0;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Graph checkers (the reference drivers' own sequential checkers).         */
/* ------------------------------------------------------------------------ */

/* BFS levels, algorithms/bfs/bfs_traditional.hpp:11-48 semantics as        */
/* bfs_blast reports them (bfs.hpp:94-142): the frontier advances along     */
/* rows of A (vxm), source level 1, unreached 0.                            */
void orc_bfs(int64_t n, const int64_t* rp, const int32_t* col, int64_t source, int32_t* level) {
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) level[i] = 0;
    int64_t head = 0, tail = 0;
    level[source] = 1;
    q[tail++] = (int32_t)source;
    while (head < tail) {
        int32_t u = q[head++];
        for (int64_t j = rp[u]; j < rp[u + 1]; ++j) {
            int32_t v = col[j];
            if (level[v] == 0) {
                level[v] = level[u] + 1;
                q[tail++] = v;
            }
        }
    }
    free(q);
}

/* Dijkstra with a binary heap, algorithms/sssp/sssp_traditional.hpp:9-77,  */
/* int32 weights; unreachable = INT32_MAX.                                  */
typedef struct { int64_t d; int32_t v; } heap_item;
static void heap_push(heap_item* h, int64_t* sz, heap_item it) {
    int64_t i = (*sz)++;
    h[i] = it;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (h[p].d <= h[i].d) break;
        heap_item t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}
static heap_item heap_pop(heap_item* h, int64_t* sz) {
    heap_item top = h[0];
    h[0] = h[--(*sz)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *sz && h[l].d < h[m].d) m = l;
        if (r < *sz && h[r].d < h[m].d) m = r;
        if (m == i) break;
        heap_item t = h[m]; h[m] = h[i]; h[i] = t;
        i = m;
    }
    return top;
}

void orc_dijkstra_i32(int64_t n, const int64_t* rp, const int32_t* col, const int32_t* w, int64_t source,
                      int32_t* dist) {
    int64_t* d = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t nnz = rp[n];
    heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (size_t)(nnz + 2));
    int64_t sz = 0;
    for (int64_t i = 0; i < n; ++i) d[i] = INT64_MAX;
    d[source] = 0;
    heap_item s = {0, (int32_t)source};
    heap_push(h, &sz, s);
    while (sz > 0) {
        heap_item it = heap_pop(h, &sz);
        if (it.d != d[it.v]) continue;
        for (int64_t j = rp[it.v]; j < rp[it.v + 1]; ++j) {
            int64_t nd = it.d + w[j];
            if (nd < d[col[j]]) {
                d[col[j]] = nd;
                heap_item x = {nd, col[j]};
                heap_push(h, &sz, x);
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) dist[i] = d[i] == INT64_MAX ? INT32_MAX : (d[i] > INT32_MAX ? INT32_MAX : (int32_t)d[i]);
    free(d);
    free(h);
}

/* Float Dijkstra (sssp_traditional.hpp:9-77 semantics); unreachable FLT_MAX. */
void orc_dijkstra_f32(int64_t n, const int64_t* rp, const int32_t* col, const float* w, int64_t source, float* dist) {
    /* exact for integer-valued weights whose path sums stay below 2^24 */
    int64_t nnz = rp[n];
    int32_t* wi = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    int32_t* di = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t j = 0; j < nnz; ++j) wi[j] = (int32_t)w[j];
    orc_dijkstra_i32(n, rp, col, wi, source, di);
    for (int64_t i = 0; i < n; ++i) dist[i] = di[i] == INT32_MAX ? FLT_MAX : (float)di[i];
    free(wi);
    free(di);
}

/* Connected components by union-find; label = minimum vertex id of the     */
/* component (the fixed point of FastSV, algorithms/cc/cc.hpp:12-119; the   */
/* reference's checker cc_bfs_based_sequential compares up to a bijection,  */
/* cc_traditional.hpp:10-56).                                               */
static int64_t uf_find(int64_t* p, int64_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}
void orc_cc(int64_t n, const int64_t* rp, const int32_t* col, int32_t* label) {
    int64_t* p = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) p[i] = i;
    for (int64_t u = 0; u < n; ++u)
        for (int64_t j = rp[u]; j < rp[u + 1]; ++j) {
            int64_t a = uf_find(p, u), b = uf_find(p, col[j]);
            if (a != b) {
                if (a < b) p[b] = a;
                else p[a] = b;
            }
        }
    for (int64_t i = 0; i < n; ++i) label[i] = (int32_t)uf_find(p, i);
    free(p);
}

/* PageRank with sinks, LAGraph_page_rank_sinks (algorithms/pr/pr.hpp:20-153),  */
/* in double precision: r_new[i] = teleport + sum_{j in row i} r[j] / d[j]  */
/* with d[j] = max(coldeg[j] / damping, 1 / damping), teleport =            */
/* (1-damping)/n + damping/n * sum_{sinks} r, sinks = vertices with         */
/* coldeg 0.  Runs itermax iterations, or stops earlier when tol > 0 and    */
/* the L1 change falls below tol.  Returns the iterations run.              */
/* sink_quirk = 1 restates what the reference computes for the sink sum    */
/* when the sink vector is sparse (at most 30% sinks): its masked vector    */
/* assign rsink<struct(sink)> = r reads r at the position inside the mask's */
/* id list instead of at the id (generic_operations.h:24-33 passes (idx, i) */
/* to the lambda of operations.h:62-64), so the sum taken is                */
/* r[0] + ... + r[nsinks-1].  sink_quirk = 0 is the PageRank definition.    */
int orc_pagerank_f64_q(int64_t n, const int64_t* rp, const int32_t* col, int itermax, double damping, double tol,
                       double* r, int sink_quirk);
int orc_pagerank_f64(int64_t n, const int64_t* rp, const int32_t* col, int itermax, double damping, double tol,
                     double* r) {
    return orc_pagerank_f64_q(n, rp, col, itermax, damping, tol, r, 0);
}
int orc_pagerank_f64_q(int64_t n, const int64_t* rp, const int32_t* col, int itermax, double damping, double tol,
                       double* r, int sink_quirk) {
    int64_t nnz = rp[n];
    int64_t* cdeg = (int64_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
    for (int64_t j = 0; j < nnz; ++j) cdeg[col[j]]++;
    double* t = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) r[i] = 1.0 / (double)n;
    int it;
    for (it = 0; it < itermax; ++it) {
        double teleport = (1.0 - damping) / (double)n, sinks = 0.0;
        int64_t nsinks = 0;
        for (int64_t i = 0; i < n; ++i)
            if (cdeg[i] == 0) { sinks += r[i]; ++nsinks; }
        if (sink_quirk && nsinks <= (int64_t)((double)n * 0.3)) {
            sinks = 0.0;
            for (int64_t k = 0; k < nsinks; ++k) sinks += r[k];
        }
        teleport += damping / (double)n * sinks;
        for (int64_t i = 0; i < n; ++i) {
            double d = (double)cdeg[i] / damping;
            if (d < 1.0 / damping) d = 1.0 / damping;
            t[i] = r[i];
            w[i] = r[i] / d;
        }
        double diff = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double s = teleport;
            for (int64_t j = rp[i]; j < rp[i + 1]; ++j) s += w[col[j]];
            r[i] = s;
            diff += fabs(t[i] - s);
        }
        if (tol > 0 && diff < tol) {
            ++it;
            break;
        }
    }
    free(cdeg);
    free(t);
    free(w);
    return it;
}

/* Personalized PageRank in FP64 on an explicit transition matrix:
 *   r' = d * A r + ((1 - d) + d * sink(r)) * p,   sink(r) = sum of r over the
 *   vertices whose column of A is empty,
 * starting from r = p, until the L1 change is below tol (tol > 0) or after
 * itermax iterations; returns the iterations run.  With p = 1/n and
 * A = 1/coldeg this is LAGraph_page_rank_sinks (algorithms/pr/pr.hpp:20-153,
 * the same recurrence as orc_pagerank_f64); with p = e_s it is the
 * personalized form the batched driver computes. */
int orc_ppr_f64(int64_t n, const int64_t* rp, const int32_t* col, const float* val, const double* p, int itermax,
                double damping, double tol, double* r) {
    int64_t nnz = rp[n];
    char* has = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
    for (int64_t j = 0; j < nnz; ++j) has[col[j]] = 1;
    double* t = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) r[i] = p[i];
    int it;
    for (it = 0; it < itermax; ++it) {
        double sink = 0.0;
        for (int64_t i = 0; i < n; ++i)
            if (!has[i]) sink += r[i];
        const double tele = (1.0 - damping) + damping * sink;
        double diff = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t j = rp[i]; j < rp[i + 1]; ++j) s += (double)val[j] * r[col[j]];
            t[i] = damping * s + tele * p[i];
        }
        for (int64_t i = 0; i < n; ++i) {
            diff += fabs(t[i] - r[i]);
            r[i] = t[i];
        }
        if (tol > 0 && diff < tol) {
            ++it;
            break;
        }
    }
    free(has);
    free(t);
    return it;
}

int orc_version(void) { return 1; }

/* ------------------------------------------------------------------------ */
/* Masked SpGEMM, SpMSpM_masked_ikj                                         */
/* (src/backend/spmspm/spmspm_masked_hash_based_gustavson.h:9-170): per row  */
/* i, the mask row's columns index a dense marker; for k in A(i,:) and j in */
/* B(k,:) with j in the mask row, the product is folded into that slot;     */
/* slots without a product hold 0 (the reference zero-fills its values).    */
/* The reference stores a row's sums in hash-map order, so they can land on */
/* other columns; here each sum stays at its own column.  A column repeated */
/* in a mask row is matched at its first occurrence.                        */
/* ------------------------------------------------------------------------ */
#define ORC_MXM(SUF, T)                                                                                     \
    static void mxm_##SUF(int64_t n, const int64_t* mrp, const int32_t* mcol, const int64_t* arp,          \
                          const int32_t* acol, const T* aval, const int64_t* brp, const int32_t* bcol,      \
                          const T* bval, int add, int mul, T* out) {                                        \
        int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));                        \
        char* hit = (char*)calloc((size_t)(mrp[n] > 0 ? mrp[n] : 1), 1);                                   \
        for (int64_t j = 0; j < n; ++j) pos[j] = -1;                                                        \
        for (int64_t i = 0; i < n; ++i) {                                                                   \
            for (int64_t p = mrp[i]; p < mrp[i + 1]; ++p)                                                   \
                if (pos[mcol[p]] < 0) pos[mcol[p]] = p;                                                     \
            for (int64_t e = arp[i]; e < arp[i + 1]; ++e) {                                                 \
                const int64_t k = acol[e];                                                                  \
                for (int64_t q = brp[k]; q < brp[k + 1]; ++q) {                                             \
                    const int64_t p = pos[bcol[q]];                                                         \
                    if (p < 0) continue;                                                                    \
                    if (!hit[p]) {                                                                          \
                        out[p] = add_id_##SUF(add);                                                         \
                        hit[p] = 1;                                                                         \
                    }                                                                                       \
                    out[p] = add_##SUF(add, out[p], mul_##SUF(mul, aval[e], bval[q]));                      \
                }                                                                                           \
            }                                                                                               \
            for (int64_t p = mrp[i]; p < mrp[i + 1]; ++p) {                                                 \
                pos[mcol[p]] = -1;                                                                          \
                if (!hit[p]) out[p] = (T)0;                                                                 \
            }                                                                                               \
        }                                                                                                   \
        free(pos);                                                                                          \
        free(hit);                                                                                          \
    }
ORC_MXM(f32, float)
ORC_MXM(f64, double)
ORC_MXM(i32, int32_t)
ORC_MXM(i64, int64_t)
ORC_MXM(u8, uint8_t)

int orc_mxm_masked(int type, int64_t n, const int64_t* mrp, const int32_t* mcol, const int64_t* arp,
                   const int32_t* acol, const void* aval, const int64_t* brp, const int32_t* bcol, const void* bval,
                   int add, int mul, void* out) {
    switch (type) {
        case T_F32: mxm_f32(n, mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul, out); return 0;
        case T_F64: mxm_f64(n, mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul, out); return 0;
        case T_I32: mxm_i32(n, mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul, out); return 0;
        case T_I64: mxm_i64(n, mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul, out); return 0;
        case T_BOOL: mxm_u8(n, mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul, out); return 0;
    }
    return -1;
}

/* C = A (+.*) B without a mask: SpMSpM_unmasked_ikj                         */
/* (src/backend/spmspm/spmspm_hash_based_gustavson.h:9-153).  Row i folds,   */
/* for every entry A(i,k) in row order and every B(k,j) in row order,        */
/* C(i,j) = add(C(i,j), mul(A(i,k), B(k,j))) from the identity; entries whose */
/* value is 0 are dropped.  Columns are written ascending (the reference     */
/* writes them in hash-map order).  Returns nnz(C); writes at most cap       */
/* entries.                                                                  */
static int cmp_i32(const void* a, const void* b) {
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}
#define ORC_MXMU(SUF, T)                                                                                    \
    static int64_t mxmu_##SUF(int64_t n, const int64_t* arp, const int32_t* acol, const T* aval,           \
                              const int64_t* brp, const int32_t* bcol, const T* bval, int add, int mul,     \
                              int64_t* crp, int32_t* ccol, T* cval, int64_t cap) {                          \
        T* acc = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                                           \
        char* seen = (char*)calloc((size_t)(n > 0 ? n : 1), 1);                                             \
        int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));                     \
        int64_t nz = 0;                                                                                     \
        crp[0] = 0;                                                                                         \
        for (int64_t i = 0; i < n; ++i) {                                                                   \
            int64_t nt = 0;                                                                                 \
            for (int64_t e = arp[i]; e < arp[i + 1]; ++e) {                                                 \
                const int64_t k = acol[e];                                                                  \
                for (int64_t q = brp[k]; q < brp[k + 1]; ++q) {                                             \
                    const int32_t j = bcol[q];                                                              \
                    if (!seen[j]) {                                                                         \
                        seen[j] = 1;                                                                        \
                        acc[j] = add_id_##SUF(add);                                                         \
                        touched[nt++] = j;                                                                  \
                    }                                                                                       \
                    acc[j] = add_##SUF(add, acc[j], mul_##SUF(mul, aval[e], bval[q]));                      \
                }                                                                                           \
            }                                                                                               \
            qsort(touched, (size_t)nt, sizeof(int32_t), cmp_i32);                                           \
            for (int64_t t = 0; t < nt; ++t) {                                                              \
                const int32_t j = touched[t];                                                               \
                seen[j] = 0;                                                                                \
                if (acc[j] != (T)0) {                                                                       \
                    if (nz < cap) {                                                                         \
                        ccol[nz] = j;                                                                       \
                        cval[nz] = acc[j];                                                                  \
                    }                                                                                       \
                    ++nz;                                                                                   \
                }                                                                                           \
            }                                                                                               \
            crp[i + 1] = nz;                                                                                \
        }                                                                                                   \
        free(acc);                                                                                          \
        free(seen);                                                                                         \
        free(touched);                                                                                      \
        return nz;                                                                                          \
    }
ORC_MXMU(f32, float)
ORC_MXMU(f64, double)
ORC_MXMU(i32, int32_t)
ORC_MXMU(i64, int64_t)
ORC_MXMU(u8, uint8_t)

int64_t orc_mxm_unmasked(int type, int64_t n, const int64_t* arp, const int32_t* acol, const void* aval,
                         const int64_t* brp, const int32_t* bcol, const void* bval, int add, int mul, int64_t* crp,
                         int32_t* ccol, void* cval, int64_t cap) {
    switch (type) {
        case T_F32: return mxmu_f32(n, arp, acol, aval, brp, bcol, bval, add, mul, crp, ccol, cval, cap);
        case T_F64: return mxmu_f64(n, arp, acol, aval, brp, bcol, bval, add, mul, crp, ccol, cval, cap);
        case T_I32: return mxmu_i32(n, arp, acol, aval, brp, bcol, bval, add, mul, crp, ccol, cval, cap);
        case T_I64: return mxmu_i64(n, arp, acol, aval, brp, bcol, bval, add, mul, crp, ccol, cval, cap);
        case T_BOOL: return mxmu_u8(n, arp, acol, aval, brp, bcol, bval, add, mul, crp, ccol, cval, cap);
    }
    return -1;
}

/* Sum over the entries of C<M> = A (plus.one) B (the triangle-count kernel  */
/* of LAGr_TriangleCount, algorithms/tc/tc.hpp:80-291).                     */
int64_t orc_masked_count(int64_t n, const int64_t* mrp, const int32_t* mcol, const int64_t* arp,
                         const int32_t* acol, const int64_t* brp, const int32_t* bcol) {
    char* mark = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t p = mrp[i]; p < mrp[i + 1]; ++p) mark[mcol[p]] = 1;
        for (int64_t e = arp[i]; e < arp[i + 1]; ++e) {
            const int64_t k = acol[e];
            for (int64_t q = brp[k]; q < brp[k + 1]; ++q) total += mark[bcol[q]];
        }
        for (int64_t p = mrp[i]; p < mrp[i + 1]; ++p) mark[mcol[p]] = 0;
    }
    free(mark);
    return total;
}
