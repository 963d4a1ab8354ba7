/*
 * gbk.h — C ABI of the MI355X (gfx950) semiring SpMV / SpMSpV library.
 *
 * Every entry point is a plain C function over opaque handles, plain pointers
 * and sizes; it never throws and returns an LA_Info-compatible code
 * (GrB_SUCCESS = 0, see gbk_info below).  Where the reference library would
 * have thrown a `const char*`, the code is returned and the message is kept
 * for gbk_last_error().
 *
 * Each declaration names the reference interface it replaces (file:line in
 * afanasyev-ilya/GB_Kun).  The C++ `lablas::` front end in include/gbk/lablas.hpp
 * lowers its template calls to these functions.
 *
 * Device memory: matrices and vectors live in HBM for the lifetime of their
 * handles.  All work is issued on the library stream (gbk_set_stream); calls
 * that return host-visible values (nvals, reductions, copies to host)
 * synchronize that stream.
 */
#ifndef GBK_H
#define GBK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Codes, types, operators                                                   */
/* ------------------------------------------------------------------------ */

/* Same numbering as LA_Info, src/cpp_graphblas/types.hpp:56-88. */
typedef enum {
    GBK_SUCCESS = 0,
    GBK_NO_VALUE = 1,
    GBK_UNINITIALIZED_OBJECT = 2,
    GBK_INVALID_OBJECT = 3,
    GBK_NULL_POINTER = 4,
    GBK_INVALID_VALUE = 5,
    GBK_INVALID_INDEX = 6,
    GBK_DOMAIN_MISMATCH = 7,
    GBK_DIMENSION_MISMATCH = 8,
    GBK_OUTPUT_NOT_EMPTY = 9,
    GBK_OUT_OF_MEMORY = 10,
    GBK_INSUFFICIENT_SPACE = 11,
    GBK_INDEX_OUT_OF_BOUNDS = 12,
    GBK_PANIC = 13
} gbk_info;

/* Element types.  BOOL is one byte holding 0 or 1. */
typedef enum {
    GBK_FP32 = 0,
    GBK_FP64 = 1,
    GBK_INT32 = 2,
    GBK_INT64 = 3,
    GBK_BOOL = 4
} gbk_type;

/* Additive monoids, src/cpp_graphblas/types.hpp:226-232 (+ MAX). */
typedef enum {
    GBK_ADD_PLUS = 0,      /* PlusMonoid, identity 0                        */
    GBK_ADD_LOR = 1,       /* LogicalOrMonoid, identity false               */
    GBK_ADD_MIN = 2,       /* MinimumMonoid, identity numeric max           */
    GBK_ADD_MAX = 3,       /* maximum, identity numeric lowest              */
    GBK_ADD_FIRST = 4,     /* FirstWinsMonoid, identity 0                   */
    GBK_ADD_SECOND = 5,    /* SecondWinsMonoid, identity 0                  */
    GBK_ADD_LESS = 6,      /* CustomLessMonoid: (l < r), identity max       */
    GBK_ADD_FIRSTMIN = 7   /* FirstMin: minimum with identity 0             */
} gbk_add_op;

/* Multiplicative operators, src/cpp_graphblas/types.hpp:105-209.
 * mul(a, x) is called with the matrix value first, as the reference kernels
 * do (spmv_csr.h:374, spmspv_atomics.h:48). */
typedef enum {
    GBK_MUL_TIMES = 0,
    GBK_MUL_LAND = 1,
    GBK_MUL_FIRST = 2,
    GBK_MUL_SECOND = 3,
    GBK_MUL_PLUS = 4,      /* saturating for integer types                   */
    GBK_MUL_ONEB = 5,      /* GrB_ONEB_T: always 1                           */
    GBK_MUL_NE = 6,        /* not_equal_to                                   */
    GBK_MUL_MINUS = 7,
    GBK_MUL_DIV = 8
} gbk_mul_op;

/* Accumulators / element-wise binary operators.  GBK_ACC_NONE is the
 * reference's NULL accum, which lowers to second<U,W,U>
 * (src/cpp_graphblas/operations.hpp:373,404). */
typedef enum {
    GBK_ACC_NONE = 0,      /* second: w = t                                  */
    GBK_ACC_FIRST = 1,
    GBK_ACC_PLUS = 2,
    GBK_ACC_MINUS = 3,
    GBK_ACC_TIMES = 4,
    GBK_ACC_DIV = 5,
    GBK_ACC_MIN = 6,
    GBK_ACC_MAX = 7,
    GBK_ACC_LOR = 8,
    GBK_ACC_LAND = 9,
    GBK_ACC_LESS = 10,
    GBK_ACC_NE = 11
} gbk_binop;

/* Semirings registered by the reference (types.hpp:274-284).  A semiring is a
 * (monoid, multiplier) pair; these are the pairs the kernels are built for. */
typedef struct { int add; int mul; } gbk_semiring;
#define GBK_PLUS_TIMES        ((gbk_semiring){GBK_ADD_PLUS, GBK_MUL_TIMES})
#define GBK_LOR_LAND          ((gbk_semiring){GBK_ADD_LOR, GBK_MUL_LAND})
#define GBK_FIRST_TIMES       ((gbk_semiring){GBK_ADD_FIRST, GBK_MUL_TIMES})
#define GBK_FIRSTMIN_TIMES    ((gbk_semiring){GBK_ADD_FIRSTMIN, GBK_MUL_TIMES})
#define GBK_PLUS_SECOND       ((gbk_semiring){GBK_ADD_PLUS, GBK_MUL_SECOND})
#define GBK_LOR_ONEB          ((gbk_semiring){GBK_ADD_LOR, GBK_MUL_ONEB})
#define GBK_MIN_PLUS          ((gbk_semiring){GBK_ADD_MIN, GBK_MUL_PLUS})
#define GBK_LESS_PLUS         ((gbk_semiring){GBK_ADD_LESS, GBK_MUL_PLUS})
#define GBK_MIN_SECOND        ((gbk_semiring){GBK_ADD_MIN, GBK_MUL_SECOND})
#define GBK_MIN_NE            ((gbk_semiring){GBK_ADD_MIN, GBK_MUL_NE})
#define GBK_PLUS_ONEB         ((gbk_semiring){GBK_ADD_PLUS, GBK_MUL_ONEB})

/* Descriptor values, numbered as Desc_value in src/cpp_graphblas/types.hpp:28-53. */
typedef enum {
    GBK_SCMP = 0,
    GBK_REPLACE = 1,
    GBK_COMP = 2,
    GBK_TRAN = 3,
    GBK_STRUCTURE = 4,
    GBK_STR_COMP = 5,
    GBK_DEFAULT = 6
} gbk_desc_value;

/* The descriptor fields the path reads (backend/descriptor/descriptor.h:10-54).
 * A NULL gbk_desc* means every field is GBK_DEFAULT. */
typedef struct {
    int mask;    /* GrB_MASK:   DEFAULT, STRUCTURE, COMP, STR_COMP, SCMP       */
    int output;  /* GrB_OUTPUT: DEFAULT or REPLACE                            */
} gbk_desc;

/* Matrix storage formats, numbered as MatrixStorageFormat in
 * src/helpers/cmd_parser/parser_options.h:15-24. */
typedef enum {
    GBK_FMT_CSR = 0,
    GBK_FMT_CSR_SEG = 3,
    GBK_FMT_LAV = 4,
    GBK_FMT_SELL_C = 6
} gbk_format;

/* Vector storage, numbered as Storage in src/cpp_graphblas/types.hpp:10-12. */
typedef enum { GBK_STORAGE_UNKNOWN = 0, GBK_STORAGE_SPARSE = 1, GBK_STORAGE_DENSE = 2 } gbk_storage;

typedef struct gbk_matrix_s* gbk_matrix;
typedef struct gbk_vector_s* gbk_vector;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

/* Selects the device and creates the library stream.  Called implicitly by
 * the first entry point that touches the GPU. */
int gbk_init(int device);
/* Uses `stream` (a hipStream_t) for all later work; NULL restores the
 * library's own stream. */
int gbk_set_stream(void* stream);
void* gbk_get_stream(void);
int gbk_synchronize(void);
/* Message of the last failing call on this thread ("" if none). */
const char* gbk_last_error(void);
/* Number of compute units and device name of the current device. */
int gbk_device_info(int* cu_count, char* name, int name_len);

/* ------------------------------------------------------------------------ */
/* Matrix — replaces lablas::Matrix<T> (src/cpp_graphblas/matrix.hpp:11-144)
 * and backend::Matrix<T> (src/backend/matrix/matrix.h:34-207).              */
/* ------------------------------------------------------------------------ */

int gbk_matrix_new(gbk_matrix* A, gbk_type type);
int gbk_matrix_free(gbk_matrix A);

/* Matrix::set_preferred_matrix_format, matrix.hpp:34-36.  As in the
 * reference it selects the optimized representation built by the next
 * build call; called on a built matrix it (re)builds that representation.
 * sell_c (32 or 64) and sigma apply to GBK_FMT_SELL_C and are 0 for the
 * defaults (C = 64, sigma = 512). */
int gbk_matrix_set_format(gbk_matrix A, int format, int sell_c, int sigma);
int gbk_matrix_get_format(gbk_matrix A, int* format);

/* Matrix::build from a COO edge list, matrix.hpp:69-91 / build.hpp:123-166:
 * n = max id + 1 (the larger of rows and columns); duplicates kept in list
 * order; CSR and CSC both built; row order inside a row = list order. */
int gbk_matrix_build_coo(gbk_matrix A, const int64_t* rows, const int64_t* cols,
                         const void* vals, int64_t nvals);

/* Matrix::build_from_csr_arrays, matrix.hpp:93-105 / build.hpp:171-210.
 * row_ptr has nrows+1 entries; the matrix is square of size
 * max(nrows, max col + 1). */
int gbk_matrix_build_csr(gbk_matrix A, const int64_t* row_ptr, const int64_t* col_ids,
                         const void* vals, int64_t nrows, int64_t nnz);
/* Same, with 32-bit column ids (avoids an 8-byte copy of large graphs). */
int gbk_matrix_build_csr32(gbk_matrix A, const int64_t* row_ptr, const int32_t* col_ids,
                           const void* vals, int64_t nrows, int64_t nnz);

/* Matrix::init_from_mtx, matrix.hpp:107-110 / read_file.hpp:556-615: reads
 * a Matrix Market (.mtx) or binary (.mtxbin) edge list (by suffix) and
 * builds CSR and CSC; every entry gets the value 1 (the reference's
 * EDGE_VAL); row order inside a row = file order; the matrix is square of
 * size max(rows, cols) of the size line. */
int gbk_matrix_read_mtx(gbk_matrix A, const char* path);

int gbk_matrix_nrows(gbk_matrix A, int64_t* n);
int gbk_matrix_ncols(gbk_matrix A, int64_t* n);
int gbk_matrix_nvals(gbk_matrix A, int64_t* nnz);
int gbk_matrix_type(gbk_matrix A, int* type);
/* Host copies of the device CSR (transposed = 0) or CSC (transposed = 1):
 * row_ptr[n+1], col_ids[nnz], vals[nnz]; any pointer may be NULL. */
int gbk_matrix_export(gbk_matrix A, int transposed, int64_t* row_ptr, int32_t* col_ids, void* vals);
/* Row / column degrees (Matrix::get_rowdegrees / get_coldegrees, matrix.hpp:61-65). */
int gbk_matrix_degrees(gbk_matrix A, int transposed, int64_t* degrees);
/* Device pointers of the CSR (transposed = 0) or CSC: int32 row_ptr[n+1],
 * int32 col[nnz], vals[nnz]. */
int gbk_matrix_device_csr(gbk_matrix A, int transposed, const int32_t** row_ptr,
                          const int32_t** col_ids, const void** vals);
/* Bytes of HBM held by the matrix (all formats and workspaces). */
int gbk_matrix_device_bytes(gbk_matrix A, int64_t* bytes);
/* Format autotuner (no reference counterpart: the reference takes the format
 * from the command line, cmd_parser.hpp:21-112): builds CSR, segmented CSR,
 * LAV and SELL-C-sigma (C = 64 and 32, sigma = 512) in turn, times `trials`
 * SpMVs of each on the device, and keeps the fastest as the matrix's format.
 * times_us (may be NULL) receives the five average times in that order. */
int gbk_matrix_autotune(gbk_matrix A, int trials, int* chosen_format, int* chosen_c, double* times_us);
/* Statistics of the optimized representation of one orientation (the
 * reference's MatrixSellC::print_stats / segment counts): stats[0] padded
 * SELL-C entries, [1] rows of problematic SELL-C chunks, [2] segments,
 * [3] (segment, row) pairs. */
int gbk_matrix_format_stats(gbk_matrix A, int transposed, int64_t* stats);

/* ------------------------------------------------------------------------ */
/* Vector — replaces lablas::Vector<T> (src/cpp_graphblas/vector.hpp:11-123)
 * and backend::Vector<T> with its sparse/dense switch
 * (src/backend/vector/vector.h:22-276).                                     */
/* ------------------------------------------------------------------------ */

int gbk_vector_new(gbk_vector* v, gbk_type type, int64_t n);   /* starts sparse, empty */
int gbk_vector_free(gbk_vector v);
int gbk_vector_size(gbk_vector v, int64_t* n);
int gbk_vector_type(gbk_vector v, int* type);
/* Vector::get_nvals: stored entries when sparse, nonzeros when dense
 * (dense_vector.hpp:44-52). */
int gbk_vector_nvals(gbk_vector v, int64_t* nvals);
int gbk_vector_storage(gbk_vector v, int* storage);
/* Vector::build(values, n) — dense (vector.h:180-186). */
int gbk_vector_build_dense(gbk_vector v, const void* vals, int64_t n);
/* Vector::build(indices, values, nvals) — sparse (vector.h:166-178). */
int gbk_vector_build_sparse(gbk_vector v, const int64_t* ids, const void* vals, int64_t nvals);
/* Vector::fill / set_constant (vector.h:37-55): 0 makes an empty sparse vector. */
int gbk_vector_fill(gbk_vector v, double value);
/* fill with a 64-bit integer value (exact for INT64) */
int gbk_vector_fill_int(gbk_vector v, int64_t value);
/* Vector::set_element (vector.h:141-144). */
int gbk_vector_set_element(gbk_vector v, double value, int64_t pos);
/* Vector::fillAscending (vector.h:215-218): v[i] = i for i < nvals. */
int gbk_vector_fill_ascending(gbk_vector v, int64_t nvals);
/* Vector::dup (vector.h:220-227). */
int gbk_vector_dup(gbk_vector dst, gbk_vector src);
/* Vector::swap (vector.h:229-233). */
int gbk_vector_swap(gbk_vector a, gbk_vector b);
/* Vector::clear (vector.h:235-240). */
int gbk_vector_clear(gbk_vector v);
/* force_to_dense / force_to_sparse / convert_if_required (vector.h:81-129). */
int gbk_vector_force_dense(gbk_vector v);
int gbk_vector_force_sparse(gbk_vector v);
int gbk_vector_convert_if_required(gbk_vector v);
/* Host copy of the dense form (forces dense). */
int gbk_vector_to_host(gbk_vector v, void* vals);
/* Host copy of the sparse form (forces sparse): ids and vals of nvals entries. */
int gbk_vector_sparse_to_host(gbk_vector v, int64_t* ids, void* vals, int64_t* nvals);
/* Device pointer of the dense form (forces dense).  Valid until the next
 * call that changes the vector's storage. */
int gbk_vector_device_dense(gbk_vector v, void** vals);

/* ------------------------------------------------------------------------ */
/* Operations — replace lablas::mxv / vxm (src/cpp_graphblas/operations.hpp:347-405),
 * backend::mxv / vxm (src/backend/operations/operations.h:151-211)        */
/* ------------------------------------------------------------------------ */

/* w = accum(w, A * u) under the mask.  mask may be NULL.  Dense u runs the
 * pull SpMV in the matrix's format (SpMV, spmv.h:20-79); sparse u runs the
 * push SpMSpV over the CSC (SpMSpV, spmspv.h:103-158).  Afterwards w is
 * made sparse or dense by the 30% rule (vector.h:117-129). */
int gbk_mxv(gbk_vector w, gbk_vector mask, int accum, gbk_semiring op,
            gbk_matrix A, gbk_vector u, const gbk_desc* desc);
/* w = accum(w, u * A): the same over the CSC / CSR (VSpM, spmv.h:83-146). */
int gbk_vxm(gbk_vector w, gbk_vector mask, int accum, gbk_semiring op,
            gbk_vector u, gbk_matrix A, const gbk_desc* desc);

/* Backend seam on raw device pointers, no vector state machine:
 * y = accum(y, A * x) (transposed = 0, CSR) or accum(y, x * A) (transposed = 1,
 * CSC), unmasked, in the matrix's current format.  Replaces backend::SpMV /
 * VSpM all-active paths (spmv.h:29-67, 92-131) and SpMV_all_active_diff_vectors
 * (spmv_csr.h:341-384).  x and y are device arrays of the matrix type and
 * must not alias (x == y runs the same-vectors path through a workspace,
 * spmv_csr.h:433-486). */
int gbk_spmv_device(gbk_matrix A, int transposed, const void* x, void* y,
                    gbk_semiring op, int accum);

/* Element-wise operations used by the drivers
 * (src/backend/operations/operations.h:15-147, 216-430). */
/* w<mask> = accum(w, value) for all i (assign with GrB_ALL, operations.h:83-112). */
int gbk_assign_scalar(gbk_vector w, gbk_vector mask, int accum, double value, const gbk_desc* desc);
/* w<mask> = accum(w, u) (operations.h:15-79). */
int gbk_assign_vector(gbk_vector w, gbk_vector mask, int accum, gbk_vector u, const gbk_desc* desc);
/* w[idx[i]] = u[i] for i < nindices (assignScatter, operations.h:116-147). */
int gbk_assign_scatter(gbk_vector w, gbk_vector mask, int accum, gbk_vector u,
                       gbk_vector indices, int64_t nindices, const gbk_desc* desc);
/* w<mask> = accum(w, op(u, v)), union (eWiseAdd, operations.h:216-240). `op`
 * is a gbk_binop; for a semiring the caller passes its monoid's binop. */
int gbk_ewise_add(gbk_vector w, gbk_vector mask, int accum, int op,
                  gbk_vector u, gbk_vector v, const gbk_desc* desc);
/* w<mask> = accum(w, op(u, v)), intersection (eWiseMult, operations.h:244-267). */
int gbk_ewise_mult(gbk_vector w, gbk_vector mask, int accum, int op,
                   gbk_vector u, gbk_vector v, const gbk_desc* desc);
/* w<mask> = accum(w, f(u)) with f in {IDENTITY, ABS, DIV-by-scalar, ...}
 * (apply, operations.h:271-341).  unary: 0 identity, 1 abs, 2 u / scalar,
 * 3 u * scalar, 4 ainv (-u). */
int gbk_apply(gbk_vector w, gbk_vector mask, int accum, int unary, double scalar,
              gbk_vector u, const gbk_desc* desc);
/* *result = accum(*result, reduce_monoid(u)) (reduce, operations.h:345-396).
 * result is a host double (exact for FP32/INT32; INT64 through result_i64). */
int gbk_reduce(double* result, int accum, int monoid, gbk_vector u, const gbk_desc* desc);
int gbk_reduce_i64(int64_t* result, int accum, int monoid, gbk_vector u, const gbk_desc* desc);
/* w[i] = u[idx[i]] (extract, operations.h:400-430). */
int gbk_extract(gbk_vector w, gbk_vector mask, int accum, gbk_vector u,
                gbk_vector indices, const gbk_desc* desc);

/* ------------------------------------------------------------------------ */
/* Graph drivers — replace algorithms/{bfs,sssp,cc,pr}                       */
/* ------------------------------------------------------------------------ */

/* Direction-optimizing BFS, replaces lablas::algorithm::bfs_blast
 * (algorithms/bfs/bfs.hpp:94-142).  levels (host, n int32): 1-based level,
 * 0 = unreached.  A is read as in the reference: the frontier advances
 * along rows of A (vxm).  stats (may be NULL): [0] levels, [1] push steps,
 * [2] pull steps, [3] edges traversed. */
int gbk_bfs(gbk_matrix A, int64_t source, int32_t* levels, int64_t* stats);
/* Same, writing the int32 levels to a device array. */
int gbk_bfs_device(gbk_matrix A, int64_t source, int32_t* levels_dev, int64_t* stats);

/* SSSP over min.plus, replaces sssp_bellman_ford_blast
 * (algorithms/sssp/sssp_blast.hpp:8-72).  INT32 weights: distances int32,
 * unreachable = INT32_MAX; FP32 weights: distances float, unreachable =
 * FLT_MAX.  dist is a host array of the matrix type. */
int gbk_sssp(gbk_matrix A, int64_t source, void* dist, int64_t* stats);
/* Same, writing the distances (n entries of the matrix type) to a device array. */
int gbk_sssp_device(gbk_matrix A, int64_t source, void* dist_dev, int64_t* stats);

/* C<M> = A (+.*) B, replaces lablas::mxm (src/cpp_graphblas/operations.hpp:
 * 411-437; backend::mxm, operations.h:436-486).  With a structural mask M
 * (masked IKJ: spmspm_masked_hash_based_gustavson.h:9-170) C gets exactly the
 * entries of M; C(i,j) = fold over k of A(i,k) (*) B(k,j), 0 where no k
 * contributes.  With M == NULL (unmasked IKJ: spmspm_hash_based_gustavson.h:
 * 9-153) C holds every (i,j) some k reaches whose fold is not 0, the fold in
 * IKJ order from the identity, columns ascending.  A, B (and M) are n x n; A
 * and B of one type, which C takes.  accum is ignored (the reference rebuilds
 * C); a complemented mask is an error. */
int gbk_mxm(gbk_matrix C, gbk_matrix M, int accum, gbk_semiring op, gbk_matrix A, gbk_matrix B,
            const gbk_desc* desc);

/* Triangle count of a symmetric matrix, replaces LAGr_TriangleCount
 * (algorithms/tc/tc.hpp:80-291).  method: 0 default (Sandia over the
 * orientation by (degree, id); the reference maps 0 to 1), 1 Burkhardt
 * sum((A^2) .* A) / 6, 2 Cohen sum((L U) .* A) / 2, 3 Sandia sum((L L) .* L),
 * 4 Sandia2 sum((U U) .* U), 5 SandiaDot sum((L U') .* L), 6 SandiaDot2
 * sum((U L') .* U), L = tril(A, -1), U = triu(A, 1).  All give the same count
 * (the reference divides the Cohen sum by 6). */
int gbk_tricount(gbk_matrix A, int method, uint64_t* ntriangles);

/* Connected components, replaces lablas::algorithm::cc (FastSV,
 * algorithms/cc/cc.hpp:12-119).  labels (host int32): the minimum vertex id
 * of each component, the fixed point FastSV converges to. A must be
 * symmetric. */
int gbk_cc(gbk_matrix A, int32_t* labels, int64_t* stats);

/* PageRank with sinks, replaces LAGraph_page_rank_sinks
 * (algorithms/pr/pr.hpp:20-153): r += A * (t ./ d) with the plus.second
 * semiring, d = max(coldegree / damping, 1 / damping).  Stops after itermax
 * iterations or, if tol > 0, when the L1 change drops below tol.
 * ranks (host float[n]).  iters receives the iterations run. */
int gbk_pagerank(gbk_matrix A, float* ranks, int itermax, double damping, double tol, int* iters);

/* The same with the ranks left in a device array of n floats (not the
 * matrix's own workspace). */
int gbk_pagerank_device(gbk_matrix A, float* ranks_dev, int itermax, double damping, double tol, int* iters);

/* ------------------------------------------------------------------------ */
/* Batched (multi-vector) operations.  The reference has no batched form: each
 * is the k-fold of one reference call, computed in one pass over the matrix. */
/* ------------------------------------------------------------------------ */

/* Y = accum(Y, A * X) for k right-hand sides at once (1 <= k <= 64): X and Y
 * are device arrays of n*k elements of the matrix type, row-major (the k
 * values of vertex v at v*k .. v*k+k-1).  Column q of Y equals
 * gbk_spmv_device on column q of X (SpMV_all_active_diff_vectors,
 * spmv_csr.h:341-384, k times).  The monoid must be commutative. */
int gbk_spmm_device(gbk_matrix A, int transposed, const void* X, void* Y, int k,
                    gbk_semiring op, int accum);

/* Multi-source BFS: nsrc <= 64 searches in one traversal (bfs_blast,
 * algorithms/bfs/bfs.hpp:94-142, once per source).  levels: nsrc*n int32,
 * source-major, 1-based, 0 = unreached.  stats (may be NULL, >= 5 entries):
 * [0] levels, [1] push steps, [2] pull steps, [3] frontier edges,
 * [4] (vertex, source) pairs reached. */
int gbk_msbfs(gbk_matrix A, const int64_t* sources, int nsrc, int32_t* levels, int64_t* stats);
/* Same, writing the levels to a device array of nsrc*n int32. */
int gbk_msbfs_device(gbk_matrix A, const int64_t* sources, int nsrc, int32_t* levels_dev, int64_t* stats);

/* Personalized PageRank for k <= 64 sources at once on an FP32 transition
 * matrix (entry (v,u) = 1/outdeg(u)): r_q = d A r_q + ((1-d) + d sink_q) e_q,
 * the per-source form of LAGraph_page_rank_sinks (algorithms/pr/pr.hpp:20-153).
 * Column q stops when its L1 change is below tol (tol > 0) or after itermax
 * iterations.  ranks: k*n floats (host), source-major; iters[q]: iterations
 * of column q. */
int gbk_ppr_multi(gbk_matrix A, const int64_t* sources, int k, float* ranks, int itermax,
                  double damping, double tol, int* iters);
/* Same, writing the ranks (k*n floats, source-major) to a device array. */
int gbk_ppr_multi_device(gbk_matrix A, const int64_t* sources, int k, float* ranks_dev, int itermax,
                         double damping, double tol, int* iters);

/* min.plus SSSP for k <= 64 sources at once on an INT32 weight matrix
 * (sssp_bellman_ford_blast, algorithms/sssp/sssp_blast.hpp:8-72, once per
 * source).  dist: k*n int32 (host), source-major, INT32_MAX = unreachable.
 * stats (may be NULL): [0] rounds, [1] edge relaxations. */
int gbk_sssp_multi(gbk_matrix A, const int64_t* sources, int k, int32_t* dist, int64_t* stats);

/* Many small graphs in one launch: BFS from one source in each of up to
 * thousands of independent graphs (bfs_blast, algorithms/bfs/bfs.hpp:94-142,
 * once per graph) by one persistent kernel captured in a HIP graph and
 * replayed on every call.  Graphs hold 1..131072 vertices each. */
typedef struct gbk_graph_batch_s* gbk_graph_batch;
/* row_ptrs[g] (nverts[g]+1 int64) and col_ids[g] (int32 local ids) are the
 * CSR of graph g. */
int gbk_graph_batch_new(gbk_graph_batch* b, int ngraphs, const int64_t* nverts,
                        const int64_t* const* row_ptrs, const int32_t* const* col_ids);
int gbk_graph_batch_free(gbk_graph_batch b);
/* levels (host, may be NULL): sum(nverts) int32, graph g's at the offset
 * sum(nverts[0..g-1]); 1-based, 0 = unreached.  device_ms (may be NULL): time
 * of the graph launch on the device. */
int gbk_graph_batch_bfs(gbk_graph_batch b, const int64_t* sources, int32_t* levels, double* device_ms);
/* info: [0] graphs, [1] vertices, [2] edges, [3] persistent workgroups. */
int gbk_graph_batch_info(gbk_graph_batch b, int64_t* info);

/* ------------------------------------------------------------------------ */
/* Query server: continuous batching of BFS / personalized PageRank / SSSP
 * queries on device-resident graphs.  No reference counterpart (the
 * reference answers one query per driver call); each query's result equals
 * the corresponding single-query call: gbk_bfs levels, gbk_ppr_multi ranks,
 * gbk_sssp distances.                                                       */
/* ------------------------------------------------------------------------ */

typedef enum { GBK_QUERY_BFS = 0, GBK_QUERY_PPR = 1, GBK_QUERY_SSSP = 2 } gbk_query_kind;
typedef struct gbk_server_s* gbk_server;

/* hbm_budget_bytes bounds the HBM the server's engines and in-flight query
 * results may hold (the graphs themselves are not counted).  ppr_slots and
 * sssp_slots (1..64) are the batch widths of those engines (BFS uses 64).
 * keep_results != 0 keeps every result vector for gbk_server_result. */
int gbk_server_new(gbk_server* s, int64_t hbm_budget_bytes, int ppr_slots, int sssp_slots,
                   double ppr_damping, double ppr_tol, int ppr_itermax, int keep_results);
int gbk_server_free(gbk_server s);
/* Registers a graph: its BFS matrix (any type, read as a pattern), its FP32
 * PageRank transition matrix and its INT32 SSSP weight matrix; a NULL matrix
 * means the graph takes no queries of that kind.  All must have one size. */
int gbk_server_add_graph(gbk_server s, gbk_matrix bfs_A, gbk_matrix ppr_A, gbk_matrix sssp_A, int* graph_id);
/* Queues a query that arrives arrival_s seconds after gbk_server_run starts. */
int gbk_server_submit(gbk_server s, int graph_id, int kind, int64_t source, double arrival_s, int64_t* query_id);
/* Serves every queued query (one worker thread and stream per graph and
 * kind), admitting each when it has arrived, its engine has a free slot and
 * the budget holds its footprint; returns when all have completed. */
int gbk_server_run(gbk_server s);
/* times: [0] arrival, [1] admission, [2] completion (seconds from the start
 * of gbk_server_run), [3] checksum (sum of the reached entries); info:
 * [0] reached entries, [1] batch occupancy at admission, [2] slot, [3] status. */
int gbk_server_query_info(gbk_server s, int64_t query_id, double* times, int64_t* info);
/* Copies a kept result: n int32 levels (BFS), n float ranks (PPR) or n int32
 * distances (SSSP). */
int gbk_server_result(gbk_server s, int64_t query_id, void* host_out);
/* stats (6 entries): [0] wall seconds of the last run, [1] queries, [2] peak
 * bytes of the budget in use, [3] engines created, [4] most engines alive at
 * once, [5] largest single engine reservation (workspace + one result). */
int gbk_server_stats(gbk_server s, double* stats);

/* ------------------------------------------------------------------------ */
/* Seeded graph generation (host) — replaces the unseeded
 * GraphGenerationAPI::RMAT / random_uniform
 * (src/helpers/graph_generation/graph_generation.hpp:37-190).
 * The result is a deduplicated CSR without self-loops; row_ptr (n+1 int64)
 * and col_ids (nnz int32) are allocated by the library and released with
 * gbk_host_free.  Columns are ascending inside each row.                    */
/* ------------------------------------------------------------------------ */

int gbk_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 int symmetric, int permute, int64_t* n, int64_t* nnz,
                 int64_t** row_ptr, int32_t** col_ids);
int gbk_gen_uniform(int scale, int edge_factor, uint64_t seed, int symmetric,
                    int64_t* n, int64_t* nnz, int64_t** row_ptr, int32_t** col_ids);
void gbk_host_free(void* p);

/* Edge-list files on the host.  gbk_mtx_read replaces the readers of
 * read_file.hpp (read_mtx_file_pipelined 52-170, binary_read_mtx_file
 * 378-555): .mtx = '%' comment lines, a "rows cols entries" line, then one
 * "row col [value]" line per entry (1-based, values ignored); .mtxbin =
 * int64 rows, cols, entries, then int64 (row, col) pairs, 1-based.  Returns
 * 0-based rows / cols in file order, allocated by the library (release with
 * gbk_host_free).  Ids outside the declared size or a short file are errors.
 * gbk_mtx_write writes either format (by suffix) in the given order, as
 * EdgeListContainer::save_as_mtx (edges_container.hpp:33-88) and
 * benchmark/scripts/mtx_api.py:61-90 lay them out. */
int gbk_mtx_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t** rows, int64_t** cols);
int gbk_mtx_write(const char* path, int64_t nrows, int64_t ncols, const int64_t* rows, const int64_t* cols,
                  int64_t nnz);
/* out[i] = uniform float in [0,1) from (seed, i). */
void gbk_fill_uniform_f32(uint64_t seed, int64_t count, float* out);
/* vals[e] = 1 / (number of entries in column col[e]) — PR transition values. */
void gbk_fill_inv_coldeg_f32(int64_t n, int64_t nnz, const int32_t* col_ids, float* vals);
/* Symmetric integer weights in [lo, hi]: w(u,v) = w(v,u), keyed by (seed, min, max). */
void gbk_fill_sym_weights_i32(int64_t n, const int64_t* row_ptr, const int32_t* col_ids,
                              uint64_t seed, int32_t lo, int32_t hi, int32_t* vals);

#ifdef __cplusplus
}
#endif

#endif /* GBK_H */
