// ref_bridge.cpp — C entry points over the COMPILED reference (GB_Kun headers
// included from /root/reference at build time; nothing of the reference is
// copied here).  Built by oracle/build_ref.sh into oracle/_ref/ and used only
// by tests/golden/make_golden.py (fixture generation) and the oracle's own
// pinning tests, in the container where /root/reference exists.
//
// It feeds the reference the same CSR the GPU path gets, through
// lablas::Matrix::build_from_csr_arrays (src/cpp_graphblas/matrix.hpp:93-105),
// and returns the results of lablas::mxv / vxm and of the algorithm drivers.
//
// Compile-time variant: with -DREF_DISABLE_SPMSPV this TU is the CC bridge
// (cc.cpp defines __DISABLE_SPMSPV__ and __REQUIRES_UNDIRECTED_GRAPHS__).
#ifdef REF_DISABLE_SPMSPV
#define __DISABLE_SPMSPV__
#define __REQUIRES_UNDIRECTED_GRAPHS__
#endif
#include "src/gb_kun.h"

#ifdef REF_DISABLE_SPMSPV
#include "algorithms/cc/cc.hpp"
#else
#include "algorithms/bfs/bfs.hpp"
#include "algorithms/pr/pr.hpp"
#include "algorithms/sssp/sssp_blast.hpp"
#endif

#include <cstdint>

namespace {

template <typename T>
void build_matrix(lablas::Matrix<T>& A, int64_t n, const int64_t* rp, const int64_t* col, const T* vals, int fmt) {
    A.set_preferred_matrix_format((MatrixStorageFormat)fmt);
    A.build_from_csr_arrays((const Index*)rp, (const Index*)col, vals, (Index)n, (Index)rp[n]);
}

template <typename T>
void load_vector(lablas::Vector<T>& v, int kind, int64_t n, const T* dense, const int64_t* ids, const T* svals,
                 int64_t nvals) {
    if (kind == 1) {
        v.build(dense, (Index)n);
    } else if (kind == 2) {
        // SparseVector::build never sets nvals (sparse_vector.h: build writes
        // ids/vals only, after fill_with_zeros), so a vector built that way
        // is empty.  Build the dense form and compact it instead, which is
        // how the reference's own operations produce sparse vectors.
        std::vector<T> d(n, T(0));
        for (int64_t k = 0; k < nvals; ++k) d[ids[k]] = svals[k];
        v.build(d.data(), (Index)n);
        v.get_vector()->force_to_sparse();
    }
}

}  // namespace

#ifndef REF_DISABLE_SPMSPV

// Runs w = accum(w, A u) (vxm = 0) or accum(w, u A) (vxm = 1) through
// lablas::mxv / vxm with a registered semiring.
//   semiring: 0 PlusMultiplies, 1 LogicalOrAnd, 2 MinimumPlus, 3 MinimumSelectSecond,
//             4 PlusSecond, 5 StructuralBool
//   accum:    0 NULL (second), 1 plus, 2 minimum
//   u_kind / w_kind / mask_kind: 0 absent (mask) / empty, 1 dense, 2 sparse
// Outputs w (dense view) and its storage after the call (1 sparse, 2 dense).
template <typename T>
static int run_mxv(int vxm, int64_t n, const int64_t* rp, const int64_t* col, const T* vals, int fmt, int semiring,
                   int accum, int u_kind, const T* u_dense, const int64_t* u_ids, const T* u_vals, int64_t u_nvals,
                   int w_kind, const T* w_dense, const int64_t* w_ids, const T* w_vals, int64_t w_nvals, int mask_kind,
                   const T* m_dense, const int64_t* m_ids, const T* m_vals, int64_t m_nvals, int desc_mask, T* w_out,
                   int* w_storage) {
    lablas::Matrix<T> A;
    build_matrix(A, n, rp, col, vals, fmt);
    lablas::Vector<T> u(n), w(n), m(n);
    load_vector(u, u_kind, n, u_dense, u_ids, u_vals, u_nvals);
    load_vector(w, w_kind, n, w_dense, w_ids, w_vals, w_nvals);
    load_vector(m, mask_kind, n, m_dense, m_ids, m_vals, m_nvals);
    const lablas::Vector<T>* mask = mask_kind ? &m : (const lablas::Vector<T>*)NULL;
    lablas::Descriptor desc;
    if (desc_mask != GrB_DEFAULT) desc.set(GrB_MASK, (Desc_value)desc_mask);
    LA_Info info = GrB_SUCCESS;
    auto go = [&](auto op) {
        auto call = [&](auto acc) {
            if (vxm) info = lablas::vxm(&w, mask, acc, op, &u, &A, &desc);
            else info = lablas::mxv(&w, mask, acc, op, &A, &u, &desc);
        };
        if (accum == 0) call(GrB_NULL);
        else if (accum == 1) call(lablas::plus<T>());
        else call(lablas::minimum<T>());
    };
    switch (semiring) {
        case 0: go(lablas::PlusMultipliesSemiring<T>()); break;
        case 1: go(lablas::LogicalOrAndSemiring<T>()); break;
        case 2: go(lablas::MinimumPlusSemiring<T>()); break;
        case 3: go(lablas::MinimumSelectSecondSemiring<T>()); break;
        case 4: go(lablas::PlusSecondSemiring<T>()); break;
        default: go(lablas::StructuralBool<T>()); break;
    }
    *w_storage = w.get_vector()->is_dense() ? 2 : 1;
    const T* wv = w.get_vector()->getDense()->get_vals();
    for (int64_t i = 0; i < n; ++i) w_out[i] = wv[i];
    return (int)info;
}

extern "C" int ref_mxv_f32(int vxm, int64_t n, const int64_t* rp, const int64_t* col, const float* vals, int fmt,
                           int semiring, int accum, int u_kind, const float* u_dense, const int64_t* u_ids,
                           const float* u_vals, int64_t u_nvals, int w_kind, const float* w_dense,
                           const int64_t* w_ids, const float* w_vals, int64_t w_nvals, int mask_kind,
                           const float* m_dense, const int64_t* m_ids, const float* m_vals, int64_t m_nvals,
                           int desc_mask, float* w_out, int* w_storage) {
    try {
        return run_mxv<float>(vxm, n, rp, col, vals, fmt, semiring, accum, u_kind, u_dense, u_ids, u_vals, u_nvals,
                              w_kind, w_dense, w_ids, w_vals, w_nvals, mask_kind, m_dense, m_ids, m_vals, m_nvals,
                              desc_mask, w_out, w_storage);
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

extern "C" int ref_mxv_i32(int vxm, int64_t n, const int64_t* rp, const int64_t* col, const int* vals, int fmt,
                           int semiring, int accum, int u_kind, const int* u_dense, const int64_t* u_ids,
                           const int* u_vals, int64_t u_nvals, int w_kind, const int* w_dense, const int64_t* w_ids,
                           const int* w_vals, int64_t w_nvals, int mask_kind, const int* m_dense,
                           const int64_t* m_ids, const int* m_vals, int64_t m_nvals, int desc_mask, int* w_out,
                           int* w_storage) {
    try {
        return run_mxv<int>(vxm, n, rp, col, vals, fmt, semiring, accum, u_kind, u_dense, u_ids, u_vals, u_nvals,
                            w_kind, w_dense, w_ids, w_vals, w_nvals, mask_kind, m_dense, m_ids, m_vals, m_nvals,
                            desc_mask, w_out, w_storage);
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

// bfs_blast (algorithms/bfs/bfs.hpp:94-142) on Matrix<float> with values 1.
extern "C" int ref_bfs(int64_t n, const int64_t* rp, const int64_t* col, int64_t source, float* levels) {
    try {
        std::vector<float> ones(rp[n], 1.0f);
        lablas::Matrix<float> A;
        build_matrix(A, n, rp, col, ones.data(), CSR);
        lablas::Vector<float> v(n);
        lablas::Descriptor desc;
        lablas::algorithm::bfs_blast(&v, &A, (Index)source, &desc);
        const float* lv = v.get_vector()->getDense()->get_vals();
        for (int64_t i = 0; i < n; ++i) levels[i] = lv[i];
        return 0;
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

// sssp_bellman_ford_blast (algorithms/sssp/sssp_blast.hpp:8-72), float weights.
extern "C" int ref_sssp(int64_t n, const int64_t* rp, const int64_t* col, const float* w, int64_t source, float* dist) {
    try {
        lablas::Matrix<float> A;
        build_matrix(A, n, rp, col, w, CSR);
        lablas::Vector<float> v(n);
        lablas::Descriptor desc;
        lablas::algorithm::sssp_bellman_ford_blast(&v, &A, (Index)source, &desc);
        const float* dv = v.get_vector()->getDense()->get_vals();
        for (int64_t i = 0; i < n; ++i) dist[i] = dv[i];
        return 0;
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

// LAGraph_page_rank_sinks (algorithms/pr/pr.hpp:20-153), itermax iterations.
extern "C" int ref_pagerank(int64_t n, const int64_t* rp, const int64_t* col, int itermax, float* ranks) {
    try {
        std::vector<float> ones(rp[n], 1.0f);
        lablas::Matrix<float> A;
        build_matrix(A, n, rp, col, ones.data(), CSR);
        LAGraph_Graph<float> G(A);
        lablas::Vector<float> r(n);
        int iters = 0;
        LAGraph_page_rank_sinks(&r, &G, &iters, itermax);
        const float* rv = r.get_vector()->getDense()->get_vals();
        for (int64_t i = 0; i < n; ++i) ranks[i] = rv[i];
        return iters;
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

// Matrix::init_from_mtx (read_file.hpp:556-615) on a file; the CSR is kept
// until the next call and copied out by ref_mtx_csr.
static lablas::Matrix<float>* g_mtx = nullptr;

extern "C" int ref_mtx_load(const char* path, int64_t* nrows, int64_t* nnz, double* seconds) {
    try {
        delete g_mtx;
        g_mtx = new lablas::Matrix<float>;
        const double t0 = omp_get_wtime();
        g_mtx->init_from_mtx(path);
        *seconds = omp_get_wtime() - t0;
        const auto* csr = g_mtx->get_matrix()->get_csr();
        *nrows = csr->get_num_rows();
        *nnz = csr->get_nnz();
        return 0;
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

extern "C" int ref_mtx_csr(int64_t* rp, int64_t* col) {
    if (!g_mtx) return -1;
    const auto* csr = g_mtx->get_matrix()->get_csr();
    const int64_t n = csr->get_num_rows();
    for (int64_t i = 0; i <= n; ++i) rp[i] = csr->get_row_ptr()[i];
    for (int64_t e = 0; e < rp[n]; ++e) col[e] = csr->get_col_ids()[e];
    return 0;
}

extern "C" int ref_set_threads(int t) {
    omp_set_num_threads(t);
    return omp_get_max_threads();
}

#else  // CC bridge

// FastSV cc (algorithms/cc/cc.hpp:12-119) on Matrix<int> with values 1.
extern "C" int ref_cc(int64_t n, const int64_t* rp, const int64_t* col, int* labels) {
    try {
        std::vector<int> ones(rp[n], 1);
        lablas::Matrix<int> A;
        build_matrix(A, n, rp, col, ones.data(), CSR);
        lablas::Vector<int> v(n);
        lablas::Descriptor desc;
        lablas::algorithm::cc(&v, &A, 0, &desc);
        const int* lv = v.get_vector()->getDense()->get_vals();
        for (int64_t i = 0; i < n; ++i) labels[i] = lv[i];
        return 0;
    } catch (const char* e) {
        fprintf(stderr, "reference threw: %s\n", e);
        return -1;
    }
}

#endif
