// A driver in the style of the reference's spmv.cpp / bfs.cpp / sssp.cpp /
// cc.cpp / pr.cpp, written against the C++ lablas front end
// (include/gbk/lablas.hpp) only: it builds a seeded R-MAT through
// lablas::Matrix::build_from_csr_arrays, runs the operation on the GPU and
// checks the result with a sequential checker, printing "error_count: k/n"
// like the reference's drivers (spmv.cpp:102-122, bfs.cpp:50-65, ...).
//
// usage: lablas_driver <spmv|vxm|bfs|sssp|cc|pr|masked> <scale> [format]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "gbk/lablas.hpp"

static void gen(int scale, bool sym, std::vector<Index>& rp, std::vector<Index>& col) {
    int64_t n = 0, nnz = 0, *r = nullptr;
    int32_t* c = nullptr;
    if (gbk_gen_rmat(scale, 16, 0.57, 0.19, 0.19, 7, sym ? 1 : 0, 1, &n, &nnz, &r, &c) != GBK_SUCCESS) {
        fprintf(stderr, "generator failed: %s\n", gbk_last_error());
        exit(2);
    }
    rp.assign(r, r + n + 1);
    col.assign(c, c + nnz);
    gbk_host_free(r);
    gbk_host_free(c);
}

static int report(long long errors, long long n) {
    printf("error_count: %lld/%lld\n", errors, n);
    return errors == 0 ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <spmv|vxm|bfs|sssp|cc|pr|masked> <scale> [format]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const int scale = atoi(argv[2]);
    const int fmt = argc > 3 ? atoi(argv[3]) : 0;
    try {
        const bool sym = (mode == "bfs" || mode == "sssp" || mode == "cc");
        std::vector<Index> rp, col;
        gen(scale, sym, rp, col);
        const Index n = (Index)rp.size() - 1, nnz = (Index)col.size();
        lablas::Descriptor desc;

        if (mode == "spmv" || mode == "vxm" || mode == "masked") {
            std::vector<float> vals(nnz), x(n);
            gbk_fill_uniform_f32(11, nnz, vals.data());
            gbk_fill_uniform_f32(12, n, x.data());
            lablas::Matrix<float> A;
            A.set_preferred_matrix_format((MatrixStorageFormat)fmt);
            A.build_from_csr_arrays(rp.data(), col.data(), vals.data(), n, nnz);
            lablas::Vector<float> u(n), w(n), m(n);
            u.build(x.data(), n);
            std::vector<double> want(n, 0.0), absum(n, 0.0);
            std::vector<float> mask(n);
            for (Index i = 0; i < n; ++i) mask[i] = (i % 3 == 0) ? 1.0f : 0.0f;
            if (mode == "spmv") {
                lablas::mxv<float, float, float, float>(&w, NULL, GrB_NULL, lablas::PlusMultipliesSemiring<float>(),
                                                        &A, &u, &desc);
                for (Index r = 0; r < n; ++r)
                    for (Index j = rp[r]; j < rp[r + 1]; ++j) {
                        want[r] += (double)vals[j] * x[col[j]];
                        absum[r] += std::fabs((double)vals[j] * x[col[j]]);
                    }
            } else if (mode == "vxm") {
                lablas::vxm<float, float, float, float>(&w, NULL, GrB_NULL, lablas::PlusMultipliesSemiring<float>(),
                                                        &u, &A, &desc);
                for (Index r = 0; r < n; ++r)
                    for (Index j = rp[r]; j < rp[r + 1]; ++j) {
                        want[col[j]] += (double)vals[j] * x[r];
                        absum[col[j]] += std::fabs((double)vals[j] * x[r]);
                    }
            } else {
                m.build(mask.data(), n);
                lablas::mxv<float, float, float, float>(&w, &m, GrB_NULL, lablas::PlusMultipliesSemiring<float>(), &A,
                                                        &u, &lablas::GrB_DESC_SC);
                for (Index r = 0; r < n; ++r) {
                    if (mask[r] != 0.0f) continue;  // complemented mask: masked rows are 0
                    for (Index j = rp[r]; j < rp[r + 1]; ++j) {
                        want[r] += (double)vals[j] * x[col[j]];
                        absum[r] += std::fabs((double)vals[j] * x[col[j]]);
                    }
                }
            }
            std::vector<float> got = w.to_host();
            long long err = 0;
            for (Index i = 0; i < n; ++i)
                if (std::fabs((double)got[i] - want[i]) > 1e-5 * absum[i] + 1e-30) ++err;
            return report(err, n);
        }

        if (mode == "bfs") {
            std::vector<float> ones(nnz, 1.0f);
            lablas::Matrix<float> A;
            A.build_from_csr_arrays(rp.data(), col.data(), ones.data(), n, nnz);
            Index s = 0;
            while (rp[s + 1] == rp[s]) ++s;
            lablas::Vector<float> levels(n);
            lablas::algorithm::bfs_blast(&levels, &A, s, &desc);
            // bfs_traditional.hpp:11-48: queue BFS, levels 1-based
            std::vector<int> lv(n, 0);
            std::queue<Index> q;
            lv[s] = 1;
            q.push(s);
            while (!q.empty()) {
                Index u = q.front();
                q.pop();
                for (Index j = rp[u]; j < rp[u + 1]; ++j)
                    if (!lv[col[j]]) {
                        lv[col[j]] = lv[u] + 1;
                        q.push(col[j]);
                    }
            }
            std::vector<float> got = levels.to_host();
            long long err = 0;
            for (Index i = 0; i < n; ++i) err += (got[i] != (float)lv[i]);
            return report(err, n);
        }

        if (mode == "sssp") {
            std::vector<int32_t> c32(col.begin(), col.end());
            std::vector<int64_t> rp64(rp.begin(), rp.end());
            std::vector<int> w(nnz);
            gbk_fill_sym_weights_i32(n, rp64.data(), c32.data(), 13, 1, 255, w.data());
            lablas::Matrix<int> A;
            A.build_from_csr_arrays(rp.data(), col.data(), w.data(), n, nnz);
            Index s = 0;
            while (rp[s + 1] == rp[s]) ++s;
            lablas::Vector<int> dist(n);
            lablas::algorithm::sssp_bellman_ford_blast(&dist, &A, s, &desc);
            // Dijkstra (sssp_traditional.hpp:9-77)
            std::vector<long long> d(n, (long long)INT32_MAX);
            std::priority_queue<std::pair<long long, Index>, std::vector<std::pair<long long, Index>>, std::greater<>> pq;
            d[s] = 0;
            pq.push({0, s});
            while (!pq.empty()) {
                auto [du, u] = pq.top();
                pq.pop();
                if (du != d[u]) continue;
                for (Index j = rp[u]; j < rp[u + 1]; ++j)
                    if (du + w[j] < d[col[j]]) {
                        d[col[j]] = du + w[j];
                        pq.push({d[col[j]], col[j]});
                    }
            }
            std::vector<int> got = dist.to_host();
            long long err = 0;
            for (Index i = 0; i < n; ++i) err += ((long long)got[i] != d[i]);
            return report(err, n);
        }

        if (mode == "cc") {
            std::vector<int> ones(nnz, 1);
            lablas::Matrix<int> A;
            A.build_from_csr_arrays(rp.data(), col.data(), ones.data(), n, nnz);
            lablas::Vector<int> labels(n);
            lablas::algorithm::cc(&labels, &A, 0, &desc);
            // label = minimum id of the component (BFS over components)
            std::vector<int> lab(n, -1);
            for (Index v = 0; v < n; ++v) {
                if (lab[v] >= 0) continue;
                std::vector<Index> st{v};
                lab[v] = (int)v;
                while (!st.empty()) {
                    Index u = st.back();
                    st.pop_back();
                    for (Index j = rp[u]; j < rp[u + 1]; ++j)
                        if (lab[col[j]] < 0) {
                            lab[col[j]] = (int)v;
                            st.push_back(col[j]);
                        }
                }
            }
            std::vector<int> got = labels.to_host();
            long long err = 0;
            for (Index i = 0; i < n; ++i) err += (got[i] != lab[i]);
            return report(err, n);
        }

        if (mode == "pr") {
            std::vector<float> ones(nnz, 1.0f);
            lablas::Matrix<float> A;
            A.build_from_csr_arrays(rp.data(), col.data(), ones.data(), n, nnz);
            lablas::Vector<float> ranks(n);
            int iters = 0;
            lablas::algorithm::page_rank_sinks(&ranks, &A, &iters, 20);
            std::vector<float> got = ranks.to_host();
            // ranks sum to one and are the 20-step power iteration (FP64 here)
            std::vector<double> r(n, 1.0 / n), w(n), cdeg(n, 0);
            for (Index j = 0; j < nnz; ++j) cdeg[col[j]] += 1;
            for (int it = 0; it < 20; ++it) {
                double sinks = 0;
                for (Index i = 0; i < n; ++i)
                    if (cdeg[i] == 0) sinks += r[i];
                double tele = 0.15 / n + 0.85 / n * sinks;
                for (Index i = 0; i < n; ++i) w[i] = r[i] / std::max(cdeg[i] / 0.85, 1 / 0.85);
                for (Index i = 0; i < n; ++i) {
                    double s = tele;
                    for (Index j = rp[i]; j < rp[i + 1]; ++j) s += w[col[j]];
                    r[i] = s;
                }
            }
            double l1 = 0;
            for (Index i = 0; i < n; ++i) l1 += std::fabs((double)got[i] - r[i]);
            printf("pagerank L1 vs fp64: %.3e after %d iterations\n", l1, iters);
            return report(l1 <= 1e-6 ? 0 : 1, 1);
        }
        fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    } catch (const char* e) {
        fprintf(stderr, "exception: %s\n", e);
        return 3;
    }
}
