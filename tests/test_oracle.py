"""Pins the CPU oracle before it is trusted as the GPU checker.

1. Golden vectors written by the compiled reference (tests/golden/*.npz,
   make_golden.py): the oracle's restatements must reproduce them bit for bit
   (fp32 included: same left fold, same IEEE operations).
2. Closed forms and brute force where no golden vector reaches: the FP64
   SpMV on integer-valued data equals exact integer arithmetic; BFS levels
   obey |level(u) - level(v)| <= 1 across edges; Dijkstra distances satisfy
   the triangle inequality with equality on a shortest-path tree; CC labels
   are constant across edges and equal the component minimum; PageRank
   ranks sum to 1 and are a fixed point.
Each check is chosen so that a dropped term, a wrong index or a transposed
operand in the oracle fails it.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
A_PLUS, A_LOR, A_MIN = 0, 1, 2
M_TIMES, M_LAND, M_SECOND, M_PLUS = 0, 1, 3, 4


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False))


def transpose(rp, col, vals):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    order = np.lexsort((rows, col))
    trp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(trp, col.astype(np.int64) + 1, 1)
    return np.cumsum(trp), rows[order].astype(np.int32), vals[order]


def test_golden_spmv_f32(oracle):
    d = load("spmv_rmat9_f32")
    rp, col, vals, x, w0 = d["rp"], d["col"], d["vals"], d["x"], d["w0"]
    n = len(rp) - 1
    z = np.zeros(n, np.float32)
    # CSR plus.times: bit-exact left fold
    assert np.array_equal(oracle.spmv(rp, col, vals, x, z, A_PLUS, M_TIMES), d["y_plus_times_csr"])
    trp, tcol, tval = transpose(rp, col, vals)
    # build_from_csr_arrays fills the CSC with an atomic fetch-add transpose
    # (build.hpp:171-210), so the fold order inside a column is not fixed:
    # vxm in fp32 is pinned to rounding, not to bits
    np.testing.assert_allclose(oracle.spmv(trp, tcol, tval, x, z, A_PLUS, M_TIMES), d["y_plus_times_vxm"], rtol=2e-6)
    assert np.array_equal(oracle.spmv(rp, col, vals, x, z, A_LOR, M_LAND), d["y_lor_land"])
    assert np.array_equal(oracle.spmv(rp, col, vals, x, z, A_MIN, M_PLUS), d["y_min_plus"])
    assert np.array_equal(oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_SECOND, accum=2), d["y_plus_second_accum_plus"])
    m = d["mdense"]
    assert np.array_equal(oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES, mask=m), d["y_mask_dense_default"])
    assert np.array_equal(oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES, mask=m, desc_mask=5),
                          d["y_mask_dense_strcomp"])
    ids = d["msparse_ids"]
    assert np.array_equal(oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES, mask_ids=ids), d["y_mask_sparse_default"])
    assert np.array_equal(oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES, mask_ids=ids, desc_mask=5),
                          d["y_mask_sparse_strcomp"])


def test_golden_spmspv_f32(oracle):
    d = load("spmv_rmat9_f32")
    rp, col, vals, w0 = d["rp"], d["col"], d["vals"], d["w0"]
    n = len(rp) - 1
    xi, xv = d["xs_ids"], d["xs_vals"]
    trp, tcol, tval = transpose(rp, col, vals)
    z = np.zeros(n, np.float32)
    # mxv pushes along the CSC, vxm along the CSR (spmspv.h:126-146);
    # the plus kernel's order differs from ours only for fp32 rounding: the
    # reference is sequential here when run with one frontier order
    got = oracle.spmspv(trp, tcol, tval, xi, xv, z, A_PLUS, M_TIMES)
    np.testing.assert_allclose(got, d["y_spmspv_plus"], rtol=1e-6, atol=0)
    got = oracle.spmspv(rp, col, vals, xi, xv, z, A_PLUS, M_TIMES)
    np.testing.assert_allclose(got, d["y_spmspv_plus_vxm"], rtol=1e-6, atol=0)
    got = oracle.spmspv(rp, col, vals, xi, xv, w0, A_LOR, M_LAND, mask=d["mdense"], desc_mask=5)
    assert np.array_equal(got, d["y_spmspv_or_masked_strcomp"])
    got = oracle.spmspv(trp, tcol, tval, xi, xv, z, A_MIN, M_PLUS)
    assert np.array_equal(got, d["y_spmspv_min_plus"])
    got = oracle.spmspv(trp, tcol, tval, xi, xv, w0, A_PLUS, M_TIMES, accum=2, mask_ids=d["msparse_ids"])
    np.testing.assert_allclose(got, d["y_spmspv_plus_mask_sparse"], rtol=1e-6, atol=0)


def test_golden_storage_state():
    """The 30% rule as the reference applied it (vector.h:117-129)."""
    d = load("spmv_rmat9_f32")
    n = len(d["rp"]) - 1
    for tag in d["cases"]:
        y, st = d["y_" + tag], int(d["st_" + tag])
        assert st == (2 if np.count_nonzero(y) > int(n * 0.3) else 1), tag


def test_golden_spmv_i32(oracle):
    d = load("spmv_rmat9_i32")
    rp, col, vals, x = d["rp"], d["col"], d["vals"], d["x"]
    n = len(rp) - 1
    trp, tcol, tval = transpose(rp, col, vals)
    ops = {"PlusMultiplies": (A_PLUS, M_TIMES), "MinimumPlus": (A_MIN, M_PLUS), "MinimumSelectSecond": (A_MIN, M_SECOND)}
    for tag in d["cases"]:
        sr, mode = tag.rsplit("_", 1)
        a, m = ops[sr]
        R = (trp, tcol, tval) if mode == "vxm" else (rp, col, vals)
        assert np.array_equal(oracle.spmv(*R, x, np.zeros(n, np.int32), a, m), d["y_" + tag]), tag


def test_golden_trailing_row_defect(oracle):
    """The reference's partition leaves rows 6-7 stale (balancing.h:5-33); the
    oracle reproduces it with ref_threads and computes them otherwise."""
    d = load("trailing_rows_8x8")
    rp, col, vals, x, w0 = d["rp"], d["col"], d["vals"], d["x"], d["w0"]
    got_ref = oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES, ref_threads=int(d["threads"]))
    assert np.array_equal(got_ref, d["y"])
    assert d["y"][6] == 7.0 and d["y"][7] == 7.0  # stale values: the defect
    fixed = oracle.spmv(rp, col, vals, x, w0, A_PLUS, M_TIMES)
    assert fixed[6] == 0.0 and fixed[7] == 0.0
    assert np.array_equal(fixed[:6], d["y"][:6])


def test_golden_bfs(oracle):
    d = load("bfs_rmat10_sym")
    for s, lv in zip(d["sources"], d["levels"]):
        got = oracle.bfs(d["rp"], d["col"], int(s))
        assert np.array_equal(got, lv.astype(np.int32))


def test_golden_sssp(oracle):
    d = load("sssp_rmat10_sym")
    for s, dist in zip(d["sources"], d["dist"]):
        got = oracle.dijkstra_i32(d["rp"], d["col"], d["w"], int(s))
        want = np.where(dist == np.finfo(np.float32).max, np.iinfo(np.int32).max, dist.astype(np.int64))
        assert np.array_equal(got.astype(np.int64), want)


def test_golden_cc(oracle):
    d = load("cc_two_blocks")
    got = oracle.cc(d["rp"], d["col"])
    assert np.array_equal(got, d["labels"])


def test_golden_pagerank(oracle):
    """The reference's ranks are reproduced once its sparse-mask sink sum is
    restated (sink_quirk); the PageRank definition differs from them."""
    d = load("pr_rmat10")
    r64, it = oracle.pagerank_f64(d["rp"], d["col"], itermax=int(d["iters"]), sink_quirk=True)
    assert it == int(d["iters"])
    assert np.abs(r64 - d["ranks"]).sum() < 2e-6
    r_def, _ = oracle.pagerank_f64(d["rp"], d["col"], itermax=int(d["iters"]))
    assert abs(r_def.sum() - 1.0) < 1e-12 and abs(d["ranks"].sum() - 1.0) > 1e-2


# ------------------------------------------------------------ closed forms
def _rand_csr(seed, n, m):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, m)
    cols = rng.integers(0, n, m).astype(np.int32)
    order = np.argsort(rows, kind="stable")
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp), cols[order]


def test_spmv_against_dense_bruteforce(oracle):
    rp, col = _rand_csr(1, 37, 300)
    n = 37
    rng = np.random.default_rng(2)
    vals = rng.integers(-50, 50, len(col)).astype(np.int64)
    x = rng.integers(-50, 50, n).astype(np.int64)
    D = np.zeros((n, n), dtype=np.int64)
    rows = np.repeat(np.arange(n), np.diff(rp))
    np.add.at(D, (rows, col), vals)
    assert np.array_equal(oracle.spmv(rp, col, vals, x, np.zeros(n, np.int64), A_PLUS, M_TIMES), D @ x)
    # min.plus brute force
    big = np.iinfo(np.int64).max
    want = np.full(n, big, dtype=np.int64)
    for r, c, v in zip(rows, col, vals):
        want[r] = min(want[r], v + x[c])
    assert np.array_equal(oracle.spmv(rp, col, vals, x, np.zeros(n, np.int64), A_MIN, M_PLUS), want)


def test_minplus_int32_saturates(oracle):
    rp = np.array([0, 2, 2], dtype=np.int64)
    col = np.array([0, 1], dtype=np.int32)
    vals = np.array([5, 3], dtype=np.int32)
    x = np.array([np.iinfo(np.int32).max, 10], dtype=np.int32)
    y = oracle.spmv(rp, col, vals, x, np.zeros(2, np.int32), A_MIN, M_PLUS)
    assert y[0] == 13 and y[1] == np.iinfo(np.int32).max


def test_bfs_invariants(oracle):
    rp, col = _rand_csr(3, 500, 1500)
    lv = oracle.bfs(rp, col, 0)
    rows = np.repeat(np.arange(500), np.diff(rp))
    reached = lv[rows] > 0
    assert np.all(lv[col[reached]] > 0)
    assert np.all(lv[col[reached]] <= lv[rows[reached]] + 1)
    assert lv[0] == 1


def test_dijkstra_invariants(oracle):
    rp, col = _rand_csr(4, 400, 2000)
    w = np.random.default_rng(5).integers(1, 256, len(col)).astype(np.int32)
    d = oracle.dijkstra_i32(rp, col, w, 0).astype(np.int64)
    rows = np.repeat(np.arange(400), np.diff(rp))
    inf = np.iinfo(np.int32).max
    ok = d[rows] < inf
    assert np.all(d[col[ok]] <= d[rows[ok]] + w[ok])
    assert d[0] == 0


def test_cc_invariants(oracle):
    rp, col = _rand_csr(6, 300, 250)
    rows = np.repeat(np.arange(300), np.diff(rp))
    # symmetrize
    r2 = np.concatenate([rows, col]); c2 = np.concatenate([col, rows]).astype(np.int32)
    order = np.argsort(r2, kind="stable")
    rp2 = np.zeros(301, dtype=np.int64)
    np.add.at(rp2, r2 + 1, 1)
    rp2 = np.cumsum(rp2)
    c2 = c2[order]
    lab = oracle.cc(rp2, c2)
    rr = np.repeat(np.arange(300), np.diff(rp2))
    assert np.all(lab[rr] == lab[c2])
    assert np.all(lab <= np.arange(300))
    assert np.all(lab[lab] == lab)


def test_pagerank_fixed_point(oracle):
    rp, col = _rand_csr(7, 200, 1000)
    r, it = oracle.pagerank_f64(rp, col, itermax=500, tol=1e-13)
    assert abs(r.sum() - 1.0) < 1e-9
    r2, _ = oracle.pagerank_f64(rp, col, itermax=it + 5, tol=0.0)
    assert np.abs(r2 - r).sum() < 1e-11


def _dense_ppr(rp, col, val, p, damping):
    """Closed form of the personalized PageRank fixed point on a small graph:
    (I - d A - d p s^T) r = (1 - d) p, s = indicator of the empty columns."""
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            A[i, col[j]] += float(val[j])
    s = np.zeros(n)
    s[np.setdiff1d(np.arange(n), np.unique(col))] = 1.0
    M = np.eye(n) - damping * A - damping * np.outer(p, s)
    return np.linalg.solve(M, (1.0 - damping) * p)


def test_ppr_closed_form(oracle, gbk):
    """The FP64 personalized PageRank converges to the linear-system solution,
    sums to 1, and with a uniform teleport vector equals the reference's
    PageRank with sinks (orc_pagerank_f64, an independent restatement)."""
    rp, col = _rand_csr(11, 120, 500)
    n = len(rp) - 1
    val = gbk.inv_coldeg_f32(n, col)
    for src in (0, 7, 63):
        p = np.zeros(n)
        p[src] = 1.0
        r, it = oracle.ppr_f64(rp, col, val, source=src, itermax=2000, tol=1e-14)
        want = _dense_ppr(rp, col, val, p, 0.85)
        assert np.abs(r - want).sum() < 1e-10, src
        assert abs(r.sum() - 1.0) < 1e-6  # columns of fp32 1/coldeg sum to 1 within rounding
    # uniform teleport vector: the global PageRank with sinks
    r_u, _ = oracle.ppr_f64(rp, col, val, p=np.full(n, 1.0 / n), itermax=300)
    r_g, _ = oracle.pagerank_f64(rp, col, itermax=300)
    assert np.abs(r_u - r_g).sum() < 1e-6  # fp32 1/coldeg values vs exact division


def test_spmv_f64_closed_form(oracle):
    """The FP64 restatement equals the dense FP64 product of the same fp32 data
    (summation order aside: integers below 2^24 make every partial sum exact)."""
    rp, col = _rand_csr(9, 53, 700)
    n = 53
    rng = np.random.default_rng(10)
    vals = rng.integers(1, 1000, len(col)).astype(np.float32)
    x = rng.integers(1, 1000, n).astype(np.float32)
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    np.add.at(D, (rows, col), vals.astype(np.float64))
    y, absum = oracle.spmv_f64(rp, col, vals, x)
    assert np.array_equal(y, D @ x.astype(np.float64))
    assert np.array_equal(absum, np.abs(D) @ np.abs(x.astype(np.float64)))


def test_cpu_baseline_kernel_matches_reference_bits(oracle):
    """The timed OpenMP baseline (the reference's row partition and per-row left
    fold) reproduces the compiled reference's fp32 CSR result bit for bit, and
    its partition is the reference's: contiguous, covering, nnz-balanced."""
    d = load("spmv_rmat9_f32")
    rp, col, vals, x = d["rp"], d["col"], d["vals"], d["x"]
    for t in (1, 3, 8):
        first, last = oracle.balance_rows(rp, t)
        assert first[0] == 0 and np.all(first[1:] == last[:-1])
        y = np.zeros(len(rp) - 1, np.float32)
        oracle.spmv_f32_omp(rp.astype(np.int64), col.astype(np.int32), vals, x, y, t, first, last)
        covered = np.zeros(len(rp) - 1, bool)
        for a, b in zip(first, last):
            covered[a:b] = True
        assert np.array_equal(y[covered], d["y_plus_times_csr"][covered])


def test_dijkstra_f32_matches_integer_dijkstra(oracle, gbk):
    """On integer weights below 2^24 the float Dijkstra is exact: it must equal
    the int32 one (pinned by the reference's golden SSSP)."""
    rp, col = gbk.gen_rmat(10, 8, seed=41, symmetric=True)
    w = gbk.sym_weights_i32(rp, col, 42, 1, 255)
    for s in (0, 5, 77):
        di = oracle.dijkstra_i32(rp, col, w, s)
        df = oracle.dijkstra_f32(rp, col, w.astype(np.float32), s)
        reach = di != np.iinfo(np.int32).max
        assert np.array_equal(df[reach].astype(np.int64), di[reach].astype(np.int64))
        assert np.all(df[~reach] == np.finfo(np.float32).max)


# ------------------------------------------------------------------ mxm / tc
def _dense_csr(n, density, seed, self_loops=True):
    rng = np.random.default_rng(seed)
    d = rng.random((n, n)) < density
    if not self_loops:
        np.fill_diagonal(d, False)
    rp = np.concatenate([[0], np.cumsum(d.sum(1))]).astype(np.int64)
    col = np.nonzero(d)[1].astype(np.int32)
    return d, rp, col


def test_oracle_mxm_masked_matches_dense_products():
    from oracle import oracle as O
    n = 60
    da, arp, acol = _dense_csr(n, 0.15, 1)
    db, brp, bcol = _dense_csr(n, 0.15, 2)
    dm, mrp, mcol = _dense_csr(n, 0.3, 3)
    rng = np.random.default_rng(4)
    av = rng.integers(1, 50, len(acol)).astype(np.int32)
    bv = rng.integers(1, 50, len(bcol)).astype(np.int32)
    A = np.zeros((n, n), np.int64); A[da] = av
    B = np.zeros((n, n), np.int64); B[db] = bv
    hits = (da.astype(np.int64) @ db.astype(np.int64)) > 0
    mi, mj = np.nonzero(dm)
    # plus.times
    got = O.mxm_masked(mrp, mcol, arp, acol, av, brp, bcol, bv, 0, 0)
    want = np.where(hits, A @ B, 0)[mi, mj]
    assert np.array_equal(got.astype(np.int64), want)
    # min.plus: min over the contributing k of A(i,k) + B(k,j)
    big = np.iinfo(np.int64).max // 4
    Ai = np.where(da, A, big)
    Bi = np.where(db, B, big)
    mp = np.min(Ai[:, :, None] + Bi[None, :, :], axis=1)
    got = O.mxm_masked(mrp, mcol, arp, acol, av, brp, bcol, bv, 2, 4)
    want = np.where(hits, mp, 0)[mi, mj]
    assert np.array_equal(got.astype(np.int64), want)


def _sp_csr(S):
    S = S.tocsr()
    S.sort_indices()
    return S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data


def test_oracle_mxm_unmasked_matches_scipy():
    """int64 plus.times against scipy's sparse product; values in [-3, 3]
    cancel, and the oracle drops the entries whose sum is 0 (as scipy's
    eliminate_zeros does)."""
    import scipy.sparse as sp
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    for n, dens in ((200, 0.03), (57, 0.2)):
        A = sp.random(n, n, density=dens, random_state=rng, format="csr",
                      data_rvs=lambda k: rng.integers(-3, 4, k)).astype(np.int64)
        B = sp.random(n, n, density=dens, random_state=rng, format="csr",
                      data_rvs=lambda k: rng.integers(-3, 4, k)).astype(np.int64)
        A.eliminate_zeros()
        B.eliminate_zeros()
        arp, acol, av = _sp_csr(A)
        brp, bcol, bv = _sp_csr(B)
        crp, ccol, cval = O.mxm_unmasked(arp, acol, av, brp, bcol, bv, 0, 0)
        W = (A @ B).tocsr()
        W.eliminate_zeros()
        wrp, wcol, wv = _sp_csr(W)
        assert np.array_equal(crp, wrp) and np.array_equal(ccol, wcol) and np.array_equal(cval, wv)
        cancelled = (A.astype(bool).astype(np.int64) @ B.astype(bool).astype(np.int64)).nnz - W.nnz
        assert cancelled > 0  # some sums did cancel to 0 and were dropped


def test_oracle_mxm_unmasked_identity_and_minplus():
    """A (+.*) I = I (+.*) A = A (fp32 values, one term per entry); min.plus
    against a brute-force minimum over k on small matrices."""
    from oracle import oracle as O
    n = 50
    da, arp, acol = _dense_csr(n, 0.2, 12)
    av = np.random.default_rng(13).random(len(acol)).astype(np.float32) + 0.5
    irp = np.arange(n + 1, dtype=np.int64)
    icol = np.arange(n, dtype=np.int32)
    one = np.ones(n, np.float32)
    for args in ((arp, acol, av, irp, icol, one), (irp, icol, one, arp, acol, av)):
        crp, ccol, cval = O.mxm_unmasked(*args, 0, 0)
        assert np.array_equal(crp, arp) and np.array_equal(ccol, acol) and np.array_equal(cval, av)
    db, brp, bcol = _dense_csr(n, 0.2, 14)
    rng = np.random.default_rng(15)
    ai = rng.integers(1, 50, len(acol)).astype(np.int32)
    bi = rng.integers(1, 50, len(bcol)).astype(np.int32)
    big = np.iinfo(np.int64).max // 4
    A = np.full((n, n), big, np.int64); A[da] = ai
    B = np.full((n, n), big, np.int64); B[db] = bi
    mp = np.min(A[:, :, None] + B[None, :, :], axis=1)
    hit = (da.astype(np.int64) @ db.astype(np.int64)) > 0
    crp, ccol, cval = O.mxm_unmasked(arp, acol, ai, brp, bcol, bi, 2, 4)
    wi, wj = np.nonzero(hit)
    assert np.array_equal(crp, np.concatenate([[0], np.cumsum(hit.sum(1))]))
    assert np.array_equal(ccol, wj) and np.array_equal(cval.astype(np.int64), mp[wi, wj])


def test_oracle_mxm_unmasked_folds_in_ikj_order():
    """C(0,0) = ((0 + 1e8) + 1) + (-1e8) in A's row order: 0 in fp32 (the
    entry is dropped), 1 in fp64 (kept)."""
    from oracle import oracle as O
    arp = np.array([0, 3, 3, 3], np.int64)
    acol = np.array([0, 1, 2], np.int32)
    brp = np.array([0, 1, 2, 3], np.int64)
    bcol = np.array([0, 0, 0], np.int32)
    for dt, want_nnz in ((np.float32, 0), (np.float64, 1)):
        av = np.array([1e8, 1.0, -1e8], dt)
        crp, ccol, cval = O.mxm_unmasked(arp, acol, av, brp, bcol, np.ones(3, dt), 0, 0)
        assert crp[-1] == want_nnz
        if want_nnz:
            assert ccol[0] == 0 and cval[0] == 1.0


def test_oracle_masked_count_is_the_triangle_count():
    import scipy.sparse as sp
    from oracle import oracle as O
    for seed in (5, 6, 7):
        d, _, _ = _dense_csr(120, 0.08, seed, self_loops=False)
        d = d | d.T
        S = sp.csr_matrix(d.astype(np.int64))
        tri = int((S @ S @ S).diagonal().sum()) // 6
        rp, col = S.indptr.astype(np.int64), S.indices.astype(np.int32)
        L = sp.tril(S, -1).tocsr()
        U = sp.triu(S, 1).tocsr()
        lrp, lcol = L.indptr.astype(np.int64), L.indices.astype(np.int32)
        urp, ucol = U.indptr.astype(np.int64), U.indices.astype(np.int32)
        assert O.masked_count(rp, col, rp, col, rp, col) == 6 * tri              # Burkhardt
        assert O.masked_count(rp, col, lrp, lcol, urp, ucol) == 2 * tri          # Cohen
        assert O.masked_count(lrp, lcol, lrp, lcol, lrp, lcol) == tri            # Sandia
        assert O.masked_count(urp, ucol, urp, ucol, urp, ucol) == tri            # Sandia2
