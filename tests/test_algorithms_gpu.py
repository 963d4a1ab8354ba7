"""GPU parity of SpMSpV (push), BFS, SSSP, CC and PageRank against the oracle
and the golden vectors the compiled reference wrote.

BFS levels, SSSP distances (integer weights) and CC labels are bit-exact;
PageRank ranks are within 1e-6 L1 of the FP64 oracle run for the same number
of iterations.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INT_MAX = np.iinfo(np.int32).max


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False))


def _sources(rp, k, seed):
    deg = np.diff(rp)
    cand = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, len(cand)), replace=False)


# ----------------------------------------------------------------- SpMSpV
@pytest.mark.parametrize("semiring,dtype", [("PlusMultiplies", np.int32), ("MinimumPlus", np.int32),
                                            ("LogicalOrAnd", np.float32), ("MinimumSelectSecond", np.int32),
                                            ("PlusMultiplies", np.float32)])
@pytest.mark.parametrize("transposed", [False, True])
def test_spmspv_push(gbk, oracle, semiring, dtype, transposed):
    add, mul = gbk.SEMIRINGS[semiring]
    rp, col = gbk.gen_rmat(11, 16, seed=71)
    u = gbk.uniform_f32(72, len(col))
    vals = u if dtype == np.float32 else (1 + (u * 255).astype(np.int64)).astype(dtype)
    A = gbk.Matrix.from_csr(rp, col, vals)
    n = A.n
    ids = np.flatnonzero(gbk.uniform_f32(73, n) < 0.05)
    xv = (gbk.uniform_f32(74, len(ids)) * 100 + 1).astype(dtype)
    x = gbk.Vector.sparse(n, ids, xv)
    w = gbk.Vector(n, gbk.TYPE_OF[np.dtype(dtype)])
    if transposed:
        gbk.vxm(w, None, gbk.ACC_NONE, semiring, x, A)
        R = (rp, col, vals)  # vxm pushes along the CSR
    else:
        gbk.mxv(w, None, gbk.ACC_NONE, semiring, A, x)
        R = A.export(transposed=True)  # mxv pushes along the CSC
    got = w.to_numpy()
    want = oracle.spmspv(*R, ids, xv, np.zeros(n, dtype), add, mul)
    if dtype == np.float32 and add == gbk.ADD_PLUS:
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    else:
        assert np.array_equal(got, want)


@pytest.mark.parametrize("desc_mask", [6, 5])
@pytest.mark.parametrize("sparse_mask", [False, True])
def test_spmspv_masked(gbk, oracle, desc_mask, sparse_mask):
    rp, col = gbk.gen_rmat(10, 16, seed=81)
    vals = (1 + (gbk.uniform_f32(82, len(col)) * 9).astype(np.int32)).astype(np.int32)
    A = gbk.Matrix.from_csr(rp, col, vals)
    n = A.n
    ids = np.flatnonzero(gbk.uniform_f32(83, n) < 0.05)
    xv = (gbk.uniform_f32(84, len(ids)) * 100 + 1).astype(np.int32)
    w0 = (gbk.uniform_f32(85, n) * 50 + 1).astype(np.int32)
    mask = (gbk.uniform_f32(86, n) < (0.1 if sparse_mask else 0.6)).astype(np.int32)
    mids = np.flatnonzero(mask)
    m = gbk.Vector.sparse(n, mids, mask[mids]) if sparse_mask else gbk.Vector.dense(mask)
    w = gbk.Vector.dense(w0)
    gbk.vxm(w, m, gbk.ACC_PLUS, "PlusMultiplies", gbk.Vector.sparse(n, ids, xv), A, desc_mask)
    got = w.to_numpy()
    if sparse_mask:
        want = oracle.spmspv(rp, col, vals, ids, xv, w0, 0, 0, accum=2, mask_ids=mids, desc_mask=desc_mask)
    else:
        want = oracle.spmspv(rp, col, vals, ids, xv, w0, 0, 0, accum=2, mask=mask, desc_mask=desc_mask)
    assert np.array_equal(got, want)


def test_golden_spmspv(gbk):
    """The reference's own SpMSpV outputs (compiled reference, golden fixture)."""
    d = load("spmv_rmat9_f32")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], d["vals"])
    n = A.n
    x = gbk.Vector.sparse(n, d["xs_ids"], d["xs_vals"])
    w = gbk.Vector(n, gbk.FP32)
    gbk.mxv(w, None, gbk.ACC_NONE, "MinimumPlus", A, x)
    assert np.array_equal(w.to_numpy(), d["y_spmspv_min_plus"])
    w = gbk.Vector(n, gbk.FP32)
    gbk.vxm(w, None, gbk.ACC_NONE, "PlusMultiplies", x, A)
    np.testing.assert_allclose(w.to_numpy(), d["y_spmspv_plus_vxm"], rtol=1e-5, atol=1e-7)
    w = gbk.Vector.dense(d["w0"])
    gbk.vxm(w, gbk.Vector.dense(d["mdense"]), gbk.ACC_NONE, "LogicalOrAnd", x, A, gbk.STR_COMP)
    assert np.array_equal(w.to_numpy(), d["y_spmspv_or_masked_strcomp"])


def test_golden_pull(gbk):
    d = load("spmv_rmat9_f32")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], d["vals"])
    n = A.n
    for tag, kw in [("mask_dense_default", dict(mask=("dense", d["mdense"]), desc=6)),
                    ("mask_dense_strcomp", dict(mask=("dense", d["mdense"]), desc=5)),
                    ("mask_sparse_default", dict(mask=("sparse", d["msparse_ids"]), desc=6)),
                    ("mask_sparse_strcomp", dict(mask=("sparse", d["msparse_ids"]), desc=5))]:
        kind, m = kw["mask"]
        mv = gbk.Vector.dense(m) if kind == "dense" else gbk.Vector.sparse(n, m, np.ones(len(m), np.float32))
        w = gbk.Vector.dense(d["w0"])
        gbk.mxv(w, mv, gbk.ACC_NONE, "PlusMultiplies", A, gbk.Vector.dense(d["x"]), kw["desc"])
        want = d["y_" + tag]
        assert w.storage == int(d["st_" + tag]), tag  # before to_numpy(), which makes w dense
        np.testing.assert_allclose(w.to_numpy(), want, rtol=1e-5, atol=1e-7, err_msg=tag)
    w = gbk.Vector(n, gbk.FP32)
    gbk.mxv(w, None, gbk.ACC_NONE, "MinimumPlus", A, gbk.Vector.dense(d["x"]))
    assert np.array_equal(w.to_numpy(), d["y_min_plus"])


def test_golden_i32(gbk):
    d = load("spmv_rmat9_i32")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], d["vals"])
    for tag in d["cases"]:
        sr, mode = tag.rsplit("_", 1)
        w = gbk.Vector(A.n, gbk.INT32)
        u = gbk.Vector.dense(d["x"])
        if mode == "vxm":
            gbk.vxm(w, None, gbk.ACC_NONE, sr, u, A)
        else:
            gbk.mxv(w, None, gbk.ACC_NONE, sr, A, u)
        assert np.array_equal(w.to_numpy(), d["y_" + tag]), tag


# -------------------------------------------------------------------- BFS
def test_bfs_golden(gbk):
    d = load("bfs_rmat10_sym")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], np.ones(len(d["col"]), np.float32))
    for s, lv in zip(d["sources"], d["levels"]):
        got, _ = gbk.bfs(A, int(s))
        assert np.array_equal(got, lv.astype(np.int32))


@pytest.mark.parametrize("scale,symmetric", [(14, True), (14, False), (16, True)])
def test_bfs_vs_oracle(gbk, oracle, scale, symmetric):
    rp, col = gbk.gen_rmat(scale, 16, seed=91, symmetric=symmetric)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    for s in _sources(rp, 4, 92):
        got, st = gbk.bfs(A, int(s))
        want = oracle.bfs(rp, col, int(s))
        assert np.array_equal(got, want), f"source {s} stats {st}"


def test_bfs_both_directions_used(gbk, oracle):
    rp, col = gbk.gen_rmat(16, 16, seed=93, symmetric=True)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    s = int(_sources(rp, 1, 94)[0])
    got, st = gbk.bfs(A, s)
    assert st[1] > 0 and st[2] > 0, st  # push and pull steps both ran
    assert np.array_equal(got, oracle.bfs(rp, col, s))


# ------------------------------------------------------------------- SSSP
def test_sssp_golden(gbk):
    d = load("sssp_rmat10_sym")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], d["w"])
    for s, dist in zip(d["sources"], d["dist"]):
        got, _ = gbk.sssp(A, int(s))
        want = np.where(dist == np.finfo(np.float32).max, INT_MAX, dist.astype(np.int64))
        assert np.array_equal(got.astype(np.int64), want)
    Af = gbk.Matrix.from_csr(d["rp"], d["col"], d["w"].astype(np.float32))
    for s, dist in zip(d["sources"], d["dist"]):
        got, _ = gbk.sssp(Af, int(s))
        assert np.array_equal(got, dist)


@pytest.mark.parametrize("scale", [13, 15])
def test_sssp_vs_dijkstra(gbk, oracle, scale):
    rp, col = gbk.gen_rmat(scale, 16, seed=95, symmetric=True)
    w = gbk.sym_weights_i32(rp, col, 96, 1, 255)
    A = gbk.Matrix.from_csr(rp, col, w)
    for s in _sources(rp, 3, 97):
        got, _ = gbk.sssp(A, int(s))
        assert np.array_equal(got, oracle.dijkstra_i32(rp, col, w, int(s)))


def test_sssp_and_bfs_device_outputs(gbk, oracle):
    from gb_kun_amd.device import DeviceArray
    rp, col = gbk.gen_rmat(13, 16, seed=91, symmetric=True)
    n = len(rp) - 1
    w = gbk.sym_weights_i32(rp, col, 92, 1, 255)
    A = gbk.Matrix.from_csr(rp, col, w)
    Ab = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    buf = DeviceArray.zeros(n, np.int32)
    try:
        for s in _sources(rp, 2, 93):
            gbk.sssp_device(A, int(s), buf.ptr)
            assert np.array_equal(buf.numpy(), oracle.dijkstra_i32(rp, col, w, int(s)))
            gbk.bfs_device(Ab, int(s), buf.ptr)
            assert np.array_equal(buf.numpy(), gbk.bfs(Ab, int(s))[0])
    finally:
        buf.free()


# --------------------------------------------------------------------- CC
def test_cc_golden(gbk):
    d = load("cc_two_blocks")
    A = gbk.Matrix.from_csr(d["rp"], d["col"], np.ones(len(d["col"]), np.int32))
    got, _ = gbk.cc(A)
    assert np.array_equal(got, d["labels"])


@pytest.mark.parametrize("scale,ef", [(14, 1), (16, 4)])
def test_cc_vs_union_find(gbk, oracle, scale, ef):
    rp, col = gbk.gen_rmat(scale, ef, seed=98, symmetric=True)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.int32))
    got, _ = gbk.cc(A)
    assert np.array_equal(got, oracle.cc(rp, col))


# --------------------------------------------------------------------- PR
@pytest.mark.parametrize("scale", [10, 14])
def test_pagerank_vs_fp64(gbk, oracle, scale):
    rp, col = gbk.gen_rmat(scale, 16, seed=99)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    r, it = gbk.pagerank(A, itermax=200, tol=1e-6)
    r64, it64 = oracle.pagerank_f64(rp, col, itermax=it)
    assert it64 == it
    assert np.abs(r.astype(np.float64) - r64).sum() <= 1e-6
    assert abs(r.astype(np.float64).sum() - 1.0) < 1e-5


def test_pagerank_tol_stop_equals_fixed_run(gbk):
    """With tol > 0 the host reads iteration k's L1 change while iteration
    k + 1 runs and drops k + 1 once k is below tol: the ranks and the count
    equal a fixed run of that many iterations, whose L1 change is the first
    below tol.  Odd and even counts (the result in either rank buffer), zero
    iterations and the device entry point give the same bits."""
    from gb_kun_amd.device import DeviceArray
    rp, col = gbk.gen_rmat(12, 16, seed=101)
    n = len(rp) - 1
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    r, it = gbk.pagerank(A, itermax=200, tol=1e-6)
    assert 3 < it < 200
    r_fixed, it_fixed = gbk.pagerank(A, itermax=it, tol=0.0)
    assert it_fixed == it and np.array_equal(r, r_fixed)
    r_prev, _ = gbk.pagerank(A, itermax=it - 1)
    r_prev2, _ = gbk.pagerank(A, itermax=it - 2)
    l1 = np.abs(r.astype(np.float64) - r_prev).sum()
    l1_prev = np.abs(r_prev.astype(np.float64) - r_prev2).sum()
    assert l1 < 1e-6 <= l1_prev
    r0, it0 = gbk.pagerank(A, itermax=0)
    assert it0 == 0 and np.all(r0 == np.float32(1.0 / n))
    d = DeviceArray.zeros(n, np.float32)
    for k in (1, 2, 3):
        rh, _ = gbk.pagerank(A, itermax=k)
        assert gbk.pagerank_device(A, d.ptr, itermax=k) == k
        assert np.array_equal(d.numpy(), rh), k
    assert gbk.pagerank_device(A, d.ptr, itermax=200, tol=1e-6) == it
    assert np.array_equal(d.numpy(), r)


def test_bfs_pull_heavy_vertex(gbk, oracle):
    """Two hubs over the same 5000 leaves: the pull steps meet an unvisited hub
    with more in-edges than the per-group limit (the one-workgroup-per-vertex
    path), once with no parent in the frontier and once with 5000."""
    from tests.graphs import csr_from_edges
    L = 5000
    h2 = L + 1
    edges = []
    for leaf in range(1, L + 1):
        edges += [(0, leaf), (leaf, 0), (h2, leaf), (leaf, h2)]
    rp, col = csr_from_edges(L + 2, edges)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    got, st = gbk.bfs(A, 0)
    want = oracle.bfs(rp, col, 0)
    assert np.array_equal(got, want)
    assert got[0] == 1 and got[h2] == 3 and np.all(got[1:L + 1] == 2)
    assert st[2] >= 1, st  # pull steps ran
