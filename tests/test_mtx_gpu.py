"""Matrix::init_from_mtx on the device: a .mtx and a .mtxbin file of the
same edge list build the same matrix (the reference's test_file_formats.cpp
check), whose CSR keeps file order inside a row, whose CSC lists rows in
ascending order, whose values are all 1, and whose SpMV equals the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _expected_csr(n, rows, cols):
    order = np.argsort(rows, kind="stable")
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp), cols[order]


@pytest.mark.parametrize("fmt", [None, 3])
def test_mtx_and_mtxbin_build_the_same_matrix(gbk, oracle, tmp_path, fmt):
    rng = np.random.default_rng(21)
    nr, nc, m = 3000, 2500, 40_000
    rows = rng.integers(0, nr, size=m)
    cols = rng.integers(0, nc, size=m)
    pt, pb = str(tmp_path / "g.mtx"), str(tmp_path / "g.mtxbin")
    gbk.mtx_write(pt, nr, nc, rows, cols)
    gbk.mtx_write(pb, nr, nc, rows, cols)
    A = gbk.Matrix.from_mtx(pt, gbk.FP32, fmt=fmt)
    B = gbk.Matrix.from_mtx(pb, gbk.FP32, fmt=fmt)
    n = max(nr, nc)
    assert A.n == n and B.n == n and A.nnz == m
    rp, col = _expected_csr(n, rows, cols)
    for M in (A, B):
        arp, acol, aval = M.export()
        assert np.array_equal(arp, rp) and np.array_equal(acol, col) and np.all(aval == 1.0)
        trp, tcol, _ = M.export(transposed=True)
        want_rp, _ = _expected_csr(n, cols, rows)
        assert np.array_equal(trp, want_rp)
        for c in range(0, n, 97):  # rows ascending inside each column
            seg = tcol[trp[c]:trp[c + 1]]
            assert np.array_equal(seg, np.sort(rows[cols == c]))
    x = gbk.uniform_f32(22, n)
    w = gbk.Vector(n, gbk.FP32)
    gbk.mxv(w, None, gbk.ACC_NONE, "PlusMultiplies", A, gbk.Vector.dense(x))
    y64, absum = oracle.spmv_f64(rp, col, np.ones(m, np.float32), x)
    assert np.all(np.abs(w.to_numpy() - y64) <= 1e-5 * absum + 1e-30)


def test_int_matrix_from_mtx_and_missing_file(gbk, tmp_path):
    p = tmp_path / "k.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n4 4 4\n1 2\n2 3\n3 4\n4 1\n")
    A = gbk.Matrix.from_mtx(str(p), gbk.INT32)
    rp, col, val = A.export()
    assert rp.tolist() == [0, 1, 2, 3, 4] and col.tolist() == [1, 2, 3, 0] and val.tolist() == [1, 1, 1, 1]
    lv, _ = gbk.bfs(A, 0)
    assert lv.tolist() == [1, 2, 3, 4]
    with pytest.raises(gbk.GbkError, match="Can not open"):
        gbk.Matrix.from_mtx(str(tmp_path / "none.mtx"))
