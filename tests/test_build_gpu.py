"""Device-side construction: the CSC built by a stable sort on the device and
the segmented format's (segment, row) pairs, both against numpy restatements
of the same definitions (CSC = the stable counting-sort transpose, rows
ascending inside each column and list order among duplicates; pairs = the
distinct (column segment, row) keys)."""
import numpy as np
import pytest

from tests.graphs import special_graphs

pytestmark = pytest.mark.gpu


def _stable_transpose(rp, col, val):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    order = np.argsort(col, kind="stable")
    trp = np.zeros(n + 1, np.int64)
    np.add.at(trp, np.asarray(col, np.int64) + 1, 1)
    return np.cumsum(trp), rows[order], val[order]


def _graphs(gbk):
    gs = {"rmat14": gbk.gen_rmat(14, 16, seed=71), "er12": gbk.gen_uniform(12, 8, seed=72)}
    gs.update(special_graphs())
    return gs


def test_device_csc_is_the_stable_transpose(gbk):
    for name, (rp, col) in _graphs(gbk).items():
        if len(col) == 0:
            continue
        vals = (np.arange(len(col)) % 251 + 1).astype(np.int32)
        A = gbk.Matrix.from_csr(rp, col, vals)
        trp, tcol, tval = A.export(transposed=True)
        wrp, wcol, wval = _stable_transpose(rp, col, vals)
        assert np.array_equal(trp, wrp) and np.array_equal(tcol, wcol) and np.array_equal(tval, wval), name


def test_device_csc_with_duplicates(gbk):
    rng = np.random.default_rng(73)
    rows = rng.integers(0, 500, 20_000)
    cols = rng.integers(0, 500, 20_000)
    rows[:10] = 7
    cols[:10] = 9  # ten copies of (7, 9)
    vals = np.arange(20_000, dtype=np.float64)
    A = gbk.Matrix(gbk.FP64)
    A.build_coo(rows, cols, vals)
    rp, col, val = A.export()
    trp, tcol, tval = A.export(transposed=True)
    wrp, wcol, wval = _stable_transpose(rp, col, val)
    assert np.array_equal(trp, wrp) and np.array_equal(tcol, wcol) and np.array_equal(tval, wval)
    seg = tval[trp[9]:trp[10]]
    assert list(seg[np.asarray(tcol[trp[9]:trp[10]]) == 7]) == list(range(10))


@pytest.mark.parametrize("scale", [10, 16, 19])
def test_segmented_pairs_match_definition(gbk, oracle, scale):
    rp, col = gbk.gen_rmat(scale, 8, seed=74)
    n = len(rp) - 1
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.int32), fmt=3)
    _, _, nseg, npairs = A.format_stats()
    want_nseg = 8 * max(1, -(-n // (8 * 512 * 1024)))
    w = max(1, -(-n // want_nseg))
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    keys = np.unique((np.asarray(col, np.int64) // w) * n + rows)
    assert nseg == want_nseg and npairs == len(keys)
    x = (gbk.uniform_f32(75, n) * 1000).astype(np.int32)
    for sr in ("PlusMultiplies", "MinimumPlus"):
        add, mul = gbk.SEMIRINGS[sr]
        w_ = gbk.Vector(n, gbk.INT32)
        gbk.mxv(w_, None, gbk.ACC_NONE, sr, A, gbk.Vector.dense(x))
        want = oracle.spmv(rp, col, np.ones(len(col), np.int32), x, np.zeros(n, np.int32), add, mul)
        assert np.array_equal(w_.to_numpy(), want), sr
