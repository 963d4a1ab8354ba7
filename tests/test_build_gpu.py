"""Device-side construction against numpy restatements of the same
definitions: the CSC (the stable counting-sort transpose, rows ascending
inside each column and list order among duplicates), the segmented format's
(segment, row) pairs (the distinct (column segment, row) keys), SELL-C-sigma's
stored slots and listed rows (sigma-window sort, chunks of C, problematic
chunks as sell_c/build.hpp:147-155 defines them), and LAV's segments (columns
by descending in-degree, the prefix covering 80% of the entries in 512 Ki
column segments, lav/build.hpp:103-212).  SpMV through each format is checked
against the oracle on the same graphs."""
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


def _sell_stats(rp, C, sigma):
    """(stored slots, listed rows): rows sorted by length, descending and
    stable, inside windows of sigma rows (the trailing window only when it
    holds more than C rows); chunks of C sorted rows; rows longer than 4C are
    listed; a chunk whose padded size is at least twice its kept entries and
    which keeps more than 1024 entries stores nothing and lists its rows."""
    n = len(rp) - 1
    L = np.diff(np.asarray(rp, np.int64))
    win = np.arange(n) // sigma
    nwin = n // sigma
    sort_key = np.where((win < nwin) | (n % sigma > C), -L, 0)
    perm = np.lexsort((sort_key, np.minimum(win, nwin)))
    nch = -(-n // C)
    Ls = np.full(nch * C, -1, np.int64)
    Ls[:n] = L[perm]
    Ls = Ls.reshape(nch, C)
    long_thr = 4 * C
    is_row = Ls >= 0
    is_long = Ls > long_thr
    kept = np.where(is_row & ~is_long, Ls, 0)
    mx, tot = kept.max(axis=1), kept.sum(axis=1)
    prob = (mx * C >= 2 * tot) & (tot > 1024)
    padded = int((np.where(prob, 0, mx) * C).sum())
    listed = int(is_long.sum() + (is_row & ~is_long & prob[:, None]).sum())
    return padded, listed


@pytest.mark.parametrize("C,sigma", [(64, 512), (32, 512), (64, 1), (64, 100), (32, 4096)])
def test_sell_c_slots_match_definition(gbk, oracle, C, sigma):
    gs = {"rmat16": gbk.gen_rmat(16, 16, seed=76), "er12": gbk.gen_uniform(12, 8, seed=77)}
    gs.update(special_graphs())
    for name, (rp, col) in gs.items():
        n = len(rp) - 1
        vals = (np.arange(len(col)) % 13 + 1).astype(np.int32)
        A = gbk.Matrix.from_csr(rp, col, vals, fmt=6, sell_c=C, sigma=sigma)
        padded, listed, _, _ = A.format_stats()
        assert (padded, listed) == _sell_stats(rp, C, sigma), name
        x = (np.arange(n) % 29).astype(np.int32)
        w = gbk.Vector(n, gbk.INT32)
        gbk.mxv(w, None, gbk.ACC_NONE, "PlusMultiplies", A, gbk.Vector.dense(x))
        assert np.array_equal(w.to_numpy(), oracle.spmv(rp, col, vals, x, np.zeros(n, np.int32), 0, 0)), name
        A.free()


def _lav_stats(rp, col):
    n = len(rp) - 1
    col = np.asarray(col, np.int64)
    deg = np.bincount(col, minlength=n)
    n2o = np.argsort(-deg, kind="stable")
    o2n = np.empty(n, np.int64)
    o2n[n2o] = np.arange(n)
    sdeg = deg[n2o]
    excl = np.cumsum(sdeg) - sdeg
    K = int((excl < 0.8 * len(col)).sum())
    w = 512 * 1024
    ndense = -(-K // w)
    nseg = ndense + (1 if K < n else 0)
    c2 = o2n[col]
    key = np.where(c2 < K, c2 // w, ndense)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    return nseg, len(np.unique(key * n + rows))


@pytest.mark.parametrize("graph", ["rmat16", "er20", "rmat20"])
def test_lav_segments_match_definition(gbk, oracle, graph):
    rp, col = {"rmat16": lambda: gbk.gen_rmat(16, 16, seed=78), "er20": lambda: gbk.gen_uniform(20, 4, seed=79),
               "rmat20": lambda: gbk.gen_rmat(20, 8, seed=80)}[graph]()
    n = len(rp) - 1
    vals = (np.arange(len(col)) % 7 + 1).astype(np.int32)
    A = gbk.Matrix.from_csr(rp, col, vals, fmt=4)
    _, _, nseg, npairs = A.format_stats()
    assert (nseg, npairs) == _lav_stats(rp, col)
    x = (gbk.uniform_f32(81, n) * 1000).astype(np.int32)
    for sr in ("PlusMultiplies", "MinimumPlus"):
        add, mul = gbk.SEMIRINGS[sr]
        w = gbk.Vector(n, gbk.INT32)
        gbk.mxv(w, None, gbk.ACC_NONE, sr, A, gbk.Vector.dense(x))
        want = oracle.spmv(rp, col, vals, x, np.zeros(n, np.int32), add, mul)
        assert np.array_equal(w.to_numpy(), want), sr
    A.free()
