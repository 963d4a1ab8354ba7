"""GPU parity of the optimized formats (SELL-C-sigma, segmented CSR, LAV).

Every format must give the CSR result: bit-exact for integer semirings; for
fp32 plus.times every format is within 1e-5 relative of the FP64 oracle.
SELL-C folds every row of at most 4C entries sequentially in CSR order (its
chunk rows and the rows of problematic chunks), so on a graph without longer
rows it reproduces the reference's CSR bits (checked on the golden vector of
the compiled reference, whose longest row has 212 entries).
"""
import numpy as np
import pytest

from tests.graphs import special_graphs

pytestmark = pytest.mark.gpu

FORMATS = [("sell64", 6, 64, 512), ("sell32", 6, 32, 512), ("sell64_s1", 6, 64, 1), ("seg", 3, 0, 0),
           ("lav", 4, 0, 0)]


def _graphs(gbk):
    gs = {"rmat12": gbk.gen_rmat(12, 16, seed=31), "er12": gbk.gen_uniform(12, 8, seed=32),
          "rmat16": gbk.gen_rmat(16, 16, seed=33)}
    gs.update(special_graphs())
    return gs


@pytest.mark.parametrize("name,fmt,C,sigma", FORMATS)
def test_format_int_exact(gbk, oracle, name, fmt, C, sigma):
    for gname, (rp, col) in _graphs(gbk).items():
        vals = (1 + (gbk.uniform_f32(41, len(col)) * 255).astype(np.int64)).astype(np.int32)
        A = gbk.Matrix.from_csr(rp, col, vals, fmt=fmt, sell_c=C, sigma=sigma)
        n = A.n
        x = (gbk.uniform_f32(42, n) * 1000).astype(np.int32)
        for sr in ("PlusMultiplies", "MinimumPlus"):
            add, mul = gbk.SEMIRINGS[sr]
            w = gbk.Vector(n, gbk.INT32)
            gbk.mxv(w, None, gbk.ACC_NONE, sr, A, gbk.Vector.dense(x))
            rp2 = np.concatenate([rp, np.full(n + 1 - len(rp), rp[-1])])
            want = oracle.spmv(rp2, col, vals, x, np.zeros(n, np.int32), add, mul)
            assert np.array_equal(w.to_numpy(), want), f"{name} {gname} {sr}"
        if fmt != 4:  # vxm through the transposed representation (LAV has none, as in the reference)
            w = gbk.Vector(n, gbk.INT32)
            gbk.vxm(w, None, gbk.ACC_NONE, "PlusMultiplies", gbk.Vector.dense(x), A)
            trp, tcol, tval = A.export(transposed=True)
            want = oracle.spmv(trp, tcol, tval, x, np.zeros(n, np.int32), 0, 0)
            assert np.array_equal(w.to_numpy(), want), f"{name} {gname} vxm"


@pytest.mark.parametrize("name,fmt,C,sigma", FORMATS)
def test_format_fp32(gbk, oracle, name, fmt, C, sigma):
    for gname, (rp, col) in _graphs(gbk).items():
        vals = gbk.uniform_f32(51, len(col))
        A = gbk.Matrix.from_csr(rp, col, vals, fmt=fmt, sell_c=C, sigma=sigma)
        n = A.n
        x = gbk.uniform_f32(52, n)
        w0 = gbk.uniform_f32(53, n)
        w = gbk.Vector.dense(w0)
        gbk.mxv(w, None, gbk.ACC_PLUS, "PlusMultiplies", A, gbk.Vector.dense(x))
        got = w.to_numpy()
        rp2 = np.concatenate([rp, np.full(n + 1 - len(rp), rp[-1])])
        y64, absum = oracle.spmv_f64(rp2, col, vals, x)
        err = np.abs(got.astype(np.float64) - (w0.astype(np.float64) + y64))
        assert np.all(err <= 1e-5 * (absum + np.abs(w0)) + 1e-30), f"{name} {gname}"


def test_sell_c_matches_reference_csr_bits(gbk):
    """Golden: the compiled reference's fp32 CSR SpMV, reproduced bit for bit by SELL-C."""
    import os
    d = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "spmv_rmat9_f32.npz")))
    A = gbk.Matrix.from_csr(d["rp"], d["col"], d["vals"], fmt=6, sell_c=64, sigma=512)
    w = gbk.Vector(A.n, gbk.FP32)
    gbk.mxv(w, None, gbk.ACC_NONE, "PlusMultiplies", A, gbk.Vector.dense(d["x"]))
    assert np.array_equal(w.to_numpy(), d["y_plus_times_csr"])


def test_problematic_chunks_not_identity(gbk, oracle):
    """A chunk padded 2x with > 1024 entries goes to the row kernel; its rows
    get the real fold (the reference returns the identity, spmv_sell_c.h:133-138)."""
    from tests.graphs import csr_from_edges
    n = 256
    # five full rows (1280 entries) in a 64-row chunk whose other rows hold one entry
    edges = [(r, c) for r in range(5) for c in range(n)] + [(r, (r * 7) % n) for r in range(5, 64)]
    rp, col = csr_from_edges(n, edges)
    vals = np.ones(len(col), np.int32)
    A = gbk.Matrix.from_csr(rp, col, vals, fmt=6, sell_c=64, sigma=512)
    assert A.format_stats()[1] > 0, "expected a problematic chunk"
    x = np.arange(n, dtype=np.int32)
    w = gbk.Vector(n, gbk.INT32)
    gbk.mxv(w, None, gbk.ACC_NONE, "PlusMultiplies", A, gbk.Vector.dense(x))
    assert np.array_equal(w.to_numpy(), oracle.spmv(rp, col, vals, x, np.zeros(n, np.int32), 0, 0))


def _seg_case(gbk, oracle, rp, col, vals, x, sr, exact=True):
    A = gbk.Matrix.from_csr(rp, col, vals, fmt=3)
    n = A.n
    add, mul = gbk.SEMIRINGS[sr]
    w = gbk.Vector(n, gbk.TYPE_OF[vals.dtype])
    gbk.mxv(w, None, gbk.ACC_NONE, sr, A, gbk.Vector.dense(x))
    rp2 = np.concatenate([rp, np.full(n + 1 - len(rp), rp[-1])])
    want = oracle.spmv(rp2, col, vals, x, np.zeros(n, vals.dtype), add, mul)
    got = w.to_numpy()
    A.free()
    return got, want


def test_seg_hot_columns_every_type(gbk, oracle):
    """The segmented format's hot-column kernel (x of a segment's highest
    in-degree columns in LDS) for every value type; order-sensitive monoids
    take the sequential path.  R-MAT 19 has more columns per segment (65,536)
    than LDS slots for 4- and 8-byte values (40,960 and 20,480), so hot and
    L2 columns mix in every tile."""
    graphs = {"rmat19": gbk.gen_rmat(19, 8, seed=61), "rmat12": gbk.gen_rmat(12, 16, seed=62)}
    graphs.update(special_graphs())
    for gname, (rp, col) in graphs.items():
        nnz = len(col)
        n = len(rp) - 1
        u = gbk.uniform_f32(63, nnz)
        ux = gbk.uniform_f32(64, n)
        for dt, sems in ((np.int32, ("PlusMultiplies", "MinimumPlus", "CustomLessPlus", "MinimumSelectSecond",
                                     "PlusOne", "MinimumNotEqualTo")),
                         (np.int64, ("PlusMultiplies", "MinimumPlus")),
                         (np.uint8, ("LogicalOrAnd", "StructuralBool"))):
            if dt == np.uint8:
                vals = (u < 0.7).astype(np.uint8)
                x = (ux < 0.3).astype(np.uint8)
            else:
                vals = (1 + (u * 255).astype(np.int64)).astype(dt)
                x = (ux * 1000).astype(dt)
            for sr in sems:
                got, want = _seg_case(gbk, oracle, rp, col, vals, x, sr)
                assert np.array_equal(got, want), f"{gname} {dt.__name__} {sr}"
        for dt in (np.float32, np.float64):
            vals = u.astype(dt)
            x = ux.astype(dt)
            got, _ = _seg_case(gbk, oracle, rp, col, vals, x, "PlusMultiplies")
            rp2 = np.concatenate([rp, np.full(n + 1 - len(rp), rp[-1])])
            y64, absum = oracle.spmv_f64(rp2, col, vals.astype(np.float32), x.astype(np.float32))
            # fp64 inputs are fp32-representable, so fp64 products are exact and
            # the bound is fp64 summation rounding
            tol = 1e-5 if dt == np.float32 else 1e-10
            err = np.abs(got.astype(np.float64) - y64)
            assert np.all(err <= tol * absum + 1e-30), f"{gname} {dt.__name__}"
