"""GPU parity of the pull SpMV (dense x) through the C ABI against the oracle.

Integer and boolean semirings must be bit-exact; fp32 plus.times / plus.second
must be within 1e-5 relative (to the row's sum of |terms|) of the FP64 oracle.
"""
import numpy as np
import pytest

from tests.graphs import special_graphs

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _graphs(gbk):
    gs = {"rmat10": gbk.gen_rmat(10, 16, seed=1), "er10": gbk.gen_uniform(10, 16, seed=2),
          "rmat12_sym": gbk.gen_rmat(12, 8, seed=3, symmetric=True)}
    gs.update(special_graphs())
    return gs


def _values(gbk, nnz, dtype, seed):
    u = gbk.uniform_f32(seed, nnz)
    if dtype == np.float32:
        return u
    if dtype == np.uint8:
        return (u < 0.7).astype(np.uint8)
    return (1 + (u * 255).astype(np.int64)).astype(dtype)


def _xvec(gbk, n, dtype, seed):
    u = gbk.uniform_f32(seed, n)
    if dtype == np.float32:
        return u
    if dtype == np.uint8:
        return (u < 0.5).astype(np.uint8)
    return (u * 1000).astype(dtype)


def _run_mxv(gbk, A, u_np, w_np, semiring, accum=0, mask=None, desc_mask=6, vxm=False, mask_sparse=False):
    u = gbk.Vector.dense(u_np)
    w = gbk.Vector.dense(w_np) if w_np is not None else gbk.Vector(len(u_np), u.type)
    m = None
    if mask is not None:
        if mask_sparse:
            ids = np.flatnonzero(mask)
            m = gbk.Vector.sparse(len(mask), ids, mask[ids]) if len(ids) else gbk.Vector(len(mask), gbk.TYPE_OF[mask.dtype])
        else:
            m = gbk.Vector.dense(mask)
    if vxm:
        gbk.vxm(w, m, accum, semiring, u, A, desc_mask)
    else:
        gbk.mxv(w, m, accum, semiring, A, u, desc_mask)
    storage = w.storage  # read first: to_numpy() makes the vector dense
    return w.to_numpy(), storage


EXACT_CASES = [
    ("PlusMultiplies", np.int32),
    ("MinimumPlus", np.int32),
    ("MinimumSelectSecond", np.int32),
    ("LogicalOrAnd", np.float32),
    ("StructuralBool", np.uint8),
    ("PlusOne", np.int64),
    ("MinimumNotEqualTo", np.int32),
    ("FirstWins", np.float32),
    ("FirstMin", np.int32),
    ("CustomLessPlus", np.float32),
    ("PlusMultiplies", np.int64),
    ("MinimumPlus", np.float32),
]


@pytest.mark.parametrize("semiring,dtype", EXACT_CASES)
@pytest.mark.parametrize("transposed", [False, True])
def test_exact_semirings(gbk, oracle, semiring, dtype, transposed):
    add, mul = gbk.SEMIRINGS[semiring]
    for name, (rp, col) in _graphs(gbk).items():
        n = len(rp) - 1
        vals = _values(gbk, len(col), dtype, 11)
        A = gbk.Matrix.from_csr(rp, col, vals)
        x = _xvec(gbk, A.n, dtype, 12)
        got, _ = _run_mxv(gbk, A, x, None, semiring, vxm=transposed)
        trp, tcol, tval = A.export(transposed=True) if transposed else (rp, col, vals)
        if A.n != n:
            trp = np.concatenate([trp, np.full(A.n - n, trp[-1])])
        want = oracle.spmv(trp, tcol, tval, x, np.zeros(A.n, dtype), add, mul)
        assert np.array_equal(got, want), f"{semiring} {dtype} {name} transposed={transposed}"


@pytest.mark.parametrize("semiring", ["PlusMultiplies", "PlusSecond"])
def test_fp32_plus_against_fp64(gbk, oracle, semiring):
    for name, (rp, col) in _graphs(gbk).items():
        vals = gbk.uniform_f32(21, len(col))
        A = gbk.Matrix.from_csr(rp, col, vals)
        x = gbk.uniform_f32(22, A.n)
        got, _ = _run_mxv(gbk, A, x, None, semiring)
        rp2 = np.concatenate([rp, np.full(A.n + 1 - len(rp), rp[-1])])
        v64 = vals if semiring == "PlusMultiplies" else np.ones_like(vals)
        want, absum = oracle.spmv_f64(rp2, col, v64, x)
        err = np.abs(got.astype(np.float64) - want)
        assert np.all(err <= REL_TOL * absum + 1e-30), f"{semiring} {name}: max rel {np.max(err / np.maximum(absum, 1e-30))}"


@pytest.mark.parametrize("accum", [2, 3, 6])  # plus, minus, min
def test_accumulators(gbk, oracle, accum):
    rp, col = gbk.gen_rmat(11, 16, seed=5)
    vals = _values(gbk, len(col), np.int32, 31)
    A = gbk.Matrix.from_csr(rp, col, vals)
    x = _xvec(gbk, A.n, np.int32, 32)
    w0 = _xvec(gbk, A.n, np.int32, 33) + 1
    got, _ = _run_mxv(gbk, A, x, w0, "PlusMultiplies", accum=accum)
    want = oracle.spmv(rp, col, vals, x, w0, 0, 0, accum=accum)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("desc_mask", [6, 4, 2, 5])  # DEFAULT, STRUCTURE, COMP, STR_COMP
@pytest.mark.parametrize("sparse_mask", [False, True])
def test_masked_pull(gbk, oracle, desc_mask, sparse_mask):
    rp, col = gbk.gen_rmat(11, 16, seed=6)
    vals = _values(gbk, len(col), np.int32, 41)
    A = gbk.Matrix.from_csr(rp, col, vals)
    n = A.n
    x = _xvec(gbk, n, np.int32, 42)
    w0 = _xvec(gbk, n, np.int32, 43) + 1
    # sparse masks must have <= 30% entries to stay sparse, dense ones more
    dens = 0.1 if sparse_mask else 0.6
    mask = (gbk.uniform_f32(44, n) < dens).astype(np.float32) * 3.0
    got, _ = _run_mxv(gbk, A, x, w0, "MinimumPlus", accum=0, mask=mask, desc_mask=desc_mask, mask_sparse=sparse_mask)
    if sparse_mask:
        want = oracle.spmv(rp, col, vals, x, w0, 2, 4, mask_ids=np.flatnonzero(mask), desc_mask=desc_mask)
    else:
        want = oracle.spmv(rp, col, vals, x, w0, 2, 4, mask=mask, desc_mask=desc_mask)
    assert np.array_equal(got, want)


def test_storage_after_mxv(gbk, oracle):
    """convert_if_required: dense iff nvals > 0.3 n (vector.h:117-129)."""
    rp, col = gbk.gen_rmat(10, 4, seed=7)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    x = (gbk.uniform_f32(51, A.n) < 0.9).astype(np.float32)
    got, storage = _run_mxv(gbk, A, x, None, "LogicalOrAnd")
    nv = int(np.count_nonzero(got))
    assert storage == (2 if nv > int(A.n * 0.3) else 1)


def test_spmv_device_repeat_deterministic(gbk, oracle):
    """The raw-pointer seam: repeated calls give bit-identical fp32 results."""
    from gb_kun_amd.device import DeviceArray
    rp, col = gbk.gen_rmat(13, 16, seed=8)
    vals = gbk.uniform_f32(61, len(col))
    A = gbk.Matrix.from_csr(rp, col, vals)
    xh = gbk.uniform_f32(62, A.n)
    x = DeviceArray.from_numpy(xh)
    y = DeviceArray.zeros(A.n, np.float32)
    outs = []
    for _ in range(3):
        gbk.spmv_device(A, x.ptr, y.ptr)
        gbk.synchronize()
        outs.append(y.numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    want, absum = oracle.spmv_f64(rp, col, vals, xh)
    assert np.all(np.abs(outs[0] - want) <= REL_TOL * absum + 1e-30)
