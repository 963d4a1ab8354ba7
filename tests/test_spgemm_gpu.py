"""SpGEMM on the GPU: the masked product C<M> = A (+.*) B against the oracle's
masked IKJ product (pinned by dense products), the unmasked product against
the oracle's unmasked IKJ product (pinned by scipy, A I = A and brute-force
min.plus) bit for bit, fp32 included (the device keeps the IKJ fold order),
and triangle counting against trace(A^3) / 6 computed by scipy."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _csr_of(S):
    S = S.tocsr()
    S.sort_indices()
    return S.indptr.astype(np.int64), S.indices.astype(np.int32)


def _graph_pair(gbk, seed):
    rp, col = gbk.gen_rmat(11, 8, seed=seed)
    n = len(rp) - 1
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=0.01, random_state=seed, format="csr")
    mrp, mcol = _csr_of(M)
    # shuffle the mask's columns inside every row (the GPU sorts them)
    mcol = mcol.copy()
    for i in range(0, n, 7):
        seg = mcol[mrp[i]:mrp[i + 1]]
        rng.shuffle(seg)
    return rp, col, mrp, mcol, n


@pytest.mark.parametrize("dt,sr,exact", [(np.int32, "PlusMultiplies", True), (np.int32, "MinimumPlus", True),
                                         (np.int64, "PlusMultiplies", True), (np.float32, "PlusMultiplies", False),
                                         (np.uint8, "LogicalOrAnd", True)])
def test_mxm_masked_matches_oracle(gbk, oracle, dt, sr, exact):
    rp, col, mrp, mcol, n = _graph_pair(gbk, 41)
    brp, bcol = gbk.gen_rmat(11, 8, seed=42)
    rng = np.random.default_rng(43)
    if dt == np.uint8:
        av = np.ones(len(col), np.uint8)
        bv = np.ones(len(bcol), np.uint8)
    elif dt == np.float32:
        av = rng.random(len(col)).astype(np.float32)
        bv = rng.random(len(bcol)).astype(np.float32)
    else:
        av = rng.integers(1, 100, len(col)).astype(dt)
        bv = rng.integers(1, 100, len(bcol)).astype(dt)
    A = gbk.Matrix.from_csr(rp, col, av)
    B = gbk.Matrix.from_csr(brp, bcol, bv)
    Mm = gbk.Matrix.from_csr(mrp, mcol, np.ones(len(mcol), np.int32))
    Cm = gbk.Matrix(gbk.TYPE_OF[np.dtype(dt)])
    gbk.mxm(Cm, Mm, gbk.ACC_NONE, sr, A, B)
    crp, ccol, cval = Cm.export()
    assert np.array_equal(crp, mrp) and np.array_equal(ccol, mcol)  # C has M's entries, in M's order
    add, mul = gbk.SEMIRINGS[sr]
    want = oracle.mxm_masked(mrp, mcol, rp, col, av, brp, bcol, bv, add, mul)
    if exact:
        assert np.array_equal(cval, want)
    else:
        assert np.allclose(cval, want, rtol=1e-5, atol=1e-6)
    assert np.count_nonzero(want) > 100  # the mask meets the product


def test_mxm_rejects_complemented_mask_and_aliasing(gbk):
    rp, col = gbk.gen_rmat(8, 4, seed=44)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.int32))
    Cm = gbk.Matrix(gbk.INT32)
    with pytest.raises(gbk.GbkError, match="complementary mask"):
        gbk.mxm(Cm, A, gbk.ACC_NONE, "PlusMultiplies", A, A, desc_mask=gbk.STR_COMP)
    with pytest.raises(gbk.GbkError, match="aliases"):
        gbk.mxm(A, None, gbk.ACC_NONE, "PlusMultiplies", A, A)


@pytest.mark.parametrize("kind,scale", [("rmat", 12), ("rmat", 14), ("er", 12)])
def test_tricount_all_methods_equal_trace(gbk, kind, scale):
    rp, col = (gbk.gen_rmat(scale, 16, seed=45, symmetric=True) if kind == "rmat"
               else gbk.gen_uniform(scale, 16, seed=46, symmetric=True))
    n = len(rp) - 1
    S = sp.csr_matrix((np.ones(len(col), np.int64), col, rp), shape=(n, n))
    want = int((S @ S @ S).diagonal().sum()) // 6
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.int32))
    got = [gbk.tricount(A, m) for m in range(7)]
    assert got == [want] * 7 and want > 0


def test_tricount_small_cases(gbk):
    # K4 has 4 triangles; a bipartite graph has none
    k4 = [(i, j) for i in range(4) for j in range(4) if i != j]
    bip = [(i, j) for i in range(3) for j in range(3, 6)]
    bip += [(j, i) for i, j in bip]
    for edges, want in ((k4, 4), (bip, 0)):
        A = gbk.Matrix(gbk.INT32)
        r = np.array([e[0] for e in edges]); c = np.array([e[1] for e in edges])
        A.build_coo(r, c, np.ones(len(edges), np.int32))
        assert [gbk.tricount(A, m) for m in range(7)] == [want] * 7


def _vals(dt, m, rng, signed=False):
    if dt == np.uint8:
        return (rng.random(m) < 0.8).astype(np.uint8)
    if dt in (np.float32, np.float64):
        return (rng.random(m) - (0.5 if signed else 0.0)).astype(dt)
    lo = -3 if signed else 1
    return rng.integers(lo, 50 if not signed else 4, m).astype(dt)


@pytest.mark.parametrize("dt,sr", [(np.int32, "PlusMultiplies"), (np.int32, "MinimumPlus"),
                                   (np.int64, "PlusMultiplies"), (np.float32, "PlusMultiplies"),
                                   (np.float64, "PlusMultiplies"), (np.uint8, "LogicalOrAnd")])
def test_mxm_unmasked_matches_oracle(gbk, oracle, dt, sr):
    """C = A (+.*) B without a mask: structure, columns ascending and values
    equal to the oracle's bit for bit; signed values make some sums cancel,
    and those entries are dropped on both sides."""
    add, mul = gbk.SEMIRINGS[sr]
    rng = np.random.default_rng(51)
    cases = [(gbk.gen_rmat(10, 8, seed=52), gbk.gen_rmat(10, 8, seed=53)),
             (gbk.gen_uniform(11, 4, seed=54), gbk.gen_uniform(11, 4, seed=55))]
    for (arp, acol), (brp, bcol) in cases:
        signed = sr == "PlusMultiplies" and dt != np.uint8
        av = _vals(dt, len(acol), rng, signed)
        bv = _vals(dt, len(bcol), rng, signed)
        A = gbk.Matrix.from_csr(arp, acol, av)
        B = gbk.Matrix.from_csr(brp, bcol, bv)
        Cm = gbk.Matrix(gbk.TYPE_OF[np.dtype(dt)])
        gbk.mxm(Cm, None, gbk.ACC_NONE, sr, A, B)
        crp, ccol, cval = Cm.export()
        wrp, wcol, wval = oracle.mxm_unmasked(arp, acol, av, brp, bcol, bv, add, mul)
        assert np.array_equal(crp, wrp) and np.array_equal(ccol, wcol), f"{dt.__name__} {sr}"
        assert np.array_equal(cval, wval), f"{dt.__name__} {sr}"
        assert len(wcol) > 1000
        for M in (A, B, Cm):
            M.free()


def test_mxm_unmasked_edge_cases(gbk, oracle):
    """An empty operand, a product with no terms, a result that replaces a
    built C, and A (+.*) A on the special graphs (empty rows, a hub)."""
    from tests.graphs import special_graphs
    n = 64
    Z = gbk.Matrix.from_csr(np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.int32))
    rp, col = gbk.gen_rmat(6, 4, seed=56)
    A = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.int32))
    Cm = gbk.Matrix(gbk.INT32)
    gbk.mxm(Cm, None, gbk.ACC_NONE, "PlusMultiplies", A, A)
    first = Cm.export()
    assert first[0][-1] > 0
    for X, Y in ((Z, A), (A, Z), (Z, Z)):
        gbk.mxm(Cm, None, gbk.ACC_NONE, "PlusMultiplies", X, Y)
        crp, ccol, _ = Cm.export()
        assert crp[-1] == 0 and len(ccol) == 0 and len(crp) == n + 1
    for name, (grp, gcol) in special_graphs().items():
        if len(gcol) == 0:
            continue
        v = (np.arange(len(gcol)) % 5 + 1).astype(np.int32)
        G = gbk.Matrix.from_csr(grp, gcol, v)
        gbk.mxm(Cm, None, gbk.ACC_NONE, "PlusMultiplies", G, G)
        crp, ccol, cval = Cm.export()
        wrp, wcol, wval = oracle.mxm_unmasked(grp, gcol, v, grp, gcol, v, 0, 0)
        assert np.array_equal(crp, wrp) and np.array_equal(ccol, wcol) and np.array_equal(cval, wval), name
        G.free()
