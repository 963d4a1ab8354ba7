"""Writes the golden fixtures of tests/golden/*.npz by running the COMPILED
reference (oracle/_ref/libgbkref*.so, built by oracle/build_ref.sh from
/root/reference) on small seeded graphs.

Each fixture holds the inputs (CSR, values, vectors) and the reference's
outputs, so the tests need neither the reference nor the generator to check
against them.  Run in the container that has /root/reference:

    python tests/golden/make_golden.py

The reference writes perf_stats.txt into the working directory, so the
script runs it from a temporary directory.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
import gb_kun_amd as g  # noqa: E402

CSR_F, SELL_C_F, SEG_F, LAV_F = 0, 6, 3, 4


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def spmv_fixtures():
    rp, col = g.gen_rmat(9, 16, seed=101)
    n = len(rp) - 1
    vals = g.uniform_f32(102, len(col))
    x = g.uniform_f32(103, n)
    w0 = g.uniform_f32(104, n)
    mdense = (g.uniform_f32(105, n) < 0.6).astype(np.float32) * 2.0
    msparse_ids = np.flatnonzero(g.uniform_f32(106, n) < 0.1).astype(np.int64)
    xs_ids = np.flatnonzero(g.uniform_f32(107, n) < 0.05).astype(np.int64)
    xs_vals = g.uniform_f32(108, len(xs_ids)) + 0.5
    out = dict(rp=rp, col=col, vals=vals, x=x, w0=w0, mdense=mdense, msparse_ids=msparse_ids, xs_ids=xs_ids,
               xs_vals=xs_vals)
    cases = []

    def run(tag, **kw):
        y, st, rc = O.ref_mxv(rp, col, vals, **kw)
        assert rc == 0, (tag, rc)
        out["y_" + tag] = y
        out["st_" + tag] = np.int32(st)
        cases.append(tag)

    for fmt_name, fmt in (("csr", CSR_F), ("sellc", SELL_C_F), ("seg", SEG_F), ("lav", LAV_F)):
        run(f"plus_times_{fmt_name}", u=("dense", x), semiring="PlusMultiplies", fmt=fmt)
    run("plus_times_vxm", u=("dense", x), semiring="PlusMultiplies", vxm=True)
    run("lor_land", u=("dense", x), semiring="LogicalOrAnd")
    run("min_plus", u=("dense", x), semiring="MinimumPlus")
    run("plus_second_accum_plus", u=("dense", x), w=("dense", w0), semiring="PlusSecond", accum=1)
    run("mask_dense_default", u=("dense", x), w=("dense", w0), semiring="PlusMultiplies",
        mask=("dense", mdense))
    run("mask_dense_strcomp", u=("dense", x), w=("dense", w0), semiring="PlusMultiplies",
        mask=("dense", mdense), desc_mask=5)
    run("mask_sparse_default", u=("dense", x), w=("dense", w0), semiring="PlusMultiplies",
        mask=("sparse", msparse_ids, np.ones(len(msparse_ids), np.float32)))
    run("mask_sparse_strcomp", u=("dense", x), w=("dense", w0), semiring="PlusMultiplies",
        mask=("sparse", msparse_ids, np.ones(len(msparse_ids), np.float32)), desc_mask=5)
    run("spmspv_plus", u=("sparse", xs_ids, xs_vals), semiring="PlusMultiplies")
    run("spmspv_plus_vxm", u=("sparse", xs_ids, xs_vals), semiring="PlusMultiplies", vxm=True)
    run("spmspv_or_masked_strcomp", u=("sparse", xs_ids, xs_vals), w=("dense", w0), semiring="LogicalOrAnd",
        mask=("dense", mdense), desc_mask=5, vxm=True)
    run("spmspv_min_plus", u=("sparse", xs_ids, xs_vals), semiring="MinimumPlus")
    run("spmspv_plus_mask_sparse", u=("sparse", xs_ids, xs_vals), w=("dense", w0), semiring="PlusMultiplies",
        mask=("sparse", msparse_ids, np.ones(len(msparse_ids), np.float32)), accum=1)
    out["cases"] = np.array(cases)
    save("spmv_rmat9_f32", **out)

    # int32 semirings
    vi = (1 + (g.uniform_f32(111, len(col)) * 255).astype(np.int32)).astype(np.int32)
    xi = (g.uniform_f32(112, n) * 1000).astype(np.int32)
    outi = dict(rp=rp, col=col, vals=vi, x=xi)
    casesi = []
    for sr in ("PlusMultiplies", "MinimumPlus", "MinimumSelectSecond"):
        for vxm in (False, True):
            tag = f"{sr}_{'vxm' if vxm else 'mxv'}"
            y, st, rc = O.ref_mxv(rp, col, vi, u=("dense", xi), semiring=sr, vxm=vxm)
            assert rc == 0
            outi["y_" + tag] = y
            casesi.append(tag)
    outi["cases"] = np.array(casesi)
    save("spmv_rmat9_i32", **outi)


def trailing_row_fixture():
    """8 x 8, rows 6 and 7 empty, nnz = 8 = threads: the reference partition never visits rows 6-7."""
    rows = [0, 0, 1, 2, 3, 4, 5, 5]
    cols = [1, 2, 3, 0, 4, 5, 0, 1]
    rp = np.zeros(9, dtype=np.int64)
    for r in rows:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array(cols, dtype=np.int32)
    vals = np.ones(8, dtype=np.float32)
    x = np.arange(1, 9, dtype=np.float32)
    w0 = np.full(8, 7.0, dtype=np.float32)
    O.ref_set_threads(8)
    y, st, rc = O.ref_mxv(rp, col, vals, u=("dense", x), w=("dense", w0), semiring="PlusMultiplies")
    assert rc == 0
    save("trailing_rows_8x8", rp=rp, col=col, vals=vals, x=x, w0=w0, y=y, threads=np.int32(8))


def algorithm_fixtures():
    rp, col = g.gen_rmat(10, 8, seed=201, symmetric=True)
    deg = np.diff(rp)
    sources = np.flatnonzero(deg > 0)[[0, 7, 31]]
    levels = np.stack([O.ref_bfs(rp, col, int(s)) for s in sources])
    save("bfs_rmat10_sym", rp=rp, col=col, sources=sources, levels=levels)

    w = g.sym_weights_i32(rp, col, 202, 1, 255)
    dists = np.stack([O.ref_sssp(rp, col, w.astype(np.float32), int(s)) for s in sources])
    save("sssp_rmat10_sym", rp=rp, col=col, w=w, sources=sources, dist=dists)

    # several components: two R-MAT blocks and isolated vertices
    rp1, col1 = g.gen_rmat(8, 2, seed=203, symmetric=True)
    rp2, col2 = g.gen_rmat(8, 2, seed=204, symmetric=True)
    n1 = len(rp1) - 1
    rpc = np.concatenate([rp1, rp2[1:] + rp1[-1], np.full(16, rp1[-1] + rp2[-1])]).astype(np.int64)
    colc = np.concatenate([col1, col2 + n1]).astype(np.int32)
    labels = O.ref_cc(rpc, colc)
    save("cc_two_blocks", rp=rpc, col=colc, labels=labels)

    rpd, cold = g.gen_rmat(10, 16, seed=205)
    ranks, iters = O.ref_pagerank(rpd, cold, 20)
    save("pr_rmat10", rp=rpd, col=cold, ranks=ranks, iters=np.int32(iters))


if __name__ == "__main__":
    if not O.have_ref():
        O.build_ref()
    os.chdir(tempfile.mkdtemp(prefix="gbk_golden_"))
    spmv_fixtures()
    trailing_row_fixture()
    algorithm_fixtures()
