"""CPU checks of the C ABI library: it loads without a GPU, exports every
function include/gbk.h declares, and its host-side pieces (seeded graph
generation, value fillers, argument validation) behave as documented."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gbk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbk_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(gbk):
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(gbk.lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_library_is_gfx950():
    """The embedded device code objects are built for gfx950 only."""
    lib = os.path.join(ROOT, "gb_kun_amd", "libgbk.so")
    data = open(lib, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}, targets


def test_rmat_deterministic_and_simple(gbk):
    rp1, c1 = gbk.gen_rmat(12, 16, seed=5)
    rp2, c2 = gbk.gen_rmat(12, 16, seed=5)
    assert np.array_equal(rp1, rp2) and np.array_equal(c1, c2)
    rp3, c3 = gbk.gen_rmat(12, 16, seed=6)
    assert not (len(c3) == len(c1) and np.array_equal(c3, c1))
    n = len(rp1) - 1
    assert n == 4096 and rp1[0] == 0 and rp1[-1] == len(c1)
    rows = np.repeat(np.arange(n), np.diff(rp1))
    assert np.all(rows != c1)  # no self loops
    # strictly ascending columns inside each row: no duplicates
    same_row = rows[1:] == rows[:-1]
    assert np.all(c1[1:][same_row] > c1[:-1][same_row])
    assert 0 < len(c1) <= 16 * n


def test_rmat_skew_and_empty_rows(gbk):
    rp, col = gbk.gen_rmat(16, 16, seed=1)
    deg = np.diff(rp)
    # R-MAT (.57,.19,.19,.05) leaves many empty rows and a heavy tail
    assert 0.3 < np.mean(deg == 0) < 0.7
    assert deg.max() > 50 * deg.mean()


def test_symmetric_generator(gbk):
    rp, col = gbk.gen_rmat(11, 8, seed=3, symmetric=True)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    fwd = set(zip(rows.tolist(), col.tolist()))
    assert all((c, r) in fwd for r, c in fwd)


def test_uniform_generator(gbk):
    rp, col = gbk.gen_uniform(12, 16, seed=2)
    n = len(rp) - 1
    deg = np.diff(rp)
    assert abs(deg.mean() - 16) < 0.1 and deg.max() < 50


def test_value_fillers(gbk):
    u = gbk.uniform_f32(7, 100000)
    assert u.dtype == np.float32 and u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    assert np.array_equal(u, gbk.uniform_f32(7, 100000))
    rp, col = gbk.gen_rmat(10, 8, seed=4, symmetric=True)
    w = gbk.sym_weights_i32(rp, col, 11, 1, 255)
    assert w.min() >= 1 and w.max() <= 255
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    wd = {(r, c): x for r, c, x in zip(rows.tolist(), col.tolist(), w.tolist())}
    assert all(wd[(c, r)] == x for (r, c), x in wd.items())
    v = gbk.inv_coldeg_f32(n, col)
    cnt = np.bincount(col, minlength=n)
    assert np.allclose(v, 1.0 / cnt[col])


def test_error_codes_without_device(gbk):
    """Argument errors are reported as LA_Info codes before any device work."""
    h = C.c_void_p()
    assert gbk.lib.gbk_matrix_new(C.byref(h), 99) == gbk.DOMAIN_MISMATCH
    assert gbk.lib.gbk_matrix_new(None, 0) == gbk.NULL_POINTER
    assert gbk.lib.gbk_mxv(None, None, 0, gbk.Semiring(0, 0), None, None, None) == gbk.UNINITIALIZED_OBJECT
    assert b"uninitialized" in gbk.lib.gbk_last_error()
    assert gbk.lib.gbk_matrix_new(C.byref(h), gbk.FP32) == gbk.SUCCESS
    assert gbk.lib.gbk_matrix_set_format(h, 5, 0, 0) == gbk.INVALID_VALUE
    assert gbk.lib.gbk_matrix_set_format(h, gbk.FMT_SELL_C, 48, 0) == gbk.INVALID_VALUE
    rp = np.array([0, 1, 0], dtype=np.int64)  # not monotone
    col = np.array([1], dtype=np.int32)
    vals = np.array([1.0], dtype=np.float32)
    rc = gbk.lib.gbk_matrix_build_csr32(h, gbk._ptr(rp, C.c_int64), gbk._ptr(col, C.c_int32), gbk._ptr(vals), 2, 1)
    assert rc == gbk.INVALID_OBJECT
    gbk.lib.gbk_matrix_free(h)


def test_graph_batch_rejects_oversized_graph(gbk):
    """Validated on the host before any device call: a graph above the
    per-graph vertex limit, and a column id out of range."""
    n = (1 << 17) + 1
    rp = np.zeros(n + 1, dtype=np.int64)
    with pytest.raises(gbk.GbkError) as e:
        gbk.GraphBatch([(rp, np.zeros(0, dtype=np.int32))])
    assert e.value.code == gbk.INVALID_VALUE
    with pytest.raises(gbk.GbkError) as e:
        gbk.GraphBatch([(np.array([0, 1], dtype=np.int64), np.array([5], dtype=np.int32))])
    assert e.value.code == gbk.INDEX_OUT_OF_BOUNDS
