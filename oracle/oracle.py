"""ctypes access to the CPU oracle (liboracle.so, built from gbk_oracle.c) and,
where it was built, to the compiled reference (oracle/_ref/libgbkref*.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

F32, F64, I32, I64, BOOL = 0, 1, 2, 3, 4
NP = {F32: np.float32, F64: np.float64, I32: np.int32, I64: np.int64, BOOL: np.uint8}
TYPE_OF = {np.dtype(v): k for k, v in NP.items()}


def build():
    """Compiles gbk_oracle.c into liboracle.so (plain gcc, no fast-math: IEEE fp32 like the reference)."""
    src = os.path.join(HERE, "gbk_oracle.c")
    tmp = LIB_PATH + ".tmp"
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=gnu11", src, "-o", tmp, "-lm"])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_ref():
    """Compiles the reference bridge (only where the reference sources exist)."""
    subprocess.check_call(["bash", os.path.join(HERE, "build_ref.sh")])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_spmv.restype = C.c_int
        _lib.orc_spmspv.restype = C.c_int
        _lib.orc_pagerank_f64.restype = C.c_int
        _lib.orc_pagerank_f64_q.restype = C.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _csr(rp, col):
    return np.ascontiguousarray(rp, dtype=np.int64), np.ascontiguousarray(col, dtype=np.int32)


def spmv(rp, col, val, x, y, add, mul, accum=0, mask=None, mask_ids=None, desc_mask=6, ref_threads=0):
    """y <- accum(y, A (+.*) x) restating SpMV_all_active_diff_vectors / SpMV_dense / SpMV_sparse.

    mask: dense mask array (any type) or None; mask_ids: sparse mask ids (int32) or None.
    Returns the new y (a copy)."""
    rp, col = _csr(rp, col)
    val = np.ascontiguousarray(val)
    t = TYPE_OF[val.dtype]
    x = np.ascontiguousarray(x, dtype=val.dtype)
    y = np.array(y, dtype=val.dtype, copy=True)
    n = len(rp) - 1
    kind, mt, mids, mn = 0, 0, None, 0
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        kind, mt = 1, TYPE_OF[mask.dtype]
    elif mask_ids is not None:
        mids = np.ascontiguousarray(mask_ids, dtype=np.int32)
        kind, mn = 2, len(mids)
    rc = lib().orc_spmv(t, C.c_int64(n), _p(rp), _p(col), _p(val), _p(x), _p(y), add, mul, accum, kind, _p(mask), mt,
                        _p(mids), C.c_int64(mn), desc_mask, ref_threads)
    assert rc == 0
    return y


def spmspv(rp, col, val, x_ids, x_vals, y, add, mul, accum=0, mask=None, mask_ids=None, desc_mask=6):
    """SpMSpV + apply_mask restated sequentially; rp/col/val are the rows pushed along."""
    rp, col = _csr(rp, col)
    val = np.ascontiguousarray(val)
    t = TYPE_OF[val.dtype]
    xi = np.ascontiguousarray(x_ids, dtype=np.int32)
    xv = np.ascontiguousarray(x_vals, dtype=val.dtype)
    y = np.array(y, dtype=val.dtype, copy=True)
    n = len(rp) - 1
    kind, mt, mids, mn = 0, 0, None, 0
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        kind, mt = 1, TYPE_OF[mask.dtype]
    elif mask_ids is not None:
        mids = np.ascontiguousarray(mask_ids, dtype=np.int32)
        kind, mn = 2, len(mids)
    rc = lib().orc_spmspv(t, C.c_int64(n), _p(rp), _p(col), _p(val), _p(xi), _p(xv), C.c_int64(len(xi)), _p(y), add,
                          mul, accum, kind, _p(mask), mt, _p(mids), C.c_int64(mn), desc_mask)
    assert rc == 0
    return y


def spmv_f64(rp, col, val, x):
    """FP64 plus.times of fp32 data: (y, sum |terms|) per row."""
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    y = np.empty(n, dtype=np.float64)
    a = np.empty(n, dtype=np.float64)
    lib().orc_spmv_f64_of_f32(C.c_int64(n), _p(rp), _p(col), _p(np.ascontiguousarray(val, dtype=np.float32)),
                              _p(np.ascontiguousarray(x, dtype=np.float32)), _p(y), _p(a))
    return y, a


def balance_rows(rp, nthreads):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    first = np.empty(nthreads, dtype=np.int64)
    last = np.empty(nthreads, dtype=np.int64)
    lib().orc_balance_rows(C.c_int64(len(rp) - 1), _p(rp), nthreads, _p(first), _p(last))
    return first, last


def spmv_f32_omp(rp, col, val, x, y, nthreads, first, last):
    """The CPU baseline kernel (OpenMP, reference partition); writes y in place."""
    lib().orc_spmv_plus_times_f32_omp(C.c_int64(len(rp) - 1), _p(rp), _p(col), _p(val), _p(x), _p(y), nthreads,
                                      _p(first), _p(last))


def bfs(rp, col, source):
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    lv = np.empty(n, dtype=np.int32)
    lib().orc_bfs(C.c_int64(n), _p(rp), _p(col), C.c_int64(source), _p(lv))
    return lv


def dijkstra_i32(rp, col, w, source):
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    d = np.empty(n, dtype=np.int32)
    lib().orc_dijkstra_i32(C.c_int64(n), _p(rp), _p(col), _p(np.ascontiguousarray(w, dtype=np.int32)),
                           C.c_int64(source), _p(d))
    return d


def dijkstra_f32(rp, col, w, source):
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    d = np.empty(n, dtype=np.float32)
    lib().orc_dijkstra_f32(C.c_int64(n), _p(rp), _p(col), _p(np.ascontiguousarray(w, dtype=np.float32)),
                           C.c_int64(source), _p(d))
    return d


def cc(rp, col):
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    lab = np.empty(n, dtype=np.int32)
    lib().orc_cc(C.c_int64(n), _p(rp), _p(col), _p(lab))
    return lab


def pagerank_f64(rp, col, itermax=100, damping=0.85, tol=0.0, sink_quirk=False):
    """FP64 PageRank with sinks; sink_quirk restates the reference's sparse-mask sink sum."""
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    r = np.empty(n, dtype=np.float64)
    it = lib().orc_pagerank_f64_q(C.c_int64(n), _p(rp), _p(col), itermax, C.c_double(damping), C.c_double(tol),
                                  _p(r), int(sink_quirk))
    return r, it


# ------------------------------------------------------------- reference
_ref = None
_ref_cc = None


def ppr_f64(rp, col, val, source=None, p=None, itermax=100, damping=0.85, tol=0.0):
    """FP64 personalized PageRank (orc_ppr_f64): teleport to `source` (one-hot) or to the vector p."""
    rp, col = _csr(rp, col)
    n = len(rp) - 1
    val = np.ascontiguousarray(val, dtype=np.float32)
    if p is None:
        p = np.zeros(n, dtype=np.float64)
        p[source] = 1.0
    p = np.ascontiguousarray(p, dtype=np.float64)
    r = np.empty(n, dtype=np.float64)
    it = lib().orc_ppr_f64(C.c_int64(n), _p(rp), _p(col), _p(val), _p(p), C.c_int(itermax), C.c_double(damping),
                           C.c_double(tol), _p(r))
    return r, it


def have_ref():
    return os.path.exists(os.path.join(REF_DIR, "libgbkref.so"))


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(os.path.join(REF_DIR, "libgbkref.so"))
    return _ref


def ref_cc_lib():
    global _ref_cc
    if _ref_cc is None:
        _ref_cc = C.CDLL(os.path.join(REF_DIR, "libgbkref_cc.so"))
    return _ref_cc


REF_SEMIRING = {"PlusMultiplies": 0, "LogicalOrAnd": 1, "MinimumPlus": 2, "MinimumSelectSecond": 3,
                "PlusSecond": 4, "StructuralBool": 5}


def ref_mxv(rp, col, val, u, semiring, accum=0, vxm=False, fmt=0, w=None, mask=None, desc_mask=6):
    """Runs lablas::mxv / vxm of the compiled reference.

    u, w, mask: None, ("dense", array) or ("sparse", ids, vals).  accum: 0 NULL, 1 plus, 2 minimum.
    Returns (w dense values, storage 1 sparse / 2 dense, LA_Info)."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col64 = np.ascontiguousarray(col, dtype=np.int64)
    val = np.ascontiguousarray(val)
    dt = val.dtype
    n = len(rp) - 1
    f = ref().ref_mxv_f32 if dt == np.float32 else ref().ref_mxv_i32

    def unpack(v):
        if v is None:
            return 0, None, None, None, 0
        if v[0] == "dense":
            return 1, np.ascontiguousarray(v[1], dtype=dt), None, None, 0
        ids = np.ascontiguousarray(v[1], dtype=np.int64)
        return 2, None, ids, np.ascontiguousarray(v[2], dtype=dt), len(ids)

    uk, ud, ui, uv, un = unpack(u)
    wk, wd, wi, wv, wn = unpack(w)
    mk, md, mi, mv, mn = unpack(mask)
    out = np.empty(n, dtype=dt)
    st = C.c_int(0)
    rc = f(int(vxm), C.c_int64(n), _p(rp), _p(col64), _p(val), fmt, REF_SEMIRING[semiring], accum,
           uk, _p(ud), _p(ui), _p(uv), C.c_int64(un), wk, _p(wd), _p(wi), _p(wv), C.c_int64(wn),
           mk, _p(md), _p(mi), _p(mv), C.c_int64(mn), desc_mask, _p(out), C.byref(st))
    return out, st.value, rc


def ref_bfs(rp, col, source):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col64 = np.ascontiguousarray(col, dtype=np.int64)
    n = len(rp) - 1
    lv = np.empty(n, dtype=np.float32)
    rc = ref().ref_bfs(C.c_int64(n), _p(rp), _p(col64), C.c_int64(source), _p(lv))
    assert rc == 0
    return lv


def mxm_masked(mrp, mcol, arp, acol, aval, brp, bcol, bval, add, mul):
    """C<M> = A (+.*) B (the masked IKJ product): values in the mask's entry order."""
    mrp, mcol = _csr(mrp, mcol)
    arp, acol = _csr(arp, acol)
    brp, bcol = _csr(brp, bcol)
    aval = np.ascontiguousarray(aval)
    bval = np.ascontiguousarray(bval, dtype=aval.dtype)
    n = len(mrp) - 1
    out = np.zeros(max(len(mcol), 1), dtype=aval.dtype)
    rc = lib().orc_mxm_masked(TYPE_OF[aval.dtype], C.c_int64(n), _p(mrp), _p(mcol), _p(arp), _p(acol), _p(aval),
                              _p(brp), _p(bcol), _p(bval), add, mul, _p(out))
    assert rc == 0
    return out[:len(mcol)]


def mxm_unmasked(arp, acol, aval, brp, bcol, bval, add, mul):
    """C = A (+.*) B without a mask (the unmasked IKJ product): (row_ptr, col, val),
    columns ascending, entries whose value is 0 dropped."""
    arp, acol = _csr(arp, acol)
    brp, bcol = _csr(brp, bcol)
    aval = np.ascontiguousarray(aval)
    bval = np.ascontiguousarray(bval, dtype=aval.dtype)
    n = len(arp) - 1
    cap = int(np.diff(brp)[acol].sum()) if len(acol) else 0
    crp = np.zeros(n + 1, dtype=np.int64)
    ccol = np.zeros(max(cap, 1), dtype=np.int32)
    cval = np.zeros(max(cap, 1), dtype=aval.dtype)
    f = lib().orc_mxm_unmasked
    f.restype = C.c_int64
    nz = int(f(TYPE_OF[aval.dtype], C.c_int64(n), _p(arp), _p(acol), _p(aval), _p(brp), _p(bcol), _p(bval),
               add, mul, _p(crp), _p(ccol), _p(cval), C.c_int64(cap)))
    assert 0 <= nz <= cap
    return crp, ccol[:nz], cval[:nz]


def masked_count(mrp, mcol, arp, acol, brp, bcol):
    """Sum over the entries of C<M> = A (plus.one) B."""
    mrp, mcol = _csr(mrp, mcol)
    arp, acol = _csr(arp, acol)
    brp, bcol = _csr(brp, bcol)
    f = lib().orc_masked_count
    f.restype = C.c_int64
    return int(f(C.c_int64(len(mrp) - 1), _p(mrp), _p(mcol), _p(arp), _p(acol), _p(brp), _p(bcol)))


def ref_read_mtx(path):
    """The compiled reference's Matrix::init_from_mtx: (row_ptr, col, seconds) of its CSR."""
    nr, nnz, sec = C.c_int64(0), C.c_int64(0), C.c_double(0)
    rc = ref().ref_mtx_load(os.fsencode(path), C.byref(nr), C.byref(nnz), C.byref(sec))
    assert rc == 0
    rp = np.empty(nr.value + 1, dtype=np.int64)
    col = np.empty(max(nnz.value, 1), dtype=np.int64)
    assert ref().ref_mtx_csr(_p(rp), _p(col)) == 0
    return rp, col[:nnz.value], sec.value


def ref_sssp(rp, col, w, source):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col64 = np.ascontiguousarray(col, dtype=np.int64)
    n = len(rp) - 1
    d = np.empty(n, dtype=np.float32)
    rc = ref().ref_sssp(C.c_int64(n), _p(rp), _p(col64), _p(np.ascontiguousarray(w, dtype=np.float32)),
                        C.c_int64(source), _p(d))
    assert rc == 0
    return d


def ref_pagerank(rp, col, itermax):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col64 = np.ascontiguousarray(col, dtype=np.int64)
    n = len(rp) - 1
    r = np.empty(n, dtype=np.float32)
    it = ref().ref_pagerank(C.c_int64(n), _p(rp), _p(col64), itermax, _p(r))
    return r, it


def ref_cc(rp, col):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col64 = np.ascontiguousarray(col, dtype=np.int64)
    n = len(rp) - 1
    lab = np.empty(n, dtype=np.int32)
    rc = ref_cc_lib().ref_cc(C.c_int64(n), _p(rp), _p(col64), _p(lab))
    assert rc == 0
    return lab


def ref_set_threads(t):
    return ref().ref_set_threads(t)
