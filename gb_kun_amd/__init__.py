"""Python binding of libgbk.so, the MI355X (gfx950) semiring SpMV / SpMSpV library.

The binding is ctypes over the C ABI declared in include/gbk.h; it adds numpy
conversion and object lifetimes, nothing else.  Every computation runs in the
HIP kernels of libgbk.so: there is no CPU fallback, and importing this package
fails if the library has not been built (python -m gb_kun_amd.build_lib, or
__graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgbk.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run python -m gb_kun_amd.build_lib")

lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------- constants
SUCCESS, NO_VALUE, UNINITIALIZED_OBJECT, INVALID_OBJECT, NULL_POINTER = 0, 1, 2, 3, 4
INVALID_VALUE, INVALID_INDEX, DOMAIN_MISMATCH, DIMENSION_MISMATCH = 5, 6, 7, 8
OUT_OF_MEMORY, INDEX_OUT_OF_BOUNDS, PANIC = 10, 12, 13

FP32, FP64, INT32, INT64, BOOL = 0, 1, 2, 3, 4
NP_TYPES = {FP32: np.float32, FP64: np.float64, INT32: np.int32, INT64: np.int64, BOOL: np.uint8}
TYPE_OF = {np.dtype(v): k for k, v in NP_TYPES.items()}

ADD_PLUS, ADD_LOR, ADD_MIN, ADD_MAX, ADD_FIRST, ADD_SECOND, ADD_LESS, ADD_FIRSTMIN = range(8)
(MUL_TIMES, MUL_LAND, MUL_FIRST, MUL_SECOND, MUL_PLUS, MUL_ONEB, MUL_NE, MUL_MINUS,
 MUL_DIV) = range(9)
(ACC_NONE, ACC_FIRST, ACC_PLUS, ACC_MINUS, ACC_TIMES, ACC_DIV, ACC_MIN, ACC_MAX, ACC_LOR,
 ACC_LAND, ACC_LESS, ACC_NE) = range(12)

SCMP, REPLACE, COMP, TRAN, STRUCTURE, STR_COMP, DEFAULT = 0, 1, 2, 3, 4, 5, 6
FMT_CSR, FMT_CSR_SEG, FMT_LAV, FMT_SELL_C = 0, 3, 4, 6
STORAGE_SPARSE, STORAGE_DENSE = 1, 2


class Semiring(C.Structure):
    _fields_ = [("add", C.c_int), ("mul", C.c_int)]


class Desc(C.Structure):
    _fields_ = [("mask", C.c_int), ("output", C.c_int)]


# the semirings registered by the reference (src/cpp_graphblas/types.hpp:274-284)
SEMIRINGS = {
    "PlusMultiplies": (ADD_PLUS, MUL_TIMES),
    "LogicalOrAnd": (ADD_LOR, MUL_LAND),
    "FirstWins": (ADD_FIRST, MUL_TIMES),
    "FirstMin": (ADD_FIRSTMIN, MUL_TIMES),
    "PlusSecond": (ADD_PLUS, MUL_SECOND),
    "StructuralBool": (ADD_LOR, MUL_ONEB),
    "MinimumPlus": (ADD_MIN, MUL_PLUS),
    "CustomLessPlus": (ADD_LESS, MUL_PLUS),
    "MinimumSelectSecond": (ADD_MIN, MUL_SECOND),
    "MinimumNotEqualTo": (ADD_MIN, MUL_NE),
    "PlusOne": (ADD_PLUS, MUL_ONEB),
}

_vp = C.c_void_p
_i64 = C.c_int64
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("gbk_init", C.c_int, C.c_int)
_sig("gbk_set_stream", C.c_int, _vp)
_sig("gbk_get_stream", _vp)
_sig("gbk_synchronize", C.c_int)
_sig("gbk_last_error", C.c_char_p)
_sig("gbk_device_info", C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int)
_sig("gbk_matrix_new", C.c_int, C.POINTER(_vp), C.c_int)
_sig("gbk_matrix_free", C.c_int, _vp)
_sig("gbk_matrix_set_format", C.c_int, _vp, C.c_int, C.c_int, C.c_int)
_sig("gbk_matrix_get_format", C.c_int, _vp, C.POINTER(C.c_int))
_sig("gbk_matrix_build_coo", C.c_int, _vp, _pi64, _pi64, _vp, _i64)
_sig("gbk_matrix_read_mtx", C.c_int, _vp, C.c_char_p)
_sig("gbk_mxm", C.c_int, _vp, _vp, C.c_int, Semiring, _vp, _vp, C.POINTER(Desc))
_sig("gbk_tricount", C.c_int, _vp, C.c_int, C.POINTER(C.c_uint64))
_sig("gbk_mtx_read", C.c_int, C.c_char_p, _pi64, _pi64, _pi64, C.POINTER(_pi64), C.POINTER(_pi64))
_sig("gbk_mtx_write", C.c_int, C.c_char_p, _i64, _i64, _pi64, _pi64, _i64)
_sig("gbk_matrix_build_csr", C.c_int, _vp, _pi64, _pi64, _vp, _i64, _i64)
_sig("gbk_matrix_build_csr32", C.c_int, _vp, _pi64, _pi32, _vp, _i64, _i64)
_sig("gbk_matrix_nrows", C.c_int, _vp, _pi64)
_sig("gbk_matrix_nvals", C.c_int, _vp, _pi64)
_sig("gbk_matrix_export", C.c_int, _vp, C.c_int, _pi64, _pi32, _vp)
_sig("gbk_matrix_degrees", C.c_int, _vp, C.c_int, _pi64)
_sig("gbk_matrix_device_bytes", C.c_int, _vp, _pi64)
_sig("gbk_matrix_format_stats", C.c_int, _vp, C.c_int, _pi64)
_sig("gbk_matrix_autotune", C.c_int, _vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double))
_sig("gbk_vector_new", C.c_int, C.POINTER(_vp), C.c_int, _i64)
_sig("gbk_vector_free", C.c_int, _vp)
_sig("gbk_vector_nvals", C.c_int, _vp, _pi64)
_sig("gbk_vector_storage", C.c_int, _vp, C.POINTER(C.c_int))
_sig("gbk_vector_build_dense", C.c_int, _vp, _vp, _i64)
_sig("gbk_vector_build_sparse", C.c_int, _vp, _pi64, _vp, _i64)
_sig("gbk_vector_fill", C.c_int, _vp, C.c_double)
_sig("gbk_vector_fill_int", C.c_int, _vp, _i64)
_sig("gbk_vector_set_element", C.c_int, _vp, C.c_double, _i64)
_sig("gbk_vector_fill_ascending", C.c_int, _vp, _i64)
_sig("gbk_vector_dup", C.c_int, _vp, _vp)
_sig("gbk_vector_swap", C.c_int, _vp, _vp)
_sig("gbk_vector_clear", C.c_int, _vp)
_sig("gbk_vector_force_dense", C.c_int, _vp)
_sig("gbk_vector_force_sparse", C.c_int, _vp)
_sig("gbk_vector_convert_if_required", C.c_int, _vp)
_sig("gbk_vector_to_host", C.c_int, _vp, _vp)
_sig("gbk_vector_sparse_to_host", C.c_int, _vp, _pi64, _vp, _pi64)
_sig("gbk_vector_device_dense", C.c_int, _vp, C.POINTER(_vp))
_sig("gbk_mxv", C.c_int, _vp, _vp, C.c_int, Semiring, _vp, _vp, C.POINTER(Desc))
_sig("gbk_vxm", C.c_int, _vp, _vp, C.c_int, Semiring, _vp, _vp, C.POINTER(Desc))
_sig("gbk_spmv_device", C.c_int, _vp, C.c_int, _vp, _vp, Semiring, C.c_int)
_sig("gbk_assign_scalar", C.c_int, _vp, _vp, C.c_int, C.c_double, C.POINTER(Desc))
_sig("gbk_assign_vector", C.c_int, _vp, _vp, C.c_int, _vp, C.POINTER(Desc))
_sig("gbk_assign_scatter", C.c_int, _vp, _vp, C.c_int, _vp, _vp, _i64, C.POINTER(Desc))
_sig("gbk_ewise_add", C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(Desc))
_sig("gbk_ewise_mult", C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.POINTER(Desc))
_sig("gbk_apply", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double, _vp, C.POINTER(Desc))
_sig("gbk_reduce", C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, _vp, C.POINTER(Desc))
_sig("gbk_reduce_i64", C.c_int, _pi64, C.c_int, C.c_int, _vp, C.POINTER(Desc))
_sig("gbk_extract", C.c_int, _vp, _vp, C.c_int, _vp, _vp, C.POINTER(Desc))
_sig("gbk_bfs", C.c_int, _vp, _i64, _pi32, _pi64)
_sig("gbk_bfs_device", C.c_int, _vp, _i64, _vp, _pi64)
_sig("gbk_sssp_device", C.c_int, _vp, _i64, _vp, _pi64)
_sig("gbk_sssp", C.c_int, _vp, _i64, _vp, _pi64)
_sig("gbk_cc", C.c_int, _vp, _pi32, _pi64)
_sig("gbk_pagerank", C.c_int, _vp, C.POINTER(C.c_float), C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int))
_sig("gbk_pagerank_device", C.c_int, _vp, _vp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int))
_sig("gbk_spmm_device", C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, Semiring, C.c_int)
_sig("gbk_msbfs", C.c_int, _vp, _pi64, C.c_int, _pi32, _pi64)
_sig("gbk_msbfs_device", C.c_int, _vp, _pi64, C.c_int, _vp, _pi64)
_sig("gbk_ppr_multi", C.c_int, _vp, _pi64, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_double, C.c_double,
     C.POINTER(C.c_int))
_sig("gbk_ppr_multi_device", C.c_int, _vp, _pi64, C.c_int, _vp, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int))
_sig("gbk_sssp_multi", C.c_int, _vp, _pi64, C.c_int, _pi32, _pi64)
_sig("gbk_graph_batch_new", C.c_int, C.POINTER(_vp), C.c_int, _pi64, C.POINTER(_pi64), C.POINTER(_pi32))
_sig("gbk_graph_batch_free", C.c_int, _vp)
_sig("gbk_graph_batch_bfs", C.c_int, _vp, _pi64, _pi32, C.POINTER(C.c_double))
_sig("gbk_graph_batch_info", C.c_int, _vp, _pi64)
_sig("gbk_server_new", C.c_int, C.POINTER(_vp), C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int)
_sig("gbk_server_free", C.c_int, _vp)
_sig("gbk_server_add_graph", C.c_int, _vp, _vp, _vp, _vp, C.POINTER(C.c_int))
_sig("gbk_server_submit", C.c_int, _vp, C.c_int, C.c_int, C.c_int64, C.c_double, _pi64)
_sig("gbk_server_run", C.c_int, _vp)
_sig("gbk_server_query_info", C.c_int, _vp, C.c_int64, C.POINTER(C.c_double), _pi64)
_sig("gbk_server_result", C.c_int, _vp, C.c_int64, _vp)
_sig("gbk_server_stats", C.c_int, _vp, C.POINTER(C.c_double))
_sig("gbk_gen_rmat", C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
     C.c_int, _pi64, _pi64, C.POINTER(_pi64), C.POINTER(_pi32))
_sig("gbk_gen_uniform", C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, _pi64, _pi64, C.POINTER(_pi64),
     C.POINTER(_pi32))
_sig("gbk_host_free", None, _vp)
_sig("gbk_fill_uniform_f32", None, C.c_uint64, _i64, _vp)
_sig("gbk_fill_inv_coldeg_f32", None, _i64, _i64, _pi32, _vp)
_sig("gbk_fill_sym_weights_i32", None, _i64, _pi64, _pi32, C.c_uint64, C.c_int32, C.c_int32, _vp)


class GbkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gbk error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != SUCCESS:
        raise GbkError(rc, lib.gbk_last_error().decode(errors="replace"))
    return rc


def _ptr(a, ctype=None):
    if a is None:
        return None
    if ctype is None:
        return a.ctypes.data_as(_vp)
    return a.ctypes.data_as(C.POINTER(ctype))


def device_info():
    cu = C.c_int(0)
    name = C.create_string_buffer(128)
    check(lib.gbk_device_info(C.byref(cu), name, 128))
    return cu.value, name.value.decode()


def synchronize():
    check(lib.gbk_synchronize())


# ------------------------------------------------------------- generators
def _take_csr(n, nnz, rp, col):
    rp_np = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    col_np = np.ctypeslib.as_array(col, shape=(max(nnz.value, 1),))[: nnz.value].copy()
    lib.gbk_host_free(C.cast(rp, _vp))
    lib.gbk_host_free(C.cast(col, _vp))
    return rp_np, col_np


def mtx_read(path):
    """(nrows, ncols, rows int64, cols int64) of a .mtx / .mtxbin file, 0-based, file order."""
    nr, nc, nnz = _i64(0), _i64(0), _i64(0)
    r, c = _pi64(), _pi64()
    check(lib.gbk_mtx_read(os.fsencode(path), C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(r), C.byref(c)))
    try:
        k = nnz.value
        rows = np.ctypeslib.as_array(r, shape=(k,)).copy() if k else np.zeros(0, np.int64)
        cols = np.ctypeslib.as_array(c, shape=(k,)).copy() if k else np.zeros(0, np.int64)
    finally:
        lib.gbk_host_free(C.cast(r, C.c_void_p))
        lib.gbk_host_free(C.cast(c, C.c_void_p))
    return nr.value, nc.value, rows, cols


def mtx_write(path, nrows, ncols, rows, cols):
    """Writes a .mtx (pattern) or .mtxbin file; rows / cols 0-based."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    check(lib.gbk_mtx_write(os.fsencode(path), nrows, ncols, _ptr(rows, C.c_int64), _ptr(cols, C.c_int64), len(rows)))


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, symmetric=False, permute=True):
    """Seeded R-MAT graph as a CSR (row_ptr int64[n+1], col int32[nnz]): duplicates and self-loops removed."""
    n, nnz = _i64(0), _i64(0)
    rp, col = _pi64(), _pi32()
    check(lib.gbk_gen_rmat(scale, edge_factor, a, b, c, seed, int(symmetric), int(permute), C.byref(n),
                           C.byref(nnz), C.byref(rp), C.byref(col)))
    return _take_csr(n, nnz, rp, col)


def gen_uniform(scale, edge_factor=16, seed=1, symmetric=False):
    """Seeded Erdos-Renyi (uniform) graph as a CSR, duplicates and self-loops removed."""
    n, nnz = _i64(0), _i64(0)
    rp, col = _pi64(), _pi32()
    check(lib.gbk_gen_uniform(scale, edge_factor, seed, int(symmetric), C.byref(n), C.byref(nnz), C.byref(rp),
                              C.byref(col)))
    return _take_csr(n, nnz, rp, col)


def uniform_f32(seed, count):
    out = np.empty(count, dtype=np.float32)
    lib.gbk_fill_uniform_f32(seed, count, _ptr(out))
    return out


def inv_coldeg_f32(n, col):
    col = np.ascontiguousarray(col, dtype=np.int32)
    out = np.empty(len(col), dtype=np.float32)
    lib.gbk_fill_inv_coldeg_f32(n, len(col), _ptr(col, C.c_int32), _ptr(out))
    return out


def sym_weights_i32(row_ptr, col, seed, lo=1, hi=255):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    out = np.empty(len(col), dtype=np.int32)
    lib.gbk_fill_sym_weights_i32(len(row_ptr) - 1, _ptr(row_ptr, C.c_int64), _ptr(col, C.c_int32), seed, lo, hi,
                                 _ptr(out))
    return out


# ---------------------------------------------------------------- objects
class Matrix:
    """Device matrix (CSR + CSC in HBM), lablas::Matrix<T> counterpart."""

    def __init__(self, type_=FP32):
        self.h = _vp()
        self.type = type_
        check(lib.gbk_matrix_new(C.byref(self.h), type_))

    @classmethod
    def from_csr(cls, row_ptr, col, vals, fmt=None, sell_c=0, sigma=0):
        vals = np.ascontiguousarray(vals)
        m = cls(TYPE_OF[vals.dtype])
        if fmt is not None:
            m.set_format(fmt, sell_c, sigma)
        m.build_csr(row_ptr, col, vals)
        return m

    @classmethod
    def from_mtx(cls, path, type_=None, fmt=None):
        """Matrix::init_from_mtx: a .mtx or .mtxbin file, every entry 1."""
        m = cls(FP32 if type_ is None else type_)
        if fmt is not None:
            m.set_format(fmt)
        check(lib.gbk_matrix_read_mtx(m.h, os.fsencode(path)))
        return m

    def set_format(self, fmt, sell_c=0, sigma=0):
        check(lib.gbk_matrix_set_format(self.h, fmt, sell_c, sigma))

    def build_csr(self, row_ptr, col, vals):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=NP_TYPES[self.type])
        nrows = len(row_ptr) - 1
        if np.asarray(col).dtype == np.int32:
            col = np.ascontiguousarray(col)
            check(lib.gbk_matrix_build_csr32(self.h, _ptr(row_ptr, C.c_int64), _ptr(col, C.c_int32), _ptr(vals),
                                             nrows, len(col)))
        else:
            col = np.ascontiguousarray(col, dtype=np.int64)
            check(lib.gbk_matrix_build_csr(self.h, _ptr(row_ptr, C.c_int64), _ptr(col, C.c_int64), _ptr(vals),
                                           nrows, len(col)))

    def build_coo(self, rows, cols, vals):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=NP_TYPES[self.type])
        check(lib.gbk_matrix_build_coo(self.h, _ptr(rows, C.c_int64), _ptr(cols, C.c_int64), _ptr(vals), len(rows)))

    @property
    def n(self):
        v = _i64(0)
        check(lib.gbk_matrix_nrows(self.h, C.byref(v)))
        return v.value

    @property
    def nnz(self):
        v = _i64(0)
        check(lib.gbk_matrix_nvals(self.h, C.byref(v)))
        return v.value

    def export(self, transposed=False):
        n, nnz = self.n, self.nnz
        rp = np.empty(n + 1, dtype=np.int64)
        col = np.empty(max(nnz, 1), dtype=np.int32)
        vals = np.empty(max(nnz, 1), dtype=NP_TYPES[self.type])
        check(lib.gbk_matrix_export(self.h, int(transposed), _ptr(rp, C.c_int64), _ptr(col, C.c_int32), _ptr(vals)))
        return rp, col[:nnz], vals[:nnz]

    def device_bytes(self):
        v = _i64(0)
        check(lib.gbk_matrix_device_bytes(self.h, C.byref(v)))
        return v.value

    def format_stats(self, transposed=False):
        """(padded SELL-C entries, problematic-chunk rows, segments, segment-row pairs)."""
        st = np.zeros(4, dtype=np.int64)
        check(lib.gbk_matrix_format_stats(self.h, int(transposed), _ptr(st, C.c_int64)))
        return tuple(int(v) for v in st)

    def autotune(self, trials=5):
        """Times CSR, segmented CSR, LAV, SELL-C (C=64, C=32) and keeps the fastest.
        Returns (format, C, {name: microseconds})."""
        f, c = C.c_int(0), C.c_int(0)
        t = (C.c_double * 5)()
        check(lib.gbk_matrix_autotune(self.h, trials, C.byref(f), C.byref(c), t))
        names = ["csr", "csr_seg", "lav", "sell_c64", "sell_c32"]
        return f.value, c.value, dict(zip(names, list(t)))

    def free(self):
        if self.h:
            lib.gbk_matrix_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Vector:
    """Device vector with the reference's sparse/dense switch (lablas::Vector<T>)."""

    def __init__(self, n, type_=FP32):
        self.h = _vp()
        self.type = type_
        self.n = n
        check(lib.gbk_vector_new(C.byref(self.h), type_, n))

    @classmethod
    def dense(cls, vals):
        vals = np.ascontiguousarray(vals)
        v = cls(len(vals), TYPE_OF[vals.dtype])
        check(lib.gbk_vector_build_dense(v.h, _ptr(vals), len(vals)))
        return v

    @classmethod
    def sparse(cls, n, ids, vals):
        vals = np.ascontiguousarray(vals)
        v = cls(n, TYPE_OF[vals.dtype])
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        check(lib.gbk_vector_build_sparse(v.h, _ptr(ids, C.c_int64), _ptr(vals), len(ids)))
        return v

    def fill(self, value):
        check(lib.gbk_vector_fill(self.h, float(value)))

    def set_element(self, value, pos):
        check(lib.gbk_vector_set_element(self.h, float(value), pos))

    @property
    def nvals(self):
        v = _i64(0)
        check(lib.gbk_vector_nvals(self.h, C.byref(v)))
        return v.value

    @property
    def storage(self):
        s = C.c_int(0)
        check(lib.gbk_vector_storage(self.h, C.byref(s)))
        return s.value

    def to_numpy(self):
        out = np.empty(self.n, dtype=NP_TYPES[self.type])
        check(lib.gbk_vector_to_host(self.h, _ptr(out)))
        return out

    def sparse_to_numpy(self):
        ids = np.empty(max(self.n, 1), dtype=np.int64)
        vals = np.empty(max(self.n, 1), dtype=NP_TYPES[self.type])
        nv = _i64(0)
        check(lib.gbk_vector_sparse_to_host(self.h, _ptr(ids, C.c_int64), _ptr(vals), C.byref(nv)))
        return ids[: nv.value], vals[: nv.value]

    def force_dense(self):
        check(lib.gbk_vector_force_dense(self.h))

    def force_sparse(self):
        check(lib.gbk_vector_force_sparse(self.h))

    def free(self):
        if self.h:
            lib.gbk_vector_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _desc(mask=DEFAULT, output=DEFAULT):
    return C.byref(Desc(mask, output))


def mxv(w, mask, accum, semiring, A, u, desc_mask=DEFAULT):
    add, mul = SEMIRINGS[semiring] if isinstance(semiring, str) else semiring
    check(lib.gbk_mxv(w.h, mask.h if mask is not None else None, accum, Semiring(add, mul), A.h, u.h,
                      _desc(desc_mask)))


def vxm(w, mask, accum, semiring, u, A, desc_mask=DEFAULT):
    add, mul = SEMIRINGS[semiring] if isinstance(semiring, str) else semiring
    check(lib.gbk_vxm(w.h, mask.h if mask is not None else None, accum, Semiring(add, mul), u.h, A.h,
                      _desc(desc_mask)))


def mxm(Cm, mask, accum, semiring, A, B, desc_mask=DEFAULT):
    """C<mask> = A (+.*) B (structural mask: C gets the mask's entries; mask None:
    the unmasked product, entries whose fold is 0 dropped)."""
    add, mul = SEMIRINGS[semiring] if isinstance(semiring, str) else semiring
    check(lib.gbk_mxm(Cm.h, mask.h if mask is not None else None, accum, Semiring(add, mul), A.h, B.h,
                      _desc(desc_mask)))


def tricount(A, method=0):
    """Triangle count of a symmetric matrix (LAGr_TriangleCount methods 0-6)."""
    v = C.c_uint64(0)
    check(lib.gbk_tricount(A.h, method, C.byref(v)))
    return v.value


def spmv_device(A, x_ptr, y_ptr, semiring=(ADD_PLUS, MUL_TIMES), accum=ACC_NONE, transposed=False):
    """y = accum(y, A (+.*) x) on raw device pointers (e.g. gb_kun_amd.device.DeviceArray.ptr)."""
    add, mul = SEMIRINGS[semiring] if isinstance(semiring, str) else semiring
    check(lib.gbk_spmv_device(A.h, int(transposed), _vp(x_ptr), _vp(y_ptr), Semiring(add, mul), accum))


# ------------------------------------------------------------ graph drivers
def bfs(A, source):
    """Direction-optimizing BFS: int32 levels (1-based, 0 unreached) and stats
    (levels, push steps, pull steps, edges traversed)."""
    lv = np.empty(A.n, dtype=np.int32)
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_bfs(A.h, source, _ptr(lv, C.c_int32), _ptr(st, C.c_int64)))
    return lv, st[:4]


def bfs_device(A, source, levels_ptr):
    """BFS writing int32 levels to a device array of n entries; returns the stats."""
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_bfs_device(A.h, source, _vp(levels_ptr), _ptr(st, C.c_int64)))
    return st[:4]


def sssp(A, source):
    """min.plus SSSP distances in the matrix type (INT32: INT32_MAX unreached, FP32: FLT_MAX)."""
    d = np.empty(A.n, dtype=NP_TYPES[A.type])
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_sssp(A.h, source, _ptr(d), _ptr(st, C.c_int64)))
    return d, st[:2]


def sssp_device(A, source, dist_ptr):
    """SSSP writing n distances of the matrix type to a device array; returns the stats."""
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_sssp_device(A.h, source, _vp(dist_ptr), _ptr(st, C.c_int64)))
    return st[:2]


def cc(A):
    """Connected-component labels (minimum vertex id of each component) of a symmetric matrix."""
    lab = np.empty(A.n, dtype=np.int32)
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_cc(A.h, _ptr(lab, C.c_int32), _ptr(st, C.c_int64)))
    return lab, st[:1]


def pagerank(A, itermax=100, damping=0.85, tol=0.0):
    """PageRank with sinks (fp32); returns (ranks, iterations run)."""
    r = np.empty(A.n, dtype=np.float32)
    it = C.c_int(0)
    check(lib.gbk_pagerank(A.h, _ptr(r, C.c_float), itermax, damping, tol, C.byref(it)))
    return r, it.value


def pagerank_device(A, ranks_ptr, itermax=100, damping=0.85, tol=0.0):
    """PageRank writing n float ranks to a device array; returns the iterations run."""
    it = C.c_int(0)
    check(lib.gbk_pagerank_device(A.h, _vp(ranks_ptr), itermax, damping, tol, C.byref(it)))
    return it.value


# ------------------------------------------------------------- batched ops
def spmm_device(A, X_ptr, Y_ptr, k, semiring=(ADD_PLUS, MUL_TIMES), accum=ACC_NONE, transposed=False):
    """Y = accum(Y, A (+.*) X) for k right-hand sides stored row-major (n x k) on the device."""
    add, mul = SEMIRINGS[semiring] if isinstance(semiring, str) else semiring
    check(lib.gbk_spmm_device(A.h, int(transposed), _vp(X_ptr), _vp(Y_ptr), int(k), Semiring(add, mul), accum))


def msbfs(A, sources):
    """Multi-source BFS (<= 64 sources): int32 levels[len(sources), n] (1-based, 0 unreached) and stats
    (levels, push steps, pull steps, frontier edges, pairs reached)."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    lv = np.empty((len(src), A.n), dtype=np.int32)
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_msbfs(A.h, _ptr(src, C.c_int64), len(src), _ptr(lv, C.c_int32), _ptr(st, C.c_int64)))
    return lv, st[:5]


def msbfs_device(A, sources, levels_ptr):
    """Multi-source BFS writing int32 levels[len(sources), n] to a device array; returns the stats."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_msbfs_device(A.h, _ptr(src, C.c_int64), len(src), _vp(levels_ptr), _ptr(st, C.c_int64)))
    return st[:5]


def ppr_multi(A, sources, itermax=100, damping=0.85, tol=1e-6):
    """Personalized PageRank for <= 64 sources on an FP32 transition matrix: ranks[len(sources), n] and the
    iterations each column ran."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    r = np.empty((len(src), A.n), dtype=np.float32)
    it = np.zeros(len(src), dtype=np.int32)
    check(lib.gbk_ppr_multi(A.h, _ptr(src, C.c_int64), len(src), _ptr(r, C.c_float), itermax, damping, tol,
                            _ptr(it, C.c_int)))
    return r, it


def ppr_multi_device(A, sources, ranks_ptr, itermax=100, damping=0.85, tol=1e-6):
    """Personalized PageRank writing ranks[len(sources), n] (float) to a device array; returns the iterations
    each column ran."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    it = np.zeros(len(src), dtype=np.int32)
    check(lib.gbk_ppr_multi_device(A.h, _ptr(src, C.c_int64), len(src), _vp(ranks_ptr), itermax, damping, tol,
                                   _ptr(it, C.c_int)))
    return it


def sssp_multi(A, sources):
    """min.plus SSSP for <= 64 sources on an INT32 weight matrix: dist[len(sources), n] (INT32_MAX unreached)
    and stats (rounds, relaxations)."""
    src = np.ascontiguousarray(sources, dtype=np.int64)
    d = np.empty((len(src), A.n), dtype=np.int32)
    st = np.zeros(8, dtype=np.int64)
    check(lib.gbk_sssp_multi(A.h, _ptr(src, C.c_int64), len(src), _ptr(d, C.c_int32), _ptr(st, C.c_int64)))
    return d, st[:2]


class GraphBatch:
    """Many small graphs packed for one persistent, HIP-graph-replayed BFS launch."""

    def __init__(self, graphs):
        """graphs: list of (row_ptr int64[n+1], col int32[nnz]) with local ids."""
        self._keep = [(np.ascontiguousarray(rp, dtype=np.int64), np.ascontiguousarray(c, dtype=np.int32))
                      for rp, c in graphs]
        G = len(self._keep)
        self.nverts = np.array([len(rp) - 1 for rp, _ in self._keep], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.nverts)]).astype(np.int64)
        rps = (_pi64 * G)(*[_ptr(rp, C.c_int64) for rp, _ in self._keep])
        cols = (_pi32 * G)(*[_ptr(c, C.c_int32) for _, c in self._keep])
        self.h = _vp()
        check(lib.gbk_graph_batch_new(C.byref(self.h), G, _ptr(self.nverts, C.c_int64), rps, cols))

    def bfs(self, sources):
        """Levels of every graph (a list of int32 arrays) and the device time of the launch in ms."""
        src = np.ascontiguousarray(sources, dtype=np.int64)
        out = np.empty(int(self.offsets[-1]), dtype=np.int32)
        ms = C.c_double(0)
        check(lib.gbk_graph_batch_bfs(self.h, _ptr(src, C.c_int64), _ptr(out, C.c_int32), C.byref(ms)))
        return [out[self.offsets[g]:self.offsets[g + 1]] for g in range(len(self.nverts))], ms.value

    def info(self):
        i = np.zeros(4, dtype=np.int64)
        check(lib.gbk_graph_batch_info(self.h, _ptr(i, C.c_int64)))
        return {"graphs": int(i[0]), "vertices": int(i[1]), "edges": int(i[2]), "workgroups": int(i[3])}

    def free(self):
        if self.h:
            lib.gbk_graph_batch_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------ query server
QUERY_BFS, QUERY_PPR, QUERY_SSSP = 0, 1, 2


class Server:
    """Continuous-batching query server over device-resident graphs (gbk_server_*)."""

    def __init__(self, hbm_budget_bytes, ppr_slots=32, sssp_slots=32, damping=0.85, tol=1e-6, itermax=100,
                 keep_results=False):
        self.h = _vp()
        check(lib.gbk_server_new(C.byref(self.h), int(hbm_budget_bytes), ppr_slots, sssp_slots, damping, tol,
                                 itermax, int(keep_results)))
        self.graph_n = []

    def add_graph(self, bfs=None, ppr=None, sssp=None):
        gid = C.c_int(0)
        check(lib.gbk_server_add_graph(self.h, bfs.h if bfs else None, ppr.h if ppr else None,
                                       sssp.h if sssp else None, C.byref(gid)))
        self.graph_n.append(next(m for m in (bfs, ppr, sssp) if m is not None).n)
        return gid.value

    def submit(self, graph, kind, source, arrival=0.0):
        q = C.c_int64(0)
        check(lib.gbk_server_submit(self.h, graph, kind, int(source), float(arrival), C.byref(q)))
        return q.value

    def run(self):
        check(lib.gbk_server_run(self.h))

    def info(self, qid):
        t = (C.c_double * 4)()
        i = (C.c_int64 * 4)()
        check(lib.gbk_server_query_info(self.h, qid, t, i))
        return {"arrival": t[0], "admit": t[1], "done": t[2], "checksum": t[3], "reached": i[0],
                "occupancy": i[1], "slot": i[2], "status": i[3]}

    def result(self, qid, graph, kind):
        out = np.empty(self.graph_n[graph], dtype=np.float32 if kind == QUERY_PPR else np.int32)
        check(lib.gbk_server_result(self.h, qid, _ptr(out)))
        return out

    def stats(self):
        st = (C.c_double * 6)()
        check(lib.gbk_server_stats(self.h, st))
        return {"wall_s": st[0], "queries": int(st[1]), "peak_bytes": int(st[2]), "engines": int(st[3]),
                "max_engines_live": int(st[4]), "largest_engine_bytes": int(st[5])}

    def free(self):
        if self.h:
            lib.gbk_server_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------- element-wise ops
def _mh(v):
    return v.h if v is not None else None


def assign_scalar(w, mask, value, desc_mask=DEFAULT, accum=ACC_NONE):
    check(lib.gbk_assign_scalar(w.h, _mh(mask), accum, float(value), _desc(desc_mask)))


def assign_vector(w, mask, u, accum=ACC_NONE, desc_mask=DEFAULT):
    check(lib.gbk_assign_vector(w.h, _mh(mask), accum, u.h, _desc(desc_mask)))


def assign_scatter(w, mask, u, indices, nindices, accum=ACC_NONE):
    check(lib.gbk_assign_scatter(w.h, _mh(mask), accum, u.h, indices.h, nindices, _desc()))


def ewise_add(w, mask, op, u, v):
    check(lib.gbk_ewise_add(w.h, _mh(mask), ACC_NONE, op, u.h, v.h, _desc()))


def ewise_mult(w, mask, op, u, v):
    check(lib.gbk_ewise_mult(w.h, _mh(mask), ACC_NONE, op, u.h, v.h, _desc()))


def apply(w, mask, unary, u, scalar=0.0):
    check(lib.gbk_apply(w.h, _mh(mask), ACC_NONE, unary, float(scalar), u.h, _desc()))


def reduce(u, monoid=ADD_PLUS, init=0.0, accum=ACC_PLUS):
    r = C.c_double(init)
    check(lib.gbk_reduce(C.byref(r), accum, monoid, u.h, _desc()))
    return r.value


def reduce_i64(u, monoid=ADD_PLUS, init=0, accum=ACC_PLUS):
    r = C.c_int64(init)
    check(lib.gbk_reduce_i64(C.byref(r), accum, monoid, u.h, _desc()))
    return r.value


def extract(w, mask, u, indices, accum=ACC_NONE):
    check(lib.gbk_extract(w.h, _mh(mask), accum, u.h, _mh(indices), _desc()))
