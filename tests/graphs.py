"""Seeded test graphs shared by the CPU and GPU tests."""
import numpy as np


def csr_from_edges(n, edges):
    """CSR (int64 row_ptr, int32 col) from a list of (row, col), keeping list order inside rows."""
    rows = np.array([e[0] for e in edges], dtype=np.int64)
    cols = np.array([e[1] for e in edges], dtype=np.int32)
    order = np.argsort(rows, kind="stable")
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp), cols[order]


def special_graphs():
    """Hand-made structures that stress the row decomposition."""
    g = {}
    # 8 x 8 with trailing empty rows 6, 7 (the reference's partition skips them)
    g["trailing_empty"] = csr_from_edges(8, [(0, 1), (0, 2), (1, 3), (2, 0), (3, 4), (4, 5), (5, 0), (5, 1)])
    # one hub row crossing many 2048-item tiles, between empty rows
    hub = [(3, c) for c in range(0, 9000)]
    g["hub_row"] = csr_from_edges(9001, hub + [(9000, 1)])
    # a long run of empty rows then a few entries
    g["empty_run"] = csr_from_edges(20000, [(19990, 5), (19991, 6), (5, 19999)])
    # rows of exactly the tile length and its neighbours
    e = []
    for r, L in enumerate([2047, 2048, 2049, 1, 0, 4095, 3]):
        e += [(r, (r * 7 + k) % 7000) for k in range(L)]
    g["tile_edges"] = csr_from_edges(7000, e)
    # no entries at all
    g["no_entries"] = (np.zeros(65, dtype=np.int64), np.zeros(0, dtype=np.int32))
    return g
