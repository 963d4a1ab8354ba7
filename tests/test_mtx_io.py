"""Edge-list files (.mtx text, .mtxbin binary) on the host: the library's
reader and writer against numpy's own parse of the same text, round trips,
and the error cases of read_file.hpp (missing file, unknown suffix, short
file).  No GPU: only the host functions of the library are called."""
import os

import numpy as np
import pytest

import gb_kun_amd as g


def _coo(n, m, seed):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, size=m)
    cols = rng.integers(0, n, size=m)
    rows[:3] = rows[0]  # a duplicated entry and a repeated row
    cols[:3] = [cols[0], cols[0], cols[1]]
    return rows.astype(np.int64), cols.astype(np.int64)


@pytest.mark.parametrize("suffix", ["mtx", "mtxbin"])
def test_round_trip_keeps_file_order(tmp_path, suffix):
    rows, cols = _coo(1000, 300_000, 5)
    p = str(tmp_path / f"g.{suffix}")
    g.mtx_write(p, 1000, 1200, rows, cols)
    nr, nc, r2, c2 = g.mtx_read(p)
    assert (nr, nc) == (1000, 1200)
    assert np.array_equal(r2, rows) and np.array_equal(c2, cols)


def test_text_matches_numpy_parse_and_binary(tmp_path):
    rows, cols = _coo(5000, 50_000, 6)
    pt, pb = str(tmp_path / "g.mtx"), str(tmp_path / "g.mtxbin")
    g.mtx_write(pt, 5000, 5000, rows, cols)
    g.mtx_write(pb, 5000, 5000, rows, cols)
    with open(pt) as fh:
        lines = [ln for ln in fh.read().splitlines() if ln and not ln.startswith("%")]
    assert lines[0].split() == ["5000", "5000", "50000"]
    want = np.loadtxt(lines[1:], dtype=np.int64).reshape(-1, 2) - 1
    _, _, rt, ct = g.mtx_read(pt)
    _, _, rb, cb = g.mtx_read(pb)
    assert np.array_equal(rt, want[:, 0]) and np.array_equal(ct, want[:, 1])
    assert np.array_equal(rt, rb) and np.array_equal(ct, cb)
    # the binary layout: three int64, then (row, col) int64 pairs, 1-based
    raw = np.fromfile(pb, dtype="<i8")
    assert list(raw[:3]) == [5000, 5000, 50000]
    assert np.array_equal(raw[3:].reshape(-1, 2), want + 1)


def test_text_with_comments_values_and_crlf(tmp_path):
    p = tmp_path / "k.mtx"
    p.write_bytes(b"%%MatrixMarket matrix coordinate real general\r\n% a comment\r\n"
                  b"4 3 5\r\n1 1 0.5\r\n4 3 2.0\r\n  2\t2 7\r\n\r\n3 1\r\n1 3 -1e3\r\n")
    nr, nc, r, c = g.mtx_read(str(p))
    assert (nr, nc) == (4, 3)
    assert r.tolist() == [0, 3, 1, 2, 0] and c.tolist() == [0, 2, 1, 0, 2]


def test_extra_lines_beyond_the_count_are_ignored(tmp_path):
    p = tmp_path / "e.mtx"
    p.write_text("3 3 2\n1 2\n2 3\n3 1\n")
    _, _, r, c = g.mtx_read(str(p))
    assert r.tolist() == [0, 1] and c.tolist() == [1, 2]


@pytest.mark.parametrize("text,msg", [
    ("3 3 4\n1 2\n2 3\n", "unexpected end"),
    ("3 3 1\n0 2\n", "outside"),
    ("3 3 1\n1 4\n", "outside"),
    ("3 3 1\nx y\n", "malformed"),
    ("% only comments\n", "unexpected end"),
])
def test_text_errors(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(g.GbkError, match=msg):
        g.mtx_read(str(p))


def test_file_and_suffix_errors(tmp_path):
    with pytest.raises(g.GbkError, match="Can not open"):
        g.mtx_read(str(tmp_path / "missing.mtx"))
    p = tmp_path / "g.txt"
    p.write_text("1 1 0\n")
    with pytest.raises(g.GbkError, match="Unsupported matrix file format"):
        g.mtx_read(str(p))
    b = tmp_path / "short.mtxbin"
    np.array([3, 3, 2, 1, 2, 2], dtype="<i8").tofile(str(b))  # one pair of two
    with pytest.raises(g.GbkError, match="Unexpected end of binary file"):
        g.mtx_read(str(b))


def test_empty_entry_list(tmp_path):
    p = tmp_path / "z.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n7 7 0\n")
    nr, nc, r, c = g.mtx_read(str(p))
    assert (nr, nc, len(r), len(c)) == (7, 7, 0, 0)


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).have_ref(),
                    reason="the compiled reference (oracle/_ref) is not built here")
@pytest.mark.parametrize("suffix", ["mtx", "mtxbin"])
def test_reader_matches_the_compiled_reference(tmp_path, suffix):
    """Matrix::init_from_mtx of the compiled reference on the same file: the
    text reader keeps file order inside a row (vector-of-vectors push_back),
    the binary reader fills rows from per-thread hash maps, so there only the
    multiset of each row is fixed."""
    from oracle import oracle as O
    rows, cols = _coo(2000, 60_000, 7)
    p = str(tmp_path / f"g.{suffix}")
    g.mtx_write(p, 2000, 2000, rows, cols)
    rrp, rcol, _ = O.ref_read_mtx(p)
    nr, _, r, c = g.mtx_read(p)
    order = np.argsort(r, kind="stable")
    rp = np.zeros(nr + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp)
    assert np.array_equal(rrp, rp)
    if suffix == "mtx":
        assert np.array_equal(rcol, c[order])
    else:
        for i in range(nr):
            assert np.array_equal(np.sort(rcol[rp[i]:rp[i + 1]]), np.sort(c[order][rp[i]:rp[i + 1]]))
