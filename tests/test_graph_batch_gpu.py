"""Many small graphs in one persistent, HIP-graph-replayed BFS launch: every
graph's levels equal the oracle's queue BFS and the eager single-graph
driver, on every replay of the captured graph (with eager work between the
replays)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(gbk, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        scale = int(rng.integers(3, 13))
        ef = int(rng.integers(2, 17))
        rp, col = gbk.gen_rmat(scale, ef, seed=seed * 1000 + i, symmetric=bool(i % 2))
        out.append((rp, col))
    out.append((np.array([0, 0], dtype=np.int64), np.zeros(0, dtype=np.int32)))  # one isolated vertex
    out.append((np.array([0, 1, 1], dtype=np.int64), np.array([1], dtype=np.int32)))  # 0 -> 1
    return out


def test_graph_batch_bfs_replays(gbk, oracle):
    graphs = _graphs(gbk, 60, 3)
    batch = gbk.GraphBatch(graphs)
    info = batch.info()
    assert info["graphs"] == len(graphs)
    rng = np.random.default_rng(9)
    for replay in range(3):
        src = [int(rng.integers(0, len(rp) - 1)) for rp, _ in graphs]
        levels, ms = batch.bfs(src)
        assert ms > 0
        for g, (rp, col) in enumerate(graphs):
            want = oracle.bfs(rp, col, src[g])
            assert np.array_equal(levels[g], want), f"replay {replay} graph {g}"
        # eager work between replays: the single-graph driver on a few graphs
        for g in range(0, len(graphs), 17):
            rp, col = graphs[g]
            A = gbk.Matrix.from_csr(rp, col, np.ones(max(len(col), 1), np.float32)[: len(col)])
            eager, _ = gbk.bfs(A, src[g])
            assert np.array_equal(eager, levels[g])
            A.free()


def test_graph_batch_largest_graph(gbk, oracle):
    """A graph at the 2^17-vertex limit next to tiny ones."""
    rp, col = gbk.gen_rmat(17, 8, seed=44, symmetric=True)
    graphs = [(rp, col)] + _graphs(gbk, 8, 5)
    batch = gbk.GraphBatch(graphs)
    src = [int(np.flatnonzero(np.diff(rp) > 0)[0])] + [0] * (len(graphs) - 1)
    levels, _ = batch.bfs(src)
    for g, (r, c) in enumerate(graphs):
        assert np.array_equal(levels[g], oracle.bfs(r, c, src[g])), g
