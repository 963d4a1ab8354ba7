"""GPU tests of the continuous-batching query server.

* Every query's result in a mixed, overlapping stream equals the same query
  served alone (bit for bit: the engines fold every batch column in the same
  order whatever the other columns hold), and matches the oracle: BFS levels
  and SSSP distances exactly, personalized PageRank within 1e-5 L1 of the
  FP64 fixed point (the fp32 iteration stops at an L1 change below 1e-6).
* Queries that arrive while others run are admitted into free slots of the
  running batch (continuous batching), and the HBM budget bounds what the
  engines and in-flight results hold.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph(gbk, scale, seed):
    rp, col = gbk.gen_rmat(scale, 8, seed=seed, symmetric=True)
    n = len(rp) - 1
    bfs = gbk.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    ppr = gbk.Matrix.from_csr(rp, col, gbk.inv_coldeg_f32(n, col))
    w = gbk.sym_weights_i32(rp, col, seed + 1, 1, 255)
    sssp = gbk.Matrix.from_csr(rp, col, w)
    return {"rp": rp, "col": col, "w": w, "n": n, "m": (bfs, ppr, sssp)}


def _sources(rp, k, seed):
    cand = np.flatnonzero(np.diff(rp) > 0)
    return np.random.default_rng(seed).choice(cand, size=k, replace=True)


def _serve(gbk, graphs, queries, budget=1 << 33, keep=True):
    srv = gbk.Server(budget, ppr_slots=8, sssp_slots=8, tol=1e-6, itermax=300, keep_results=keep)
    gids = [srv.add_graph(*g["m"]) for g in graphs]
    qids = [srv.submit(gids[g], kind, s, t) for (g, kind, s, t) in queries]
    srv.run()
    return srv, qids


def test_mixed_stream_matches_alone_and_oracle(gbk, oracle):
    graphs = [_graph(gbk, 10, 11), _graph(gbk, 11, 12), _graph(gbk, 9, 13)]
    rng = np.random.default_rng(5)
    queries = []
    t = 0.0
    for i in range(90):
        g = int(rng.integers(0, 3))
        kind = [gbk.QUERY_BFS, gbk.QUERY_PPR, gbk.QUERY_SSSP][i % 3]
        s = int(_sources(graphs[g]["rp"], 1, 100 + i)[0])
        t += float(rng.exponential(2e-4))
        queries.append((g, kind, s, t))
    srv, qids = _serve(gbk, graphs, queries)
    occupancy = []
    for (g, kind, s, _), q in zip(queries, qids):
        info = srv.info(q)
        assert info["status"] == 0 and info["done"] >= info["admit"] >= 0
        occupancy.append(info["occupancy"])
        got = srv.result(q, g, kind)
        G = graphs[g]
        if kind == gbk.QUERY_BFS:
            assert np.array_equal(got, oracle.bfs(G["rp"], G["col"], s)), (g, s)
        elif kind == gbk.QUERY_SSSP:
            assert np.array_equal(got, oracle.dijkstra_i32(G["rp"], G["col"], G["w"], s)), (g, s)
        else:
            want, _ = oracle.ppr_f64(G["rp"], G["col"], gbk.inv_coldeg_f32(G["n"], G["col"]), source=s,
                                     itermax=2000, tol=1e-13)
            assert np.abs(got.astype(np.float64) - want).sum() <= 1e-5, (g, s)
    assert max(occupancy) > 1, "no query joined a running batch"
    # the same queries one at a time, each in a fresh server: identical results
    for (g, kind, s, _), q in list(zip(queries, qids))[::7]:
        alone, aq = _serve(gbk, graphs, [(g, kind, s, 0.0)])
        assert np.array_equal(alone.result(aq[0], g, kind), srv.result(q, g, kind)), (g, kind, s)


def test_budget_bounds_engines_and_results(gbk):
    graphs = [_graph(gbk, 11, 21), _graph(gbk, 11, 22)]
    queries = [(i % 2, gbk.QUERY_BFS, int(_sources(graphs[i % 2]["rp"], 1, i)[0]), 0.0) for i in range(40)]
    queries += [(i % 2, gbk.QUERY_SSSP, int(_sources(graphs[i % 2]["rp"], 1, 50 + i)[0]), 0.0) for i in range(20)]
    big, _ = _serve(gbk, graphs, queries, keep=False)
    budget = int(big.stats()["largest_engine_bytes"] * 1.3)  # room for one engine at a time
    small, qids = _serve(gbk, graphs, queries, budget=budget, keep=False)
    st = small.stats()
    assert st["peak_bytes"] <= budget
    for q, bq in zip(qids, range(len(qids))):
        a, b = small.info(q), big.info(bq)
        assert a["status"] == 0
        assert a["checksum"] == b["checksum"] and a["reached"] == b["reached"]


def test_budget_too_small_is_an_error(gbk):
    graphs = [_graph(gbk, 10, 31)]
    srv = gbk.Server(1 << 12, keep_results=False)
    gid = srv.add_graph(*graphs[0]["m"])
    srv.submit(gid, gbk.QUERY_BFS, int(_sources(graphs[0]["rp"], 1, 1)[0]), 0.0)
    with pytest.raises(gbk.GbkError) as e:
        srv.run()
    assert e.value.code == gbk.OUT_OF_MEMORY
