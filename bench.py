"""Headline benchmark: semiring SpMV on a seeded R-MAT scale-22 graph on one MI355X.

One step = one plus.times fp32 SpMV y = A x over the whole graph (R-MAT scale
22, edge factor 16, (a,b,c) = (.57,.19,.19), duplicates and self-loops
removed, values 1/coldeg as in the PageRank transition matrix, x uniform in
[0,1)), in the matrix's current storage format, with A, x and y resident in
HBM.  `value` is the algorithmic HBM bandwidth of the step,

    bytes/step = 4(n+1) row_ptr + (4+4) nnz col/val + 4n x + 4n y,

divided by the mean step time measured with HIP events on the library's
stream.  GTEPS (nnz / time) is reported beside it.  The CPU baseline is the
oracle's OpenMP restatement of the reference's CSR SpMV
(spmv_csr.h:341-384 with the balancing.h partition) on the host cores, timed
on the same graph.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 22] [--format csr]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS & HBM GB/s (% peak), semiring SpMV/BFS RMAT-22; batched/served q/s, p99"
HBM_PEAK_GBS = 8000.0
KERNELS = {
    "csr": "k_spmv_merge + k_spmv_merge_fixup",
    "csr_seg": "k_hot_x + k_seg_flag (persistent, XCD-mapped segments, hot x in LDS, flagged pair ends) + "
               "k_flag_fix + k_seg_merge_rows",
    "lav": "k_gather + k_spmv_merge (XCD-mapped segment tiles) + k_spmv_merge_fixup + k_seg_merge",
    "sell_c": "k_sell_c + long-row k_spmv_merge + k_spmv_merge_fixup",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--format", default="auto", choices=["auto", "csr", "sell_c", "csr_seg", "lav"])
    ap.add_argument("--graph", default="rmat", choices=["rmat", "er"])
    ap.add_argument("--sell-c", type=int, default=0)
    ap.add_argument("--sigma", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--bfs-scale", type=int, default=23, help="scale of the symmetric R-MAT for BFS / SSSP GTEPS")
    ap.add_argument("--tc-scale", type=int, default=20, help="scale of the symmetric R-MAT for triangle counting")
    ap.add_argument("--serve-full", action="store_true",
                    help="served measurement on the full 8-graph set (R-MAT 18..24 + ER 22, 20,000 queries)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the timed loop, copy this many MiB device-to-device on the library stream "
                         "(a kernel of known traffic, to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--section", default="all", choices=["all", "batched", "served", "traversal", "io", "tc", "pr"],
                    help="run one secondary measurement alone (for profiling); prints that section as JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the multi-GPU row sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    import gb_kun_amd as g
    from gb_kun_amd.device import DeviceArray, Event

    if args.section != "all":
        if args.section == "batched":
            rp, col = g.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=1, symmetric=False, permute=True)
            res = batched_bench(g, rp, col, len(rp) - 1, args.scale)
        elif args.section == "served":
            res = served_bench(g, full=args.serve_full)
        elif args.section == "io":
            res = io_bench(g, args.scale, args.edge_factor)
        elif args.section == "tc":
            res = tc_bench(g, args.tc_scale)
        elif args.section == "pr":
            rp, col = g.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=1, symmetric=False, permute=True)
            fmt = {"csr": g.FMT_CSR, "sell_c": g.FMT_SELL_C, "csr_seg": g.FMT_CSR_SEG, "lav": g.FMT_LAV}.get(
                args.format, g.FMT_CSR_SEG)
            A = g.Matrix.from_csr(rp, col, g.inv_coldeg_f32(len(rp) - 1, col), fmt=fmt)
            res = pr_bench(g, A, None)
        else:
            res = traversal_bench(g, args.bfs_scale)
        print(json.dumps({"section": args.section, args.section: res}), flush=True)
        return

    t0 = time.time()
    if args.graph == "rmat":
        rp, col = g.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=1, symmetric=False, permute=True)
    else:
        rp, col = g.gen_uniform(args.scale, args.edge_factor, seed=1)
    n = len(rp) - 1
    nnz = len(col)
    vals = g.inv_coldeg_f32(n, col)
    x_np = g.uniform_f32(3, n)
    log(f"graph: {args.graph} scale {args.scale} ef {args.edge_factor}: n={n} nnz={nnz} "
        f"empty rows={int(np.count_nonzero(np.diff(rp) == 0))} max row={int(np.diff(rp).max())} "
        f"({time.time() - t0:.1f} s)")

    t0 = time.time()
    tuned = None
    if args.format == "auto":
        A = g.Matrix.from_csr(rp, col, vals)
        f, c, times = A.autotune(trials=5)
        fmt_name = {g.FMT_CSR: "csr", g.FMT_CSR_SEG: "csr_seg", g.FMT_LAV: "lav"}.get(f, f"sell_c{c}")
        tuned = {k: round(v, 1) for k, v in times.items()}
        log(f"autotuner (us per SpMV): {tuned} -> {fmt_name}")
    else:
        fmt = {"csr": g.FMT_CSR, "sell_c": g.FMT_SELL_C, "csr_seg": g.FMT_CSR_SEG, "lav": g.FMT_LAV}[args.format]
        A = g.Matrix.from_csr(rp, col, vals, fmt=fmt, sell_c=args.sell_c, sigma=args.sigma)
        fmt_name = args.format
    log(f"device matrix built in {time.time() - t0:.1f} s, {A.device_bytes() / 2**20:.0f} MiB of HBM")
    cu, arch = g.device_info()

    x = DeviceArray.from_numpy(x_np)
    y = DeviceArray.zeros(n, np.float32)

    def step():
        g.spmv_device(A, x.ptr, y.ptr)

    for _ in range(args.warmup):
        step()
    g.synchronize()

    # correctness spot check of the timed configuration against FP64
    from oracle import oracle as O
    y_np = y.numpy()
    rows = np.random.default_rng(0).choice(n, size=min(n, 20000), replace=False)
    want, absum = O.spmv_f64(rp, col, vals, x_np)
    err = np.max(np.abs(y_np[rows].astype(np.float64) - want[rows]) / np.maximum(absum[rows], 1e-30))
    log(f"spot check vs FP64 on {len(rows)} rows: max rel err {err:.2e}")

    # HIP events on the library stream, where the kernels run
    g.synchronize()
    ev0, ev1 = Event(), Event()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    ev1.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_ms(ev1) / args.steps

    algo_bytes = 4 * (n + 1) + 8 * nnz + 4 * n + 4 * n
    ref_bytes = 16 * nnz
    gbs = algo_bytes / (ms * 1e-3) / 1e9
    gteps = nnz / (ms * 1e-3) / 1e9
    log(f"SpMV {ms * 1e3:.1f} us/step  {gbs:.0f} GB/s algorithmic ({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s)  "
        f"{gteps:.1f} GTEPS  wall {wall * 1e3 / args.steps:.1f} us/step")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_spmv_{fmt_name}_{args.graph}{args.scale}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    if args.copy_probe > 0:
        from gb_kun_amd.device import copy_d2d
        a = DeviceArray(args.copy_probe << 18, np.float32)
        b = DeviceArray(args.copy_probe << 18, np.float32)
        copy_d2d(b, a)
        g.synchronize()
        log(f"copy probe: {args.copy_probe} MiB device-to-device")

    pr = pr_bench(g, A, ms)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(O, rp, col, vals, x_np, algo_bytes, args.cpu_seconds)

    batched = served = traversal = None
    if not args.no_batched:
        A.free()
        batched = batched_bench(g, rp, col, n, args.scale)
        served = served_bench(g, full=args.serve_full)
        traversal = traversal_bench(g, args.bfs_scale)

    out = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT, values 1/coldeg)",
        "config": {"workload": f"plus.times fp32 SpMV, {'R-MAT' if args.graph == 'rmat' else 'Erdos-Renyi'} scale "
                               f"{args.scale}, edge factor {args.edge_factor}, "
                               f"{fmt_name} format", "n": n, "nnz": nnz, "format": fmt_name, "autotuner_us": tuned,
                   "parallelism": "single GPU"},
        "gteps": round(gteps, 3),
        "ref_accounting_gbs": round(ref_bytes / (ms * 1e-3) / 1e9, 2),
        "device": {"arch": arch, "cus": cu},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": KERNELS.get(fmt_name.rstrip("0123456789"), "SpMV"),
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": cpu,
        "pagerank": pr,
        "batched": batched,
        "served": served,
        "bfs_sssp": traversal,
    }
    print(json.dumps(out), flush=True)


def pr_bench(g, A, spmv_ms):
    """PageRank with sinks (damping 0.85) on the SpMV graph, ranks left in
    HBM: a fixed 20 iterations (time per iteration beside one SpMV's), then
    to L1 < 1e-6."""
    from gb_kun_amd.device import DeviceArray
    r = DeviceArray.zeros(A.n, np.float32)
    g.pagerank_device(A, r.ptr, itermax=2, tol=0.0)
    g.synchronize()
    t0 = time.perf_counter()
    g.pagerank_device(A, r.ptr, itermax=20, tol=0.0)
    t_fixed = (time.perf_counter() - t0) / 20
    t0 = time.perf_counter()
    iters = g.pagerank_device(A, r.ptr, itermax=200, tol=1e-6)
    t_tol = time.perf_counter() - t0
    res = {"ms_per_iteration": round(t_fixed * 1e3, 4), "tol": 1e-6, "iterations_to_tol": int(iters),
           "ms_to_tol": round(t_tol * 1e3, 3), "ms_per_iteration_to_tol": round(t_tol * 1e3 / max(iters, 1), 4)}
    msg = (f"PageRank: {t_fixed * 1e3:.3f} ms/iteration (20 fixed); L1 < 1e-6 in {iters} iterations, "
           f"{t_tol * 1e3:.1f} ms")
    if spmv_ms:
        res["spmv_ms"] = round(spmv_ms, 4)
        res["iteration_over_spmv"] = round(t_fixed * 1e3 / spmv_ms, 3)
        msg += f"; iteration / SpMV = {t_fixed * 1e3 / spmv_ms:.2f}"
    log(msg)
    return res


def _sources(rp, k, seed):
    cand = np.flatnonzero(np.diff(rp) > 0)
    return np.random.default_rng(seed).choice(cand, size=k, replace=False)


def batched_bench(g, rp, col, n, scale):
    """Batched queries on the headline graph: 64-source BFS as one boolean SpMM
    traversal vs 64 single-source BFS; 32 personalized PageRanks (one-hot
    starts, L1 < 1e-6) as one SpMM iteration vs 32 single-vector runs; results
    left in HBM on both sides."""
    res = {"graph": f"R-MAT scale {scale}, edge factor 16 (directed)"}
    from gb_kun_amd.device import DeviceArray
    Ab = g.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    src = _sources(rp, 64, 7)
    # levels stay in HBM on both sides (64 x n int32 each)
    out_single = DeviceArray((64, n), np.int32)
    out_multi = DeviceArray((64, n), np.int32)
    g.bfs_device(Ab, int(src[0]), out_single.ptr)  # warm-up
    g.synchronize()
    t0 = time.perf_counter()
    for i, s_ in enumerate(src):
        g.bfs_device(Ab, int(s_), out_single.ptr + 4 * n * i)
    g.synchronize()
    t_single = time.perf_counter() - t0
    g.msbfs_device(Ab, src[:2], out_multi.ptr)
    g.synchronize()
    t0 = time.perf_counter()
    g.msbfs_device(Ab, src, out_multi.ptr)
    g.synchronize()
    t_multi = time.perf_counter() - t0
    same = bool(np.array_equal(out_single.numpy(), out_multi.numpy()))
    out_single.free()
    out_multi.free()
    res["bfs64"] = {"single_qps": round(64 / t_single, 2), "batched_qps": round(64 / t_multi, 2),
                    "speedup": round(t_single / t_multi, 2), "levels_identical": bool(same),
                    "batched_ms": round(t_multi * 1e3, 2), "single_ms_each": round(t_single / 64 * 1e3, 3)}
    log(f"batched BFS x64: {t_multi * 1e3:.1f} ms vs {t_single * 1e3:.1f} ms single ({t_single / t_multi:.1f}x), "
        f"identical={same}")
    Ab.free()
    Ap = g.Matrix.from_csr(rp, col, g.inv_coldeg_f32(n, col))
    srcp = _sources(rp, 32, 8)
    # ranks stay in HBM on both sides (32 x n floats each)
    pr_single = DeviceArray((32, n), np.float32)
    pr_multi = DeviceArray((32, n), np.float32)
    g.ppr_multi_device(Ap, srcp[:1], pr_single.ptr, itermax=100, tol=1e-6)
    g.synchronize()
    t0 = time.perf_counter()
    for i, s_ in enumerate(srcp):
        g.ppr_multi_device(Ap, [int(s_)], pr_single.ptr + 4 * n * i, itermax=100, tol=1e-6)
    g.synchronize()
    t_single = time.perf_counter() - t0
    t0 = time.perf_counter()
    its = g.ppr_multi_device(Ap, srcp, pr_multi.ptr, itermax=100, tol=1e-6)
    g.synchronize()
    t_multi = time.perf_counter() - t0
    l1 = float(np.abs(pr_multi.numpy().astype(np.float64) - pr_single.numpy()).sum(axis=1).max())
    pr_single.free()
    pr_multi.free()
    res["ppr32"] = {"single_qps": round(32 / t_single, 2), "batched_qps": round(32 / t_multi, 2),
                    "speedup": round(t_single / t_multi, 2), "max_l1_batched_vs_single": l1,
                    "iterations": [int(v) for v in its[:4]], "batched_ms": round(t_multi * 1e3, 2)}
    log(f"batched PPR x32: {t_multi * 1e3:.1f} ms vs {t_single * 1e3:.1f} ms single ({t_single / t_multi:.1f}x), "
        f"max L1 batched vs single {l1:.1e}")
    Ap.free()
    res["graphs512"] = small_graphs_bench(g)
    return res


def small_graphs_bench(g, count=512, seed=12):
    """512 independent R-MAT graphs, scale uniform in [8, 16], edge factor in
    [8, 32]: one BFS per graph by one persistent kernel replayed from a HIP
    graph, vs the single-graph driver called once per graph."""
    rng = np.random.default_rng(seed)
    graphs, srcs = [], []
    t0 = time.perf_counter()
    for i in range(count):
        sc, ef = int(rng.integers(8, 17)), int(rng.integers(8, 33))
        rp, col = g.gen_rmat(sc, ef, seed=5000 + i)
        graphs.append((rp, col))
        srcs.append(int(_sources(rp, 1, 6000 + i)[0]))
    batch = g.GraphBatch(graphs)
    info = batch.info()
    log(f"{count} graphs: {info['vertices']} vertices, {info['edges']} edges, built in {time.perf_counter() - t0:.1f} s")
    batch.bfs(srcs)  # first replay
    t0 = time.perf_counter()
    levels, dev_ms = batch.bfs(srcs)
    t_batch = time.perf_counter() - t0
    mats = [g.Matrix.from_csr(rp, col, np.ones(len(col), np.float32)) for rp, col in graphs]
    g.bfs(mats[0], srcs[0])
    t0 = time.perf_counter()
    same = True
    for i, A in enumerate(mats):
        lv, _ = g.bfs(A, srcs[i])
        same &= bool(np.array_equal(lv, levels[i]))
    t_single = time.perf_counter() - t0
    for A in mats:
        A.free()
    res = {"graphs": count, "vertices": info["vertices"], "edges": info["edges"],
           "persistent_workgroups": info["workgroups"], "batched_ms": round(t_batch * 1e3, 2),
           "batched_device_ms": round(dev_ms, 3), "single_ms": round(t_single * 1e3, 2),
           "speedup": round(t_single / t_batch, 2), "levels_identical": same,
           "gteps_batched": round(info["edges"] / (dev_ms * 1e-3) / 1e9, 2)}
    log(f"{count} graphs BFS: one HIP-graph launch {dev_ms:.2f} ms on device ({t_batch * 1e3:.1f} ms with copies) "
        f"vs {t_single * 1e3:.0f} ms one call per graph ({t_single / t_batch:.0f}x), identical={same}")
    batch.free()
    return res


def traversal_bench(g, scale, nsrc=64, seed=13):
    """Direction-optimizing BFS and int32 min.plus SSSP on a symmetrized R-MAT
    from 64 seeded non-isolated sources; GTEPS = undirected edges of the
    reached component / time (harmonic mean over sources, as Graph500)."""
    t0 = time.perf_counter()
    rp, col = g.gen_rmat(scale, 16, seed=seed, symmetric=True)
    n = len(rp) - 1
    deg = np.diff(rp)
    Ab = g.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
    w = g.sym_weights_i32(rp, col, seed + 1, 1, 255)
    As = g.Matrix.from_csr(rp, col, w)
    log(f"traversal graph: symmetric R-MAT {scale}: n={n} stored nnz={len(col)} ({time.perf_counter() - t0:.1f} s)")
    src = _sources(rp, nsrc, seed + 2)
    from gb_kun_amd.device import DeviceArray, synchronize
    out = DeviceArray(n, np.int32)  # levels / distances stay on the device while timed
    g.bfs_device(Ab, int(src[0]), out.ptr)
    teps, times, levels, dirs = [], [], [], [0, 0]
    for s_ in src:
        t1 = time.perf_counter()
        st = g.bfs_device(Ab, int(s_), out.ptr)
        synchronize()
        dt = time.perf_counter() - t1
        lv = out.numpy()
        m = float(deg[lv > 0].sum()) / 2.0
        teps.append(m / dt)
        times.append(dt)
        levels.append(int(st[0]))
        dirs[0] += int(st[1])
        dirs[1] += int(st[2])
    bfs = {"gteps_hmean": round(len(teps) / sum(1.0 / t for t in teps) / 1e9, 3),
           "ms_median": round(float(np.median(times)) * 1e3, 3), "levels_median": int(np.median(levels)),
           "push_steps": dirs[0], "pull_steps": dirs[1], "sources": len(src)}
    g.sssp_device(As, int(src[0]), out.ptr)
    teps, times = [], []
    for s_ in src[:16]:
        t1 = time.perf_counter()
        g.sssp_device(As, int(s_), out.ptr)
        synchronize()
        dt = time.perf_counter() - t1
        d = out.numpy()
        m = float(deg[d != np.iinfo(np.int32).max].sum()) / 2.0
        teps.append(m / dt)
        times.append(dt)
    out.free()
    sssp = {"gteps_hmean": round(len(teps) / sum(1.0 / t for t in teps) / 1e9, 3),
            "ms_median": round(float(np.median(times)) * 1e3, 3), "sources": 16}
    log(f"BFS symmetric R-MAT {scale}: {bfs['gteps_hmean']} GTEPS (median {bfs['ms_median']} ms); "
        f"SSSP {sssp['gteps_hmean']} GTEPS (median {sssp['ms_median']} ms)")
    Ab.free()
    As.free()
    return {"graph": f"symmetric R-MAT scale {scale}, edge factor 16", "n": n, "stored_nnz": int(len(col)),
            "sources": nsrc, "bfs": bfs, "sssp": sssp}


def served_bench(g, full=False, seed=11):
    """Poisson stream of BFS (50%), personalized PageRank (30%) and SSSP (20%)
    queries over device-resident graphs, served with continuous batching;
    p50/p99 latency and q/s, beside the plain per-query GPU path (each query
    run alone by the single-query drivers) on a sample of the same mix."""
    specs = ([("rmat", s) for s in range(18, 25)] + [("er", 22)]) if full else \
        [("rmat", 18), ("rmat", 19), ("rmat", 20), ("er", 18)]
    nq = 20000 if full else 3000
    graphs = []
    t0 = time.perf_counter()
    for i, (kind, sc) in enumerate(specs):
        if kind == "rmat":
            rp, col = g.gen_rmat(sc, 16, seed=100 + i, symmetric=True)
        else:
            rp, col = g.gen_uniform(sc, 16, seed=100 + i, symmetric=True)
        n = len(rp) - 1
        mb = g.Matrix.from_csr(rp, col, np.ones(len(col), np.float32))
        mp = g.Matrix.from_csr(rp, col, g.inv_coldeg_f32(n, col))
        ms = g.Matrix.from_csr(rp, col, g.sym_weights_i32(rp, col, 200 + i, 1, 255))
        graphs.append({"name": f"{kind}{sc}", "rp": rp, "m": (mb, mp, ms), "n": n})
    log(f"serving graphs built in {time.perf_counter() - t0:.1f} s: " +
        ", ".join(f"{G['name']} n={G['n']} nnz={G['m'][0].nnz}" for G in graphs))
    rng = np.random.default_rng(seed)
    kinds = rng.choice([g.QUERY_BFS, g.QUERY_PPR, g.QUERY_SSSP], size=nq, p=[0.5, 0.3, 0.2])
    gsel = rng.integers(0, len(graphs), size=nq)
    srcs = [int(_sources(graphs[gi]["rp"], 1, 1000 + i)[0]) for i, gi in enumerate(gsel)]

    # plain per-query path on a sample: every query alone through the single-query drivers
    sample = min(150, nq)
    t0 = time.perf_counter()
    for i in range(sample):
        G = graphs[gsel[i]]
        if kinds[i] == g.QUERY_BFS:
            g.bfs(G["m"][0], srcs[i])
        elif kinds[i] == g.QUERY_PPR:
            g.ppr_multi(G["m"][1], [srcs[i]], itermax=100, tol=1e-6)
        else:
            g.sssp(G["m"][2], srcs[i])
    plain_qps = sample / (time.perf_counter() - t0)

    def run(rate):
        srv = g.Server(64 << 30, ppr_slots=32, sssp_slots=32, tol=1e-6, itermax=100, keep_results=False)
        gids = [srv.add_graph(*G["m"]) for G in graphs]
        arr = np.cumsum(rng.exponential(1.0 / rate, size=nq)) if rate else np.zeros(nq)
        qids = [srv.submit(gids[gsel[i]], int(kinds[i]), srcs[i], float(arr[i])) for i in range(nq)]
        srv.run()
        st = srv.stats()
        lat = np.array([srv.info(q)["done"] - srv.info(q)["arrival"] for q in qids])
        occ = np.array([srv.info(q)["occupancy"] for q in qids])
        return st, lat, occ

    st, lat, occ = run(None)  # everything at t = 0: capacity
    cap_qps = nq / st["wall_s"]
    res = {"graphs": [G["name"] for G in graphs], "queries": nq, "mix": "50% BFS, 30% PPR (L1 1e-6), 20% SSSP",
           "plain_per_query_qps": round(plain_qps, 1), "plain_sample": sample,
           "capacity_qps": round(cap_qps, 1), "speedup_vs_plain": round(cap_qps / plain_qps, 2),
           "peak_hbm_bytes": st["peak_bytes"], "poisson": []}
    for frac in (0.25, 0.5):
        rate = frac * cap_qps
        st2, lat2, occ2 = run(rate)
        res["poisson"].append({"rate_qps": round(rate, 1), "served_qps": round(nq / st2["wall_s"], 1),
                               "p50_ms": round(float(np.percentile(lat2, 50)) * 1e3, 2),
                               "p99_ms": round(float(np.percentile(lat2, 99)) * 1e3, 2),
                               "mean_batch_occupancy": round(float(occ2.mean()), 1)})
        log(f"served: Poisson {rate:.0f} q/s -> p50 {res['poisson'][-1]['p50_ms']} ms "
            f"p99 {res['poisson'][-1]['p99_ms']} ms")
    log(f"served: capacity {cap_qps:.0f} q/s ({cap_qps / plain_qps:.1f}x plain {plain_qps:.0f} q/s)")
    for G in graphs:
        for m in G["m"]:
            m.free()
    return res


def tc_bench(g, scale):
    """Triangle counting (LAGr_TriangleCount) on a symmetric R-MAT: time of the
    default (Sandia over the degree orientation), Burkhardt (A^2 .* A) and
    Sandia (L L .* L) masked products, and the oracle's sequential masked
    count of Sandia on the host."""
    from oracle import oracle as O
    import scipy.sparse as sp
    rp, col = g.gen_rmat(scale, 16, seed=47, symmetric=True)
    n = len(rp) - 1
    A = g.Matrix.from_csr(rp, col, np.ones(len(col), np.int32))
    res = {"graph": f"symmetric R-MAT scale {scale}, edge factor 16", "n": n, "stored_nnz": len(col)}
    for name, m in (("default_degree_oriented", 0), ("burkhardt", 1), ("sandia", 3)):
        g.tricount(A, m)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            tri = g.tricount(A, m)
            times.append(time.perf_counter() - t0)
        res[name] = {"triangles": tri, "ms_median": round(float(np.median(times)) * 1e3, 2)}
        log(f"tricount {name}: {tri} triangles in {np.median(times) * 1e3:.1f} ms")
    S = sp.csr_matrix((np.ones(len(col), np.int8), col, rp), shape=(n, n))
    L = sp.tril(S, -1).tocsr()
    t0 = time.perf_counter()
    cpu = O.masked_count(L.indptr, L.indices, L.indptr, L.indices, L.indptr, L.indices)
    tc = time.perf_counter() - t0
    res["cpu_sandia"] = {"triangles": cpu, "seconds": round(tc, 3), "cores": 1, "kind": "port"}
    log(f"oracle (1 core) Sandia: {cpu} triangles in {tc:.2f} s")
    A.free()
    res["mxm_unmasked"] = spgemm_bench(g, O)
    return res


def spgemm_bench(g, O, scale=14):
    """Unmasked C = A (+.*) A (int32 plus.times) on a symmetric R-MAT: device
    time, products expanded and entries of C, beside the oracle's sequential
    IKJ product on one host core."""
    rp, col = g.gen_rmat(scale, 16, seed=48, symmetric=True)
    n = len(rp) - 1
    v = (np.arange(len(col)) % 7 + 1).astype(np.int32)
    A = g.Matrix.from_csr(rp, col, v)
    Cm = g.Matrix(g.INT32)
    g.mxm(Cm, None, g.ACC_NONE, "PlusMultiplies", A, A)
    times = []
    for _ in range(3):
        g.synchronize()
        t0 = time.perf_counter()
        g.mxm(Cm, None, g.ACC_NONE, "PlusMultiplies", A, A)
        times.append(time.perf_counter() - t0)
    products = int(np.diff(rp)[col].sum())
    cnnz = Cm.nnz
    t0 = time.perf_counter()
    _, ccol, _ = O.mxm_unmasked(rp, col, v, rp, col, v, 0, 0)
    tc = time.perf_counter() - t0
    ms = float(np.median(times)) * 1e3
    log(f"mxm unmasked symmetric R-MAT {scale}: {products} products -> {cnnz} entries in {ms:.1f} ms "
        f"({products / (ms * 1e-3) / 1e9:.2f} G products/s); oracle (1 core) {tc:.2f} s")
    assert len(ccol) == cnnz
    A.free()
    Cm.free()
    return {"graph": f"symmetric R-MAT scale {scale}, edge factor 16", "n": n, "stored_nnz": len(col),
            "products": products, "c_nnz": cnnz, "ms_median": round(ms, 2),
            "cpu_ikj": {"seconds": round(tc, 3), "cores": 1, "kind": "port"}}


def io_bench(g, scale, ef):
    """Edge-list files: the R-MAT graph written as .mtx and .mtxbin, then read
    back on the host (gbk_mtx_read) and read into a device matrix with CSR and
    CSC (Matrix::init_from_mtx)."""
    import tempfile
    rp, col = g.gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed=1, symmetric=False, permute=True)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    cols = col.astype(np.int64)
    res = {"graph": f"R-MAT scale {scale}, edge factor {ef}", "entries": len(cols)}
    with tempfile.TemporaryDirectory() as d:
        for suf in ("mtx", "mtxbin"):
            p = os.path.join(d, "g." + suf)
            t0 = time.perf_counter()
            g.mtx_write(p, n, n, rows, cols)
            tw = time.perf_counter() - t0
            size = os.path.getsize(p)
            t0 = time.perf_counter()
            g.mtx_read(p)
            tr = time.perf_counter() - t0
            t0 = time.perf_counter()
            A = g.Matrix.from_mtx(p, g.FP32)
            tb = time.perf_counter() - t0
            ok = A.nnz == len(cols)
            A.free()
            os.remove(p)
            res[suf] = {"bytes": size, "write_s": round(tw, 3), "read_s": round(tr, 3),
                        "read_to_device_matrix_s": round(tb, 3),
                        "read_Mentries_per_s": round(len(cols) / tr / 1e6, 1), "nnz_ok": ok}
            log(f"{suf}: {size / 1e6:.0f} MB, write {tw:.2f} s, host read {tr:.2f} s "
                f"({len(cols) / tr / 1e6:.0f} M entries/s), read into a device CSR + CSC {tb:.2f} s")
    return res


def cpu_baseline(O, rp, col, vals, x, algo_bytes, seconds):
    """Times the oracle's OpenMP CSR SpMV (the reference's partition and fold) on the host."""
    nthreads = int(os.environ.get("GBK_CPU_THREADS", "0")) or len(os.sched_getaffinity(0))
    nthreads = min(nthreads, 16)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    first, last = O.balance_rows(rp, nthreads)
    rp64 = np.ascontiguousarray(rp, dtype=np.int64)
    col32 = np.ascontiguousarray(col, dtype=np.int32)
    y = np.zeros(len(x), dtype=np.float32)
    O.spmv_f32_omp(rp64, col32, vals, x, y, nthreads, first, last)  # warm-up
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end and len(times) < 200:
        t0 = time.perf_counter()
        O.spmv_f32_omp(rp64, col32, vals, x, y, nthreads, first, last)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    gbs = algo_bytes / med / 1e9
    log(f"CPU baseline ({nthreads} threads): median {med * 1e3:.2f} ms over {len(times)} runs, {gbs:.1f} GB/s")
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"full graph, median of {len(times)} SpMV runs ({med * 1e3:.2f} ms each)",
            "ms_per_spmv": round(med * 1e3, 3)}


if __name__ == "__main__":
    main()
