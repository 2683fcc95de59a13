#!/usr/bin/env python3
"""Benchmark of the gfx950 traversal engine on RMAT-26 (one MI355X).

One step = one BFSLVLv run (source: highest-degree vertex) + one CCv run +
one PRv run of 20 FP32 power iterations, on a symmetric RMAT graph of scale
26, edge factor 16 (a=.57 b=.19 c=.19 d=.05), generated and cleaned on the
GPU before timing.  `value` is whole-job GTEPS: (BFS edges of the reached
component + m for CC + 20*m for PR) / wall time of the K timed steps.  The
timed runs leave their results in HBM, as Graptor's driver times run()
alone (driver.C:245-255); one untimed run afterwards downloads them for the
parity check against the CPU run.

Also reported: per-algorithm GTEPS, the HBM roofline of the dominant step
(the PR edgemap: its kernels timed with HIP events on the engine's own
stream), BF16/FP16
PR per-iteration GTEPS, and the CPU baseline (the OpenMP restatement of
Graptor's algorithm in oracle/, timed on the host on a bounded sample).
`apps` times the other two benchmarks on the same graph, each beside its
CPU oracle, which also checks the results at full size: APR_vary (delta
PageRank, variant 2046) and KC_bucket (coreness).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS (BFS, CC, PR iter) on RMAT-26, 1 GPU; % HBM roofline; vs host CPU"
HBM_PEAK_GBS = 8000.0
SEED = 0x47524150


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pr_pull_bytes(n, m, s):
    """Algorithmic bytes of one PR edgemap y = A^T contrib: CSC offsets
    8(n+1), neighbour IDs 4m, one contribution of s bytes per edge, the y
    write 4n."""
    return 8 * (n + 1) + 4 * m + s * m + 4 * n


# The kernels of one PR edgemap, by profile scope: the cold-source write
# (with its mass reduction), the pull over the pulled prefix, the long-row
# fold and the cold-source accumulation.  Without the split only the pull
# and the long-row fold run.
PR_SCOPES = ("pr_cold_write", "pr_pull", "pr_long", "pr_cold_accum")


# HBM bytes of one PR edgemap (FP32, every edge binned, the previous
# iteration's tail fused into its write stage; averaged over a 20-iteration
# run) on relabelled RMAT-26 ef16,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over the edgemap's kernels
# (k_cold_write, k_pr_update_gaps, k_cold_accum, k_cold_combine_groups,
# k_cold_combine).  Each counter is scaled by what it reads for
# k_pr_update<fp32>, a plain stream of 12n bytes in and 8n out: FETCH_SIZE
# x 2.00, WRITE_SIZE x 1.00.  The summary with the commands is
# profiles/pr_edgemap_rmat26_pmc.md.  Other configurations report null.
MEASURED_TRAFFIC = {(26, 16, 0.0): 12.43e9}


def run_bytes(trace, n, algo):
    """Algorithmic HBM bytes of one BFSLVLv or CCv run from its per-step
    trace: sum over steps of 8*V_scan + 4*E_scan + w*E_scan + 4*V_upd.
    Push: the frontier's vertices (offsets) and out-edges (neighbour ID plus
    the destination's 4 B value), w = 4.  Pull: the destinations that pass
    the filter (BFS: the unvisited ones; CC: every vertex) and the in-edges
    scanned up to early exit; w = 1/8 for BFS (a frontier-bit test) and 4
    for CC (a label gather).  A fused step counts its relaxations as edges.
    V_upd: the vertices the step activates (4 B each)."""
    total = 0.0
    reached = 1
    for st in trace:
        e = st["edges_scanned"]
        if st["direction"] == 2:   # pull
            v = (n - reached) if algo == "bfs" else n
            w = 0.125 if algo == "bfs" else 4.0
        else:
            v = st["nactv_in"]
            w = 4.0
        total += 8.0 * v + (4.0 + w) * e + 4.0 * st["nactv_out"]
        reached += st["nactv_out"]
    return total


def pr_edgemap_ms(g):
    """Per-scope average ms of the profiled PR edgemaps and their sum per
    edgemap (each scope runs at most once per edgemap)."""
    per = {}
    launches = 0
    for sc in PR_SCOPES:
        ms, k = g.profile_get(sc)
        if k:
            per[sc] = ms / k
        launches = max(launches, k)
    return sum(per.values()), per, launches


def run(args):
    import numpy as np
    import graptor_amd as G

    if args.pr_split_bytes is not None:
        G.tune_set("pr_split_bytes", args.pr_split_bytes)
    t0 = time.time()
    g = G.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, SEED)
    n, m = g.n, g.m
    src = g.highest_degree_vertex()
    log("graph scale %d: n=%d m=%d source=%d (%.1fs)" % (args.scale, n, m, src, time.time() - t0))
    if not args.no_reorder:
        # preprocessing outside the timed region, as Graptor's VEBO
        # relabelling is (driver.C:209-216 before the rounds loop)
        t0 = time.time()
        g.reorder()
        log("degree-sorted relabelling (%.1fs)" % (time.time() - t0))

    # the timed runs leave their results on the device, as Graptor's driver
    # times run() alone (driver.C:245-255); one untimed run afterwards
    # downloads them for the parity check
    algos = set(args.algos.split(","))
    none = G.RunInfo()

    def step(download=False):
        lv, ib = g.bfslvl(src, download=download, fusion=args.fusion) if "bfs" in algos else (None, none)
        lab, ic = g.ccv(download=download, fusion=args.fusion) if "cc" in algos else (None, none)
        x, ip = g.prv(max_iter=args.pr_iters, tol=0.0, download=download) if "pr" in algos else (None, none)
        return (lv, ib), (lab, ic), (x, ip)

    for w in range(args.warmup):
        step()
        log("warmup %d done" % w)
    g.profile(True)
    g.profile_reset()
    tb = tc = tp = 0.0
    eb = ec = ep = 0
    t_start = time.perf_counter()
    for k in range(args.steps):
        (lv, ib), (lab, ic), (x, ip) = step()
        tb += ib.ms
        tc += ic.ms
        tp += ip.ms
        eb += ib.edges
        ec += ic.edges
        ep += ip.edges
        log("step %d: bfs %.2f ms (%d iters), cc %.2f ms (%d iters), pr %.2f ms" % (k, ib.ms, ib.iters, ic.ms, ic.iters,
                                                                                 ip.ms))
    wall = time.perf_counter() - t_start
    avg_pull_ms, per_scope, pull_launches = pr_edgemap_ms(g)
    g.profile(False)
    (lv, _), (lab, _), (x, _) = step(download=True)
    if algos != {"bfs", "cc", "pr"}:
        # a subset of the step (for counter runs): timings only
        print(json.dumps({"algos": sorted(algos), "device_ms": {"bfs": tb / args.steps, "cc": tc / args.steps,
                                                                 "pr": tp / args.steps}}), flush=True)
        g.destroy()
        return None
    # the BFS and CC runs once more with per-step traces, for their
    # algorithmic bytes
    g.bfslvl(src, download=False, trace=True, fusion=args.fusion)
    b_bfs = run_bytes(g.trace(), n, "bfs")
    g.ccv(download=False, trace=True, fusion=args.fusion)
    b_cc = run_bytes(g.trace(), n, "cc")
    total_edges = eb + ec + ep
    value = total_edges / wall / 1e9
    # with every edge binned, each iteration but the first carries the
    # previous iteration's tail in its write stage (gpr_pr_iterate_fused):
    # 8 B per vertex with out-edges (y and the degree read; the ranks are
    # not written, GPR_PR_SKIP_RANKS), and in the last fused iteration 12 B
    # for every vertex (x written too, also for the vertices without
    # out-edges), averaged over the profiled launches
    bytes_edgemap = pr_pull_bytes(n, m, 4)
    split = G.tune_get("pr_split_bytes")
    index = np.empty(n + 1, np.uint64)
    G._check(G.lib().gpr_graph_download(g.h, index.ctypes.data, None))
    n_linked = int(np.count_nonzero(np.diff(index)))
    del index
    k = args.pr_iters
    tail_bytes = 8 * (k - 2) * n_linked + 12 * n if split == 0 and k > 1 else 0
    bytes_pull = bytes_edgemap + tail_bytes / max(1, k)
    achieved = bytes_pull / (avg_pull_ms * 1e-3) / 1e9
    gteps = {
        "bfs": eb / (tb * 1e-3) / 1e9,
        "cc": ec / (tc * 1e-3) / 1e9,
        "pr_iter_fp32": ep / (tp * 1e-3) / 1e9,
    }
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (BFS levels, CC labels); f32 store / 64-bit fixed-point accumulate (PR)",
        "data": "synthetic RMAT generated on the GPU from a fixed seed",
        "config": {
            "workload": "RMAT-%d ef%d a=.57 b=.19 c=.19 d=.05 symmetric, cleaned: BFSLVLv from the highest-degree "
                        "vertex + CCv + PRv %d FP32 iterations" % (args.scale, args.edge_factor, args.pr_iters),
            "scale": args.scale, "edge_factor": args.edge_factor, "n": n, "m": m, "seed": SEED,
            "bfs_source": src, "parallelism": "single GPU",
            "vertex_order": "original" if args.no_reorder else "degree-sorted relabelling (untimed preprocessing)",
        },
        "gteps": gteps,
        "device_ms": {"bfs": tb / args.steps, "cc": tc / args.steps, "pr": tp / args.steps},
        "roofline": {
            "kernel": ("PR edgemap <fp32>: k_cold_write + k_pr_pull + k_pr_long + k_cold_accum" if "pr_cold_accum" in
                       per_scope else "PR edgemap <fp32>: k_pr_pull + k_pr_long"),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (MEASURED_TRAFFIC.get((args.scale, args.edge_factor, split))
                        if not args.no_reorder else None),
            "algorithmic_bytes_per_launch": bytes_pull,
            "edgemap_bytes": bytes_edgemap,
            "fused_tail_bytes_per_launch": tail_bytes / max(1, k),
            "vertices_with_out_edges": n_linked,
            "avg_launch_ms": avg_pull_ms,
            "kernels_avg_ms": per_scope,
            "launches": pull_launches,
            "pr_split_bytes": split,
        },
        # BFSLVLv and CCv against the same peak, on their algorithmic bytes
        # per run (run_bytes) over the run's device time
        "roofline_bfs": {"bound": "hbm", "achieved": b_bfs / (tb / args.steps * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": b_bfs / (tb / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_run": b_bfs, "avg_run_ms": tb / args.steps},
        "roofline_cc": {"bound": "hbm", "achieved": b_cc / (tc / args.steps * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": b_cc / (tc / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_run": b_cc, "avg_run_ms": tc / args.steps},
    }
    # reduced-precision PR, per iteration (PR_vary contribution formats)
    if not args.no_precision:
        prec = {}
        for fmt, s in (("bf16", 2), ("fp16", 2), ("cfp16", 2)):
            g.profile(True)
            g.profile_reset()
            g.pr_vary([(fmt, 0.0)], max_iter=args.pr_iters)
            xr, info = g.pr_vary([(fmt, 0.0)], max_iter=args.pr_iters)
            pm, _, _ = pr_edgemap_ms(g)
            g.profile(False)
            prec[fmt] = {"gteps_iter": info.edges / (info.ms * 1e-3) / 1e9,
                         "edgemap_avg_ms": pm,
                         "edgemap_GBs": pr_pull_bytes(n, m, s) / (pm * 1e-3) / 1e9,
                         "speedup_vs_fp32_iter": (info.edges / info.ms) / (ep / tp)}
        res["pr_precision"] = prec
    outs = None
    if not args.no_apps:
        res["apps"], outs = gpu_apps(g, args)
    if not args.no_cpu:
        res["cpu_baseline"], res["parity"] = cpu_baseline(g, args, src, lv, lab, x, res.get("apps"), outs)
    g.destroy()
    return res


def gpu_apps(g, args):
    """APR_vary (variant 2046: FP32 contributions to L1 < 1e-1, then
    scustomfp<true,6,9> to 1e-7, eps2 = 1e-2) and KC_bucket on the graph:
    one warm-up run, `steps` timed runs leaving the results in HBM, then one
    run that downloads them.  GTEPS counts edges as the reference does
    (every edge of an all-sources iteration, the frontier's edges otherwise;
    KC: the edges of the peeled vertices, i.e. m)."""
    out = {}
    res = {}
    g.apr_vary(download=False)
    ms = []
    for _ in range(args.steps):
        _, info = g.apr_vary(download=False)
        ms.append(info.ms)
    x, info = g.apr_vary()
    res["apr_vary"] = {"device_ms": sum(ms) / len(ms), "gteps": info.edges / (sum(ms) / len(ms) * 1e-3) / 1e9,
                       "iters": info.iters, "push_iters": info.push_iters, "edges": info.edges,
                       "l1_final": info.delta}
    out["apr"] = x
    g.kcore(download=False)
    ms = []
    for _ in range(args.steps):
        _, _, info = g.kcore(download=False)
        ms.append(info.ms)
    core, kmax, info = g.kcore()
    res["kc_bucket"] = {"device_ms": sum(ms) / len(ms), "gteps": info.edges / (sum(ms) / len(ms) * 1e-3) / 1e9,
                        "levels": info.iters, "rounds": info.push_iters, "largest_core": kmax}
    out["core"] = (core, kmax)
    log("apps: apr %.2f ms, kcore %.2f ms" % (res["apr_vary"]["device_ms"], res["kc_bucket"]["device_ms"]))
    return res, out


def cpu_baseline(g, args, src, lv_gpu, lab_gpu, x_gpu, apps=None, outs=None):
    """The OpenMP restatement of Graptor's algorithm (oracle/) on the host
    cores: one BFSLVLv run, one CCv run and a few PRv iterations on the same
    graph.  Its results double as a full-size parity check of the GPU run."""
    import numpy as np
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t0 = time.time()
    # the CPU runs on the same vertex layout as the GPU (internal IDs)
    index, nbrs = g.download()
    order = g.order()
    log("downloaded graph for the CPU baseline (%.1fs)" % (time.time() - t0))
    n, m = len(index) - 1, len(nbrs)
    src_int = g.remap_id(src)
    rounds = args.cpu_rounds

    def median_time(f):
        """Median wall time of `rounds` runs (driver.C:195 -rounds) and the
        last run's result."""
        ts, out = [], None
        for _ in range(rounds):
            t = time.perf_counter()
            out = f()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)), out

    t_bfs, (lv, bi) = median_time(lambda: O.graptor_bfs(index, nbrs, src_int, threads))
    deg = np.diff(index)
    e_bfs = int(deg[lv <= n].sum())
    log("cpu bfs %.2fs (median of %d)" % (t_bfs, rounds))
    t_cc, (lab, ci) = median_time(lambda: O.graptor_cc(index, nbrs, threads))
    log("cpu cc %.2fs (median of %d)" % (t_cc, rounds))
    pr_it = args.cpu_pr_iters
    t_pr, (xc, it, _) = median_time(lambda: O.graptor_pr(index, nbrs, pr_it, 0.0, threads))
    log("cpu pr %d iters %.2fs (median of %d)" % (pr_it, t_pr, rounds))
    tot_e = e_bfs + m + pr_it * m
    tot_t = t_bfs + t_cc + t_pr
    base = {
        "value": tot_e / tot_t / 1e9,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": "one BFSLVLv run + one CCv run + %d PRv iterations on the same RMAT-%d graph, each the median "
                  "of %d runs (OpenMP restatement of Graptor's push/pull engine, oracle/oracle.c)"
                  % (pr_it, args.scale, rounds),
        "gteps": {"bfs": e_bfs / t_bfs / 1e9, "cc": m / t_cc / 1e9, "pr_iter_fp32": pr_it * m / t_pr / 1e9},
        "seconds": {"bfs": t_bfs, "cc": t_cc, "pr": t_pr},
        "cpu": _cpu_model(),
    }
    # CPU results are in internal IDs: bring the GPU's (original IDs) over
    # and canonicalise the CPU labels to the minimum original ID per
    # component (CCv.C:121-169)
    minorig = np.full(n, np.iinfo(np.uint32).max, np.uint32)
    np.minimum.at(minorig, lab, order)
    parity = {"bfs_levels_equal": bool(np.array_equal(lv, lv_gpu[order])),
              "cc_labels_equal": bool(np.array_equal(minorig[lab], lab_gpu[order]))}
    if apps is not None and outs is not None:
        # APR_vary: the OpenMP restatement on the same graph; KC_bucket: the
        # textbook bucket algorithm (one core)
        t = time.perf_counter()
        xa, it, l1, ed, _ = O.apr_vary(index, nbrs, threads=threads)
        t_apr = time.perf_counter() - t
        log("cpu apr_vary %d iters %.2fs" % (it, t_apr))
        apps["apr_vary"]["cpu"] = {"seconds": t_apr, "gteps": ed / t_apr / 1e9, "cores": threads, "kind": "port",
                                   "iters": it}
        parity["apr_vary_l1_vs_oracle"] = float(np.abs(outs["apr"][order] - xa).sum())
        t = time.perf_counter()
        core, kmax = O.kcore(index, nbrs)
        t_kc = time.perf_counter() - t
        log("cpu kcore %.2fs" % t_kc)
        apps["kc_bucket"]["cpu"] = {"seconds": t_kc, "gteps": m / t_kc / 1e9, "cores": 1, "kind": "port",
                                    "algorithm": "sequential bucket peeling (Batagelj-Zaversnik)"}
        parity["kc_coreness_equal"] = bool(np.array_equal(outs["core"][0][order], core)
                                           and outs["core"][1] == kmax)
    return base, parity


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--cpu-pr-iters", type=int, default=5)
    ap.add_argument("--cpu-rounds", type=int, default=5, help="CPU baseline: median of this many runs")
    ap.add_argument("--fusion", type=int, default=0,
                    help="BFSLVLv/CCv fused steps: 0 = each benchmark's reference default, -1 = off, 1, 2, 3 = small")
    ap.add_argument("--algos", default="bfs,cc,pr",
                    help="run a subset of the step (for counter runs; prints timings only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-reorder", action="store_true")
    ap.add_argument("--no-apps", action="store_true", help="skip the APR_vary and KC_bucket timings")
    ap.add_argument("--pr-split-bytes", type=float, default=None,
                    help="override the tunable pr_split_bytes (FP32 bytes of pulled sources; 0 = no split)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is in the "
              "documentation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    res = run(args)
    if res is not None:
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
