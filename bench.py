#!/usr/bin/env python3
"""Benchmark: static PageRank on one MI355X, printed as one JSON line.

A step is one complete static solve (tolerance 1e-10 on the L-inf change, damping 0.85, FP64 ranks, at most 500
iterations) of a synthetic graph that is already resident in HBM, writing its ranks to a device buffer
(prh_solve_device). Default workload: the web-scale graph of the dynamic configuration, directed RMAT scale 26 with
edge factor 30 (67,108,864 vertices, 2,013,265,920 samples, a/b/c/d = 0.57/0.19/0.19/0.05, seed 42, duplicates
removed, one self-loop per vertex: 2.04B edges); --scale 24 --edge-factor 16 gives RMAT-24. Before timing, the
autotuner picks the launch plan (prh_autotune; --no-autotune keeps the defaults). `value` is GEdges/s per
iteration over the whole timed region (|E| x iterations / wall time of the K solves, including initialisation
and the ranks' return to the caller's vertex order). config.gedges_per_s_iter_host_buffers is the same solve with
host buffers (ranks copied over PCIe), reported beside it.

roofline: the dominant work is the rank update, once per iteration: one rank_update_kernel launch with the row
plan, one window_update_kernel launch with the banded plan (tuning.window), or expand_kernel + accumulate_kernel
(+ finish_bins_kernel) with the propagation-blocked plan (tuning.blocking), timed together. Its algorithmic bytes per iteration are 12|E| + 24|V| (per edge a 4-byte source
id and an 8-byte FP64 contribution; per vertex a 4-byte offset, 4-byte out-degree, 8-byte old-rank read and 8-byte
new-rank write; the L-inf reduction is fused), whatever the plan moves in addition. Its duration is measured with HIP
events recorded on the engine's stream around every iteration's update that did work inside the timed region.

roofline.traffic: HBM bytes per iteration's update, from rocprofv3 counter runs of this script on the same workload
and plan (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor that the same counter needs to give the
known byte count of a plain device copy, measured with --copy-calibration). The records, with the commands that
produced them, are in profiles/rank_update_traffic.json; one is reported only when its workload and plan strings
equal this run's.

cpu_baseline: the oracle's OpenMP restatement of the reference's pagerankStaticOmp (bit-identical results to the
compiled reference on the test graphs), run on this host's cores on the same graph for a bounded number of
iterations (about 10 s).

--workload dfp: the dynamic side instead. Each step applies a fresh seeded batch of --batch-fraction |E| (80%
insertions, 20% deletions) to the device graph in place (prh_graph_update: the out-CSR rewritten, the blocked plan
patched, the in-CSR left to be built when read) and times a DF-P solve warm-started from the previous graph's static ranks (engine-measured time:
the update, initialisation, marking and iterations, as PagerankResult.time), beside a GPU static solve of the same
updated graph with the update added; config.dfp_ms_without_update gives the solve alone. Its cpu_baseline is the oracle's port of the reference's asynchronous OpenMP DF-P
(pagerankPruneDynamicFrontierOmp<true>, the CPU solver the reference's experiment runs; bit-identical to the
compiled reference on one thread) on the same batch, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "pagerank-cuda-dynamic_amd"))
sys.path.insert(0, str(REPO / "oracle"))

HBM_PEAK_GBS = 8000.0
METRIC = "GEdges/s/iter + %HBM peak at τ=1e-10; batched & adjoint vs plain solve"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_graph(args):
    from pagerank_amd import HostGraph
    if args.workload == "dfp":
        ef = args.edge_factor
        g = HostGraph.rmat(args.scale, ef, 0.57, 0.19, 0.19, seed=42)
        desc = (f"DF-P on directed RMAT scale {args.scale} (edge factor {ef:g}, a/b/c/d=.57/.19/.19/.05, seed 42, "
                f"dedup, self-loops), batches of {args.batch_fraction:g}|E| (80% insertions) applied in place")
        return g, desc
    if args.workload == "rmat":
        ef = args.edge_factor
        g = HostGraph.rmat(args.scale, ef, 0.57, 0.19, 0.19, seed=42)
        desc = (f"static PageRank, directed RMAT scale {args.scale} (edge factor {ef:g}, a/b/c/d=.57/.19/.19/.05, "
                f"seed 42, dedup, self-loops)")
    else:
        g = HostGraph.grid(args.rows, args.cols, 0.62, seed=42)
        desc = f"static PageRank, road-like grid {args.rows}x{args.cols} (edges kept p=0.62, symmetric, self-loops)"
    return g, desc


def cpu_baseline(g, budget_s: float, kind: str = "port"):
    """Static solves (tolerance 1e-10) of the same graph on this host's cores, bounded to about budget_s seconds:
    iterations of one solve are cut at the budget, and a solve that converges early is repeated. kind "port": the
    oracle's OpenMP restatement of pagerankStaticOmp (bit-identical to the compiled reference on the test graphs);
    "reference": the reference's own pagerankStaticOmp compiled from its headers (oracle/_ref), its time being
    PagerankResult.time (graph loading into its DiGraph excluded), where that build exists."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if kind == "reference" and not oracle.ref_available():
        log("cpu baseline: oracle/_ref is not built here; timing the port instead")
        kind = "port"

    def run(k):
        if kind == "reference":
            _, it, ms = oracle.ref_pagerank(0, g, threads=threads, max_iterations=k)
        else:
            _, it, ms = oracle.static_omp(g, max_iterations=k, threads=threads)
        return it, ms
    it, ms = run(2)
    per_iter = ms / max(it, 1) / 1e3
    k = int(max(2, min(500, budget_s / max(per_iter, 1e-6))))
    iters, total_ms, solves = 0, 0.0, 0
    while total_ms / 1e3 < 0.8 * budget_s or solves == 0:
        it, ms = run(k)
        iters += it
        total_ms += ms
        solves += 1
        if it < 2:
            break
    value = g.M * iters / (total_ms / 1e3) / 1e9
    what = ("the reference's pagerankStaticOmp (compiled from its headers)" if kind == "reference"
            else "the OpenMP port of pagerankStaticOmp")
    return {"value": round(value, 4), "unit": "GEdges/s/iter", "cores": threads, "kind": kind,
            "sample": f"{solves} static solve(s) of {what} on the same graph, {iters} synchronous iterations in "
                      f"{total_ms / 1e3:.1f} s"}


def bench_c1(args) -> int:
    """The CPU configuration: static PageRank on RMAT scale 16 (edge factor 16, seed 42, dedup, self-loops; fixed) by the
    reference's own pagerankStaticOmp (compiled from its headers; PagerankResult.time averaged over 5 repeats) and by
    the OpenMP port, on this host's cores, beside the GPU static solve of the same graph."""
    import oracle
    from pagerank_amd import STATIC, Context, HostGraph, PagerankOptions, solve
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scale, ef = 16, 16.0
    g = HostGraph.rmat(scale, ef, 0.57, 0.19, 0.19, seed=42)
    rows = {}
    if oracle.ref_available():
        r_ref, it, ms = oracle.ref_pagerank(0, g, threads=threads, repeat=5)
        rows["reference"] = (it, ms)
    r_port, it, ms = oracle.static_omp(g, threads=threads)
    rows["port"] = (it, ms)
    with Context(0) as ctx:
        yg = ctx.upload(g, with_out=False)
        solve(yg, STATIC, PagerankOptions())  # the first solve builds the layout's plan
        res = solve(yg, STATIC, PagerankOptions(repeat=5))
        yg.close()
    rows["gpu"] = (res.iterations, res.time)
    for k, (it, ms) in rows.items():
        log(f"{k}: {it} iterations, {ms:.3f} ms, {g.M * it / (ms / 1e3) / 1e9:.3f} GEdges/s/iter")
    kind = "reference" if "reference" in rows else "port"
    it, ms = rows[kind]
    out = {
        "metric": "static PageRank on the CPU path (GEdges/s/iter), with the GPU solve of the same graph",
        "value": round(g.M * it / (ms / 1e3) / 1e9, 4),
        "unit": "GEdges/s/iter",
        "n_gpus": 1, "steps": 5, "warmup": 0, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"static PageRank, directed RMAT scale {scale} (edge factor {ef:g}, "
                               f"a/b/c/d=.57/.19/.19/.05, seed 42, dedup, self-loops), on {threads} host threads",
                   "vertices": g.N, "edges": g.M, "kind": kind,
                   "iterations": {k: v[0] for k, v in rows.items()},
                   "ms": {k: round(v[1], 3) for k, v in rows.items()},
                   "gpu_speedup_over_cpu": round(rows[kind][1] / rows["gpu"][1], 2)},
        "roofline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_dfp(args, ctx, g, desc, overridden: bool = False) -> int:
    """GPU DF-P per update batch against GPU static on the same updated graph; CPU: the async DF-P port.

    Each step applies a fresh seeded batch to the device graph in place (prh_graph_update: the out-CSR rewritten,
    the propagation-blocked plan patched, the in-CSR left to be built when read) and solves DF-P from the previous graph's static ranks, then Static on the same
    graph. Both times include the update (DF-P's solve reports it as preparation; it is added to Static's)."""
    from pagerank_amd import PRUNE_DYNAMIC_FRONTIER, STATIC, Batch, PagerankOptions, apply_batch, solve
    o = PagerankOptions()
    xg = ctx.upload(g)
    autotuned = False
    if not args.no_autotune and not overridden:
        t, ms, hit = ctx.autotune(xg, args.tuning_cache, iterations=10)
        autotuned = True
        log(f"autotune: {t} at {ms:.3f} ms/iteration")
    q = solve(xg, STATIC, o).ranks
    size = max(1, int(args.batch_fraction * g.M))
    dfp_ms, st_ms, upd_ms, its, st_its, last, solve_ms = [], [], [], [], [], None, []
    x = g
    for i in range(args.warmup + args.steps):
        b = Batch.random(x, size, 0.8, seed=1000 + i)
        y = apply_batch(x, b)  # host copy: the next batch's base and the CPU baseline's input
        u = xg.update(b)
        r = solve(xg, PRUNE_DYNAMIC_FRONTIER, o, batch=b, q=q)
        s = solve(xg, STATIC, o)
        log(f"batch {i}: -{b.deletions}/+{b.insertions}: update {u:.2f} ms, DF-P {r.time:.2f} ms with it "
            f"({r.iterations} it), static {s.time + u:.2f} ms with it ({s.iterations} it), L-inf DF-P vs static "
            f"{abs(r.ranks - s.ranks).max():.2e}")
        if i >= args.warmup:
            dfp_ms.append(r.time)
            solve_ms.append(r.time - r.preparationTime)
            st_ms.append(s.time + u)
            upd_ms.append(u)
            its.append(r.iterations)
            st_its.append(s.iterations)
            last = (y, b, q)
        q = s.ranks
        x = y
    xg.close()
    value = sum(dfp_ms) / len(dfp_ms)
    static = sum(st_ms) / len(st_ms)
    tuning = ctx.get_tuning()
    out = {
        "metric": f"DF-P solve time per {args.batch_fraction:g}|E| update batch (ms), and speedup over GPU static",
        "value": round(value, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "vertices": g.N, "edges": g.M, "batch_edges": size,
                   "dfp_iterations": its, "static_ms": round(static, 3), "static_iterations": st_its,
                   "update_ms": round(sum(upd_ms) / len(upd_ms), 3),
                   "dfp_ms_without_update": round(sum(solve_ms) / len(solve_ms), 3),
                   "speedup_over_gpu_static": round(static / value, 3), "plan": plan_name(tuning),
                   "autotuned": autotuned},
        "roofline": None,
    }
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        y, b, q = last
        _, it, ms = oracle.pagerank_async(PRUNE_DYNAMIC_FRONTIER, y, b, q, threads=threads)
        out["cpu_baseline"] = {"value": round(ms, 3), "unit": "ms", "cores": threads, "kind": "port",
                               "sample": f"one asynchronous OpenMP DF-P solve (the reference's "
                                         f"pagerankPruneDynamicFrontierOmp<true>, ported) of the last timed batch, "
                                         f"{it} sweeps"}
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(gib: float) -> int:
    """A plain device-to-device copy of a known size, for scaling the HBM byte counters."""
    import torch
    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 copies of {n * 8} bytes (read {n * 8}, write {n * 8} per copy)")
    return 0


def measured_traffic(desc: str, plan: str):
    path = REPO / "profiles" / "rank_update_traffic.json"
    try:
        recs = json.loads(path.read_text()).get("records", [])
    except (OSError, ValueError, AttributeError):
        return None
    for rec in recs:
        if rec.get("workload") == desc and rec.get("plan") == plan:
            return rec.get("hbm_bytes_per_launch")
    return None


def plan_name(t) -> str:
    """The rank-update plan a static solve runs with tuning t, and the kernels of one iteration."""
    order = ", rows in the caller's vertex order" if t.layout == 1 else ""
    if t.blocking == 0 and t.window == 1 and t.layout == 1:
        return "banded plan (rows in the caller's vertex order, a window of contributions in LDS): window_update_kernel"
    if t.blocking > 0:
        return (f"propagation blocking {t.blocking}{order}: expand_kernel<{t.blocking}> + "
                f"accumulate_kernel<{t.blocking}> + finish_bins_kernel")
    if t.layout == 1:
        return f"row plan{order} (block_edges {t.block_edges}): rank_update_kernel<0>"
    if t.source_slices > 0:
        return (f"source slices {t.source_slices} (huge_degree {t.huge_degree}, block_edges {t.block_edges}): "
                f"sliced_gather_kernel + sliced_finish_kernel")
    return (f"row plan (switch_degree {t.switch_degree}, huge_degree {t.huge_degree}, block_edges {t.block_edges}): "
            f"rank_update_kernel<0>")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["rmat", "grid", "dfp", "c1"], default="rmat")
    ap.add_argument("--batch-fraction", type=float, default=1e-5, help="dfp workload: batch size as a fraction of |E|")
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=float, default=30.0, help="rmat workload: edge samples per vertex")
    ap.add_argument("--rows", type=int, default=6000)
    ap.add_argument("--cols", type=int, default=4000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-kind", choices=["port", "reference"], default="port",
                    help="cpu baseline: the oracle's OpenMP port, or the reference's compiled pagerankStaticOmp")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch plan instead of timing the autotuner's candidates first")
    ap.add_argument("--tuning-cache", default=None, help="JSON cache file for the autotuner's choice")
    ap.add_argument("--source-slices", type=int, default=None, help="override the tuning's source_slices")
    ap.add_argument("--huge-degree", type=int, default=None, help="override the tuning's huge_degree")
    ap.add_argument("--blocking", type=int, default=None, help="override the tuning's blocking (propagation blocking)")
    ap.add_argument("--layout", type=int, default=None, help="override the tuning's layout (1: the caller's order)")
    ap.add_argument("--window", type=int, default=None, help="override the tuning's window (1: the banded plan, "
                                                            "with --layout 1 --blocking 0)")
    ap.add_argument("--copy-calibration", type=float, metavar="GIB", default=0.0,
                    help="only run a plain device copy of GIB GiB (for scaling the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported (the engine has no multi-GPU "
              "path: a solve runs on one device).", file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)

    import torch  # device buffer for the ranks; imported before the engine library (shared HIP runtime)

    if args.workload == "c1":
        return bench_c1(args)

    from pagerank_amd import STATIC, Context, PagerankOptions, pagerankStaticHip, solve_device

    t0 = time.time()
    g, desc = make_graph(args)
    log(f"graph N={g.N} M={g.M} built in {time.time() - t0:.1f}s")
    ctx = Context(0)
    overridden = any(v is not None for v in (args.source_slices, args.huge_degree, args.blocking, args.layout,
                                              args.window))
    if overridden:
        t = ctx.get_tuning()
        if args.source_slices is not None:
            t.source_slices = args.source_slices
        if args.huge_degree is not None:
            t.huge_degree = args.huge_degree
        if args.blocking is not None:
            t.blocking = args.blocking
        if args.layout is not None:
            t.layout = args.layout
        if args.window is not None:
            t.window = args.window
        elif t.layout != 1:
            t.window = 0
        ctx.set_tuning(t)
    if args.workload == "dfp":
        return bench_dfp(args, ctx, g, desc, overridden)
    yg = ctx.upload(g, with_out=False)
    autotuned = False
    if not args.no_autotune and not overridden:
        t0 = time.time()
        t, ms, hit = ctx.autotune(yg, args.tuning_cache, iterations=10)
        autotuned = True
        log(f"autotune: {t} at {ms:.3f} ms/iteration ({'cache' if hit else f'{time.time() - t0:.1f} s'})")
    tuning = ctx.get_tuning()
    o = PagerankOptions()
    ranks = torch.empty(g.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for i in range(args.warmup):
        st = solve_device(yg, STATIC, o, None, ranks.data_ptr())
        log(f"warmup {i}: {st.iterations} iterations, {st.time:.2f} ms")
    ctx.kernel_timing(True)
    ctx.kernel_timing_read()
    iters = 0
    tcomp = 0.0
    torch.cuda.synchronize()
    t_start = time.perf_counter()  # every solve returns after the engine's stream has drained
    for _ in range(args.steps):
        st = solve_device(yg, STATIC, o, None, ranks.data_ptr())
        iters += st.iterations
        tcomp += st.computation_time
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    launches, kms = ctx.kernel_timing_read()
    ctx.kernel_timing(False)
    log(f"timed {args.steps} solves: {iters} iterations, {t_total * 1e3:.1f} ms, rank-update launches {launches}")
    # The same solve through host buffers (ranks copied back over PCIe, as the C++ drop-in does): reported beside
    # the headline, never as it.
    t0 = time.perf_counter()
    r = pagerankStaticHip(yg, o)
    host_rate = g.M * r.iterations / (time.perf_counter() - t0) / 1e9

    plan = plan_name(tuning)
    value = g.M * iters / t_total / 1e9
    kavg_s = kms / max(launches, 1) / 1e3
    bytes_alg = 12 * g.M + 24 * g.N
    achieved = bytes_alg / kavg_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GEdges/s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "vertices": g.N, "edges": g.M, "tolerance": 1e-10, "damping": 0.85,
                   "tuning": {"switch_degree": tuning.switch_degree, "huge_degree": tuning.huge_degree,
                              "source_slices": tuning.source_slices, "block_edges": tuning.block_edges,
                              "blocking": tuning.blocking, "layout": tuning.layout, "window": tuning.window,
                              "autotuned": autotuned},
                   "iterations_per_solve": iters / args.steps,
                   "gedges_per_s_iter_host_buffers": round(host_rate, 4),
                   "gedges_per_s_iter_compute_only": round(g.M * iters / (tcomp / 1e3) / 1e9, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(desc, plan),
                     "kernel": plan, "kernel_avg_us": round(kavg_s * 1e6, 2),
                     "bytes_per_launch": bytes_alg},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds, args.cpu_kind)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
