#!/usr/bin/env python3
"""Benchmark of the GAS engine on one MI355X at RMAT scale 24.

Workload (one step): a plain SSSP run of the reference's vertex program
(sssp.cu:69-111: all vertices active first, Bellman-Ford-style min-gather)
over RMAT scale 24 (16,777,216 vertices, edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 6, self-loops and duplicate arcs removed, uint32
weights uniform on [1, 64]) from the vertex of maximum out-degree.  The graph,
weights and vertex data are resident in HBM before timing; a step restores
the vertex data on the device, runs the engine loop to an empty frontier and
leaves the result in HBM.

value = GTEPS = sum of out-degrees of reached vertices / step time (the
reference's TEPS, sssp.cu:195-213).  ms_per_step_with_get_results adds the
reference's getResults (the distances copied back to the host, bfs.cu:113-126).
roofline = the dominant kernel, the blocked gather (k_gather_blocked) that
runs the full gathers: algorithmic bytes 12 B per gathered in-edge + 16 B per
active vertex, divided by its measured duration (hipEvents on the engine's
stream).  cpu_baseline = the CPU oracle (restated refgas.h engine, one
thread) on the same graph, which also checks the GPU result bit for bit.
Extra fields report plain BFS, CC and PageRank at the same scale, SSSP with
the edge arrays streamed from host memory, 64-source batches, the query
server, CC on the 4096^2 grid and the PageRank precision form at RMAT-22.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# The query server runs its three pools (BFS, SSSP, PPR) on three streams.
# With HIP's default of 4 hardware queues per process, streams created
# earlier in the process (the engines' own) can leave two pools on one
# queue, whose kernels then run one after another: a BFS level waits behind
# a PPR pass.  8 queues give every pool its own.  Set before HIP starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "GTEPS + HBM GB/s fraction at RMAT-24; batched vs plain aggregate TEPS, p99 lat"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_loop(eng, all_active_n=None, source=None):
    """The reference app loop on an engine whose graph is bound."""
    eng.resetVertexData()
    if all_active_n is not None:
        eng.setActive(0, all_active_n)
    else:
        eng.setActive(source, source + 1)
    it = 0
    eng.setIterationCount(0)
    while eng.countActive():
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        it += 1
        eng.setIterationCount(it)
    return it


def run_engine(eng, n):
    """The reference's CC and PageRank apps: every vertex active, then the
    engine's own loop (connected_component.cu:75-77, pagerank.cu:95-97:
    setActive + run(), gpugas.h:1792-1807)."""
    eng.resetVertexData()
    eng.setActive(0, n)
    eng.setIterationCount(0)
    return eng.run()


def timed(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--seed", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="all",
                    help="comma list of extra workloads (bfs,pagerank,cc,out_of_core,batched,"
                         "serving,cc_grid,pr_precision) or 'all'")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (reference point for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU sharding is a design, not a runnable path", file=sys.stderr)
        return 2

    import graphreduce_amd as gr

    if args.calibrate_copy:
        gbs = gr.device_copy_gbs(1 << 30, repeats=3)
        print(json.dumps({"copy_bytes": 1 << 30, "copy_read_plus_write_gbs": round(gbs, 1)}))
        return 0

    info = gr.device_info()
    log(f"device: {info}")
    n = 1 << args.scale
    t = time.time()
    g = gr.Graph.rmat(args.scale, 16, seed=args.seed)
    gr.synchronize()
    log(f"rmat{args.scale}: V={g.n_vertices} E={g.n_edges} built on device in {time.time() - t:.2f}s")
    src, src_deg = g.max_out_degree_vertex()
    off_out, dsts = g.csr()
    outdeg = np.diff(off_out).astype(np.int64)
    w_dev = g.hash_weights_device(seed=args.seed)

    # ---------------------------------------------------------- SSSP (value)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP)
    eng.bindGraph(g, vd, w_dev)
    from graphreduce_amd import tune

    tuned = tune.apply_tuning(eng, gr.SSSP)
    last = {"trips": 0}

    def step():
        last["trips"] = run_loop(eng, all_active_n=n)
        return last["trips"]

    for _ in range(args.warmup):
        step()
    sec = timed(step, args.steps, gr.synchronize)
    # the reached set of the timed runs (every run reaches the same vertices)
    trips = last["trips"]
    dist = eng.readVertexData()
    reached = dist < gr.SSSP_INF
    e_reached = int(outdeg[reached].sum())
    gteps = e_reached / sec / 1e9
    log(f"SSSP: trips={trips} reached={int(reached.sum())} E_reached={e_reached} "
        f"{sec * 1e3:.3f} ms/step {gteps:.2f} GTEPS")

    # the reference's timed region also holds getResults (bfs.cu:113-126)
    def step_with_results():
        step()
        eng.getResults()

    sec_results = timed(step_with_results, max(2, args.steps // 2), gr.synchronize)

    # kernel timing pass: one more step with hipEvents around every phase
    eng.resetStats()
    eng.enableTiming(True)
    step()
    st = eng.stats()
    eng.enableTiming(False)
    eng.resetStats()
    # Dominant kernel: the blocked gather, which runs the full gathers (the
    # dense iterations, counted with the merge-path tile gathers in tile_*);
    # the first, all-active gather and the sparse iterations are gathered by
    # pushing out-edges (push_*).
    tile_bytes = 12 * st["tile_gather_edges"] + 16 * st["tile_gather_rows"]
    gather_s = st["gather_ms"] / 1e3
    achieved = tile_bytes / gather_s / 1e9 if gather_s > 0 else 0.0
    ref_bytes = 12 * st["gathered_edges"] + 16 * st["active_vertices"]
    phase_s = (st["gather_ms"] + st["push_ms"]) / 1e3
    phases = {k: round(st[k + "_ms"], 4) for k in ("gather", "apply", "scatter", "next")}
    phases["gather"] = round(st["gather_ms"] + st["push_ms"], 4)
    phases["gather_tiles"] = round(st["gather_ms"], 4)
    phases["gather_push"] = round(st["push_ms"], 4)
    launches = {k: st[k + "_launches"] for k in ("gather", "apply", "scatter", "next")}
    step_alg_gbs = ref_bytes / sec / 1e9
    log(f"  phases ms {phases} launches {launches}; tile kernel {achieved:.1f} GB/s algorithmic")

    # The random-read floor of a full gather: one bare pass of dist[srcs[e]]
    # over every CSC slot of the same degree-ordered layout (a full gather of
    # this run covers > 99 % of the slots).
    floor_ms = g.source_read_ms(ordered=True)
    # launches that gathered: scatterActivate launches the fused blocked
    # gather wherever the pull may run, and the kernel stands down on the
    # device where it does not (a few microseconds each, left in gather_ms)
    idle = st["idle_gather_launches"]
    working = max(1, launches["gather"] - idle)
    avg_launch = st["gather_ms"] / working

    # HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # summary (calibrated against a plain copy kernel)
    prof_traffic = None
    tfile = ROOT / "profiles" / "gather_blocked_sssp_rmat24_traffic.md"
    if tfile.exists():
        import re

        m = re.search(r"hbm_bytes_per_launch: (\d+)", tfile.read_text())
        prof_traffic = int(m.group(1)) if m else None

    result = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "ms_per_step_with_get_results": round(sec_results * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (device RMAT generator, counter-based hash streams)",
        "config": {
            "workload": f"plain SSSP (sssp.cu loop, all-active start) on RMAT-{args.scale}, "
                        f"edge factor 16, a/b/c/d=.57/.19/.19/.05, seed {args.seed}, "
                        "dedup + no self-loops, weights U[1,64], source = max out-degree",
            "n_vertices": g.n_vertices,
            "n_edges": int(g.n_edges),
            "source": src,
            "source_out_degree": src_deg,
            "trips": trips,
            "edges_reached": e_reached,
        },
        "roofline": {
            "kernel": "k_gather_blocked<SSSPProgram> (the full gathers of the run; those "
                      "launched by scatterActivate in place of the pull also mark the next "
                      "active set)",
            "blocked_launches": st["blocked_gathers"],
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": prof_traffic,
            "bytes_per_launch": round(tile_bytes / working),
            "avg_launch_ms": round(avg_launch, 5),
            "launches": working,
            "idle_launches": idle,
            "fused_activations": st["fused_activations"],
            "algorithmic_bytes": "12 B per gathered in-edge (source id, weight, source "
                                 "distance) + 16 B per active vertex",
            "push_gathers": {
                "launches": st["push_gathers"],
                "fused_with_scatter": st["fused_gathers"],
                "edges": st["push_edges"],
                "ms": round(st["push_ms"], 4),
                "alg_gbs": round(12 * st["push_edges"] / (st["push_ms"] / 1e3) / 1e9, 1)
                if st["push_ms"] > 0 else None,
                "algorithmic_bytes": "12 B per pushed out-edge (target id, weight, candidate)",
            },
            "gather_phase_reference_gbs": round(ref_bytes / phase_s / 1e9, 1) if phase_s else None,
            "gather_phase_reference_frac": round(ref_bytes / phase_s / 1e9 / HBM_PEAK_GBS, 4)
            if phase_s else None,
            "reference_bytes": "12 B per in-edge of every active vertex + 16 B per active "
                               "vertex over all 9 gathers, as the reference's full gathers "
                               "would move them",
            "random_read_floor_ms": round(floor_ms, 4),
            "kernel_over_floor": round(avg_launch / floor_ms, 3) if floor_ms > 0 else None,
            "random_read_floor": "one bare pass of value[srcs[e]] over all CSC slots in CSC "
                                 "order (gr_graph_source_read_probe): the random 4-byte reads "
                                 "of a row-ordered gather, without edge data, reduction or "
                                 "writes; the blocked gather reads sources in ascending order "
                                 "per row block and runs below it",
        },
        "phase_ms_per_step": phases,
        "tuning": "graphreduce_amd/tuned.json (autotuner)" if tuned else "defaults",
        "step_algorithmic_gbs": round(step_alg_gbs, 1),
    }

    extra = {}
    if not args.no_extras:
        try:
            extra.update(run_extras(gr, g, args, src, outdeg, w_dev))
        except Exception as exc:  # extras never hide the headline
            extra["error"] = repr(exc)
    result["extra"] = extra

    # ---------------------------------------------------------- CPU baseline
    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(gr, g, args, src, off_out, dsts, dist, e_reached)
        except Exception as exc:
            result["cpu_baseline"] = {"error": repr(exc)}
    # release device objects while the runtime is up (no finalizer at exit)
    eng.close()
    w_dev.free()
    g.close()
    print(json.dumps(result), flush=True)
    return 0


def _frac(nbytes, sec):
    return round(nbytes / sec / 1e9 / HBM_PEAK_GBS, 4) if sec > 0 else None


def run_extras(gr, g, args, src, outdeg, w_dev):
    """The other workloads of the RMAT-24 config and the other configs.  Each
    plain program reports its step time, GTEPS, the reference's algorithmic
    bytes per step (BFS 4 E_reached + 16 V; CC and
    PageRank 8 E_in(A) + 16 #A summed over the run) over the step time as a
    fraction of HBM peak, and the same for its gather phase."""
    from graphreduce_amd import tune

    want = set(args.extras.split(",")) if args.extras != "all" else None
    on = lambda name: want is None or name in want  # noqa: E731
    out = {}
    n = g.n_vertices
    if on("bfs"):  # plain BFS (bfs.cu loop)
        vd = np.zeros(n, gr.PROGRAMS[gr.BFS]["vertex"])
        vd["depth"] = -1
        eng = gr.GASEngine(gr.BFS)
        eng.bindGraph(g, vd)
        tune.apply_tuning(eng, gr.BFS)
        step = lambda: run_loop(eng, source=src)  # noqa: E731
        trips = step()
        depth = eng.readVertexData()["depth"]
        e_r = int(outdeg[depth >= 0].sum())
        sec = timed(step, max(3, args.steps), gr.synchronize)
        eng.resetStats()
        eng.enableTiming(True)
        step()
        st = eng.stats()
        eng.close()
        out["bfs"] = dict(gteps=round(e_r / sec / 1e9, 3), ms=round(sec * 1e3, 4), trips=trips,
                          edges_reached=e_r,
                          alg_gbs=round((4 * e_r + 16 * n) / sec / 1e9, 1),
                          step_frac=_frac(4 * e_r + 16 * n, sec),
                          phase_ms={k: round(st[k + "_ms"], 4)
                                    for k in ("apply", "scatter", "next")},
                          pull_iterations=st["pull_iterations"])
        log(f"BFS: {out['bfs']}")
    if on("pagerank"):  # plain PageRank (pagerank.cu: setActive + run, activation-driven)
        vd = np.zeros(n, gr.PROGRAMS[gr.PAGERANK]["vertex"])
        vd["rank"] = np.float32(0.15)
        vd["numOutEdges"] = outdeg.astype(np.int32)
        eng = gr.GASEngine(gr.PAGERANK)
        eng.bindGraph(g, vd)
        tune.apply_tuning(eng, gr.PAGERANK)
        step = lambda: run_engine(eng, n)  # noqa: E731
        trips = step()
        sec = timed(step, 2, gr.synchronize)
        # one more run with the phase events for the gather figures
        eng.resetStats()
        eng.enableTiming(True)
        step()
        st = eng.stats()
        eng.close()
        ein = st["gathered_edges"]
        pb = 8 * ein + 16 * st["active_vertices"]
        gs_ = (st["gather_ms"] + st["push_ms"]) / 1e3
        out["pagerank"] = dict(gteps=round(ein / sec / 1e9, 3), ms=round(sec * 1e3, 3),
                               trips=trips, gathered_edges=ein,
                               step_frac=_frac(pb, sec), gather_ms=round(gs_ * 1e3, 3),
                               gather_alg_gbs=round(pb / gs_ / 1e9, 1),
                               gather_frac=_frac(pb, gs_))
        log(f"PR: {out['pagerank']}")
    if on("cc"):  # plain CC on the symmetrised graph (connected_component.cu: setActive + run)
        gs = gr.Graph.rmat(args.scale, 16, seed=args.seed, symmetrize=True)
        vd = np.arange(n, dtype=np.int32)
        eng = gr.GASEngine(gr.CC)
        eng.bindGraph(gs, vd)
        tune.apply_tuning(eng, gr.CC)
        step = lambda: run_engine(eng, n)  # noqa: E731
        trips = step()
        sec = timed(step, max(3, args.steps), gr.synchronize)
        eng.resetStats()
        eng.enableTiming(True)
        step()
        st = eng.stats()
        eng.close()
        sym_edges = int(gs.n_edges)
        gs.close()
        ein = st["gathered_edges"]
        cb = 8 * ein + 16 * st["active_vertices"]
        gs_ = (st["gather_ms"] + st["push_ms"]) / 1e3
        tile_s = st["gather_ms"] / 1e3
        tb = 8 * st["tile_gather_edges"] + 16 * st["tile_gather_rows"]
        out["cc"] = dict(gteps=round(ein / sec / 1e9, 3), ms=round(sec * 1e3, 3), trips=trips,
                         n_edges=sym_edges, gathered_edges=ein, step_frac=_frac(cb, sec),
                         gather_ms=round(gs_ * 1e3, 3), gather_alg_gbs=round(cb / gs_ / 1e9, 1),
                         gather_frac=_frac(cb, gs_), blocked_gathers=st["blocked_gathers"],
                         full_gather_kernel_frac=_frac(tb, tile_s))
        log(f"CC: {out['cc']}")
    if on("out_of_core"):
        out["sssp_out_of_core"] = run_out_of_core(gr, g, args, src, outdeg, w_dev)
        log(f"SSSP out of core: {out['sssp_out_of_core']}")
    if on("batched"):
        out["batched"] = run_batched(gr, g, args, outdeg, w_dev)
    if on("serving"):
        out["serving"] = run_serving(gr, g, args, outdeg, w_dev)
    if on("cc_grid"):
        out["cc_grid_4096"] = run_cc_grid(gr)
    if on("pr_precision") and not args.no_cpu_baseline:
        out["pagerank_precision_rmat22"] = run_pr_precision(gr)
    return out


def run_batched(gr, g, args, outdeg, w_dev):
    """64-source BFS / SSSP in one batch vs 64 sequential plain runs of the
    same sources (seeded random vertices with out-degree > 0), every run on
    the same clock: host wall time around the call, synchronised on both
    sides.  The plain engines carry the tuned settings and are bound before
    timing; the batches run on a context (gr_batch) created before timing
    and leave their results in a preallocated device array; the one-shot
    calls, which create and free their context inside, are timed beside
    them.  All 64 rows are compared."""
    from graphreduce_amd import tune

    n = g.n_vertices
    rng = np.random.default_rng(args.seed)
    cand = np.flatnonzero(outdeg > 0)
    srcs = rng.choice(cand, 64, replace=False).astype(np.int32).tolist()
    res = {"sources": "64 seeded random vertices with out-degree > 0",
           "clock": "host wall time around each call, device synchronised before and after"}
    out = gr.device_empty(4 * n * 64)
    # --- BFS: one tuned engine, reset per source (the vertex init is the same)
    vd = np.zeros(n, gr.PROGRAMS[gr.BFS]["vertex"])
    vd["depth"] = -1
    eng = gr.GASEngine(gr.BFS)
    eng.bindGraph(g, vd)
    tune.apply_tuning(eng, gr.BFS)
    run_loop(eng, source=srcs[0])  # warm
    t_plain, er_tot, plain_rows = 0.0, 0, []
    for s_ in srcs:
        gr.synchronize()
        t0 = time.perf_counter()
        run_loop(eng, source=s_)
        gr.synchronize()
        t_plain += time.perf_counter() - t0
        row = eng.readVertexData()["depth"]
        er_tot += int(outdeg[row >= 0].sum())
        plain_rows.append(row)
    eng.close()
    # the batch context is created before timing, as the plain engine is bound
    # before timing; the one-shot call (context created and freed inside) is
    # timed beside it
    bctx = gr.Batch(g, w_dev)
    bctx.bfs(srcs, out=out)  # warm
    gr.synchronize()
    t0 = time.perf_counter()
    _, info = bctx.bfs(srcs, out=out)
    gr.synchronize()
    t_b = time.perf_counter() - t0
    depth = gr.download(out, np.int32, (64, n))
    same = all(np.array_equal(depth[i], plain_rows[i]) for i in range(64))
    gr.msbfs(g, srcs, out=out)  # warm
    gr.synchronize()
    t0 = time.perf_counter()
    gr.msbfs(g, srcs, out=out)
    gr.synchronize()
    t_once = time.perf_counter() - t0
    res["bfs"] = dict(plain_sequential_ms=round(t_plain * 1e3, 3), batched_ms=round(t_b * 1e3, 3),
                      batched_device_ms=round(info["ms"], 3),
                      plain_aggregate_gteps=round(er_tot / t_plain / 1e9, 3),
                      batched_aggregate_gteps=round(er_tot / t_b / 1e9, 3),
                      speedup=round(t_plain / t_b, 2), levels=info["levels"],
                      rows_compared=64, identical_to_plain=bool(same),
                      one_shot_ms=round(t_once * 1e3, 3),
                      one_shot_speedup=round(t_plain / t_once, 2))
    log(f"batched BFS: {res['bfs']}")
    plain_bfs = plain_rows
    del depth
    # the same batch as ONE persistent cooperative launch of every level
    # (GR_MSBFS_PERSISTENT): measured beside the default per-level launches
    bctx.bfs(srcs, persistent=True, out=out)  # warm
    gr.synchronize()
    t0 = time.perf_counter()
    _, pinfo = bctx.bfs(srcs, persistent=True, out=out)
    gr.synchronize()
    t_p = time.perf_counter() - t0
    pdepth = gr.download(out, np.int32, (64, n))
    res["bfs"]["persistent_launch"] = dict(ms=round(t_p * 1e3, 3),
                                           device_ms=round(pinfo["ms"], 3), grid=pinfo["grid"],
                                           speedup=round(t_plain / t_p, 2),
                                           identical_to_plain=bool(all(
                                               np.array_equal(pdepth[i], plain_bfs[i])
                                               for i in range(64))))
    del pdepth, plain_bfs
    log(f"batched BFS persistent: {res['bfs']['persistent_launch']}")
    # --- SSSP: a tuned engine per source (the start vertex is part of the
    # bound vertex data), all 64 timed
    t_plain, er_tot, rows = 0.0, 0, []
    for i, s_ in enumerate(srcs):
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[s_] = 0
        eng = gr.GASEngine(gr.SSSP)
        eng.bindGraph(g, vd, w_dev)
        tune.apply_tuning(eng, gr.SSSP)
        if i == 0:
            run_loop(eng, all_active_n=n)  # warm
        gr.synchronize()
        t0 = time.perf_counter()
        run_loop(eng, all_active_n=n)
        gr.synchronize()
        t_plain += time.perf_counter() - t0
        row = eng.readVertexData()
        er_tot += int(outdeg[row < gr.SSSP_INF].sum())
        rows.append(row)
        eng.close()
    bctx.sssp(srcs, out=out)  # warm
    gr.synchronize()
    t0 = time.perf_counter()
    _, info = bctx.sssp(srcs, out=out)
    gr.synchronize()
    t_b = time.perf_counter() - t0
    dist = gr.download(out, np.int32, (64, n))
    same = all(np.array_equal(dist[i], rows[i]) for i in range(64))
    bctx.close()
    gr.mssp(g, w_dev, srcs, out=out)  # warm
    gr.synchronize()
    t0 = time.perf_counter()
    gr.mssp(g, w_dev, srcs, out=out)
    gr.synchronize()
    t_once = time.perf_counter() - t0
    res["sssp"] = dict(plain_sequential_ms=round(t_plain * 1e3, 3), batched_ms=round(t_b * 1e3, 3),
                       batched_device_ms=round(info["ms"], 3),
                       plain_aggregate_gteps=round(er_tot / t_plain / 1e9, 3),
                       batched_aggregate_gteps=round(er_tot / t_b / 1e9, 3),
                       speedup=round(t_plain / t_b, 2), rounds=info["levels"],
                       rows_compared=64, identical_to_plain=bool(same),
                       one_shot_ms=round(t_once * 1e3, 3),
                       one_shot_speedup=round(t_plain / t_once, 2))
    log(f"batched SSSP: {res['sssp']}")
    out.free()
    return res


def _serve_once(gr, g, w_dev, rate, verts, kinds, names):
    srv = gr.Server(g, w_dev)
    t0 = time.perf_counter()
    ids = srv.load_test(rate, verts, kinds, keep=False)
    lat = {k: [] for k in names}
    cs_mix = []
    in_order = []  # latency per request, in arrival order
    for i, rid in enumerate(ids):
        _, cs, ms, _ = srv.wait(int(rid), want_result=False)
        lat[int(kinds[i])].append(ms)
        cs_mix.append(cs)
        in_order.append(ms)
    wall = time.perf_counter() - t0
    st = srv.stats()
    q = max(1, len(in_order) // 4)
    first_q, last_q = float(np.median(in_order[:q])), float(np.median(in_order[-q:]))
    row = {
        "offered_rate_rps": rate,
        "completed": st["completed"],
        "throughput_rps": round(st["throughput_rps"], 1),
        "wall_s": round(wall, 3),
        "p50_ms": round(st["p50_ms"], 2),
        "p99_ms": round(st["p99_ms"], 2),
        "max_ms": round(st["max_ms"], 2),
        "max_lanes_in_use": st["max_lanes_in_use"],
        "peak_lane_bytes": st["peak_bytes"],
        "steps": st["steps"],
        "first_quarter_p50_ms": round(first_q, 2),
        "last_quarter_p50_ms": round(last_q, 2),
    }
    for k, name in names.items():
        if lat[k]:
            a = np.sort(np.array(lat[k]))
            row[f"{name}_p50_ms"] = round(float(np.percentile(a, 50)), 2)
            row[f"{name}_p99_ms"] = round(float(np.percentile(a, 99)), 2)
    return srv, row, cs_mix


def run_serving(gr, g, args, outdeg, w_dev, n_req=384,
                rates=(50.0, 100.0, 200.0, 400.0, 1000.0, 2000.0), per_kind_sample=2):
    """Poisson streams of mixed BFS / SSSP / PPR queries (uniform kinds, seeded
    random sources with out-degree > 0) into the continuous-batching server at
    several mean rates; latency is submit -> result ready.  The configured
    rate is 2,000 req/s; `sustained` is the highest rate whose throughput
    reaches 90 % of the offer.  Throughput is completions over first submit
    -> last completion, so it includes the last requests' latency (a short
    stream at a high rate ends before its queue shows; `first_quarter_p50_ms`
    and `last_quarter_p50_ms` show the growth within the stream).  After the 2,000 req/s stream a sample of its
    queries is resubmitted one at a time on the idle server (alone) and their
    checksums must equal the ones from the mix; one BFS and one SSSP alone
    result are also compared with plain engine runs."""
    rng = np.random.default_rng(args.seed + 101)
    cand = np.flatnonzero(outdeg > 0)
    verts = rng.choice(cand, n_req).astype(np.int32)
    kinds = rng.integers(1, 4, n_req).astype(np.int32)
    names = {gr.QUERY_BFS: "bfs", gr.QUERY_SSSP: "sssp", gr.QUERY_PPR: "ppr"}
    res = {"graph": f"RMAT-{args.scale} seed {args.seed}", "requests_per_stream": n_req,
           "kinds": "uniform over BFS, SSSP, PPR (10 power iterations)", "by_rate": []}
    sustained = None
    for rate in rates:
        srv, row, cs_mix = _serve_once(gr, g, w_dev, rate, verts, kinds, names)
        res["by_rate"].append(row)
        if row["throughput_rps"] >= 0.9 * rate:
            sustained = row
        if rate == rates[-1]:
            sample = []
            for k in names:
                sample += [i for i in range(n_req) if kinds[i] == k][:per_kind_sample]
            equal = True
            alone_rows = {}
            for i in sample:
                rid = srv.submit(int(kinds[i]), int(verts[i]), keep=True)
                r_, cs, _, _ = srv.wait(rid, int(kinds[i]))
                equal &= cs == cs_mix[i]
                alone_rows.setdefault(int(kinds[i]), (int(verts[i]), r_))
        srv.close()
        log(f"serving @ {rate}: {row}")
    res["configured"] = res["by_rate"][-1]
    res["sustained"] = sustained
    # service time without other traffic (one request on an idle server) and
    # each pool's throughput with 64 requests of one kind at once: the
    # latencies of the mixed streams above come from the three pools' kernels
    # sharing the device, and the pools' throughputs bound the mix's capacity
    solo, pool = {}, {}
    kinds_ = (gr.QUERY_BFS, gr.QUERY_SSSP, gr.QUERY_PPR)
    srv = gr.Server(g, w_dev)
    for k in kinds_:
        lat_k = []
        for v in verts[:3]:
            rid = srv.submit(k, int(v), keep=False)
            lat_k.append(srv.wait(rid, want_result=False)[2])
        solo[names[k]] = round(float(np.median(lat_k)), 2)
    srv.close()
    for k in kinds_:
        srv = gr.Server(g, w_dev)
        t0 = time.perf_counter()
        ids = srv.load_test(1e9, verts[:64], np.full(64, k, np.int32), keep=False)
        for rid in ids:
            srv.wait(int(rid), want_result=False)
        pool[names[k]] = round(64 / (time.perf_counter() - t0), 1)
        srv.close()
    mix_capacity = 3.0 / sum(1.0 / pool[names[k]] for k in kinds_)
    res["solo_latency_ms"] = solo
    res["pool_throughput_rps_64_at_once"] = pool
    res["mix_capacity_bound_rps"] = round(mix_capacity, 1)
    res["alone_equals_mix"] = bool(equal)
    res["alone_vs_mix_sample"] = len(sample)
    v, row = alone_rows[gr.QUERY_BFS]
    plain, _ = gr.run_bfs(g, v)
    res["bfs_equals_plain_engine"] = bool(np.array_equal(row, plain))
    v, row = alone_rows[gr.QUERY_SSSP]
    plain, _ = gr.run_sssp(g, w_dev, v)
    res["sssp_equals_plain_engine"] = bool(np.array_equal(row, plain))
    return res


def run_out_of_core(gr, g, args, src, outdeg, w_dev, shard_edges=1 << 25):
    """Plain SSSP with the edge arrays in pinned host memory, cut into vertex
    ranges of <= 32M edges and streamed every iteration (the reference's
    out-of-core shards, gpugas.h:1398-1619); distances must equal the
    in-memory run's."""
    n = g.n_vertices
    s, d = g.edges()
    w = gr.hash_weights(len(s), seed=args.seed)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP)
    eng.setStreaming(shard_edges)
    t0 = time.perf_counter()
    eng.setGraph(n, vd, len(s), w, s, d)
    setup = time.perf_counter() - t0
    del s, d, w
    trips = run_loop(eng, all_active_n=n)  # warm-up
    eng.resetStats()
    eng.enableTiming(True)
    gr.synchronize()
    t0 = time.perf_counter()
    trips = run_loop(eng, all_active_n=n)
    gr.synchronize()
    sec = time.perf_counter() - t0
    st = eng.stats()
    dist = eng.readVertexData()
    eng.close()
    ref = gr.GASEngine(gr.SSSP)
    ref.bindGraph(g, vd, w_dev)
    run_loop(ref, all_active_n=n)
    same = bool(np.array_equal(dist, ref.readVertexData()))
    ref.close()
    e_r = int(outdeg[dist < gr.SSSP_INF].sum())
    return dict(shard_edges=shard_edges, trips=trips, ms=round(sec * 1e3, 2),
                gteps=round(e_r / sec / 1e9, 3), streamed_gb=round(st["streamed_bytes"] / 1e9, 2),
                shard_copies=st["streamed_shards"],
                h2d_gbs=round(st["streamed_bytes"] / sec / 1e9, 1),
                phase_ms={k: round(st[k + "_ms"], 2) for k in ("gather", "scatter", "apply", "next")},
                setup_s=round(setup, 1), equals_in_memory=same)


def run_cc_grid(gr):
    """CC on the 4096x4096 grid with 1 % of its edges deleted (seed 5)."""
    gg = gr.Graph.grid(4096, 0.01, seed=5)
    n = gg.n_vertices
    vd = np.arange(n, dtype=np.int32)
    eng = gr.GASEngine(gr.CC)
    eng.bindGraph(gg, vd)
    from graphreduce_amd import tune

    tuned = tune.apply_tuning(eng, gr.CC, uniform=True)  # chosen on the seed-6 grid
    gr.synchronize()
    t0 = time.perf_counter()
    trips = run_engine(eng, n)
    gr.synchronize()
    sec = time.perf_counter() - t0
    lab = eng.readVertexData()
    st = eng.stats()
    eng.close()
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    s, d = gg.edges()
    nc, lbl = cg.connected_components(sp.csr_matrix((np.ones(len(s)), (s, d)), shape=(n, n)),
                                      directed=False)
    mins = np.full(nc, n)
    np.minimum.at(mins, lbl, np.arange(n))
    ok = bool(np.array_equal(lab, mins[lbl]))
    out = dict(ms=round(sec * 1e3, 2), trips=trips, n_edges=int(gg.n_edges),
               components=int(nc), gathered_edges=st["gathered_edges"],
               gteps=round(st["gathered_edges"] / sec / 1e9, 3),
               us_per_trip=round(sec / max(1, trips) * 1e6, 1),
               labels_equal_scipy_min_id=ok,
               tuning="graphreduce_amd/tuned.json, uniform (grid seed 6)" if tuned else "defaults")
    gg.close()
    log(f"CC grid: {out}")
    return out


def run_pr_precision(gr):
    """PageRank precision form on RMAT-22 seed 3 vs the FP64 CPU oracle."""
    import oracle

    g = gr.Graph.rmat(22, 16, seed=3)
    out = {}
    ranks = {}
    for comp in (True, False):
        gr.pagerank_precise(g, max_iters=2, tol=0.0, compensated=comp)  # warm
        rk, info = gr.pagerank_precise(g, max_iters=50, tol=1e-7, compensated=comp)
        ranks[comp] = (rk, info)
    iters = ranks[True][1]["iterations"]
    off_out, _ = g.csr()
    off_in, srcs = g.csc()
    t0 = time.perf_counter()
    ref, l1 = oracle.pagerank_fp64_csr(g.n_vertices, off_out, off_in, srcs, iters)
    cpu_s = time.perf_counter() - t0
    # Error budget: the same recurrence with exact (FP64) in-edge sums but the
    # kernel's FP32 per-vertex operations and FP32 storage (0.15 + 0.85 * g
    # rounded per operation, rank / out-degree rounded), so that what is left
    # of the GPU's error beyond it is the summation's.
    import scipy.sparse as sp

    n = g.n_vertices
    deg = np.diff(off_out).astype(np.float32)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off_in))
    a = sp.csr_matrix((np.ones(len(srcs), np.float64), (rows, srcs.astype(np.int64))),
                      shape=(n, n))
    del rows
    damp = np.float32(1.0) - np.float32(0.15)
    r32 = np.full(n, np.float32(0.15), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        c32 = np.where(deg > 0, r32 / deg, np.float32(0.0)).astype(np.float32)
        for _ in range(iters):
            gsum = (a @ c32.astype(np.float64)).astype(np.float32)
            r32 = (np.float32(0.15) + damp * gsum).astype(np.float32)
            c32 = np.where(deg > 0, r32 / deg, np.float32(0.0)).astype(np.float32)
    del a
    rel_ops = np.abs(r32.astype(np.float64) - ref) / ref
    out["fp32_vertex_ops_exact_sums"] = dict(
        max_rel_err_vs_fp64=float(rel_ops.max()),
        p99_rel_err_vs_fp64=float(np.percentile(rel_ops, 99)),
        note="the recurrence with FP64 in-edge sums and the kernel's FP32 per-vertex "
             "operations: the error the per-vertex FP32 arithmetic alone leaves")
    out["fp32_rounding_of_fp64_result"] = dict(
        max_rel_err=float((np.abs(ref.astype(np.float32).astype(np.float64) - ref) / ref).max()),
        note="the error of storing the FP64 ranks in FP32: the floor for any FP32 result")
    for comp, name in ((True, "compensated"), (False, "plain_fp32")):
        rk, info = ranks[comp]
        rel = np.abs(rk.astype(np.float64) - ref) / ref
        out[name] = dict(iterations=info["iterations"], residual_fp64=info["residual"],
                         relative_residual=info["residual"] / float(rk.astype(np.float64).sum()),
                         residual_drop=info["residual"] / info["residuals"][0],
                         ms=round(info["ms"], 3),
                         gteps=round(info["gathered_edges"] / (info["ms"] / 1e3) / 1e9, 2),
                         gather_alg_gbs=round((8 * info["gathered_edges"] + 16 * g.n_vertices *
                                               info["iterations"]) / (info["gather_ms"] / 1e3)
                                              / 1e9, 1),
                         max_rel_err_vs_fp64=float(rel.max()),
                         p99_rel_err_vs_fp64=float(np.percentile(rel, 99)))
    out["stop_rule"] = "absolute FP64 L1 residual < 1e-7 or 50 iterations"
    out["fp64_oracle_cpu_s"] = round(cpu_s, 2)
    out["n_edges"] = int(g.n_edges)
    g.close()
    log(f"PR precision: {out}")
    return out


def cpu_baseline(gr, g, args, src, off_out, dsts, dist_gpu, e_reached):
    import oracle

    off_in, srcs = g.csc()
    _, csc_eid = g.perm()
    w = gr.hash_weights(g.n_edges, seed=args.seed)
    w_csc = w[csc_eid]
    log("cpu baseline: oracle SSSP (1 thread) on the same graph ...")
    dist, trips, st = oracle.sssp_csr(g.n_vertices, off_out, dsts, off_in, srcs, w_csc, src)
    sec = st["seconds"]
    exact = bool(np.array_equal(dist, dist_gpu))
    log(f"cpu baseline: {sec:.2f}s trips={trips} bit-exact={exact}")
    return {
        "value": round(e_reached / sec / 1e9, 5),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"one full SSSP run on the same RMAT-{args.scale} graph (loop only, "
                  f"{sec:.2f} s); CPU {cpu_model()}, nproc {os.cpu_count()}",
        "seconds": round(sec, 3),
        "gpu_result_bit_exact": exact,
    }


if __name__ == "__main__":
    sys.exit(main())
