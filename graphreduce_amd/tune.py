"""Autotuner of the engine's launch parameters.

Chooses, per vertex program and graph, the gather tile shape (merge-path
items per thread: 4, 8 or 16 in a 256-thread tile), the push/pull switch
point alpha of scatterActivate (pull when the out-edges of the changed
vertices exceed E / alpha) and, for the delta-gather programs, the share of E
up to which activation and the next push gather are fused and the share of
changed in-edges below which a gather pushes, by timing complete runs of the
reference loop on the device graph.  The default graph is RMAT-24 seed 7, so
that the benchmark's graph (seed 6) is not the one the settings were chosen
on.  Results are written as JSON and applied with
:func:`apply_tuning`.  Graphs of uniform degree (grids, road maps) have
settings of their own under "uniform", chosen with ``--grid`` on the 4096 x
4096 grid with 1 % of its edges deleted, seed 6 (the benchmark's grid is seed
5), by timing ``run()``.

Usage:  python -m graphreduce_amd.tune --scale 24 --seed 7 --out graphreduce_amd/tuned.json
        python -m graphreduce_amd.tune --grid 4096 --grid-seed 6
"""
from __future__ import annotations

import argparse
import itertools
import json
import time
from pathlib import Path

import numpy as np

import graphreduce_amd as gr

DEFAULT_FILE = Path(__file__).resolve().parent / "tuned.json"
TILES = (4, 8, 16)
ALPHAS = (1.0, 2.0, 4.0, 8.0, 15.0, 30.0, 60.0)
FUSE_SHARES = (0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.8)  # SSSP and CC only
PUSH_SHARES = (0.1, 0.2, 0.3, 0.45, 0.6, 0.8)  # SSSP and CC only
# the uniform-degree sweep (CC on a grid: thousands of trips per run)
GRID_ALPHAS = (4.0, 15.0, 60.0)
GRID_FUSE_SHARES = (0.0, 0.05, 0.1, 0.2, 0.4)
GRID_PUSH_SHARES = (0.1, 0.6)


def _loop(eng, n, source=None):
    eng.resetVertexData()
    if source is None:
        eng.setActive(0, n)
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


def _engine(program, g, weights=None):
    n = g.n_vertices
    if program == gr.BFS:
        vd = np.zeros(n, gr.PROGRAMS[gr.BFS]["vertex"])
        vd["depth"] = -1
    elif program == gr.SSSP:
        vd = np.full(n, gr.SSSP_INF, np.int32)
    elif program == gr.CC:
        vd = np.arange(n, dtype=np.int32)
    else:
        off, _ = g.csr()
        vd = np.zeros(n, gr.PROGRAMS[gr.PAGERANK]["vertex"])
        vd["rank"] = np.float32(0.15)
        vd["numOutEdges"] = np.diff(off)
    eng = gr.GASEngine(program)
    return eng, vd


def _run(eng, n):
    """The engine's own loop over every vertex (connected_component.cu:75-77)."""
    eng.resetVertexData()
    eng.setActive(0, n)
    eng.setIterationCount(0)
    return eng.run()


def tune_program(program, g, source, weights=None, repeats=2, log=print, uniform=False):
    """Returns {"gather_tile": vt, "alpha": a, ["fuse_share": f,] "ms": best}
    for one program.  uniform: the reduced sweep for graphs of uniform degree,
    timing run() over every vertex."""
    eng, vd = _engine(program, g)
    if program == gr.SSSP:
        vd[source] = 0
    eng.bindGraph(g, vd, weights)
    src = source if program == gr.BFS else None
    if uniform:
        def loop():
            return _run(eng, g.n_vertices)
    else:
        def loop():
            return _loop(eng, g.n_vertices, src)
    loop()  # warm
    best = None
    tiles = TILES if program != gr.BFS else (8,)
    alphas = ALPHAS
    fuses = FUSE_SHARES if program in (gr.SSSP, gr.CC) else (None,)
    pushes = PUSH_SHARES if program in (gr.SSSP, gr.CC) else (None,)
    if uniform:
        tiles, alphas = (8,), GRID_ALPHAS
        if program in (gr.SSSP, gr.CC):
            fuses, pushes = GRID_FUSE_SHARES, GRID_PUSH_SHARES
    for vt, alpha, fuse, push in itertools.product(tiles, alphas, fuses, pushes):
        eng.setGatherTile(vt)
        eng.setDirection(0, alpha)
        if fuse is not None:
            eng.setFuseShare(fuse)
            eng.setPushShare(push)
        ts = []
        for _ in range(repeats):
            gr.synchronize()
            t0 = time.perf_counter()
            loop()
            gr.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = min(ts) * 1e3
        log(f"  {gr.PROGRAMS[program]['name']} tile={vt} alpha={alpha} fuse={fuse} "
            f"push={push}: {ms:.3f} ms")
        if best is None or ms < best["ms"]:
            best = {"gather_tile": vt, "alpha": alpha, "ms": round(ms, 4)}
            if fuse is not None:
                best["fuse_share"] = fuse
                best["push_share"] = push
    eng.close()
    return best


def apply_tuning(eng: "gr.GASEngine", program: int, path: Path | str = DEFAULT_FILE,
                 uniform: bool = False) -> bool:
    """Applies the tuned parameters of `program` from `path` (if present);
    uniform: those chosen for graphs of uniform degree."""
    p = Path(path)
    if not p.exists():
        return False
    data = json.loads(p.read_text())
    if uniform:
        data = data.get("uniform", {})
    entry = data.get("programs", {}).get(gr.PROGRAMS[program]["name"])
    if not entry:
        return False
    eng.setGatherTile(int(entry["gather_tile"]))
    eng.setDirection(0, float(entry["alpha"]))
    if "fuse_share" in entry:
        eng.setFuseShare(float(entry["fuse_share"]))
    if "push_share" in entry:
        eng.setPushShare(float(entry["push_share"]))
    return True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=str(DEFAULT_FILE))
    ap.add_argument("--programs", default="BFS,SSSP,CC,PageRank")
    ap.add_argument("--grid", type=int, default=0,
                    help="tune CC on a grid of this side (1 %% of edges deleted) into 'uniform'")
    ap.add_argument("--grid-seed", type=int, default=6)
    args = ap.parse_args()
    if args.grid:
        tune_uniform(args.grid, args.grid_seed, args.out)
        return
    g = gr.Graph.rmat(args.scale, 16, seed=args.seed)
    src, _ = g.max_out_degree_vertex()
    w = g.hash_weights_device(seed=args.seed)
    out = {"graph": f"rmat{args.scale} seed {args.seed}", "device": gr.device_info()["name"],
           "programs": {}}
    names = {v["name"]: k for k, v in gr.PROGRAMS.items()}
    for name in args.programs.split(","):
        prog = names[name]
        gg = g
        if prog == gr.CC:
            gg = gr.Graph.rmat(args.scale, 16, seed=args.seed, symmetrize=True)
        print(f"tuning {name}", flush=True)
        out["programs"][name] = tune_program(prog, gg, src, w if prog == gr.SSSP else None)
        print(f"  -> {out['programs'][name]}", flush=True)
    Path(args.out).write_text(json.dumps(out, indent=1))
    print(f"wrote {args.out}")


def tune_uniform(side, seed, out_path):
    """CC on the grid: the settings go under "uniform" of out_path, the rest
    of the file is kept."""
    g = gr.Graph.grid(side, 0.01, seed=seed)
    p = Path(out_path)
    data = json.loads(p.read_text()) if p.exists() else {}
    uni = {"graph": f"grid {side} x {side}, 1 % of edges deleted, seed {seed}",
           "device": gr.device_info()["name"], "programs": {}}
    print("tuning CC (uniform)", flush=True)
    uni["programs"]["CC"] = tune_program(gr.CC, g, None, repeats=1, uniform=True)
    print(f"  -> {uni['programs']['CC']}", flush=True)
    g.close()
    data["uniform"] = uni
    p.write_text(json.dumps(data, indent=1))
    print(f"wrote {out_path}")


if __name__ == "__main__":
    main()
