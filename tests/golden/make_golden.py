"""Writes tests/golden/golden.{json,npz}: small graphs with expected outputs
computed by scipy (BFS levels, Dijkstra, connected components, an FP64 sparse
mat-vec of the all-active PageRank recurrence).  The oracle is never called
here, so the fixtures pin it from outside.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

OUT = Path(__file__).resolve().parent
INF = 2**31 - 1 - 100000


def adj(n, s, d, w=None):
    data = np.ones(len(s)) if w is None else w.astype(np.float64)
    return sp.csr_matrix((data, (s, d)), shape=(n, n))


def graphs():
    rng = np.random.default_rng(20261203)
    out = []
    # a path, a cycle with a chord, a star and two disjoint triangles
    out.append(("path8", 8, np.arange(7), np.arange(1, 8)))
    out.append(("cycle_chord", 6, np.array([0, 1, 2, 3, 4, 5, 0]), np.array([1, 2, 3, 4, 5, 0, 3])))
    out.append(("star_out", 9, np.zeros(8, int), np.arange(1, 9)))
    out.append(("triangles", 7, np.array([0, 1, 2, 3, 4, 5]), np.array([1, 2, 0, 4, 5, 3])))
    for i in range(4):
        n = int(rng.integers(20, 200))
        m = int(rng.integers(n, 6 * n))
        out.append((f"rand{i}", n, rng.integers(0, n, m), rng.integers(0, n, m)))
    return out


def main():
    rng = np.random.default_rng(7)
    cases, arrays = [], {}
    for name, n, s, d in graphs():
        s = np.asarray(s, np.int32)
        d = np.asarray(d, np.int32)
        outdeg = np.bincount(s, minlength=n)
        # BFS from the first vertex of maximum out-degree
        src = int(np.argmax(outdeg))
        lv = cg.shortest_path(adj(n, s, d), unweighted=True, indices=src)
        lv = np.where(np.isinf(lv), -1, lv).astype(np.int32)
        mx = lv.max()
        trips = int(mx + 1 + (1 if outdeg[lv == mx].sum() > 0 else 0))
        key = f"{name}_bfs"
        arrays.update({key + "_src": s, key + "_dst": d, key + "_out": lv})
        cases.append(dict(name=key, kind="bfs", n=n, source=src, trips=trips))
        # SSSP with weights in [1, 64]
        w = rng.integers(1, 65, len(s)).astype(np.int32)
        # parallel arcs: scipy would add their weights, the shortest path takes the lightest
        best = {}
        for a, b, c in zip(s.tolist(), d.tolist(), w.tolist()):
            best[(a, b)] = min(c, best.get((a, b), c))
        bs = np.array([k[0] for k in best], np.int32)
        bd = np.array([k[1] for k in best], np.int32)
        bw = np.array(list(best.values()), np.int32)
        dist = cg.dijkstra(adj(n, bs, bd, bw), indices=src)
        dist = np.where(np.isinf(dist), INF, dist).astype(np.int32)
        key = f"{name}_sssp"
        arrays.update({key + "_src": s, key + "_dst": d, key + "_w": w, key + "_out": dist})
        cases.append(dict(name=key, kind="sssp", n=n, source=src))
        # CC on the symmetrised graph: min vertex id per component
        ss, dd = np.concatenate([s, d]), np.concatenate([d, s])
        nc, lbl = cg.connected_components(adj(n, ss, dd), directed=False)
        mins = np.full(nc, n)
        np.minimum.at(mins, lbl, np.arange(n))
        key = f"{name}_cc"
        arrays.update({key + "_src": ss, key + "_dst": dd, key + "_out": mins[lbl].astype(np.int32)})
        cases.append(dict(name=key, kind="cc", n=n))
        # all-active FP64 PageRank, 12 steps
        A = adj(n, s, d).T.tocsr()
        r = np.full(n, 0.15)
        for _ in range(12):
            c = np.divide(r, outdeg, out=np.zeros(n), where=outdeg > 0)
            r = 0.15 + 0.85 * (A @ c)
        key = f"{name}_pr64"
        arrays.update({key + "_src": s, key + "_dst": d, key + "_out": r})
        cases.append(dict(name=key, kind="pr64", n=n, iterations=12))
    np.savez_compressed(OUT / "golden.npz", **arrays)
    (OUT / "golden.json").write_text(json.dumps(dict(
        generator="tests/golden/make_golden.py (scipy %s)" % __import__("scipy").__version__,
        cases=cases), indent=1))
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
