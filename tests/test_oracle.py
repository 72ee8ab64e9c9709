"""Pins of the CPU oracle (oracle/) against independent references.

Each function of the oracle is checked against something that does not share
its code: scipy's graph routines (BFS levels, Dijkstra, connected
components), closed forms of PageRank on cycles and stars, an FP64 sparse
mat-vec of the all-active PageRank recurrence, brute-force enumeration of the
GAS loop on tiny graphs written separately below, and the committed golden
fixtures of tests/golden/.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

import graphreduce_amd as gr
import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
INF = oracle.SSSP_INF


def _adj(n, s, d, w=None):
    data = np.ones(len(s)) if w is None else np.asarray(w, dtype=np.float64)
    return sp.csr_matrix((data, (s, d)), shape=(n, n))


def _bfs_levels(n, s, d, src):
    lv = cg.shortest_path(_adj(n, s, d), unweighted=True, indices=src)
    return np.where(np.isinf(lv), -1, lv).astype(np.int32)


def _bfs_trips(n, s, depth):
    """Loop trips of bfs.cu's loop: one per level, plus one more when a vertex of
    the deepest level has an out-edge (its neighbours are re-activated and
    every apply returns false)."""
    mx = depth.max()
    outdeg = np.bincount(s, minlength=n)
    return int(mx + 1 + (1 if outdeg[depth == mx].sum() > 0 else 0))


def _cc_min_labels(n, s, d):
    nc, lbl = cg.connected_components(_adj(n, s, d), directed=False)
    mins = np.full(nc, n, np.int64)
    np.minimum.at(mins, lbl, np.arange(n))
    return mins[lbl].astype(np.int32)


@pytest.mark.parametrize("scale,seed", [(10, 1), (12, 1), (14, 3), (16, 1)])
def test_bfs_vs_scipy(scale, seed):
    n = 1 << scale
    s, d = gr.rmat_edges(scale, 16, seed=seed)
    src = int(np.argmax(np.bincount(s, minlength=n)))
    depth, trips, st = oracle.bfs(n, s, d, src)
    assert np.array_equal(depth, _bfs_levels(n, s, d, src))
    assert trips == _bfs_trips(n, s, depth)
    # E_reached of bfs.cu:219-238 equals the out-edges pushed by changed vertices
    outdeg = np.bincount(s, minlength=n)
    assert st["out_edges"] == outdeg[depth >= 0].sum()


@pytest.mark.parametrize("seed", range(8))
def test_bfs_random_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    m = int(rng.integers(0, 5 * n))
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    src = int(rng.integers(0, n))
    depth, trips, _ = oracle.bfs(n, s, d, src)
    assert np.array_equal(depth, _bfs_levels(n, s, d, src))
    assert trips == _bfs_trips(n, s, depth)


@pytest.mark.parametrize("scale,seed", [(10, 2), (14, 2), (16, 2)])
def test_sssp_vs_dijkstra(scale, seed):
    n = 1 << scale
    s, d = gr.rmat_edges(scale, 16, seed=seed)
    w = gr.hash_weights(len(s), seed=seed)
    src = int(np.argmax(np.bincount(s, minlength=n)))
    dist, _, _ = oracle.sssp(n, s, d, w, src)
    ref = cg.dijkstra(_adj(n, s, d, w), indices=src)
    ref = np.where(np.isinf(ref), INF, ref).astype(np.int64)
    assert np.array_equal(dist, ref)


def test_sssp_multigraph_takes_min_parallel_arc():
    s = np.array([0, 0, 1, 1], np.int32)
    d = np.array([1, 1, 2, 2], np.int32)
    w = np.array([9, 4, 7, 1], np.int32)
    dist, trips, _ = oracle.sssp(3, s, d, w, 0)
    assert dist.tolist() == [0, 4, 5]


@pytest.mark.parametrize("scale,seed", [(10, 4), (14, 4), (16, 4)])
def test_cc_vs_scipy(scale, seed):
    n = 1 << scale
    s, d = gr.rmat_edges(scale, 16, seed=seed, symmetrize=True)
    lab, _, _ = oracle.cc(n, s, d)
    assert np.array_equal(lab, _cc_min_labels(n, s, d))


def test_cc_grid_vs_scipy():
    side = 64
    s, d = gr.grid_edges(side, 0.05, seed=5)
    n = side * side
    lab, trips, _ = oracle.cc(n, s, d)
    assert np.array_equal(lab, _cc_min_labels(n, s, d))
    assert trips > side  # label propagation needs about one trip per hop


# ---------------------------------------------------------------- PageRank

def test_pagerank_cycle_closed_form():
    """Directed n-cycle: after t trips every vertex holds 1 - 0.85^(t+1); trip k
    changes ranks by 0.15 * 0.85^(k+1) and the loop ends after the first trip
    whose change falls below 0.01."""
    n = 50
    s = np.arange(n, dtype=np.int32)
    d = ((s + 1) % n).astype(np.int32)
    rank, trips, _ = oracle.pagerank(n, s, d)
    expected_trips = next(t for t in range(1, 100) if 0.15 * 0.85**t < 0.01)
    assert trips == expected_trips
    closed = 1.0 - 0.85 ** (trips + 1)
    assert np.allclose(rank, closed, rtol=2e-6, atol=0)
    assert np.all(rank == rank[0])


def test_pagerank_star_and_sources():
    k = 40
    s = np.arange(1, k + 1, dtype=np.int32)
    d = np.zeros(k, np.int32)
    rank, trips, _ = oracle.pagerank(k + 1, s, d)
    assert trips == 1
    assert rank[1:].tolist() == [np.float32(0.15)] * k  # zero in-degree
    assert abs(rank[0] - (0.15 + 0.85 * 0.15 * k)) <= 1e-5 * rank[0]


def test_pagerank64_close_to_fp32_on_small_graphs():
    n = 1 << 10
    s, d = gr.rmat_edges(10, 16, seed=3)
    r32, t32, _ = oracle.pagerank(n, s, d)
    r64, t64, _ = oracle.pagerank64(n, s, d)
    assert t32 == t64
    assert np.max(np.abs(r32 - r64) / r64) < 1e-5


def test_pagerank_fp64_all_active_vs_scipy():
    n = 1 << 12
    s, d = gr.rmat_edges(12, 16, seed=3)
    off_out, _ = oracle.build_csr(n, s, d)
    off_in, srcs = oracle.build_csr(n, d, s)
    iters = 20
    rank, l1 = oracle.pagerank_fp64_csr(n, off_out, off_in, srcs, iters)
    outdeg = np.diff(off_out).astype(np.float64)
    A = _adj(n, s, d).T.tocsr()  # A[v, u] = 1 for arc u -> v
    r = np.full(n, 0.15)
    l1_ref = []
    for _ in range(iters):
        c = np.divide(r, outdeg, out=np.zeros(n), where=outdeg > 0)
        nr = 0.15 + 0.85 * (A @ c)
        l1_ref.append(np.abs(nr - r).sum())
        r = nr
    assert np.allclose(rank, r, rtol=1e-12, atol=0)
    assert np.allclose(l1, l1_ref, rtol=1e-9)


# ------------------------------------------------ brute force, tiny graphs

def _brute_gas(n, s, d, vd, gather_map, reduce_fn, zero, apply_fn, active, it0=0,
               scatter_fn=None):
    """Independent Python statement of the bulk-synchronous loop of
    refgas.h:220-229 (separate gather and apply sweeps, flag-based
    activation with Program::scatter on every out-edge of a changed vertex,
    ascending next active list)."""
    it = it0
    trips = 0
    while active:
        g = {}
        for v in active:
            acc = zero
            for e in range(len(s)):
                if d[e] == v:
                    acc = reduce_fn(acc, gather_map(vd, int(s[e]), e))
            g[v] = acc
        changed = [v for v in active if apply_fn(vd, v, g[v], it)]
        flags = set()
        for u in changed:
            for e in range(len(s)):
                if s[e] == u:
                    flags.add(int(d[e]))
                    if scatter_fn is not None:
                        scatter_fn(vd, u, int(d[e]), e)
        active = sorted(flags)
        it += 1
        trips += 1
    return trips


@pytest.mark.parametrize("seed", range(25))
def test_brute_force_tiny(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 9))
    m = int(rng.integers(0, 16))
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(1, 65, m).astype(np.int32)
    src = int(rng.integers(0, n))

    dep = [-1] * n

    def bfs_apply(vd, v, g, it):
        if vd[v] == -1:
            vd[v] = it
            return True
        return False

    trips = _brute_gas(n, s, d, dep, lambda vd, u, e: 0, lambda a, b: 0, 0, bfs_apply, [src])
    got, t, _ = oracle.bfs(n, s, d, src)
    assert got.tolist() == dep and t == trips

    dist = [INF] * n
    dist[src] = 0

    def min_apply(vd, v, g, it):
        ch = g < vd[v]
        vd[v] = min(vd[v], g)
        return ch

    trips = _brute_gas(n, s, d, dist, lambda vd, u, e: vd[u] + int(w[e]), min, INF, min_apply,
                       list(range(n)))
    got, t, _ = oracle.sssp(n, s, d, w, src)
    assert got.tolist() == dist and t == trips

    lab = list(range(n))
    trips = _brute_gas(n, s, d, lab, lambda vd, u, e: vd[u], min, 2**31 - 1, min_apply,
                       list(range(n)))
    got, t, _ = oracle.cc(n, s, d)
    assert got.tolist() == lab and t == trips

    outdeg = np.bincount(s, minlength=n)
    rk = [np.float32(0.15)] * n

    def pr_apply(vd, v, g, it):
        nr = np.float32(np.float32(0.15) + np.float32(np.float32(1.0) - np.float32(0.15)) * g)
        ret = abs(np.float32(nr - vd[v])) >= np.float32(0.01)
        vd[v] = nr
        return bool(ret)

    trips = _brute_gas(n, s, d, rk, lambda vd, u, e: np.float32(vd[u] / np.float32(outdeg[u])),
                       lambda a, b: np.float32(a + b), np.float32(0.0), pr_apply, list(range(n)))
    got, t, _ = oracle.pagerank(n, s, d)
    assert t == trips
    assert np.array_equal(got, np.array(rk, np.float32))

    # a program whose scatter changes edge data: SSSP on congesting links
    dist = [INF] * n
    dist[src] = 0
    wc = [int(x) for x in w]

    def congest(vd, u, v, e):
        if vd[v] > vd[u]:
            wc[e] += 1

    trips = _brute_gas(n, s, d, dist, lambda vd, u, e: vd[u] + wc[e], min, INF, min_apply,
                       list(range(n)), scatter_fn=congest)
    got, t, _, wout = oracle.sssp_congestion(n, s, d, w, src)
    assert got.tolist() == dist and t == trips and wout.tolist() == wc


def test_congestion_program_by_hand():
    """0->1 (1), 0->2 (5), 1->3 (1), 2->3 (1), source 0, all active: trip 1
    sets d1 = 1, d2 = 5 and both raise their edge to the still-unreached 3
    to weight 2; trip 2 gives d3 = min(1 + 2, 5 + 2) = 3; 3 has no out-edge."""
    s, d, w = [0, 0, 1, 2], [1, 2, 3, 3], [1, 5, 1, 1]
    dist, trips, _, wout = oracle.sssp_congestion(4, s, d, w, 0)
    assert dist.tolist() == [0, 1, 5, 3] and trips == 2 and wout.tolist() == [1, 5, 2, 2]


# ------------------------------------------------------------ golden files

def test_golden_fixtures():
    """Fixtures written by tests/golden/make_golden.py from scipy."""
    meta = json.loads((GOLDEN / "golden.json").read_text())
    data = np.load(GOLDEN / "golden.npz")
    for case in meta["cases"]:
        name, n = case["name"], case["n"]
        s, d = data[name + "_src"], data[name + "_dst"]
        if case["kind"] == "bfs":
            got, trips, _ = oracle.bfs(n, s, d, case["source"])
            assert np.array_equal(got, data[name + "_out"]), name
            assert trips == case["trips"], name
        elif case["kind"] == "sssp":
            got, _, _ = oracle.sssp(n, s, d, data[name + "_w"], case["source"])
            assert np.array_equal(got, data[name + "_out"]), name
        elif case["kind"] == "cc":
            got, _, _ = oracle.cc(n, s, d)
            assert np.array_equal(got, data[name + "_out"]), name
        elif case["kind"] == "pr64":
            off_out, _ = oracle.build_csr(n, s, d)
            off_in, srcs = oracle.build_csr(n, d, s)
            got, _ = oracle.pagerank_fp64_csr(n, off_out, off_in, srcs, case["iterations"])
            assert np.allclose(got, data[name + "_out"], rtol=1e-12), name


# ---------------------------------------------------------------- builders

def test_build_csr_vs_scipy():
    rng = np.random.default_rng(3)
    n = 300
    s = rng.integers(0, n, 3000).astype(np.int32)
    d = rng.integers(0, n, 3000).astype(np.int32)
    off, adj = oracle.build_csr(n, s, d)
    order = np.argsort(s, kind="stable")
    assert np.array_equal(adj, d[order])
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(s, minlength=n))]))


# ------------------------------------------------- CSR-input entry points

def _min_weight_adj(n, s, d, w):
    """Weighted adjacency keeping the lightest of parallel arcs (scipy would
    add duplicate entries together)."""
    key = s.astype(np.int64) * n + d
    order = np.lexsort((w, key))
    first = np.ones(len(order), bool)
    first[1:] = key[order][1:] != key[order][:-1]
    keep = order[first]
    return _adj(n, s[keep], d[keep], w[keep])


def _random_multigraph(seed):
    """Random graph with duplicate arcs, self-loops and isolated vertices."""
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(2, 300))
    m = int(rng.integers(0, 6 * n))
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    if m:
        k = max(1, m // 10)
        s[:k] = d[:k]                       # self-loops
        s[k:2 * k], d[k:2 * k] = s[2 * k:3 * k], d[2 * k:3 * k]  # duplicates
    w = rng.integers(1, 65, m).astype(np.int32)
    return n, s, d, w


@pytest.mark.parametrize("seed", range(12))
def test_csr_entry_points_equal_edge_list_twins(seed):
    """oracle_*_csr (prebuilt CSR + CSC, edge data in CSC order) give the same
    values, trip counts and work counters as the edge-list drivers, which
    tests above pin against scipy.  The CSC is the oracle builder's, stable in
    edge-list order within a row, so PageRank's FP32 sums run in the same
    order and must agree bit for bit."""
    n, s, d, w = _random_multigraph(seed)
    off_out, dsts = oracle.build_csr(n, s, d)
    off_in, srcs = oracle.build_csr(n, d, s)
    w_csc = w[np.argsort(d, kind="stable")]
    src = int(np.argmax(np.bincount(s, minlength=n))) if len(s) else 0

    a, ta, sa = oracle.bfs(n, s, d, src)
    b, tb, sb = oracle.bfs_csr(n, off_out, dsts, off_in, srcs, src)
    assert np.array_equal(a, b) and ta == tb
    assert {k: sa[k] for k in sa if k != "seconds"} == {k: sb[k] for k in sb if k != "seconds"}

    a, ta, sa = oracle.sssp(n, s, d, w, src)
    b, tb, sb = oracle.sssp_csr(n, off_out, dsts, off_in, srcs, w_csc, src)
    assert np.array_equal(a, b) and ta == tb and sa["in_edges"] == sb["in_edges"]
    ref = cg.dijkstra(_min_weight_adj(n, s, d, w), indices=src)
    assert np.array_equal(b, np.where(np.isinf(ref), INF, ref).astype(np.int64))

    # CC on the symmetric closure
    ss, dd = np.concatenate([s, d]), np.concatenate([d, s])
    so, sd_ = oracle.build_csr(n, ss, dd)
    si, sc = oracle.build_csr(n, dd, ss)
    a, ta, sa = oracle.cc(n, ss, dd)
    b, tb, sb = oracle.cc_csr(n, so, sd_, si, sc)
    assert np.array_equal(a, b) and ta == tb and sa["in_edges"] == sb["in_edges"]
    assert np.array_equal(b, _cc_min_labels(n, ss, dd))

    a, ta, sa = oracle.pagerank(n, s, d)
    b, tb, sb = oracle.pagerank_csr(n, off_out, dsts, off_in, srcs)
    assert np.array_equal(a, b) and ta == tb and sa["in_edges"] == sb["in_edges"]


def test_csr_entry_points_on_rmat():
    n = 1 << 12
    s, d = gr.rmat_edges(12, 16, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    off_out, dsts = oracle.build_csr(n, s, d)
    off_in, srcs = oracle.build_csr(n, d, s)
    src = int(np.argmax(np.bincount(s, minlength=n)))
    b, _, _ = oracle.sssp_csr(n, off_out, dsts, off_in, srcs, w[np.argsort(d, kind="stable")], src)
    ref = cg.dijkstra(_adj(n, s, d, w), indices=src)
    assert np.array_equal(b, np.where(np.isinf(ref), INF, ref).astype(np.int64))
    b, _, _ = oracle.bfs_csr(n, off_out, dsts, off_in, srcs, src)
    assert np.array_equal(b, _bfs_levels(n, s, d, src))


def _pagerank_activation_fp64(n, s, d):
    """pagerank.cu's activation-driven loop (all active first; rank starts at
    0.15f; gather = sum over in-edges in CSC order of rank/outdeg; apply sets
    rank = 0.15 + 0.85 g and activates on |change| >= 0.01) written in plain
    Python floats (IEEE double), one in-edge at a time."""
    c015 = float(np.float32(0.15))
    tol = float(np.float32(0.01))
    outdeg = np.bincount(s, minlength=n).tolist()
    order = np.argsort(d, kind="stable")
    ins = [[] for _ in range(n)]
    for e in order.tolist():
        ins[int(d[e])].append(int(s[e]))
    outs = [[] for _ in range(n)]
    for e in range(len(s)):
        outs[int(s[e])].append(int(d[e]))
    rank = [c015] * n
    active = list(range(n))
    trips = 0
    while active:
        g = {}
        for v in active:
            acc = 0.0
            for u in ins[v]:
                acc = acc + rank[u] / outdeg[u]
            g[v] = acc
        changed = []
        for v in active:
            nr = c015 + (1.0 - c015) * g[v]
            if abs(nr - rank[v]) >= tol:
                changed.append(v)
            rank[v] = nr
        nxt = set()
        for u in changed:
            nxt.update(outs[u])
        active = sorted(nxt)
        trips += 1
    return np.array(rank), trips


@pytest.mark.parametrize("scale,seed", [(8, 3), (10, 3)])
def test_pagerank64_equals_python_double_loop(scale, seed):
    """oracle_pagerank64 computes in double: it equals a plain-Python double
    statement of the same loop bit for bit (a float accumulation or a float
    division would differ by ~1e-8 relative and fail)."""
    n = 1 << scale
    s, d = gr.rmat_edges(scale, 16, seed=seed)
    r64, t64, _ = oracle.pagerank64(n, s, d)
    ref, trips = _pagerank_activation_fp64(n, s, d)
    assert t64 == trips
    assert np.array_equal(r64, ref)
    r32, _, _ = oracle.pagerank(n, s, d)
    assert not np.array_equal(r32.astype(np.float64), r64)


# ------------------------------------------------ widest paths (open Program)

def _widest_dijkstra(n, s, d, cap, src):
    """Maximum-bottleneck path widths by a max-heap Dijkstra (independent of
    the GAS loop): INT_MAX at the source, 0 where unreachable."""
    import heapq

    adj = [[] for _ in range(n)]
    for u, v, c in zip(s.tolist(), d.tolist(), cap.tolist()):
        adj[u].append((v, c))
    best = [0] * n
    best[src] = 2**31 - 1
    heap = [(-best[src], src)]
    done = [False] * n
    while heap:
        negw, u = heapq.heappop(heap)
        if done[u]:
            continue
        done[u] = True
        for v, c in adj[u]:
            w = min(-negw, c)
            if w > best[v]:
                best[v] = w
                heapq.heappush(heap, (-w, v))
    return np.array(best, np.int64)


@pytest.mark.parametrize("seed", range(10))
def test_widest_vs_bottleneck_dijkstra(seed):
    n, s, d, w = _random_multigraph(seed)
    src = int(np.argmax(np.bincount(s, minlength=n))) if len(s) else 0
    got, trips, _ = oracle.widest(n, s, d, w, src)
    assert np.array_equal(got, _widest_dijkstra(n, s, d, w, src))
    # brute force of the loop: same trips
    width = [0] * n
    width[src] = 2**31 - 1

    def wapply(vd, v, g, it):
        if g > vd[v]:
            vd[v] = g
            return True
        return False

    btrips = _brute_gas(n, s, d, width, lambda vd, u, e: min(vd[u], int(w[e])), max, 0, wapply,
                        list(range(n))) if n * len(s) <= 200_000 else None
    assert btrips is None or (trips == btrips and got.tolist() == width)


def test_widest_on_rmat():
    n = 1 << 12
    s, d = gr.rmat_edges(12, 16, seed=5)
    cap = gr.hash_weights(len(s), seed=5)
    src = int(np.argmax(np.bincount(s, minlength=n)))
    got, _, _ = oracle.widest(n, s, d, cap, src)
    assert np.array_equal(got, _widest_dijkstra(n, s, d, cap, src))
