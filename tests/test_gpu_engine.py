"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Integer programs (BFS depth, SSSP distance, CC label) must match bit for bit,
together with the loop-trip count and the per-run work counters; PageRank
(activation-driven, FP32) must match within the tolerance stated in each test.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIRECTIONS = [(0, "auto"), (1, "push"), (2, "pull")]


def _rmat(gr, scale, seed, sym=False):
    s, d = gr.rmat_edges(scale, 16, seed=seed, symmetrize=sym)
    return 1 << scale, s, d


def _max_out_src(n, s):
    deg = np.bincount(s, minlength=n)
    return int(np.argmax(deg))  # first on ties, bfs.cu:187


@pytest.mark.parametrize("scale", [10, 16])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_bfs_rmat(gpu, scale, direction):
    gr = gpu
    n, s, d = _rmat(gr, scale, seed=1)
    src = _max_out_src(n, s)
    ref, ref_trips, _ = oracle.bfs(n, s, d, src)
    depth, trips = gr.run_bfs((n, s, d), src, direction=direction)
    assert np.array_equal(depth, ref), f"{(depth != ref).sum()} depth mismatches"
    assert trips == ref_trips


@pytest.mark.parametrize("scale", [10, 16])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_sssp_rmat(gpu, scale, direction):
    gr = gpu
    n, s, d = _rmat(gr, scale, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    src = _max_out_src(n, s)
    ref, ref_trips, _ = oracle.sssp(n, s, d, w, src)
    dist, trips = gr.run_sssp((n, s, d), w, src, direction=direction)
    assert np.array_equal(dist, ref), f"{(dist != ref).sum()} distance mismatches"
    assert trips == ref_trips


@pytest.mark.parametrize("scale", [10, 16])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_cc_rmat_symmetric(gpu, scale, direction):
    gr = gpu
    n, s, d = _rmat(gr, scale, seed=4, sym=True)
    ref, ref_trips, _ = oracle.cc(n, s, d)
    lab, trips = gr.run_cc((n, s, d), direction=direction)
    assert np.array_equal(lab, ref), f"{(lab != ref).sum()} label mismatches"
    assert trips == ref_trips


def test_cc_grid(gpu):
    gr = gpu
    side = 128
    s, d = gr.grid_edges(side, 0.01, seed=5)
    n = side * side
    ref, ref_trips, _ = oracle.cc(n, s, d)
    lab, trips = gr.run_cc((n, s, d))
    assert np.array_equal(lab, ref)
    assert trips == ref_trips


@pytest.mark.parametrize("scale", [10, 14])
def test_pagerank_rmat(gpu, scale):
    """Activation-driven PageRank of pagerank.cu.  Sum order differs from the
    sequential CPU sum, so ranks agree to FP32 rounding.  The reference's own
    test rule (pagerank.cu:154-166) is 1e-6 absolute per vertex; it is applied
    to every vertex whose rank is below 1, where FP32 spacing is at most
    1.2e-7 and a few roundings of a reordered sum stay inside it.  Above that
    the spacing itself grows past 1e-6 (rank 8 and up: 9.5e-7 per ulp, hubs
    reach ranks in the hundreds), so one different rounding already breaks an
    absolute bound, and those vertices are held to 1e-5 relative.  The
    activation schedule must match: same trips."""
    gr = gpu
    n, s, d = _rmat(gr, scale, seed=3)
    outdeg = np.bincount(s, minlength=n).astype(np.int32)
    ref, ref_trips, _ = oracle.pagerank(n, s, d)
    rank, trips = gr.run_pagerank((n, s, d), outdeg)
    diff = np.abs(rank.astype(np.float64) - ref.astype(np.float64))
    small = ref < 1.0
    assert small.sum() > n // 2
    assert diff[small].max() <= 1e-6, f"max abs err {diff[small].max():.3g} below rank 1"
    rel = diff / np.abs(ref.astype(np.float64))
    assert rel.max() <= 1e-5, f"max rel err {rel.max():.3g}"
    assert trips == ref_trips


def test_work_counters_match_oracle(gpu):
    """sum #A and sum E_in(A) (the roofline numerator) equal the oracle's."""
    gr = gpu
    n, s, d = _rmat(gr, 12, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    src = _max_out_src(n, s)
    _, _, st = oracle.sssp(n, s, d, w, src)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP)
    eng.setGraph(n, vd, len(s), w, s, d)
    eng.setActive(0, n)
    trips = eng.run()
    got = eng.stats()
    assert trips == st["iterations"]
    assert got["active_vertices"] == st["active"]
    assert got["gathered_edges"] == st["in_edges"]
    # sum #C and sum E_out(C): the changed vertices and their out-edges
    assert got["changed_vertices"] == st["changed"]
    assert got["pushed_edges"] == st["out_edges"]
    eng.close()


# ------------------------------------------------------------ edge cases

def _random_graph(rng, n, m, dup=True, loops=True):
    s = rng.integers(0, n, m, dtype=np.int64).astype(np.int32)
    d = rng.integers(0, n, m, dtype=np.int64).astype(np.int32)
    if not loops:
        keep = s != d
        s, d = s[keep], d[keep]
    if dup and len(s) > 3:
        s = np.concatenate([s, s[:3]])
        d = np.concatenate([d, d[:3]])
    return s, d


@pytest.mark.parametrize("seed", range(6))
def test_random_multigraphs_all_programs(gpu, seed):
    """Ragged random multigraphs with duplicate arcs and self-loops."""
    gr = gpu
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3000))
    m = int(rng.integers(0, 6 * n + 1))
    s, d = _random_graph(rng, n, m)
    w = rng.integers(1, 65, len(s)).astype(np.int32)
    src = int(rng.integers(0, n))
    ref, rt, _ = oracle.bfs(n, s, d, src)
    got, t = gr.run_bfs((n, s, d), src)
    assert np.array_equal(got, ref) and t == rt
    ref, rt, _ = oracle.sssp(n, s, d, w, src)
    got, t = gr.run_sssp((n, s, d), w, src)
    assert np.array_equal(got, ref) and t == rt
    ss, dd = np.concatenate([s, d]), np.concatenate([d, s])
    ref, rt, _ = oracle.cc(n, ss, dd)
    got, t = gr.run_cc((n, ss, dd))
    assert np.array_equal(got, ref) and t == rt


def test_empty_and_single_vertex(gpu):
    gr = gpu
    e = np.zeros(0, np.int32)
    got, t = gr.run_bfs((1, e, e), 0)
    assert got.tolist() == [0] and t == 1
    got, t = gr.run_sssp((5, e, e), e, 2)
    ref, rt, _ = oracle.sssp(5, e, e, e, 2)
    assert np.array_equal(got, ref) and t == rt
    got, t = gr.run_cc((7, e, e))
    assert got.tolist() == list(range(7)) and t == 1


def test_hub_spanning_many_tiles(gpu):
    """One vertex with 50k in-edges (rows spanning many merge-path tiles) and
    one with exactly 2048 in-edges."""
    gr = gpu
    n = 60000
    hub_in = np.arange(1, 50001, dtype=np.int32)
    s = np.concatenate([hub_in, np.arange(10000, 12048, dtype=np.int32), np.arange(0, n - 1, dtype=np.int32)])
    d = np.concatenate([np.zeros(50000, np.int32), np.full(2048, 5, np.int32), np.arange(1, n, dtype=np.int32)])
    w = (np.arange(len(s)) % 64 + 1).astype(np.int32)
    for src in (0, 7):
        ref, rt, _ = oracle.sssp(n, s, d, w, src)
        got, t = gr.run_sssp((n, s, d), w, src)
        assert np.array_equal(got, ref) and t == rt
    ss, dd = np.concatenate([s, d]), np.concatenate([d, s])
    ref, rt, _ = oracle.cc(n, ss, dd)
    got, t = gr.run_cc((n, ss, dd))
    assert np.array_equal(got, ref) and t == rt
    # PageRank: the CPU's sequential FP32 sum over the hub's 50k in-edges is
    # off by ~(k-1)*2^-24 relative (0.3 %); the GPU's tree sum is checked
    # against the same activation-driven loop run in FP64 instead.
    outdeg = np.bincount(s, minlength=n).astype(np.int32)
    ref64, rt, _ = oracle.pagerank64(n, s, d)
    got, t = gr.run_pagerank((n, s, d), outdeg)
    assert t == rt
    rel = np.abs(got - ref64) / ref64
    assert rel.max() <= 2e-6, f"max rel err vs FP64 {rel.max():.3g}"


def test_invalid_arguments_raise(gpu):
    gr = gpu
    eng = gr.GASEngine(gr.BFS)
    with pytest.raises(gr.GRError):
        eng.setActive(0, 1)  # no graph yet
    s = np.array([0, 1], np.int32)
    d = np.array([1, 5], np.int32)  # 5 out of range
    vd = np.zeros(3, gr.PROGRAMS[gr.BFS]["vertex"])
    with pytest.raises(gr.GRError):
        eng.setGraph(3, vd, 2, None, s, d)
    d = np.array([1, 2], np.int32)
    eng.setGraph(3, vd, 2, None, s, d)
    with pytest.raises(gr.GRError):
        eng.setActive(2, 4)
    eng.close()
    with pytest.raises(gr.GRError):
        gr.GASEngine(gr.SSSP).setGraph(3, np.zeros(3, np.int32), 2, None, s, d)  # needs weights


# ------------------------------------------------------ device graph build

@pytest.mark.parametrize("sym", [False, True])
def test_device_rmat_equals_host(gpu, sym):
    gr = gpu
    s, d = gr.rmat_edges(14, 16, seed=6, symmetrize=sym)
    g = gr.Graph.rmat(14, 16, seed=6, symmetrize=sym)
    assert g.n_edges == len(s)
    gs, gd = g.edges()
    assert np.array_equal(gs, s) and np.array_equal(gd, d)
    off, adj = g.csr()
    roff, radj = oracle.build_csr(1 << 14, s, d)
    assert np.array_equal(off, roff) and np.array_equal(adj, radj)
    off, adj = g.csc()
    roff, radj = oracle.build_csr(1 << 14, d, s)
    assert np.array_equal(off, roff) and np.array_equal(adj, radj)
    w_host = gr.hash_weights(len(s), seed=6)
    wd = g.hash_weights_device(seed=6)
    back = np.empty(len(s), np.int32)
    gr.lib().gr_memcpy_d2h(back.ctypes.data, wd.ptr, back.nbytes)
    assert np.array_equal(back, w_host)
    g.close()


def test_graph_from_unsorted_edges_with_duplicates(gpu):
    gr = gpu
    rng = np.random.default_rng(7)
    n = 500
    s = rng.integers(0, n, 4000).astype(np.int32)
    d = rng.integers(0, n, 4000).astype(np.int32)
    g = gr.Graph.from_edges(n, s, d)

    def rows_sorted(off, adj):  # the order inside a row is free (util.cuh:68 is unstable)
        row = np.repeat(np.arange(n), np.diff(off))
        return adj[np.lexsort((adj, row))]

    off, adj = g.csr()
    roff, radj = oracle.build_csr(n, s, d)
    assert np.array_equal(off, roff)
    assert np.array_equal(adj, rows_sorted(off, adj)), "device rows are sorted by column"
    assert np.array_equal(adj, rows_sorted(roff, radj))
    off, adj = g.csc()
    roff, radj = oracle.build_csr(n, d, s)
    assert np.array_equal(off, roff)
    assert np.array_equal(adj, rows_sorted(roff, radj))
    es, ed = g.edges()
    assert np.array_equal(es, s) and np.array_equal(ed, d)
    g.close()


def test_bound_graph_shared_by_engines(gpu):
    """Several engines on one device graph give the plain results."""
    gr = gpu
    g = gr.Graph.rmat(12, 16, seed=2)
    n = g.n_vertices
    s, d = g.edges()
    w = g.hash_weights_device(seed=2)
    wh = gr.hash_weights(len(s), seed=2)
    src, _ = g.max_out_degree_vertex()
    a, ta = gr.run_sssp(g, w, src)
    b, tb = gr.run_sssp((n, s, d), wh, src)
    ref, rt, _ = oracle.sssp(n, s, d, wh, src)
    assert np.array_equal(a, ref) and np.array_equal(b, ref) and ta == tb == rt
    dep, t = gr.run_bfs(g, src)
    ref, rt, _ = oracle.bfs(n, s, d, src)
    assert np.array_equal(dep, ref) and t == rt
    g.close()


# ------------------------------------------------------------ delta gather

def _engine_run(gr, prog, n, vd, s, d, w=None, start=None, delta=1):
    eng = gr.GASEngine(prog)
    eng.setDeltaGather(delta)
    # getResults writes back into the array given to setGraph (the engine owns
    # it until then, refgas.h:69-72), so each run gets its own copy
    eng.setGraph(n, vd.copy(), len(s), w, s, d)
    if start is None:
        eng.setActive(0, n)
    else:
        eng.setActive(start, start + 1)
    trips = eng.run()
    eng.getResults()
    out = eng.vertexData.copy()
    st = eng.stats()
    eng.close()
    return out, trips, st


@pytest.mark.parametrize("case", ["sssp_rmat14", "cc_sym_rmat14", "cc_grid96", "sssp_grid96"])
def test_delta_gather_equals_full_gather(gpu, case):
    """Delta gather reads fewer source values but gives the same values, trips
    and work counters as the full gather of refgas.h:141-160."""
    gr = gpu
    if case.endswith("rmat14"):
        n, s, d = _rmat(gr, 14, seed=3, sym=case.startswith("cc"))
    else:
        n, (s, d) = 96 * 96, gr.grid_edges(96, 0.02, seed=5)
    if case.startswith("sssp"):
        w = gr.hash_weights(len(s), seed=3)
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[_max_out_src(n, s)] = 0
        prog = gr.SSSP
        ref, ref_trips, _ = oracle.sssp(n, s, d, w, _max_out_src(n, s))
    else:
        w, vd, prog = None, np.arange(n, dtype=np.int32), gr.CC
        ref, ref_trips, _ = oracle.cc(n, s, d)
    on, t_on, st_on = _engine_run(gr, prog, n, vd, s, d, w, delta=2)  # pull-delta if valid
    pu, t_pu, st_pu = _engine_run(gr, prog, n, vd, s, d, w, delta=3)  # push if valid
    ad, t_ad, st_ad = _engine_run(gr, prog, n, vd, s, d, w, delta=1)  # adaptive
    off, t_off, st_off = _engine_run(gr, prog, n, vd, s, d, w, delta=0)
    for got in (on, pu, ad, off):
        assert np.array_equal(got, ref)
    assert t_on == t_pu == t_ad == t_off == ref_trips
    for k in ("iterations", "active_vertices", "gathered_edges", "changed_vertices",
              "pushed_edges"):
        assert st_on[k] == st_pu[k] == st_ad[k] == st_off[k], k
    assert st_off["delta_gathers"] == st_off["inert_gathers"] == 0
    assert st_on["delta_gathers"] == t_on - 1 and st_on["push_gathers"] == 0
    assert st_pu["delta_gathers"] == t_on - 1
    assert st_ad["delta_gathers"] <= t_on - 1
    # SSSP's first gather skips the unreached (inert) sources, by push when
    # forced; CC has no inert values
    inert = 1 if prog == gr.SSSP else 0
    assert st_on["inert_gathers"] == st_pu["inert_gathers"] == st_ad["inert_gathers"] == inert
    assert st_pu["push_gathers"] == t_on - 1 + inert


def test_delta_gather_not_used_after_partial_start(gpu):
    """With setActive on a sub-range, a vertex may be activated whose other
    in-neighbours it never gathered from: vertex 1 must see source 2 (value 0,
    never changed), so its distance is 1, not 6 as a delta gather would give."""
    gr = gpu
    s = np.array([3, 0, 2], np.int32)
    d = np.array([0, 1, 1], np.int32)
    w = np.array([1, 5, 1], np.int32)
    vd = np.array([gr.SSSP_INF, gr.SSSP_INF, 0, 0], np.int32)
    for delta in (3, 2, 1, 0):
        out, trips, st = _engine_run(gr, gr.SSSP, 4, vd, s, d, w, start=0, delta=delta)
        assert out.tolist() == [1, 1, 0, 0]
        assert trips == 2
        assert st["delta_gathers"] == 0
        # every gather of the broken chain may still skip the inert sources
        assert st["inert_gathers"] == (trips if delta else 0)
        assert st["push_gathers"] == 0  # never pushes with a partial active set
    # the same graph from an all-active start: delta is valid and used
    for delta in (2, 3):
        out, trips, st = _engine_run(gr, gr.SSSP, 4, vd, s, d, w, delta=delta)
        assert out.tolist() == [1, 1, 0, 0]
        assert st["delta_gathers"] == trips - 1


@pytest.mark.parametrize("bad", [None, 100000, 100001, -1])
def test_inert_sources_need_bounded_weights(gpu, bad):
    """SSSP's first gather skips sources at gatherZero only while every weight
    is in [0, 100000] (gatherZero + w then neither overflows nor drops below
    gatherZero); with another weight bound, every gather is the plain one and
    the values are those of the full-gather run."""
    gr = gpu
    n, s, d = _rmat(gr, 12, seed=8)
    w = gr.hash_weights(len(s), seed=8)
    if bad is not None:
        w[len(w) // 3] = bad
    src = _max_out_src(n, s)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    full, t_full, _ = _engine_run(gr, gr.SSSP, n, vd, s, d, w, delta=0)
    for delta in (1, 2, 3):
        got, t, st = _engine_run(gr, gr.SSSP, n, vd, s, d, w, delta=delta)
        assert np.array_equal(got, full) and t == t_full
        assert st["inert_gathers"] == (1 if bad in (None, 100000) else 0)
    if bad in (None, 100000):
        ref, ref_trips, _ = oracle.sssp(n, s, d, w, src)
        assert np.array_equal(full, ref) and t_full == ref_trips


# ------------------------------------------------------- vertex order

def _ordered_run(gr, prog, n, vd, s, d, w=None, start=None, order=1):
    eng = gr.GASEngine(prog)
    eng.setVertexOrder(order)
    eng.setGraph(n, vd.copy(), len(s), w, s, d)
    if start is None:
        eng.setActive(0, n)
    else:
        eng.setActive(*start)
    trips = 0
    while eng.countActive():
        eng.setIterationCount(trips)
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        trips += 1
    first = eng.readVertexData()
    eng.getResults()
    out = eng.vertexData.copy()
    st = eng.stats()
    eng.close()
    assert np.array_equal(first.view(np.uint8), out.view(np.uint8))
    return out, trips, st


@pytest.mark.parametrize("case", ["bfs", "sssp", "sssp_range", "cc", "cc_grid"])
def test_vertex_order_gives_the_same_results(gpu, case):
    """The degree-ordered layout is internal: values in the caller's ids,
    trips and work counters equal those of the layout as given, and of the
    automatic choice (order 2: the RMAT graphs are skewed, the grid is not)."""
    gr = gpu
    if case == "cc_grid":
        s, d = gr.grid_edges(96, 0.01, seed=5)
        n = 96 * 96
    else:
        n, s, d = _rmat(gr, 14, seed=8, sym=(case == "cc"))
    w = gr.hash_weights(len(s), seed=8)
    src = _max_out_src(n, s)
    start = None
    if case == "bfs":
        prog, vd = gr.BFS, np.zeros(n, gr.PROGRAMS[gr.BFS]["vertex"])
        vd["depth"] = -1
        start, w = (src, src + 1), None
        ref, ref_trips, _ = oracle.bfs(n, s, d, src)
    elif case.startswith("sssp"):
        prog, vd = gr.SSSP, np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
        ref, ref_trips, _ = oracle.sssp(n, s, d, w, src)
        if case == "sssp_range":  # sources 100..163 at distance 0, start on them only
            vd[100:164] = 0
            start, ref, ref_trips = (100, 164), None, None
    else:
        prog, vd, w = gr.CC, np.arange(n, dtype=np.int32), None
        ref, ref_trips, _ = oracle.cc(n, s, d)
    a, ta, sa = _ordered_run(gr, prog, n, vd, s, d, w, start, order=0)
    b, tb, sb = _ordered_run(gr, prog, n, vd, s, d, w, start, order=1)
    c, tc, sc = _ordered_run(gr, prog, n, vd, s, d, w, start, order=2)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.array_equal(a.view(np.int32), c.view(np.int32))
    assert ta == tb == tc
    for k in ("iterations", "active_vertices", "gathered_edges", "changed_vertices",
              "pushed_edges"):
        assert sa[k] == sb[k] == sc[k], k
    if ref is not None:
        got = b["depth"] if case == "bfs" else b
        assert np.array_equal(got, ref)
        assert tb == ref_trips


def test_vertex_order_pagerank_within_tolerance(gpu):
    gr = gpu
    n, s, d = _rmat(gr, 12, seed=3)
    outdeg = np.bincount(s, minlength=n).astype(np.int32)
    vd = np.zeros(n, gr.PROGRAMS[gr.PAGERANK]["vertex"])
    vd["rank"] = 1.0
    vd["numOutEdges"] = outdeg
    a, ta, _ = _ordered_run(gr, gr.PAGERANK, n, vd, s, d, order=0)
    b, tb, _ = _ordered_run(gr, gr.PAGERANK, n, vd, s, d, order=1)
    assert np.array_equal(a["numOutEdges"], b["numOutEdges"])
    assert np.allclose(a["rank"], b["rank"], rtol=1e-5, atol=1e-6)


# --------------------------------------------- Program::scatter on edge data

def _congestion_run(gr, n, s, d, w, src, order=1, have_scatter=True, bind=False):
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP_CONGESTION)
    eng.setVertexOrder(order)
    w_io = np.array(w, np.int32)
    if bind:
        g = gr.Graph.from_edges(n, s, d)
        eng.bindGraph(g, vd, w_io)
    else:
        eng.setGraph(n, vd, len(s), w_io, s, d)
    eng.setActive(0, n)
    trips = 0
    while eng.countActive():
        eng.gather()
        eng.apply()
        eng.scatterActivate(have_scatter)
        eng.nextIter()
        trips += 1
    eng.getResults()
    out = eng.vertexData.copy()
    w_dev = eng.readEdgeData()
    eng.close()
    return out, trips, w_io, w_dev


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("case", ["rmat12", "grid40", "multigraph"])
def test_scatter_program_matches_oracle(gpu, case, order):
    """A program whose scatter writes edge data (refgas.h:192-196): distances,
    trips and the final edge data equal the oracle's; getResults writes the
    edge data back into the caller's array, readEdgeData gives the same."""
    gr = gpu
    rng = np.random.default_rng(4)
    if case == "rmat12":
        n, s, d = _rmat(gr, 12, seed=5)
    elif case == "grid40":
        n, (s, d) = 1600, gr.grid_edges(40, 0.05, seed=2)
    else:
        n, m = 300, 2000
        s = rng.integers(0, n, m).astype(np.int32)
        d = rng.integers(0, n, m).astype(np.int32)
        s = np.concatenate([s, s[:40], np.arange(5, dtype=np.int32)])  # duplicates, self-loops
        d = np.concatenate([d, d[:40], np.arange(5, dtype=np.int32)])
    w = gr.hash_weights(len(s), seed=6)
    src = _max_out_src(n, s)
    ref, rtrips, _, wref = oracle.sssp_congestion(n, s, d, w, src)
    assert not np.array_equal(wref, w)  # the scatter did change edges
    got, trips, w_io, w_dev = _congestion_run(gr, n, s, d, w, src, order=order)
    assert np.array_equal(got, ref) and trips == rtrips
    assert np.array_equal(w_io, wref) and np.array_equal(w_dev, wref)
    got, trips, w_io, w_dev = _congestion_run(gr, n, s, d, w, src, order=order, bind=True)
    assert np.array_equal(got, ref) and trips == rtrips and np.array_equal(w_dev, wref)


def test_scatter_flag_off_is_plain_sssp(gpu):
    """scatterActivate(false) skips Program::scatter, as the reference's CPU
    engine does (refgas.h:192): the edges stay as given and the distances are
    plain SSSP's."""
    gr = gpu
    n, s, d = _rmat(gr, 11, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    src = _max_out_src(n, s)
    got, trips, w_io, w_dev = _congestion_run(gr, n, s, d, w, src, have_scatter=False)
    ref, rtrips, _ = oracle.sssp(n, s, d, w, src)
    assert np.array_equal(got, ref) and trips == rtrips
    assert np.array_equal(w_io, w) and np.array_equal(w_dev, w)


def test_scatter_program_through_cpp_header(gpu, tmp_path):
    """GASEngineGPU.h: getResults() copies the scattered edge data back into
    the caller's EdgeData array (hand case: see oracle test)."""
    import subprocess
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    exe = tmp_path / "scatter_header_test"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{root / 'include'}",
                    str(root / "tests" / "cpp" / "scatter_header_test.cpp"), "-o", str(exe),
                    f"-L{root / 'graphreduce_amd' / 'lib'}", "-lgraphreduce_amd",
                    f"-Wl,-rpath,{root / 'graphreduce_amd' / 'lib'}"], check=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout
    assert "trips 2" in out and "dist 0 1 5 3" in out and "w 1 5 2 2" in out


# ------------------------------------------------------- out-of-core shards

def _stream_run(gr, prog, n, vd, s, d, w=None, start=None, max_edges=0, order=1):
    eng = gr.GASEngine(prog)
    eng.setVertexOrder(order)
    eng.setStreaming(max_edges)
    eng.setGraph(n, vd.copy(), len(s), w, s, d)
    if start is None:
        eng.setActive(0, n)
    else:
        eng.setActive(start, start + 1)
    trips = 0
    while eng.countActive():
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        trips += 1
        eng.setIterationCount(trips)
    eng.getResults()
    out = eng.vertexData.copy()
    st = eng.stats()
    eng.close()
    return out, trips, st


@pytest.mark.parametrize("prog", ["bfs", "sssp", "cc", "pagerank"])
@pytest.mark.parametrize("max_edges", [3000, 40000])
def test_streaming_equals_in_memory(gpu, prog, max_edges):
    """Out-of-core shards (gpugas.h:1398-1619): the edge arrays stay in host
    memory, cut into vertex ranges of <= max_edges edges (hubs exceed 3000:
    shards of their own) and are streamed every iteration; values, trips and
    work counters equal the in-memory engine's (PageRank within FP32 noise:
    other tiles)."""
    gr = gpu
    n, s, d = _rmat(gr, 13, seed=7, sym=(prog == "cc"))
    src = _max_out_src(n, s)
    w, start = None, None
    if prog == "bfs":
        p = gr.BFS
        vd = np.zeros(n, gr.PROGRAMS[gr.BFS]["vertex"])
        vd["depth"] = -1
        start = src
    elif prog == "sssp":
        p = gr.SSSP
        w = gr.hash_weights(len(s), seed=7)
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
    elif prog == "cc":
        p, vd = gr.CC, np.arange(n, dtype=np.int32)
    else:
        p = gr.PAGERANK
        vd = np.zeros(n, gr.PROGRAMS[gr.PAGERANK]["vertex"])
        vd["rank"] = 0.15
        vd["numOutEdges"] = np.bincount(s, minlength=n)
    ref, rt, rst = _stream_run(gr, p, n, vd, s, d, w, start, max_edges=0)
    got, t, st = _stream_run(gr, p, n, vd, s, d, w, start, max_edges=max_edges)
    assert t == rt
    for k in ("active_vertices", "changed_vertices"):
        assert st[k] == rst[k], k
    if prog == "pagerank":
        assert np.allclose(got["rank"], ref["rank"], rtol=1e-5, atol=1e-6)
    else:
        assert np.array_equal(got, ref)
    assert st["streamed_shards"] > t and st["streamed_bytes"] > 0
    assert rst["streamed_shards"] == 0
    if prog == "sssp":
        oref, otrips, _ = oracle.sssp(n, s, d, w, src)
        assert np.array_equal(got, oref) and t == otrips


def test_streaming_refuses_scatter_programs(gpu):
    gr = gpu
    n, s, d = _rmat(gr, 8, seed=1)
    eng = gr.GASEngine(gr.SSSP_CONGESTION)
    eng.setStreaming(1000)
    with pytest.raises(gr.GRError, match="scatter"):
        eng.setGraph(n, np.zeros(n, np.int32), len(s), np.ones(len(s), np.int32), s, d)
    eng.close()


def test_streaming_edge_cases(gpu):
    """Streaming mode on an edgeless graph, isolated vertices and a graph whose
    every vertex is its own shard; edge data reads back in edge-list order."""
    gr = gpu
    e = np.zeros(0, np.int32)
    got, t, st = _stream_run(gr, gr.CC, 5, np.arange(5, dtype=np.int32), e, e, max_edges=10)
    assert got.tolist() == list(range(5)) and t == 1 and st["streamed_shards"] == 0
    s = np.array([0, 1, 2, 2, 5], np.int32)
    d = np.array([1, 2, 0, 3, 4], np.int32)
    w = np.array([3, 1, 4, 1, 5], np.int32)
    vd = np.full(8, gr.SSSP_INF, np.int32)
    vd[0] = 0
    for max_edges in (1, 2, 100):
        got, t, _ = _stream_run(gr, gr.SSSP, 8, vd, s, d, w, max_edges=max_edges)
        ref, rt, _ = oracle.sssp(8, s, d, w, 0)
        assert np.array_equal(got, ref) and t == rt
    eng = gr.GASEngine(gr.SSSP)
    eng.setStreaming(2)
    eng.setGraph(8, vd, len(s), w, s, d)
    assert eng.readEdgeData().tolist() == w.tolist()
    eng.close()


@pytest.mark.parametrize("delta", [1, 3])
def test_fused_scatter_gather(gpu, delta):
    """When the changed out-edges are few, scatterActivate also relaxes the
    next gather's candidates (one pass); values, trips and counters stay the
    oracle's.  Restarting right after such a scatter (resetVertexData +
    setActive) must not leak the pending candidates into the new run."""
    gr = gpu
    n, s, d = _rmat(gr, 13, seed=9)
    w = gr.hash_weights(len(s), seed=9)
    src = _max_out_src(n, s)
    ref, rtrips, rst = oracle.sssp(n, s, d, w, src)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP)
    eng.setDeltaGather(delta)
    eng.setGraph(n, vd, len(s), w, s, d)

    def step():
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()

    eng.setActive(0, n)
    for _ in range(rtrips - 1):  # stop with a fused gather still pending
        step()
    eng.resetVertexData()
    eng.setActive(0, n)
    eng.resetStats()
    trips = 0
    while eng.countActive():
        step()
        trips += 1
    eng.getResults()
    st = eng.stats()
    assert np.array_equal(eng.vertexData, ref) and trips == rtrips
    assert st["fused_gathers"] > 0 and st["fused_edges"] > 0
    assert st["active_vertices"] == rst["active"] and st["gathered_edges"] == rst["in_edges"]
    eng.close()


@pytest.mark.parametrize("prog_name", ["SSSP", "CC"])
def test_fuse_share_keeps_results(gpu, prog_name):
    """setFuseShare moves work between the scatter and the next gather only:
    off (0), the default and always (1) give the oracle's values, trips and
    work counters; 0 fuses nothing, 1 fuses every eligible scatter."""
    gr = gpu
    n, s, d = _rmat(gr, 14, seed=12, sym=prog_name == "CC")
    if prog_name == "SSSP":
        prog, w = gr.SSSP, gr.hash_weights(len(s), seed=12)
        src = _max_out_src(n, s)
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
        ref, rtrips, rst = oracle.sssp(n, s, d, w, src)
    else:
        prog, w, vd = gr.CC, None, np.arange(n, dtype=np.int32)
        ref, rtrips, rst = oracle.cc(n, s, d)
    fused = {}
    for share in (0.0, 0.4, 1.0):
        eng = gr.GASEngine(prog)
        eng.setFuseShare(share)
        eng.setGraph(n, vd.copy(), len(s), w, s, d)  # getResults writes back into it
        eng.setActive(0, n)
        trips = eng.run()
        eng.getResults()
        st = eng.stats()
        assert np.array_equal(eng.vertexData, ref), share
        assert trips == rtrips, share
        assert st["active_vertices"] == rst["active"] and st["gathered_edges"] == rst["in_edges"]
        fused[share] = st["fused_gathers"]
        eng.close()
    assert fused[0.0] == 0 and fused[1.0] >= fused[0.4] and fused[1.0] > 0


@pytest.mark.parametrize("prog_name", ["SSSP", "CC"])
def test_fused_pass_marks_by_candidate_scan(gpu, prog_name):
    """A fused pass over more than 2^20 out-edges leaves the frontier marks to
    a scan of the candidates (a relaxed target ends below gatherZero): the
    next active lists, values, trips and counters stay the oracle's."""
    gr = gpu
    n, s, d = _rmat(gr, 18, seed=21, sym=prog_name == "CC")
    if prog_name == "SSSP":
        prog, w = gr.SSSP, gr.hash_weights(len(s), seed=21)
        src = _max_out_src(n, s)
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
        ref, rtrips, rst = oracle.sssp(n, s, d, w, src)
    else:
        prog, w, vd = gr.CC, None, np.arange(n, dtype=np.int32)
        ref, rtrips, rst = oracle.cc(n, s, d)
    eng = gr.GASEngine(prog)
    eng.setFuseShare(1.0)
    eng.setGraph(n, vd, len(s), w, s, d)
    eng.setActive(0, n)
    trips, biggest, prev = 0, 0, 0
    while eng.countActive():
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        trips += 1
        fe = eng.stats()["fused_edges"]
        biggest, prev = max(biggest, fe - prev), fe
    eng.getResults()
    st = eng.stats()
    assert biggest > 1 << 20, biggest
    assert np.array_equal(eng.vertexData, ref) and trips == rtrips
    assert st["active_vertices"] == rst["active"] and st["gathered_edges"] == rst["in_edges"]
    eng.close()


@pytest.mark.parametrize("direction", [0, 1])
def test_push_marks_hubs_and_tail(gpu, direction):
    """Push activation marks targets below 2^20 (the hubs of the degree order)
    with byte flags folded into the bitmap and the others with atomics: on a
    skewed graph of 2^20 + 37 vertices (a partial last bitmap word) BFS
    depths and trips equal the oracle's, pushing only or automatic."""
    gr = gpu
    n = (1 << 20) + 37
    rng = np.random.default_rng(31)
    m = 4_000_000
    s = rng.integers(0, n, m)
    d = np.minimum((n * rng.random(m) ** 3).astype(np.int64), n - 1)
    keep = s != d
    pairs = np.unique(np.stack([s[keep], d[keep]], 1), axis=0)
    s, d = pairs[:, 0].astype(np.int32), pairs[:, 1].astype(np.int32)
    src = _max_out_src(n, s)
    ref, ref_trips, _ = oracle.bfs(n, s, d, src)
    depth, trips = gr.run_bfs((n, s, d), src, direction=direction)
    assert np.array_equal(depth, ref), f"{(depth != ref).sum()} depth mismatches"
    assert trips == ref_trips
    assert (ref >= 0).sum() > n // 2  # most vertices reached, through both kinds of marks


def test_scan_marks_values_not_below_gather_zero(gpu):
    """In a fused pass above 2^20 edges, targets are found by scanning the
    candidates for values below gatherZero; an edge whose relaxed value is not
    below it is marked in the pass itself.  Vertices that start above
    gatherZero fall to it in the first apply, so their out-edges relax to
    gatherZero + w: the fused run (share 1) and the unfused one (share 0)
    must agree in values, trips and work counters."""
    gr = gpu
    n, s, d = _rmat(gr, 18, seed=41)
    w = gr.hash_weights(len(s), seed=41)
    src = _max_out_src(n, s)
    rng = np.random.default_rng(41)
    vd = np.full(n, gr.SSSP_INF, np.int32)
    high = rng.random(n) < 0.5
    vd[high] = gr.SSSP_INF + 50000  # above gatherZero, no overflow with w <= 64
    vd[src] = 0
    runs = []
    for share in (0.0, 1.0):
        eng = gr.GASEngine(gr.SSSP)
        eng.setFuseShare(share)
        eng.setGraph(n, vd.copy(), len(s), w, s, d)
        eng.setActive(0, n)
        trips, biggest, prev = 0, 0, 0
        while eng.countActive():
            eng.gather()
            eng.apply()
            eng.scatterActivate()
            eng.nextIter()
            trips += 1
            fe = eng.stats()["fused_edges"]
            biggest, prev = max(biggest, fe - prev), fe
        eng.getResults()
        st = eng.stats()
        runs.append((eng.vertexData.copy(), trips, st["active_vertices"], st["gathered_edges"],
                     biggest))
        eng.close()
    (a, ta, aa, ga, _), (b, tb, ab, gb, big) = runs
    assert big > 1 << 20, big
    assert np.array_equal(a, b) and ta == tb and aa == ab and ga == gb
    assert (a == gr.SSSP_INF).any()  # some vertices settle at gatherZero itself


@pytest.mark.parametrize("prog_name", ["SSSP", "CC"])
def test_push_gather_values_reach_apply_once(gpu, prog_name):
    """A push gather leaves the active rows' values in the candidates, which
    apply reads and resets.  Calling gather twice per iteration (the second
    call first resets the first one's candidates) gives the oracle's values,
    trips and work counters; so does a restart right after a gather."""
    gr = gpu
    n, s, d = _rmat(gr, 13, seed=51, sym=prog_name == "CC")
    if prog_name == "SSSP":
        prog, w = gr.SSSP, gr.hash_weights(len(s), seed=51)
        src = _max_out_src(n, s)
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
        ref, rtrips, rst = oracle.sssp(n, s, d, w, src)
    else:
        prog, w, vd = gr.CC, None, np.arange(n, dtype=np.int32)
        ref, rtrips, rst = oracle.cc(n, s, d)
    eng = gr.GASEngine(prog)
    eng.setGraph(n, vd.copy(), len(s), w, s, d)
    eng.setActive(0, n)
    for _ in range(3):  # a few iterations, then a restart right after a gather
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
    eng.gather()
    eng.resetVertexData()
    eng.setActive(0, n)
    eng.resetStats()
    trips = 0
    while eng.countActive():
        eng.gather()
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        trips += 1
    eng.getResults()
    st = eng.stats()
    assert np.array_equal(eng.vertexData, ref) and trips == rtrips
    assert st["push_gathers"] > 0
    eng.close()


# ------------------------------------------------------------ blocked gather

def _run_min_program(gr, prog, graph, n, w_dev, src, share, delta, order=1):
    if prog == gr.SSSP:
        vd = np.full(n, gr.SSSP_INF, np.int32)
        vd[src] = 0
    else:
        vd = np.arange(n, dtype=np.int32)
    eng = gr.GASEngine(prog)
    eng.setVertexOrder(order)
    eng.setDeltaGather(delta)
    eng.setBlockedGather(share)
    eng.bindGraph(graph, vd, w_dev if prog == gr.SSSP else None)
    eng.setActive(0, n)
    trips = eng.run()
    st = eng.stats()
    out = eng.readVertexData().copy()
    eng.close()
    return out, trips, st


@pytest.mark.parametrize("delta", [0, 1, 2])
@pytest.mark.parametrize("order", [0, 1])
def test_blocked_gather_equals_oracle(gpu, delta, order):
    """The source-listed row-block gather gives the oracle's distances,
    labels, trip counts and work counters: off (share 0), for the dense
    gathers (0.5) and for every gather that reaches the tile path (1e-9)."""
    gr = gpu
    n, s, d = _rmat(gr, 15, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    src = _max_out_src(n, s)
    ref, rtrips, rst = oracle.sssp(n, s, d, w, src)
    g = gr.Graph.from_edges(n, s, d)
    wd = gr.upload_i32(w)
    for share in (0.0, 0.5, 1e-9):
        dist, trips, st = _run_min_program(gr, gr.SSSP, g, n, wd, src, share, delta, order)
        assert np.array_equal(dist, ref) and trips == rtrips, f"share {share}"
        assert st["gathered_edges"] == rst["in_edges"]
        if share == 0.0:
            assert st["blocked_gathers"] == 0
        elif delta == 0 and share < 0.01:
            assert st["blocked_gathers"] == trips  # every gather is a full one
        else:
            assert st["blocked_gathers"] > 0
    g.close()
    ns, ss, ds = _rmat(gr, 14, seed=4, sym=True)
    lref, ltrips, _ = oracle.cc(ns, ss, ds)
    gs = gr.Graph.from_edges(ns, ss, ds)
    for share in (0.0, 0.5, 1e-9):
        lab, trips, st = _run_min_program(gr, gr.CC, gs, ns, None, 0, share, delta, order)
        assert np.array_equal(lab, lref) and trips == ltrips, f"CC share {share}"
        assert (st["blocked_gathers"] > 0) == (share > 0)
    gs.close()


def test_blocked_gather_giant_row_and_sparse_tail(gpu):
    """A row with more in-edges than one block holds (2^18), parallel arcs,
    self-loops, vertices without in-edges in the middle and at the end."""
    gr = gpu
    rng = np.random.default_rng(8)
    n = 400_000
    hub_in = rng.integers(1, n, 300_000).astype(np.int32)  # 300k arcs into vertex 0
    s = np.concatenate([hub_in, rng.integers(0, n // 2, 600_000).astype(np.int32),
                        np.arange(0, n // 2, 7, dtype=np.int32)])
    d = np.concatenate([np.zeros(len(hub_in), np.int32),
                        rng.integers(0, n // 2, 600_000).astype(np.int32),
                        np.arange(0, n // 2, 7, dtype=np.int32)])
    w = rng.integers(1, 65, len(s)).astype(np.int32)
    src = int(hub_in[0])
    ref, rtrips, _ = oracle.sssp(n, s, d, w, src)
    g = gr.Graph.from_edges(n, s, d)
    wd = gr.upload_i32(w)
    for order in (0, 1):
        dist, trips, st = _run_min_program(gr, gr.SSSP, g, n, wd, src, 1e-9, 0, order)
        assert np.array_equal(dist, ref) and trips == rtrips
        assert st["blocked_gathers"] == trips
    lref, ltrips, _ = oracle.cc(n, np.concatenate([s, d]), np.concatenate([d, s]))
    gs = gr.Graph.from_edges(n, np.concatenate([s, d]), np.concatenate([d, s]))
    lab, trips, _ = _run_min_program(gr, gr.CC, gs, n, None, 0, 1e-9, 1, 1)
    assert np.array_equal(lab, lref) and trips == ltrips
    g.close()
    gs.close()


def test_blocked_pagerank_fixed_point_sum(gpu):
    """PageRank's dense gathers on the row blocks, summed exactly in fixed
    point: the reference's tolerance vs the oracle (as test_pagerank_rmat), the
    oracle's trip count, and the same ranks on a second run (the sum does not
    depend on the order of the LDS additions)."""
    gr = gpu
    n, s, d = _rmat(gr, 14, seed=3)
    outdeg = np.bincount(s, minlength=n).astype(np.int32)
    ref, ref_trips, _ = oracle.pagerank(n, s, d)
    r64, _, _ = oracle.pagerank64(n, s, d)
    g = gr.Graph.from_edges(n, s, d)
    runs = []
    for share in (0.5, 0.5, 0.0):
        vd = np.zeros(n, gr.PROGRAMS[gr.PAGERANK]["vertex"])
        vd["rank"] = np.float32(0.15)
        vd["numOutEdges"] = outdeg
        eng = gr.GASEngine(gr.PAGERANK)
        eng.setBlockedGather(share)
        eng.bindGraph(g, vd)
        eng.setActive(0, n)
        trips = eng.run()
        st = eng.stats()
        runs.append((eng.readVertexData()["rank"].copy(), trips, st["blocked_gathers"]))
        eng.close()
    g.close()
    (r1, t1, b1), (r2, t2, _), (r0, t0, b0) = runs
    assert b1 > 0 and b0 == 0
    assert np.array_equal(r1, r2)
    for rank, trips in ((r1, t1), (r0, t0)):
        assert trips == ref_trips
        diff = np.abs(rank.astype(np.float64) - ref.astype(np.float64))
        small = ref < 1.0
        assert diff[small].max() <= 1e-6
        assert (diff / ref).max() <= 1e-5
        # and against the same loop in FP64 (the FP32 oracle's own sequential
        # sums drift from it by up to 2.4e-4 on hub ranks at RMAT-15)
        assert (np.abs(rank - r64) / r64).max() <= 1e-5
