"""The measurement configurations at the sizes they are stated for.

C3  PageRank precision form, RMAT-22 seed 3: compensated FP32 within 1e-6
    relative per vertex of the FP64 oracle (pagerank.cu:46-66 recurrence).
C4  CC on the 4096 x 4096 grid with 1 % of its edges deleted (seed 5):
    labels equal the minimum vertex id of each scipy component.
C5  RMAT-24 seed 6, weights U[1,64]:
    plain SSSP and BFS bit-exact against the CPU oracle (sssp.cu:69-111,
    bfs.cu:93-134), CC on the symmetrised graph bit-exact against the oracle;
    64-source batched BFS and SSSP: every row equals the plain engine's run
    for that source and four rows equal the oracle; a served stream of mixed
    BFS / SSSP / PPR queries: every result equals its solo run.

Each test prints the case it is about to run (flushed) before its GPU work.
"""
import sys
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def rmat24(gpu):
    gr = gpu
    _say("C5: building RMAT-24 seed 6 on the device")
    g = gr.Graph.rmat(24, 16, seed=6)
    off_out, dsts = g.csr()
    off_in, srcs = g.csc()
    _, csc_eid = g.perm()
    w = gr.hash_weights(g.n_edges, seed=6)
    w_csc = w[csc_eid]
    del csc_eid
    wd = g.hash_weights_device(seed=6)
    host = dict(off_out=off_out, dsts=dsts, off_in=off_in, srcs=srcs, w=w, w_csc=w_csc,
                outdeg=np.diff(off_out))
    yield gr, g, wd, host
    wd.free()
    g.close()


def _sources64(outdeg, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(np.flatnonzero(outdeg > 0), 64, replace=False).astype(np.int32).tolist()


def test_c5_sssp_rmat24_bit_exact(rmat24):
    gr, g, wd, h = rmat24
    n = g.n_vertices
    src, _ = g.max_out_degree_vertex()
    _say(f"C5 SSSP: source {src}")
    vd = np.full(n, gr.SSSP_INF, np.int32)
    vd[src] = 0
    eng = gr.GASEngine(gr.SSSP)
    eng.bindGraph(g, vd, wd)
    from graphreduce_amd import tune

    tune.apply_tuning(eng, gr.SSSP)  # the settings the bench times
    eng.setActive(0, n)
    trips = eng.run()
    st = eng.stats()
    eng.getResults()
    dist = eng.vertexData.copy()
    eng.close()
    t0 = time.time()
    ref, rtrips, rst = oracle.sssp_csr(n, h["off_out"], h["dsts"], h["off_in"], h["srcs"],
                                       h["w_csc"], src)
    _say(f"C5 SSSP: oracle {time.time() - t0:.1f}s, trips {rtrips}")
    assert np.array_equal(dist, ref), f"{int((dist != ref).sum())} distances differ"
    assert trips == rtrips
    # the roofline numerator: sum #A and sum E_in(A) of the reference loop
    assert st["active_vertices"] == rst["active"]
    assert st["gathered_edges"] == rst["in_edges"]


def test_c5_bfs_rmat24_bit_exact(rmat24):
    gr, g, wd, h = rmat24
    n = g.n_vertices
    src, _ = g.max_out_degree_vertex()
    _say(f"C5 BFS: source {src}")
    depth, trips = gr.run_bfs(g, src)
    ref, rtrips, _ = oracle.bfs_csr(n, h["off_out"], h["dsts"], h["off_in"], h["srcs"], src)
    assert np.array_equal(depth, ref) and trips == rtrips


def test_c5_cc_symmetric_rmat24_bit_exact(rmat24):
    gr = rmat24[0]
    _say("C5 CC: symmetrised RMAT-24 seed 6")
    gs = gr.Graph.rmat(24, 16, seed=6, symmetrize=True)
    n = gs.n_vertices
    lab, trips = gr.run_cc(gs)
    off_out, dsts = gs.csr()
    off_in, srcs = gs.csc()
    gs.close()
    ref, rtrips, _ = oracle.cc_csr(n, off_out, dsts, off_in, srcs)
    assert np.array_equal(lab, ref) and trips == rtrips


def test_c5_batched_bfs_64_rows(rmat24):
    gr, g, wd, h = rmat24
    n = g.n_vertices
    srcs = _sources64(h["outdeg"], 6)
    _say("C5 batched BFS: 64 sources")
    depth, info = gr.msbfs(g, srcs)
    assert depth.shape == (64, n)
    for i, s in enumerate(srcs):
        plain, _ = gr.run_bfs(g, s)
        assert np.array_equal(depth[i], plain), f"row {i} (source {s})"
    for i in (0, 21, 42, 63):
        ref, _, _ = oracle.bfs_csr(n, h["off_out"], h["dsts"], h["off_in"], h["srcs"], srcs[i])
        assert np.array_equal(depth[i], ref), f"row {i} vs oracle"
    _say("C5 batched BFS: 64 sources, one persistent launch")
    pdepth, _ = gr.msbfs(g, srcs, persistent=True)
    assert np.array_equal(pdepth, depth)


def test_c5_batched_sssp_64_rows(rmat24):
    gr, g, wd, h = rmat24
    n = g.n_vertices
    srcs = _sources64(h["outdeg"], 6)
    _say("C5 batched SSSP: 64 sources")
    dist, info = gr.mssp(g, wd, srcs)
    assert dist.shape == (64, n)
    for i, s in enumerate(srcs):
        plain, _ = gr.run_sssp(g, wd, s)
        assert np.array_equal(dist[i], plain), f"row {i} (source {s})"
    for i in (0, 21, 42, 63):
        ref, _, _ = oracle.sssp_csr(n, h["off_out"], h["dsts"], h["off_in"], h["srcs"],
                                    h["w_csc"], srcs[i])
        assert np.array_equal(dist[i], ref), f"row {i} vs oracle"


def test_c5_served_stream_equals_solo(rmat24):
    """A Poisson stream (2,000 req/s offered) of 48 mixed queries on RMAT-24:
    the server overlaps them; each BFS / SSSP result equals a plain engine
    run, each PPR result equals the same query served alone, bit for bit."""
    gr, g, wd, h = rmat24
    rng = np.random.default_rng(17)
    k = 48
    verts = rng.choice(np.flatnonzero(h["outdeg"] > 0), k).astype(np.int32)
    kinds = np.array([gr.QUERY_BFS, gr.QUERY_SSSP, gr.QUERY_PPR] * (k // 3), np.int32)
    _say(f"C5 serving: {k} requests")
    srv = gr.Server(g, wd)
    ids = srv.load_test(2000.0, verts, kinds, keep=True)
    mix = [srv.wait(int(rid), int(kinds[i])) for i, rid in enumerate(ids)]
    st = srv.stats()
    srv.close()
    assert st["completed"] == k
    assert st["max_lanes_in_use"] > 3, "requests overlapped on the device"
    solo = gr.Server(g, wd)
    for i in range(k):
        row, cs, _, _ = mix[i]
        v = int(verts[i])
        if kinds[i] == gr.QUERY_BFS:
            ref, _ = gr.run_bfs(g, v)
        elif kinds[i] == gr.QUERY_SSSP:
            ref, _ = gr.run_sssp(g, wd, v)
        else:
            rid = solo.submit(int(kinds[i]), v, keep=True)
            ref, cs_solo, _, _ = solo.wait(rid, int(kinds[i]))
            assert cs == cs_solo
        assert np.array_equal(row, ref), f"request {i} (kind {int(kinds[i])}, vertex {v})"
    solo.close()


def test_c4_cc_grid_4096(gpu):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    gr = gpu
    _say("C4: CC on the 4096^2 grid, 1 % deleted, seed 5")
    g = gr.Graph.grid(4096, 0.01, seed=5)
    n = g.n_vertices
    lab, trips = gr.run_cc(g)
    s, d = g.edges()
    g.close()
    nc, lbl = cg.connected_components(sp.csr_matrix((np.ones(len(s), np.int8), (s, d)),
                                                    shape=(n, n)), directed=False)
    mins = np.full(nc, n)
    np.minimum.at(mins, lbl, np.arange(n))
    assert np.array_equal(lab, mins[lbl])
    assert trips > 4096


def test_c3_pagerank_precision_rmat22(gpu):
    gr = gpu
    _say("C3: PageRank precision form, RMAT-22 seed 3")
    g = gr.Graph.rmat(22, 16, seed=3)
    rk, info = gr.pagerank_precise(g, max_iters=50, tol=1e-7, compensated=True)
    rp, _ = gr.pagerank_precise(g, max_iters=50, tol=1e-7, compensated=False)
    off_out, _ = g.csr()
    off_in, srcs = g.csc()
    n = g.n_vertices
    g.close()
    ref, l1 = oracle.pagerank_fp64_csr(n, off_out, off_in, srcs, info["iterations"])
    rel = np.abs(rk.astype(np.float64) - ref) / ref
    assert rel.max() <= 1e-6, f"compensated: max rel err {rel.max():.3g}"
    relp = np.abs(rp.astype(np.float64) - ref) / ref
    assert relp.max() <= 1e-4, f"plain FP32: max rel err {relp.max():.3g}"
    # the FP64 residual that decides convergence follows the FP64 recurrence
    assert np.allclose(info["residuals"], l1, rtol=1e-3)
