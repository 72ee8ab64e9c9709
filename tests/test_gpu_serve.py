"""Continuous-batching server: every result in a mixed, overlapping stream
equals the same query run alone (bit for bit: BFS/SSSP against the CPU oracle,
PPR against a solo server run), PPR is within 1e-5 relative of an FP64 power
iteration, and memory-based admission bounds the lanes in use."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle

pytestmark = pytest.mark.gpu

PPR_ITERS = 8


def _ppr_fp64(n, s, d, seed, iters):
    outdeg = np.bincount(s, minlength=n).astype(np.float64)
    A = sp.csr_matrix((np.ones(len(s)), (d, s)), shape=(n, n))  # A[v, u] for arc u -> v
    e = np.zeros(n)
    e[seed] = 1.0
    r = e.copy()
    for _ in range(iters):
        c = np.divide(r, outdeg, out=np.zeros(n), where=outdeg > 0)
        r = 0.15 * e + 0.85 * (A @ c)
    return r


@pytest.fixture(scope="module")
def setup(gpu):
    gr = gpu
    g = gr.Graph.rmat(12, 16, seed=6)
    s, d = g.edges()
    w = gr.hash_weights(len(s), seed=6)
    wd = gr.upload_i32(w)
    yield gr, g, s, d, w, wd
    g.close()


def test_mixed_stream_equals_solo(setup):
    gr, g, s, d, w, wd = setup
    n = g.n_vertices
    rng = np.random.default_rng(9)
    cand = np.flatnonzero(np.bincount(s, minlength=n) > 0)
    k = 90
    verts = rng.choice(cand, k).astype(np.int32)
    verts[6] = verts[3]  # the same BFS query twice in the mix
    kinds = np.array([1, 2, 3] * (k // 3), np.int32)
    srv = gr.Server(g, wd, ppr_iterations=PPR_ITERS)
    ids = srv.load_test(4000.0, verts, kinds, keep=True)
    results = {}
    for i, rid in enumerate(ids):
        results[i] = srv.wait(int(rid), int(kinds[i]))
    st = srv.stats()
    assert st["completed"] == k and st["submitted"] == k
    assert st["max_lanes_in_use"] > 3, "requests overlapped on the device"
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    srv.close()
    solo = gr.Server(g, wd, ppr_iterations=PPR_ITERS)
    for i in range(k):
        res, cs, lat, steps = results[i]
        v = int(verts[i])
        if kinds[i] == gr.QUERY_BFS:
            ref, _, _ = oracle.bfs(n, s, d, v)
            assert np.array_equal(res, ref), f"BFS {v}"
        elif kinds[i] == gr.QUERY_SSSP:
            ref, _, _ = oracle.sssp(n, s, d, w, v)
            assert np.array_equal(res, ref), f"SSSP {v}"
        else:
            assert steps == PPR_ITERS
            alone, cs_alone, _, _ = solo.wait(solo.submit(gr.QUERY_PPR, v), gr.QUERY_PPR)
            assert np.array_equal(res.view(np.int32), alone.view(np.int32)), f"PPR {v}"
            assert cs == cs_alone
            ref = _ppr_fp64(n, s, d, v, PPR_ITERS)
            big = ref > 1e-7
            assert np.allclose(res[big], ref[big], rtol=1e-5, atol=0)
            assert np.all(np.abs(res[~big] - ref[~big]) < 1e-7)
    # identical queries give identical checksums
    assert kinds[3] == kinds[6] and results[3][1] == results[6][1]
    solo.close()


def test_memory_budget_admission(setup):
    gr, g, s, d, w, wd = setup
    n = g.n_vertices
    # lane bytes: BFS 4n + 24n/64, SSSP 4n + 16n/64, PPR 8n + 16n/64; a budget of
    # two BFS lanes admits at most two lanes of any mix (and one PPR lane)
    lane_bytes = 4 * n + 24 * n // 64
    srv = gr.Server(g, wd, memory_budget_bytes=2 * lane_bytes, ppr_iterations=4)
    rng = np.random.default_rng(2)
    verts = rng.integers(0, n, 30).astype(np.int32)
    kinds = np.array([1, 2, 3] * 10, np.int32)
    ids = srv.load_test(10000.0, verts, kinds, keep=False)
    for rid in ids:
        srv.wait(int(rid), want_result=False)
    st = srv.stats()
    assert st["completed"] == 30
    assert st["max_lanes_in_use"] <= 2
    assert st["peak_bytes"] <= 2 * lane_bytes
    srv.close()
    tiny = gr.Server(g, wd, memory_budget_bytes=1024)
    with pytest.raises(gr.GRError):
        tiny.submit(gr.QUERY_BFS, 0)
    tiny.close()
