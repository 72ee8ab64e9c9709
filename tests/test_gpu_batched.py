"""Batched multi-source BFS / SSSP: every row must equal the plain engine's
result for that source (and the CPU oracle's), bit for bit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _sources(n, k, seed, extra=()):
    rng = np.random.default_rng(seed)
    return list(extra) + rng.integers(0, n, k - len(extra)).tolist()


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_msbfs_equals_plain_runs(gpu, direction):
    gr = gpu
    g = gr.Graph.rmat(14, 16, seed=6)
    n = g.n_vertices
    s, d = g.edges()
    hub, _ = g.max_out_degree_vertex()
    srcs = _sources(n, 64, 11, extra=(hub, hub))  # a hub and a duplicated source
    depth, info = gr.msbfs(g, srcs, direction=direction)
    assert info["levels"] >= 1
    for i, src in enumerate(srcs):
        ref, _, _ = oracle.bfs(n, s, d, src)
        assert np.array_equal(depth[i], ref), f"source {src} (row {i})"
    g.close()


@pytest.mark.skip(reason="streaming-pull MS-BFS kernel disabled: its cooperative launch did not "
                  "finish on the GPU (RMAT-14, 64 sources, automatic direction); cause not "
                  "established, a CPU lane emulation of it terminates with exact depths")
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_msbfs_streaming_pull_equals_plain_runs(gpu, direction):
    gr = gpu
    g = gr.Graph.rmat(14, 16, seed=6)
    n = g.n_vertices
    s, d = g.edges()
    hub, _ = g.max_out_degree_vertex()
    srcs = _sources(n, 64, 11, extra=(hub, hub))
    depth, _ = gr.msbfs(g, srcs, direction=direction, stream_pull=True)
    for i, src in enumerate(srcs):
        ref, _, _ = oracle.bfs(n, s, d, src)
        assert np.array_equal(depth[i], ref), f"source {src} (row {i})"
    g.close()


def test_msbfs_streaming_pull_is_refused(gpu):
    gr = gpu
    g = gr.Graph.rmat(8, 16, seed=1)
    with pytest.raises(gr.GRError, match="disabled"):
        gr.msbfs(g, [0, 1], stream_pull=True)
    g.close()


def test_msbfs_partial_batch_and_plain_engine(gpu):
    gr = gpu
    g = gr.Graph.rmat(12, 16, seed=1)
    n = g.n_vertices
    srcs = _sources(n, 5, 3)
    depth, _ = gr.msbfs(g, srcs)
    for i, src in enumerate(srcs):
        plain, _ = gr.run_bfs(g, src)
        assert np.array_equal(depth[i], plain)
    g.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_msbfs_grid_deep_levels(gpu, persistent):
    """Over 150 levels: the persistent launch's counter ring and barriers
    hold level after level."""
    gr = gpu
    g = gr.Graph.grid(96, 0.01, seed=5)
    s, d = g.edges()
    srcs = [0, 95, 96 * 96 - 1, 4000]
    depth, info = gr.msbfs(g, srcs, persistent=persistent)
    assert info["levels"] > 150
    for i, src in enumerate(srcs):
        ref, _, _ = oracle.bfs(g.n_vertices, s, d, src)
        assert np.array_equal(depth[i], ref)
    g.close()


def test_mssp_equals_plain_runs(gpu):
    gr = gpu
    g = gr.Graph.rmat(13, 16, seed=2)
    n = g.n_vertices
    s, d = g.edges()
    w = gr.hash_weights(len(s), seed=2)
    wd = gr.upload_i32(w)
    hub, _ = g.max_out_degree_vertex()
    srcs = _sources(n, 64, 5, extra=(hub,))
    dist, info = gr.mssp(g, wd, srcs)
    assert info["levels"] >= 1
    for i, src in enumerate(srcs[:16]):
        ref, _, _ = oracle.sssp(n, s, d, w, src)
        assert np.array_equal(dist[i], ref), f"source {src}"
    for i in (20, 40, 63):
        plain, _ = gr.run_sssp(g, wd, srcs[i])
        assert np.array_equal(dist[i], plain)
    g.close()


def test_batched_rejects_bad_sources(gpu):
    gr = gpu
    g = gr.Graph.rmat(8, 16, seed=1)
    with pytest.raises(gr.GRError):
        gr.msbfs(g, [0] * 65)
    with pytest.raises(gr.GRError):
        gr.msbfs(g, [g.n_vertices])
    g.close()


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_msbfs_variants_agree(gpu, direction):
    """Multi-kernel (with and without the frontier-map filter) and the single
    persistent launch give the same depths as the oracle."""
    gr = gpu
    g = gr.Graph.rmat(13, 16, seed=4)
    n = g.n_vertices
    s, d = g.edges()
    hub, _ = g.max_out_degree_vertex()
    srcs = _sources(n, 40, 7, extra=(hub,))
    outs = [gr.msbfs(g, srcs, direction=direction)[0],
            gr.msbfs(g, srcs, direction=direction, pull_filter=True)[0],
            gr.msbfs(g, srcs, direction=direction, persistent=True)[0]]
    for i, src in enumerate(srcs[:8]):
        ref, _, _ = oracle.bfs(n, s, d, src)
        assert np.array_equal(outs[0][i], ref)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    g.close()


def test_batch_context_reuse(gpu):
    """One context, several runs of both kinds and sizes in a row: every run
    equals the oracle, so no run depends on state another one left."""
    gr = gpu
    g = gr.Graph.rmat(13, 16, seed=4)
    n = g.n_vertices
    s, d = g.edges()
    w = gr.hash_weights(len(s), seed=9)
    w_dev = gr.upload_i32(w)
    hub, _ = g.max_out_degree_vertex()
    b = gr.Batch(g, w_dev)
    big = _sources(n, 64, 21, extra=(hub,))
    small = _sources(n, 5, 22)
    for srcs, kw in ((big, {}), (small, {}), (big, {"direction": 2}), (small, {"persistent": True})):
        depth, info = b.bfs(srcs, **kw)
        assert info["levels"] >= 1
        for i, src in enumerate(srcs):
            ref, _, _ = oracle.bfs(n, s, d, src)
            assert np.array_equal(depth[i], ref), f"BFS {kw} source {src} (row {i})"
    for srcs in (big, small, big[:17]):
        dist, _ = b.sssp(srcs)
        for i, src in enumerate(srcs):
            ref, _, _ = oracle.sssp(n, s, d, w, src)
            assert np.array_equal(dist[i], ref), f"SSSP source {src} (row {i})"
    # the one-shot forms give the same rows
    depth1, _ = gr.msbfs(g, small)
    depth2, _ = b.bfs(small)
    assert np.array_equal(depth1, depth2)
    b.close()
    # a context without weights runs BFS and refuses SSSP
    b2 = gr.Batch(g)
    depth3, _ = b2.bfs(small)
    assert np.array_equal(depth1, depth3)
    with pytest.raises(gr.GRError, match="without weights"):
        b2.sssp(small)
    b2.close()
    w_dev.free()
    g.close()
