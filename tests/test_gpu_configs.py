"""The measurement configurations at their stated sizes, checked bit for bit
against the CPU oracle (or scipy where the oracle's per-iteration vertex scan
would take too long)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_c1_bfs_rmat16_seed1(gpu):
    gr = gpu
    g = gr.Graph.rmat(16, 16, seed=1)
    s, d = g.edges()
    src, deg = g.max_out_degree_vertex()
    assert src == int(np.argmax(np.bincount(s, minlength=g.n_vertices)))
    depth, trips = gr.run_bfs(g, src)
    ref, rtrips, _ = oracle.bfs(g.n_vertices, s, d, src)
    assert np.array_equal(depth, ref) and trips == rtrips
    g.close()


def test_c2_sssp_rmat20_seed2(gpu):
    gr = gpu
    g = gr.Graph.rmat(20, 16, seed=2)
    n = g.n_vertices
    s, d = g.edges()
    w = gr.hash_weights(len(s), seed=2)
    src, _ = g.max_out_degree_vertex()
    dist, trips = gr.run_sssp(g, gr.upload_i32(w), src)
    ref, rtrips, _ = oracle.sssp(n, s, d, w, src)
    assert np.array_equal(dist, ref) and trips == rtrips
    g.close()


def test_c4_cc_symmetric_rmat22_seed4(gpu):
    gr = gpu
    g = gr.Graph.rmat(22, 16, seed=4, symmetrize=True)
    n = g.n_vertices
    off_out, dsts = g.csr()
    off_in, srcs = g.csc()
    lab, trips = gr.run_cc(g)
    ref, rtrips, _ = oracle.cc_csr(n, off_out, dsts, off_in, srcs)
    assert np.array_equal(lab, ref) and trips == rtrips
    g.close()


def test_c4_cc_grid_1024(gpu):
    """The grid config at 1024^2 (the 4096^2 run is in bench.py): labels equal
    the minimum vertex id of each scipy component."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    gr = gpu
    g = gr.Graph.grid(1024, 0.01, seed=5)
    n = g.n_vertices
    lab, trips = gr.run_cc(g)
    s, d = g.edges()
    nc, lbl = cg.connected_components(sp.csr_matrix((np.ones(len(s)), (s, d)), shape=(n, n)),
                                      directed=False)
    mins = np.full(nc, n)
    np.minimum.at(mins, lbl, np.arange(n))
    assert np.array_equal(lab, mins[lbl])
    assert trips > 1024
    g.close()
