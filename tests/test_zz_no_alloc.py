"""gr_pagerank_precise_noalloc (no device allocation: its buffers live in one
static device array per GPU) against gr_pagerank_precise, the entry point it
is a variant of.  Both run the same kernels in the same order, so ranks and
FP64 residuals must be identical, bit for bit, on two calls in a row, the
second on a smaller graph (the static array keeps the first call's contents).
A graph beyond the array's size is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(gr, fn, g, iters, comp):
    ranks = np.empty(g.n_vertices, np.float32)
    resid = np.zeros(iters, np.float64)
    r = gr._PRResult()
    rc = fn(g.handle, iters, 1e-7, int(comp), ranks.ctypes.data_as(C.POINTER(C.c_float)),
            resid.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r))
    gr._check(rc)
    return ranks, resid[: r.iterations], r.iterations


@pytest.mark.parametrize("comp", [True, False])
def test_noalloc_equals_allocating_entry(gpu, comp):
    gr = gpu
    L = gr.lib()
    for scale in (16, 12):  # the second call on a smaller graph
        print(f"noalloc PageRank: RMAT-{scale}, compensated={comp}", flush=True)
        g = gr.Graph.rmat(scale, 16, seed=3)
        a, ra, ia = _call(gr, L.gr_pagerank_precise, g, 30, comp)
        b, rb, ib = _call(gr, L.gr_pagerank_precise_noalloc, g, 30, comp)
        g.close()
        assert ia == ib == 30
        assert np.array_equal(a, b), f"RMAT-{scale}: {(a != b).sum()} ranks differ"
        assert np.array_equal(ra, rb)
        assert a.min() >= np.float32(0.15)


def test_noalloc_refuses_graph_beyond_scratch(gpu):
    gr = gpu
    print("noalloc PageRank: 6M-vertex graph (over the limit)", flush=True)
    big = gr.Graph.from_edges(6_000_000, np.array([0, 1], np.int32), np.array([1, 2], np.int32))
    with pytest.raises(gr.GRError, match="maximum is 83886080"):
        gr.pagerank_precise(big, max_iters=2, noalloc=True)
    big.close()
    g = gr.Graph.rmat(10, 16, seed=3)
    a, ia = gr.pagerank_precise(g, max_iters=5, noalloc=True)
    b, ib = gr.pagerank_precise(g, max_iters=5)
    g.close()
    assert np.array_equal(a, b) and ia["iterations"] == ib["iterations"] == 5
