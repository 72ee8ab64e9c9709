"""The C++ app drivers (graphreduce_amd/bin, reference command lines) against
the CPU oracle: the `-t` comparison of the reference apps (bfs.cu:246-261,
sssp.cu:221-236, pagerank.cu:154-170, connected_component.cu:139-154)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import graphreduce_amd as gr
import oracle

BIN = Path(gr.__file__).resolve().parent / "bin"
pytestmark = pytest.mark.gpu


def _run(app, args, tmp_path):
    out = tmp_path / f"{app}.txt"
    proc = subprocess.run([str(BIN / app), *args, str(out)], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    vals = np.loadtxt(out, dtype=np.float64, ndmin=2)
    return proc.stdout, vals[:, 1]


def test_bfs_app(gpu, tmp_path):
    stdout, depth = _run("bfs", ["-m", "rmat:12:1", "0"], tmp_path)
    s, d = gr.rmat_edges(12, 16, seed=1)
    src = int(np.argmax(np.bincount(s, minlength=1 << 12)))
    assert f"using vertex {src}" in stdout
    ref, trips, _ = oracle.bfs(1 << 12, s, d, src)
    assert np.array_equal(depth.astype(np.int32), ref)
    assert f"search depth (number of iterations): {trips}" in stdout
    assert "MTEPS" in stdout


def test_sssp_app(gpu, tmp_path):
    stdout, dist = _run("sssp", ["-m", "rmat:12:2", "0"], tmp_path)
    s, d = gr.rmat_edges(12, 16, seed=2)
    w = gr.hash_weights(len(s), seed=2)
    src = int(np.argmax(np.bincount(s, minlength=1 << 12)))
    ref, _, _ = oracle.sssp(1 << 12, s, d, w, src)
    assert np.array_equal(dist.astype(np.int64), ref.astype(np.int64))


def test_cc_app_grid(gpu, tmp_path):
    stdout, lab = _run("connected_component", ["grid:64:5"], tmp_path)
    s, d = gr.grid_edges(64, 0.01, seed=5)
    ref, _, _ = oracle.cc(64 * 64, s, d)
    assert np.array_equal(lab.astype(np.int32), ref)
    assert "Iterations:" in stdout and "Took" in stdout


def test_pagerank_app(gpu, tmp_path):
    stdout, rank = _run("pagerank", ["rmat:10:3"], tmp_path)
    s, d = gr.rmat_edges(10, 16, seed=3)
    ref, _, _ = oracle.pagerank(1 << 10, s, d)
    # the app prints "%f" (6 decimals), pagerank.cu:80
    assert np.allclose(rank, ref, rtol=1e-5, atol=1e-6)


def test_mtx_input(gpu, tmp_path):
    mtx = tmp_path / "g.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate integer symmetric\n% comment\n"
                   "5 5 5\n1 2 7\n2 3 3\n3 3 9\n4 5 1\n1 4 2\n")
    stdout, dist = _run("sssp", [str(mtx), "0"], tmp_path)
    # symmetric expansion, self-loop (3,3) dropped, 1-based ids (graphio.cpp:133-166)
    assert dist.astype(np.int64).tolist() == [0, 7, 10, 2, 3]


def test_gr_and_edge_file_inputs(gpu, tmp_path):
    """SSSP over a .gr file written by mtx2gr's writer, BFS over a .edge.gz
    file: the same results as the oracle on the file's edges."""
    import gzip
    s, d = gr.rmat_edges(11, 16, seed=7)
    n = 1 << 11
    w = gr.hash_weights(len(s), seed=7)
    gr.write_gr(tmp_path / "g.gr", n, s, d, w)
    n2, fs, fd, fw = gr.read_graph(tmp_path / "g.gr", True)
    assert n2 == n
    src = int(np.argmax(np.bincount(fs, minlength=n)))
    stdout, dist = _run("sssp", ["-m", str(tmp_path / "g.gr"), "0"], tmp_path)
    assert f"using vertex {src}" in stdout
    ref, _, _ = oracle.sssp(n, fs, fd, fw, src)
    assert np.array_equal(dist.astype(np.int64), ref.astype(np.int64))
    body = "# rmat 11\n" + "".join(f"{a} {b}\n" for a, b in zip(s.tolist(), d.tolist()))
    (tmp_path / "g.edge.gz").write_bytes(gzip.compress(body.encode()))
    stdout, depth = _run("bfs", ["-m", str(tmp_path / "g.edge.gz"), "0"], tmp_path)
    ne = int(max(s.max(), d.max())) + 1  # a text file has 1 + the largest id vertices
    src = int(np.argmax(np.bincount(s, minlength=ne)))
    ref, _, _ = oracle.bfs(ne, s, d, src)
    assert np.array_equal(depth.astype(np.int32), ref)


def test_regression_checkers_accept_app_output(gpu, tmp_path):
    """The regression flow of regressions/Makefile:66-110: app output vs a gold
    file (here the oracle's, with PowerGraph's 10000000 for unreachable) through
    checkDist / checkPageRank."""
    from graphreduce_amd import checkers
    s, d = gr.rmat_edges(11, 16, seed=9)
    n = 1 << 11
    w = gr.hash_weights(len(s), seed=9)
    gr.write_gr(tmp_path / "g.gr", n, s, d, w)
    _, fs, fd, fw = gr.read_graph(tmp_path / "g.gr", True)
    _run("sssp", [str(tmp_path / "g.gr"), "0"], tmp_path)
    ref, _, _ = oracle.sssp(n, fs, fd, fw, 0)
    gold = tmp_path / "sssp.gold"
    gold.write_text("".join(f"{v} {10000000 if x == gr.SSSP_INF else x}\n"
                            for v, x in enumerate(ref.tolist())))
    assert checkers.check_dist(tmp_path / "sssp.txt", gold)[0]
    bad = tmp_path / "sssp.bad"
    bad.write_text("".join(f"{v} {x + (v == 7)}\n" for v, x in enumerate(ref.tolist())))
    assert not checkers.check_dist(tmp_path / "sssp.txt", bad)[0]
    _run("pagerank", [str(tmp_path / "g.gr")], tmp_path)
    pr, _, _ = oracle.pagerank(n, fs, fd)
    pgold = tmp_path / "pr.gold"
    pgold.write_text("".join(f"{v} {x:.9f}\n" for v, x in enumerate(pr.tolist())))
    ok, msg = checkers.check_pagerank(tmp_path / "pagerank.txt", pgold)
    assert ok, msg


def test_widest_path_app_runs_its_own_program(gpu, tmp_path):
    """A vertex program outside the library (widest paths, reference-style
    two-argument apply) compiled into the app with hipcc: its results, trips
    and reached count equal the oracle's Engine<WidestPath>."""
    for graph, seed in (("rmat:12:5", 5), ("rmat:14:3", 3)):
        scale = int(graph.split(":")[1])
        stdout, width = _run("widest_path", ["-m", graph, "0"], tmp_path)
        s, d = gr.rmat_edges(scale, 16, seed=seed)
        caps = gr.hash_weights(len(s), seed=seed)
        src = int(np.argmax(np.bincount(s, minlength=1 << scale)))
        assert f"using vertex {src}" in stdout
        ref, trips, _ = oracle.widest(1 << scale, s, d, caps, src)
        assert np.array_equal(width.astype(np.int64), ref.astype(np.int64)), graph
        assert f"number of iterations: {trips}" in stdout
        assert f"vertices reached: {int((ref > 0).sum())}" in stdout
