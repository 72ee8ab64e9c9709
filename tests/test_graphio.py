"""Graph files: .edge / .mtx / .gr readers and writers (graphio.cpp:69-473),
edgeListToCSR (util.cuh:79-108) and the mtx2gr / gr2mtx tools.

The oracle (oracle/graphio.py) is pinned by hand-derived expectations on
small files that exercise every rule of the reference's readers; the native
readers are then compared with it on generated files (plain and gzip), and
the writers are checked byte for byte and by round trips.  Host code only:
no GPU is touched."""
import gzip
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import graphreduce_amd as gr
from oracle import graphio as og

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "graphreduce_amd" / "bin"


def _write(path: Path, text: str, compress: bool = False) -> Path:
    data = text.encode()
    if compress:
        data = gzip.compress(data)
    path.write_bytes(data)
    return path


def _both(path, values):
    """Native reader and oracle on the same file; they must agree."""
    n, s, d, v = gr.read_graph(path, values)
    on, os_, od, ov = og.load_graph(path, values)
    assert n == on
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    assert (v is None) == (ov is None)
    if v is not None:
        assert np.array_equal(v, ov)
    return n, s.tolist(), d.tolist(), None if v is None else v.tolist()


# ------------------------------------------------ hand-derived expectations

MTX_GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment
%another one

7 7 5
1 2 3.7
2 3 -2.5

3 1 1e1
9 9 4
4 2 0.99
"""


def test_mtx_general_rules(tmp_path):
    p = _write(tmp_path / "g.mtx", MTX_GENERAL)
    # values wanted: the 9-9 self-loop is dropped and does not count toward
    # nVertices; the size line (7 7 5) is ignored; values are truncated
    assert _both(p, True) == (4, [0, 1, 2, 3], [1, 2, 0, 1], [3, -2, 10, 0])
    # no values: the self-loop line counts toward nVertices
    assert _both(p, False) == (9, [0, 1, 2, 3], [1, 2, 0, 1], None)


def test_mtx_symmetry_kinds(tmp_path):
    body = "3 3 2\n1 2 5\n3 1 2.5\n"
    sym = _write(tmp_path / "s.mtx", "%%MatrixMarket matrix coordinate integer symmetric\n" + body)
    assert _both(sym, True) == (3, [0, 1, 2, 0], [1, 0, 0, 2], [5, 5, 2, 2])
    skew = _write(tmp_path / "k.mtx",
                  "%%MatrixMarket matrix coordinate real skew-symmetric\n" + body)
    assert _both(skew, True) == (3, [0, 1, 2, 0], [1, 0, 0, 2], [5, -5, 2, -2])
    herm = _write(tmp_path / "h.mtx", "%%MatrixMarket matrix coordinate real hermitian\n" + body)
    assert _both(herm, False) == (3, [0, 1, 2, 0], [1, 0, 0, 2], None)
    pat = _write(tmp_path / "p.mtx", "%%MatrixMarket matrix coordinate pattern general\n"
                 "3 3 2\n1 2\n3 1\n")
    assert _both(pat, True) == (3, [0, 2], [1, 0], None)  # pattern: no values even if asked


def test_edge_file_rules(tmp_path):
    p = _write(tmp_path / "a.edge", "# SNAP\n0 1\n1 2 99\n\n5 5\n# x\n2 0\n")
    # 0-based, extra columns ignored, the self-loop dropped but counted
    assert _both(p, True) == (6, [0, 1, 2], [1, 2, 0], None)


def test_gr_rules(tmp_path):
    # 4 vertices, 3 edges (odd: padded), vertex 2 has no out-edges, a self-loop kept
    img = og.gr_bytes(4, [0, 2, 2, 2, 3], [1, 3, 3], [7, -1, 9])
    (tmp_path / "a.gr").write_bytes(img)
    assert len(img) == 32 + 8 * 4 + 4 * 3 + 4 + 4 * 3
    assert _both(tmp_path / "a.gr", True) == (4, [0, 0, 3], [1, 3, 3], [7, -1, 9])
    assert _both(tmp_path / "a.gr", False) == (4, [0, 0, 3], [1, 3, 3], None)
    (tmp_path / "b.gr.gz").write_bytes(gzip.compress(img))
    assert _both(tmp_path / "b.gr.gz", True) == (4, [0, 0, 3], [1, 3, 3], [7, -1, 9])


def test_extension_rules(tmp_path):
    txt = "0 1\n1 0\n"
    assert _both(_write(tmp_path / "x.edges", txt), False)[0] == 2  # prefix ".edge"
    assert _both(_write(tmp_path / "x.edge.gz", txt, compress=True), False)[0] == 2
    with pytest.raises(gr.GRError, match="extension"):
        gr.read_graph(_write(tmp_path / "x.txt", txt))
    with pytest.raises(ValueError):
        og.load_graph(tmp_path / "x.txt")


# ------------------------------------------------------------ error paths

@pytest.mark.parametrize("name,text,values,msg", [
    ("c.mtx", "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 2 1 1\n", True,
     "complex"),
    ("a.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n", False, "coordinate"),
    ("y.mtx", "%%MatrixMarket matrix coordinate real upper\n1 1 1\n1 2 1\n", False, "symmetry"),
    ("t.mtx", "%%MatrixMarket matrix\n", False, "header"),
    ("m.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2\n", True, "value"),
    ("h.mtx", "%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 2 1\n", True,
     "hermitian"),
    ("z.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 2 1\n", False, "range"),
    ("b.edge", "0 1\n1\n", False, "src/dst"),
])
def test_text_errors(tmp_path, name, text, values, msg):
    p = _write(tmp_path / name, text)
    with pytest.raises(gr.GRError, match=msg):
        gr.read_graph(p, values)
    with pytest.raises(ValueError):
        og.load_graph(p, values)


def test_gr_errors(tmp_path):
    good = og.gr_bytes(3, [0, 1, 2, 2], [1, 2])
    (tmp_path / "nov.gr").write_bytes(good)
    with pytest.raises(gr.GRError, match="edge values"):  # graphio.cpp:292-296
        gr.read_graph(tmp_path / "nov.gr", True)
    assert gr.read_graph(tmp_path / "nov.gr", False)[0] == 3
    (tmp_path / "short.gr").write_bytes(good[:-3])
    with pytest.raises(gr.GRError, match="truncated"):
        gr.read_graph(tmp_path / "short.gr")
    wide = struct.pack("<4Q", 1, 8, 1, 0) + b"\0" * 8
    (tmp_path / "wide.gr").write_bytes(wide)
    with pytest.raises(gr.GRError, match="8 bytes"):
        gr.read_graph(tmp_path / "wide.gr")
    bad = og.gr_bytes(3, [0, 2, 1, 2], [1, 2])  # offsets decrease
    (tmp_path / "bad.gr").write_bytes(bad)
    with pytest.raises(gr.GRError, match="out of order"):
        gr.read_graph(tmp_path / "bad.gr")
    far = og.gr_bytes(2, [0, 1, 1], [5])  # destination outside [0, nV)
    (tmp_path / "far.gr").write_bytes(far)
    with pytest.raises(gr.GRError, match="out of range"):
        gr.read_graph(tmp_path / "far.gr")
    for f in ("bad.gr", "far.gr"):
        with pytest.raises(ValueError):
            og.load_graph(tmp_path / f)
    with pytest.raises(gr.GRError, match="opening"):
        gr.read_graph(tmp_path / "missing.mtx")


# ------------------------------------------------ generated files vs oracle

def _random_mtx(rng, n, m, sym, kind):
    lines = [f"%%MatrixMarket matrix coordinate {kind} {sym}", "% generated"]
    lines.append(f"{n} {n} {m}")
    for _ in range(m):
        s, d = rng.integers(1, n + 1, 2)
        if rng.random() < 0.05:
            d = s
        if kind == "pattern":
            lines.append(f"{s} {d}")
        else:
            v = rng.uniform(-100, 100) if kind == "real" else rng.integers(-50, 50)
            lines.append(f"{s}\t{d}  {v}")
        if rng.random() < 0.03:
            lines.append("%" + "c" * int(rng.integers(0, 10)))
        if rng.random() < 0.03:
            lines.append("  \t ")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", range(12))
def test_native_reader_matches_oracle(tmp_path, seed):
    rng = np.random.default_rng(seed)
    sym = ["general", "symmetric", "skew-symmetric", "hermitian"][seed % 4]
    kind = ["real", "integer", "pattern"][seed % 3]
    text = _random_mtx(rng, int(rng.integers(2, 300)), int(rng.integers(0, 2000)), sym, kind)
    name = f"r{seed}.mtx" + (".gz" if seed % 2 else "")
    p = _write(tmp_path / name, text, compress=seed % 2 == 1)
    want = sym != "hermitian"
    _both(p, want)
    _both(p, False)
    # the same edges as a .edge file (0-based, '#' comments)
    n, s, d, _ = og.load_graph(p, False)
    body = "# x\n" + "".join(f"{a} {b}\n" for a, b in zip(s, d))
    _both(_write(tmp_path / f"r{seed}.edge", body), False)


# ------------------------------------------------------------ writers

def test_edge_list_to_csr_stable():
    rng = np.random.default_rng(3)
    n = 50
    s = rng.integers(0, n, 600).astype(np.int32)
    d = rng.integers(0, n, 600).astype(np.int32)
    off, out, idx = gr.edge_list_to_csr(n, s, d)
    o2, out2, idx2 = og.edge_list_to_csr(n, s, d)
    assert np.array_equal(off, o2) and np.array_equal(out, out2) and np.array_equal(idx, idx2)
    # closed form: offsets are the prefix sums of the out-degrees, and every
    # CSR slot holds its edge's destination
    assert np.array_equal(np.diff(off), np.bincount(s, minlength=n))
    assert np.array_equal(out, d[idx]) and np.array_equal(s[idx], np.repeat(np.arange(n), np.diff(off)))
    assert sorted(idx.tolist()) == list(range(600))


def test_writers_bytes_and_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n, m = 40, 301  # odd edge count: .gr padding
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(-5, 64, m).astype(np.int32)
    gr.write_gr(tmp_path / "w.gr", n, s, d, w)
    off, out, idx = og.edge_list_to_csr(n, s, d)
    assert (tmp_path / "w.gr").read_bytes() == og.gr_bytes(n, off, out, w[idx])
    rn, rs, rd, rv = gr.read_graph(tmp_path / "w.gr", True)
    assert rn == n and np.array_equal(rs, s[idx]) and np.array_equal(rd, d[idx])
    assert np.array_equal(rv, w[idx])
    gr.write_mtx(tmp_path / "w.mtx", n, rs, rd, rv)
    assert (tmp_path / "w.mtx").read_text() == og.mtx_text(n, rs.tolist(), rd.tolist(),
                                                           rv.tolist())
    # read back: self-loops of the edge list are dropped by the text reader
    keep = rs != rd
    bn, bs, bd, bv = gr.read_graph(tmp_path / "w.mtx", True)
    assert np.array_equal(bs, rs[keep]) and np.array_equal(bd, rd[keep])
    assert np.array_equal(bv, rv[keep])
    gr.write_gr(tmp_path / "nw.gr", n, s, d)  # no values: edge-data size 0
    assert struct.unpack_from("<Q", (tmp_path / "nw.gr").read_bytes(), 8)[0] == 0


def test_mtx2gr_and_gr2mtx_tools(tmp_path):
    src = _write(tmp_path / "in.mtx", MTX_GENERAL)
    r = subprocess.run([str(BIN / "mtx2gr"), str(src), str(tmp_path / "out.gr")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    n, s, d, v = gr.read_graph(tmp_path / "out.gr", True)
    on, os_, od, ov = og.load_graph(src, True)
    off, out, idx = og.edge_list_to_csr(on, os_, od)
    assert n == on and np.array_equal(d, out) and np.array_equal(v, ov[idx])
    r = subprocess.run([str(BIN / "gr2mtx"), str(tmp_path / "out.gr"), str(tmp_path / "back.mtx")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "back.mtx").read_text() == og.mtx_text(n, s.tolist(), d.tolist(), v.tolist())
    r = subprocess.run([str(BIN / "mtx2gr"), str(tmp_path / "nope.mtx"), str(tmp_path / "o.gr")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "nope.mtx" in r.stderr
