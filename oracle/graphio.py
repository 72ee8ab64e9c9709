"""CPU oracle of the graph-file readers -- TEST INFRASTRUCTURE ONLY.

A pure-Python restatement of the reference's loadGraph family
(graphio.cpp:69-473) and of edgeListToCSR (util.cuh:79-108), used by the
tests as the checker of the native readers in gr_graphio.cpp.  Only
``tests/`` may use it; the product never imports it.

It follows the reference line by line where the reference is defined:
  - text files: lines starting with the comment character or holding only
    whitespace are skipped (graphio.cpp:101-106); MatrixMarket skips its first
    data line, the size line (:108-114); "%d%d" ids, decremented for
    MatrixMarket (:116-129); self-loops dropped, and with values requested a
    self-loop line also leaves the vertex count alone (:131-137, 160-162);
    symmetric / skew-symmetric / hermitian add the reverse arc, values copied,
    negated, or (hermitian) not supported (:139-158); values are "%f" floats
    truncated to int (:141-146); nVertices = 1 + the largest id (:164-165);
  - MatrixMarket header tokens (:207-250);
  - .gr: u64 version, edge-data size, nVertices, nEdges; u64 end offsets;
    u32 destinations padded to 8 bytes; u32 edge data (:265-362), expanded
    to an edge list in CSR order;
  - extension detection with a trailing ".gz" skipped and prefix matching
    (:441-472).
Where the reference is undefined (missing values, out-of-range ids, a
hermitian file with values, malformed offsets) or refuses a valid .gr file
without edge data when no values are wanted (:292-301), it raises ValueError
or reads the file, as the native reader does.
"""
from __future__ import annotations

import gzip
import re
import struct

import numpy as np

_INT = re.compile(r"\s*([+-]?\d+)")
_FLOAT = re.compile(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))",
                    re.IGNORECASE)


def _open(path):
    with open(path, "rb") as f:
        magic = f.read(2)
    if magic == b"\x1f\x8b":
        return gzip.open(path, "rb")
    return open(path, "rb")


def _format(path: str) -> str:
    p = str(path)
    dot = p.rfind(".")
    if dot >= 0 and p[dot:] == ".gz":
        dot = p.rfind(".", 0, dot)
    if dot < 0:
        raise ValueError("no filetype extension")
    ext = p[dot:]
    for name, n in ((".edge", 5), (".mtx", 4), (".gr", 3)):
        if ext[:n] == name:
            return name
    raise ValueError(f"unrecognized filetype extension {ext}")


def _to_int_value(text: str, negate: bool) -> int:
    m = _FLOAT.match(text)
    if not m:
        raise ValueError("missing edge value")
    f = np.float32(float(m.group(1)))
    if not np.isfinite(f):
        raise ValueError("invalid edge value")
    if negate:
        f = -f
    if not (-2147483649.0 < float(f) < 2147483648.0):
        raise ValueError("edge value out of int range")
    return int(np.trunc(float(f)))


def _load_text(lines, comment: bytes, size_line: bool, one_based: bool, sym: str,
               want_values: bool):
    srcs, dsts, vals = [], [], []
    max_id = 0
    for raw in lines:
        line = raw.rstrip(b"\n")
        if not line or line[:1] == comment or not line.strip():
            continue
        if size_line:
            size_line = False
            continue
        text = line.decode("latin-1")
        m1 = _INT.match(text)
        m2 = _INT.match(text, m1.end()) if m1 else None
        if not m2:
            raise ValueError("error parsing src/dst")
        s, d = int(m1.group(1)), int(m2.group(1))
        if one_based:
            s, d = s - 1, d - 1
        if s < 0 or d < 0 or s >= 2**31 - 1 or d >= 2**31 - 1:
            raise ValueError("vertex id out of range")
        if s == d:
            if want_values:
                continue
        else:
            srcs.append(s)
            dsts.append(d)
            if sym != "general":
                srcs.append(d)
                dsts.append(s)
            if want_values:
                rest = text[m2.end():]
                v = _to_int_value(rest, False)
                vals.append(v)
                if sym == "symmetric":
                    vals.append(v)
                elif sym == "skew-symmetric":
                    vals.append(_to_int_value(rest, True))
        max_id = max(max_id, s, d)
    return (max_id + 1, np.array(srcs, np.int32), np.array(dsts, np.int32),
            np.array(vals, np.int32) if want_values else None)


def _load_mtx(f, want_values: bool):
    header = f.readline().decode("latin-1")
    tok = header.split()
    if len(tok) < 5:
        raise ValueError("truncated MatrixMarket header")
    if tok[1] != "matrix":
        raise ValueError(f"unrecognized token {tok[1]} in header")
    if tok[2] != "coordinate":
        raise ValueError("only coordinate format is supported")
    if tok[3] == "pattern":
        want_values = False
    elif tok[3] == "complex":
        raise ValueError("complex edge values not supported")
    sym = tok[4]
    if sym not in ("general", "symmetric", "skew-symmetric", "hermitian"):
        raise ValueError(f"unrecognized symmetry type '{sym}'")
    if sym == "hermitian" and want_values:
        raise ValueError("hermitian edge values not supported")
    return _load_text(f, b"%", True, True, sym, want_values)


def _load_gr(f, want_values: bool):
    data = f.read()
    if len(data) < 32:
        raise ValueError("error reading header")
    _version, esz, nv, ne = struct.unpack_from("<4Q", data, 0)
    if esz == 0 and want_values:
        raise ValueError("file does not have edge values")
    if esz not in (0, 4):
        raise ValueError(f"file edge data is {esz} bytes wide")
    if nv >= 2**31 - 1 or ne > 2**31 - 1:
        raise ValueError("sizes exceed int32")
    pos = 32
    need = pos + 8 * nv + 4 * ne
    if len(data) < need:
        raise ValueError("truncated lonestar binary CSR file")
    ends = np.frombuffer(data, "<u8", nv, pos).astype(np.int64)
    pos += 8 * nv
    dsts = np.frombuffer(data, "<u4", ne, pos).astype(np.int64)
    pos += 4 * ne
    if nv and (np.any(np.diff(ends) < 0) or ends[0] < 0 or ends[-1] != ne):
        raise ValueError("offsets out of order")
    if not nv and ne:
        raise ValueError("offsets out of order")
    if ne and dsts.max() >= nv:
        raise ValueError("destination out of range")
    vals = None
    if want_values:
        pos += 4 * (ne % 2)
        if len(data) < pos + 4 * ne:
            raise ValueError("truncated edge data")
        vals = np.frombuffer(data, "<i4", ne, pos).astype(np.int32)
    starts = np.concatenate([[0], ends[:-1]]) if nv else np.zeros(0, np.int64)
    srcs = np.repeat(np.arange(nv, dtype=np.int32), (ends - starts).astype(np.int64))
    return nv, srcs.astype(np.int32), dsts.astype(np.int32), vals


def load_graph(path, want_values: bool = False):
    """(n_vertices, srcs, dsts, values or None) as loadGraph reads `path`."""
    fmt = _format(path)
    with _open(path) as f:
        if fmt == ".edge":
            return _load_text(f, b"#", False, False, "general", False)
        if fmt == ".mtx":
            return _load_mtx(f, want_values)
        return _load_gr(f, want_values)


def edge_list_to_csr(n_vertices: int, srcs, dsts):
    """util.cuh:79-108 with a stable order inside each source: (offsets,
    csr dsts, sort indices)."""
    srcs = np.asarray(srcs, np.int64)
    idx = np.argsort(srcs, kind="stable")
    off = np.zeros(n_vertices + 1, np.int64)
    np.add.at(off, srcs + 1, 1)
    off = np.cumsum(off)
    return off.astype(np.int32), np.asarray(dsts, np.int32)[idx], idx.astype(np.int32)


def gr_bytes(n_vertices: int, offsets, dsts, values=None) -> bytes:
    """File image of writeGraph_binaryCSR (graphio.cpp:365-413)."""
    ne = len(dsts)
    out = struct.pack("<4Q", 1, 4 if values is not None else 0, n_vertices, ne)
    out += np.asarray(offsets[1:], "<u8").tobytes()
    out += np.asarray(dsts, "<u4").tobytes()
    if ne % 2:
        out += b"\0\0\0\0"
    if values is not None:
        out += np.asarray(values, "<i4").tobytes()
    return out


def mtx_text(n_vertices: int, srcs, dsts, values=None) -> str:
    """File image of writeGraph_mtx (graphio.cpp:416-438)."""
    lines = ["%%MatrixMarket matrix coordinate Integer general",
             f"{n_vertices} {n_vertices} {len(srcs)}"]
    for i in range(len(srcs)):
        if values is None:
            lines.append(f"{srcs[i] + 1} {dsts[i] + 1}")
        else:
            lines.append(f"{srcs[i] + 1} {dsts[i] + 1} {values[i]}")
    return "\n".join(lines) + "\n"
