"""graphreduce_amd: Gather-Apply-Scatter graph engine for one AMD MI355X.

Python mirror of the reference's engine interface (``GASEngineGPU<Program>``,
gpugas.h:1025-1809) over the C ABI of ``lib/libgraphreduce_amd.so``.  The
engine methods keep the reference's names and argument meaning
(``setGraph``, ``setActive``, ``countActive``, ``gather``, ``apply``,
``scatterActivate``, ``nextIter``, ``getResults``, ``run``).  Errors raise
:class:`GRError` (the reference aborts, gpugas.h:158-165).

There is no CPU fallback: every call goes to the HIP library, and importing
the package fails loudly when the library has not been built.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

__all__ = [
    "GRError", "lib", "Graph", "GASEngine", "PROGRAMS", "BFS", "SSSP", "PAGERANK", "CC",
    "rmat_edges", "grid_edges", "hash_weights", "device_count", "device_info",
    "run_bfs", "run_sssp", "run_cc", "run_pagerank", "SSSP_INF", "pagerank_precise",
    "msbfs", "mssp", "Batch", "upload_i32", "Server", "QUERY_BFS", "QUERY_SSSP", "QUERY_PPR",
    "SSSP_CONGESTION", "read_graph", "write_gr", "write_mtx", "edge_list_to_csr",
]

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "lib" / "libgraphreduce_amd.so"

# Program ids (include/graphreduce_amd/gr_abi.h)
BFS, SSSP, PAGERANK, CC = 1, 2, 3, 4
# SSSP on congesting links: a program whose scatter changes edge data
SSSP_CONGESTION = 5

# Vertex/edge data layouts of the reference programs (bfs.cu:33-38,
# sssp.cu:28-29, pagerank.cu:32-39, connected_component.cu:28-29).
PROGRAMS = {
    BFS: dict(name="BFS", vertex=np.dtype([("depth", "<i4")]), edge=None),
    SSSP: dict(name="SSSP", vertex=np.dtype("<i4"), edge=np.dtype("<i4")),
    PAGERANK: dict(name="PageRank",
                   vertex=np.dtype([("rank", "<f4"), ("numOutEdges", "<i4")]), edge=None),
    CC: dict(name="CC", vertex=np.dtype("<i4"), edge=None),
    SSSP_CONGESTION: dict(name="SSSPCongestion", vertex=np.dtype("<i4"), edge=np.dtype("<i4")),
}

# SSSP::gatherZero = INT_MAX - maxLength (sssp.cu:32-33): the "unreached" distance.
SSSP_INF = 2**31 - 1 - 100000

RMAT_ABCD = (0.57, 0.19, 0.19)


class GRError(RuntimeError):
    """A call into libgraphreduce_amd failed (carries the status code)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[gr error {code}] {msg}")
        self.code = code


class _Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("active_vertices", C.c_int64),
        ("gathered_edges", C.c_int64),
        ("changed_vertices", C.c_int64),
        ("pushed_edges", C.c_int64),
        ("pull_iterations", C.c_int64),
        ("gather_launches", C.c_int64),
        ("gather_ms", C.c_double),
        ("scatter_launches", C.c_int64),
        ("scatter_ms", C.c_double),
        ("apply_launches", C.c_int64),
        ("apply_ms", C.c_double),
        ("next_launches", C.c_int64),
        ("next_ms", C.c_double),
        ("delta_gathers", C.c_int64),
        ("push_gathers", C.c_int64),
        ("push_edges", C.c_int64),
        ("push_ms", C.c_double),
        ("tile_gather_rows", C.c_int64),
        ("tile_gather_edges", C.c_int64),
        ("inert_gathers", C.c_int64),
        ("inert_ms", C.c_double),
        ("streamed_bytes", C.c_int64),
        ("streamed_shards", C.c_int64),
        ("fused_gathers", C.c_int64),
        ("fused_edges", C.c_int64),
        ("blocked_gathers", C.c_int64),
        ("fused_activations", C.c_int64),
        ("idle_gather_launches", C.c_int64),
    ]


class _PRResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("residual", C.c_double),
        ("ms", C.c_double),
        ("gather_ms", C.c_double),
        ("gathered_edges", C.c_int64),
    ]


class _BatchResult(C.Structure):
    _fields_ = [("levels", C.c_int32), ("ms", C.c_double), ("grid", C.c_int32)]


class _ServerConfig(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_size_t), ("max_lanes_per_kind", C.c_int),
                ("ppr_iterations", C.c_int)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("steps", C.c_int64),
                ("max_lanes_in_use", C.c_int32), ("peak_bytes", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("throughput_rps", C.c_double)]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every entry point of gr_abi.h
_SIGS = {
    "gr_last_error": (C.c_char_p, []),
    "gr_version": (C.c_char_p, []),
    "gr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gr_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "gr_device_synchronize": (C.c_int, []),
    "gr_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(_P)]),
    "gr_device_free": (C.c_int, [_P]),
    "gr_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
    "gr_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
    "gr_device_copy": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_float)]),
    "gr_graph_source_read_probe": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "gr_host_rmat_edges": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_uint64, C.c_int, _I32P, _I32P, C.c_int64, _I64P]),
    "gr_host_grid_edges": (C.c_int, [C.c_int, C.c_double, C.c_uint64, _I32P, _I32P,
                                      C.c_int64, _I64P]),
    "gr_host_hash_weights": (C.c_int, [C.c_uint64, C.c_int64, C.c_int32, C.c_int, _I32P]),
    "gr_graph_from_edges": (C.c_int, [C.c_int32, C.c_int32, _I32P, _I32P, C.POINTER(_P)]),
    "gr_graph_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_uint64, C.c_int, C.POINTER(_P)]),
    "gr_graph_grid": (C.c_int, [C.c_int, C.c_double, C.c_uint64, C.POINTER(_P)]),
    "gr_graph_destroy": (C.c_int, [_P]),
    "gr_graph_info": (C.c_int, [_P, _I32P, _I64P]),
    "gr_graph_copy_csr": (C.c_int, [_P, _I32P, _I32P]),
    "gr_graph_copy_csc": (C.c_int, [_P, _I32P, _I32P]),
    "gr_graph_copy_edges": (C.c_int, [_P, _I32P, _I32P]),
    "gr_graph_copy_perm": (C.c_int, [_P, _I32P, _I32P]),
    "gr_graph_max_out_degree_vertex": (C.c_int, [_P, _I32P, _I32P]),
    "gr_graph_hash_weights_device": (C.c_int, [_P, C.c_uint64, C.c_int32, C.c_int,
                                                C.POINTER(_P)]),
    "gr_engine_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    # for C++ callers of gas_program.hpp (an engine object from the headers)
    "gr_engine_adopt": (C.c_int, [_P, C.POINTER(_P)]),
    "gr_engine_destroy": (C.c_int, [_P]),
    "gr_engine_set_graph": (C.c_int, [_P, C.c_int32, _P, C.c_size_t, C.c_int32, _P, C.c_size_t,
                                       _I32P, _I32P]),
    "gr_engine_bind_graph": (C.c_int, [_P, _P, _P, C.c_size_t, _P, C.c_size_t, C.c_int]),
    "gr_engine_set_active": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "gr_engine_count_active": (C.c_int, [_P, _I32P]),
    "gr_engine_gather": (C.c_int, [_P, C.c_int]),
    "gr_engine_apply": (C.c_int, [_P]),
    "gr_engine_scatter_activate": (C.c_int, [_P, C.c_int]),
    "gr_engine_next_iter": (C.c_int, [_P, _I32P]),
    "gr_engine_get_results": (C.c_int, [_P]),
    "gr_engine_run": (C.c_int, [_P, _I32P]),
    "gr_engine_set_iteration_count": (C.c_int, [_P, C.c_int32]),
    "gr_engine_get_iteration_count": (C.c_int, [_P, _I32P]),
    "gr_engine_read_vertex_data": (C.c_int, [_P, _P, C.c_size_t]),
    "gr_engine_read_edge_data": (C.c_int, [_P, _P, C.c_size_t]),
    "gr_program_has_scatter": (C.c_int, [C.c_int]),
    "gr_engine_reset_vertex_data": (C.c_int, [_P]),
    "gr_engine_set_direction": (C.c_int, [_P, C.c_int, C.c_double]),
    "gr_engine_set_gather_tile": (C.c_int, [_P, C.c_int]),
    "gr_engine_set_delta_gather": (C.c_int, [_P, C.c_int]),
    "gr_engine_set_vertex_order": (C.c_int, [_P, C.c_int]),
    "gr_engine_set_streaming": (C.c_int, [_P, C.c_int64]),
    "gr_engine_set_fuse_share": (C.c_int, [_P, C.c_double]),
    "gr_engine_set_blocked_gather": (C.c_int, [_P, C.c_double]),
    "gr_engine_set_push_share": (C.c_int, [_P, C.c_double]),
    "gr_engine_enable_timing": (C.c_int, [_P, C.c_int]),
    "gr_engine_get_stats": (C.c_int, [_P, C.POINTER(_Stats)]),
    "gr_engine_reset_stats": (C.c_int, [_P]),
    "gr_msbfs": (C.c_int, [_P, _I32P, C.c_int, _P, C.c_int, C.c_double,
                            C.POINTER(_BatchResult)]),
    "gr_mssp": (C.c_int, [_P, _P, _I32P, C.c_int, _P, C.POINTER(_BatchResult)]),
    "gr_batch_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "gr_batch_bfs": (C.c_int, [_P, _I32P, C.c_int, _P, C.c_int, C.c_double,
                               C.POINTER(_BatchResult)]),
    "gr_batch_sssp": (C.c_int, [_P, _I32P, C.c_int, _P, C.POINTER(_BatchResult)]),
    "gr_batch_destroy": (C.c_int, [_P]),
    "gr_server_create": (C.c_int, [_P, _P, C.POINTER(_ServerConfig), C.POINTER(_P)]),
    "gr_server_destroy": (C.c_int, [_P]),
    "gr_server_submit": (C.c_int, [_P, C.c_int, C.c_int32, C.c_int, _I64P]),
    "gr_server_wait": (C.c_int, [_P, C.c_int64, _P, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_double), _I32P]),
    "gr_server_get_stats": (C.c_int, [_P, C.POINTER(_ServerStats)]),
    "gr_server_load_test": (C.c_int, [_P, C.c_double, C.c_int, _I32P, _I32P, C.c_int, _I64P]),
    "gr_pagerank_precise": (C.c_int, [_P, C.c_int, C.c_double, C.c_int,
                                       C.POINTER(C.c_float), C.POINTER(C.c_double),
                                       C.POINTER(_PRResult)]),
    "gr_pagerank_precise_noalloc": (C.c_int, [_P, C.c_int, C.c_double, C.c_int,
                                               C.POINTER(C.c_float), C.POINTER(C.c_double),
                                               C.POINTER(_PRResult)]),
    # gr_graphio.h
    "gr_graphfile_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P)]),
    "gr_graphfile_n_vertices": (C.c_int32, [_P]),
    "gr_graphfile_n_edges": (C.c_int64, [_P]),
    "gr_graphfile_srcs": (_I32P, [_P]),
    "gr_graphfile_dsts": (_I32P, [_P]),
    "gr_graphfile_values": (_I32P, [_P]),
    "gr_graphfile_free": (None, [_P]),
    "gr_write_graph_gr": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _I32P, _I32P, _I32P]),
    "gr_write_graph_mtx": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _I32P, _I32P, _I32P]),
    "gr_edge_list_to_csr": (C.c_int, [C.c_int32, C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P]),
}

_lib = None

# Interpreter shutdown: once the atexit hook below has run, finalizers make no
# library call.  Objects still alive then are collected during module teardown
# or not at all, possibly after the HIP runtime (a C-level exit handler) has
# been torn down, where a hipFree or stream call can fault; the process exit
# returns their memory anyway.  Live servers are stopped by the hook itself,
# while the runtime is still up, so that no worker thread outlives it.
_SHUTDOWN = False
_LIVE_SERVERS = weakref.WeakSet()


def _at_exit() -> None:
    global _SHUTDOWN
    for srv in list(_LIVE_SERVERS):
        try:
            srv.close()
        except Exception:
            pass
    _SHUTDOWN = True


atexit.register(_at_exit)


def lib() -> C.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = os.environ.get("GRAPHREDUCE_AMD_LIB", str(LIB_PATH))
        if not Path(path).exists():
            raise ImportError(
                f"graphreduce_amd native library not found at {path}; "
                "run `python -c 'import __graft_entry__; __graft_entry__.build()'`")
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise GRError(rc, lib().gr_last_error().decode(errors="replace"))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().gr_device_count(C.byref(n)))
    return n.value


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    mem = C.c_size_t(0)
    _check(lib().gr_device_info(name, 256, C.byref(cus), C.byref(mem)))
    return dict(name=name.value.decode(), compute_units=cus.value, total_bytes=mem.value)


def synchronize() -> None:
    _check(lib().gr_device_synchronize())


def device_copy_gbs(nbytes: int = 1 << 30, repeats: int = 5) -> float:
    """Best read+write bandwidth (GB/s) of a plain device copy of nbytes."""
    src, dst = C.c_void_p(), C.c_void_p()
    _check(lib().gr_device_alloc(nbytes, C.byref(src)))
    _check(lib().gr_device_alloc(nbytes, C.byref(dst)))
    try:
        best = 0.0
        for _ in range(repeats):
            ms = C.c_float(0)
            _check(lib().gr_device_copy(dst, src, nbytes, C.byref(ms)))
            best = max(best, 2 * nbytes / (ms.value * 1e-3) / 1e9)
        return best
    finally:
        lib().gr_device_free(src)
        lib().gr_device_free(dst)


# ------------------------------------------------------------------ inputs

def rmat_edges(scale: int, edge_factor: int = 16, seed: int = 1, abc=RMAT_ABCD,
               symmetrize: bool = False):
    """Host RMAT edge list (self-loops and duplicates removed, sorted by (src, dst))."""
    a, b, c = abc
    m = (edge_factor << scale) * (2 if symmetrize else 1)
    srcs = np.empty(m, np.int32)
    dsts = np.empty(m, np.int32)
    n = C.c_int64(0)
    _check(lib().gr_host_rmat_edges(scale, edge_factor, a, b, c, seed, int(symmetrize),
                                    _p32(srcs), _p32(dsts), m, C.byref(n)))
    return srcs[: n.value].copy(), dsts[: n.value].copy()


def grid_edges(side: int, delete_frac: float = 0.01, seed: int = 5):
    """Host 4-neighbour grid arcs (both directions of every kept edge)."""
    n = C.c_int64(0)
    _check(lib().gr_host_grid_edges(side, delete_frac, seed, None, None, 0, C.byref(n)))
    srcs = np.empty(n.value, np.int32)
    dsts = np.empty(n.value, np.int32)
    _check(lib().gr_host_grid_edges(side, delete_frac, seed, _p32(srcs), _p32(dsts), n.value,
                                    C.byref(n)))
    return srcs, dsts


def hash_weights(n: int, seed: int, lo: int = 1, span_log2: int = 6) -> np.ndarray:
    """Uniform integer weights on [lo, lo + 2**span_log2) per edge-list position."""
    w = np.empty(n, np.int32)
    _check(lib().gr_host_hash_weights(seed, n, lo, span_log2, _p32(w)))
    return w


# ------------------------------------------------------------- graph files

def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def read_graph(path, values: bool = False):
    """loadGraph (graphio.cpp:441-473) for .edge, .mtx and .gr files (each
    optionally gzip-compressed): returns (n_vertices, srcs, dsts, vals), vals
    None unless `values` is set and the file has edge values."""
    h = _P()
    _check(lib().gr_graphfile_load(_path(path), int(values), C.byref(h)))
    try:
        L = lib()
        n = int(L.gr_graphfile_n_vertices(h))
        m = int(L.gr_graphfile_n_edges(h))

        def take(ptr):
            if not ptr or m == 0:
                return None if not ptr else np.empty(0, np.int32)
            return np.ctypeslib.as_array(ptr, shape=(m,)).copy()

        s = take(L.gr_graphfile_srcs(h))
        d = take(L.gr_graphfile_dsts(h))
        v = take(L.gr_graphfile_values(h))
        if s is None:
            s = np.empty(0, np.int32)
            d = np.empty(0, np.int32)
        return n, s, d, v
    finally:
        lib().gr_graphfile_free(h)


def edge_list_to_csr(n_vertices: int, srcs, dsts):
    """edgeListToCSR (util.cuh:79-108), stable within a source: returns
    (offsets, csr_dsts, sort_indices)."""
    s, d = _i32(srcs), _i32(dsts)
    off = np.empty(n_vertices + 1, np.int32)
    out = np.empty(len(s), np.int32)
    idx = np.empty(len(s), np.int32)
    _check(lib().gr_edge_list_to_csr(n_vertices, len(s), _p32(s), _p32(d), _p32(off), _p32(out),
                                     _p32(idx)))
    return off, out, idx


def write_gr(path, n_vertices: int, srcs, dsts, values=None) -> None:
    """Edge list -> Lonestar binary CSR file, as the mtx2gr tool does
    (mtx2gr.cpp:37-52, writeGraph_binaryCSR graphio.cpp:365-413)."""
    off, out, idx = edge_list_to_csr(n_vertices, srcs, dsts)
    vals = None if values is None else _i32(np.asarray(values)[idx])
    _check(lib().gr_write_graph_gr(_path(path), n_vertices, len(out), _p32(off), _p32(out),
                                   None if vals is None else _p32(vals)))


def write_mtx(path, n_vertices: int, srcs, dsts, values=None) -> None:
    """writeGraph_mtx (graphio.cpp:416-438): 1-based 'src dst [value]' lines."""
    s, d = _i32(srcs), _i32(dsts)
    vals = None if values is None else _i32(values)
    _check(lib().gr_write_graph_mtx(_path(path), n_vertices, len(s), _p32(s), _p32(d),
                                    None if vals is None else _p32(vals)))


# ------------------------------------------------------------------- graph

class Graph:
    """Device-resident CSR + CSC graph shared by engines."""

    def __init__(self, handle):
        self._h = handle
        nv = C.c_int32(0)
        ne = C.c_int64(0)
        _check(lib().gr_graph_info(self._h, C.byref(nv), C.byref(ne)))
        self.n_vertices = nv.value
        self.n_edges = ne.value

    @classmethod
    def from_edges(cls, n_vertices: int, srcs, dsts) -> "Graph":
        s, d = _i32(srcs), _i32(dsts)
        h = _P()
        _check(lib().gr_graph_from_edges(n_vertices, len(s), _p32(s), _p32(d), C.byref(h)))
        return cls(h)

    @classmethod
    def from_file(cls, path):
        """Graph of a .edge / .mtx / .gr file (read_graph)."""
        n, s, d, _ = read_graph(path)
        return cls.from_edges(n, s, d)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, seed: int = 1, abc=RMAT_ABCD,
             symmetrize: bool = False) -> "Graph":
        a, b, c = abc
        h = _P()
        _check(lib().gr_graph_rmat(scale, edge_factor, a, b, c, seed, int(symmetrize), C.byref(h)))
        return cls(h)

    @classmethod
    def grid(cls, side: int, delete_frac: float = 0.01, seed: int = 5) -> "Graph":
        h = _P()
        _check(lib().gr_graph_grid(side, delete_frac, seed, C.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            lib().gr_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        if _SHUTDOWN:
            return
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def csr(self):
        off = np.empty(self.n_vertices + 1, np.int32)
        adj = np.empty(self.n_edges, np.int32)
        _check(lib().gr_graph_copy_csr(self._h, _p32(off), _p32(adj)))
        return off, adj

    def csc(self):
        off = np.empty(self.n_vertices + 1, np.int32)
        adj = np.empty(self.n_edges, np.int32)
        _check(lib().gr_graph_copy_csc(self._h, _p32(off), _p32(adj)))
        return off, adj

    def edges(self):
        s = np.empty(self.n_edges, np.int32)
        d = np.empty(self.n_edges, np.int32)
        _check(lib().gr_graph_copy_edges(self._h, _p32(s), _p32(d)))
        return s, d

    def perm(self):
        """(csr_eid, csc_eid): edge-list index of each CSR / CSC position."""
        a = np.empty(self.n_edges, np.int32)
        b = np.empty(self.n_edges, np.int32)
        _check(lib().gr_graph_copy_perm(self._h, _p32(a), _p32(b)))
        return a, b

    def source_read_ms(self, ordered: bool = True, repeats: int = 5) -> float:
        """Time (ms) of one bare pass of value[srcs[e]] reads over the CSC
        (the random reads a full gather cannot avoid)."""
        t = C.c_float(0)
        _check(lib().gr_graph_source_read_probe(self._h, int(ordered), repeats, C.byref(t)))
        return t.value

    def max_out_degree_vertex(self):
        v = C.c_int32(0)
        deg = C.c_int32(0)
        _check(lib().gr_graph_max_out_degree_vertex(self._h, C.byref(v), C.byref(deg)))
        return v.value, deg.value

    def hash_weights_device(self, seed: int, lo: int = 1, span_log2: int = 6) -> "DeviceArray":
        p = _P()
        _check(lib().gr_graph_hash_weights_device(self._h, seed, lo, span_log2, C.byref(p)))
        return DeviceArray(p, self.n_edges * 4)


class DeviceArray:
    """A raw device allocation owned by Python."""

    def __init__(self, ptr, nbytes: int):
        self.ptr = ptr
        self.nbytes = nbytes

    def free(self) -> None:
        if self.ptr:
            lib().gr_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        if _SHUTDOWN:
            return
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ engine

class GASEngine:
    """GASEngineGPU<Program> (gpugas.h:1025) for one of the compiled programs."""

    def __init__(self, program: int):
        if program not in PROGRAMS:
            raise GRError(5, f"unknown program id {program}")
        self.program = program
        self.vertex_dtype = PROGRAMS[program]["vertex"]
        self.edge_dtype = PROGRAMS[program]["edge"]
        h = _P()
        _check(lib().gr_engine_create(program, C.byref(h)))
        self._h = h
        self._vdata = None
        self._edata = None
        self._graph = None
        self._n_edges = 0
        self._scatter = lib().gr_program_has_scatter(program) == 1

    def close(self) -> None:
        if self._h:
            lib().gr_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        if _SHUTDOWN:
            return
        try:
            self.close()
        except Exception:
            pass

    def _vertex_array(self, vertex_data) -> np.ndarray:
        vd = np.ascontiguousarray(vertex_data)
        if vd.dtype != self.vertex_dtype:
            vd = np.ascontiguousarray(vd.astype(self.vertex_dtype))
        return vd

    def _edge_ptr(self, edge_data):
        if self.edge_dtype is None:
            return None, 0, None
        if edge_data is None:
            raise GRError(1, f"{PROGRAMS[self.program]['name']} needs edge data")
        ed = np.ascontiguousarray(edge_data, dtype=self.edge_dtype)
        return ed.ctypes.data_as(_P), ed.dtype.itemsize, ed

    # gpugas.h:1314-1319
    def setGraph(self, nVertices: int, vertexData, nEdges: int, edgeData, srcs, dsts) -> None:
        vd = self._vertex_array(vertexData)
        s, d = _i32(srcs), _i32(dsts)
        if len(s) != nEdges or len(d) != nEdges or len(vd) != nVertices:
            raise GRError(1, "setGraph: array lengths do not match nVertices/nEdges")
        ep, esz, keep = self._edge_ptr(edgeData)
        _check(lib().gr_engine_set_graph(self._h, nVertices, vd.ctypes.data_as(_P),
                                         vd.dtype.itemsize, nEdges, ep, esz, _p32(s), _p32(d)))
        self._vdata = vd
        # a program whose scatter changes edge data gets it written back here
        # by getResults (the reference engine works on the caller's array)
        self._edata = keep if self._scatter else None
        self._graph = None
        self._n_edges = nEdges

    def bindGraph(self, graph: Graph, vertexData, edgeData=None) -> None:
        """setGraph with a prebuilt device graph; edgeData may be a DeviceArray."""
        vd = self._vertex_array(vertexData)
        if len(vd) != graph.n_vertices:
            raise GRError(1, "bindGraph: vertex data length != graph vertices")
        _keep = None
        if isinstance(edgeData, DeviceArray):
            ep, esz, on_dev = edgeData.ptr, 4, 1
        else:
            ep, esz, _keep = self._edge_ptr(edgeData)
            on_dev = 0
        _check(lib().gr_engine_bind_graph(self._h, graph.handle, vd.ctypes.data_as(_P),
                                          vd.dtype.itemsize, ep, esz, on_dev))
        self._vdata = vd
        self._edata = _keep if (self._scatter and not on_dev) else None
        self._graph = graph
        self._n_edges = graph.n_edges

    @property
    def vertexData(self) -> np.ndarray:
        """The host array bound by setGraph (updated by getResults)."""
        return self._vdata

    def setActive(self, start: int, end: int) -> None:
        _check(lib().gr_engine_set_active(self._h, start, end))

    def countActive(self) -> int:
        n = C.c_int32(0)
        _check(lib().gr_engine_count_active(self._h, C.byref(n)))
        return n.value

    def gather(self, haveGather: bool = True) -> None:
        _check(lib().gr_engine_gather(self._h, int(haveGather)))

    def apply(self) -> None:
        _check(lib().gr_engine_apply(self._h))

    def scatterActivate(self, haveScatter: bool = True) -> None:
        _check(lib().gr_engine_scatter_activate(self._h, int(haveScatter)))

    def nextIter(self) -> int:
        n = C.c_int32(0)
        _check(lib().gr_engine_next_iter(self._h, C.byref(n)))
        return n.value

    @property
    def edgeData(self) -> np.ndarray:
        """The host edge array given to setGraph (programs with a scatter:
        updated by getResults); None otherwise."""
        return self._edata

    def getResults(self) -> None:
        _check(lib().gr_engine_get_results(self._h))
        if self._edata is not None and len(self._edata):
            _check(lib().gr_engine_read_edge_data(self._h, self._edata.ctypes.data_as(_P),
                                                  self._edata.dtype.itemsize))

    def readEdgeData(self) -> np.ndarray:
        """The engine's edge data in edge-list order (a copy)."""
        if self.edge_dtype is None:
            raise GRError(1, f"{PROGRAMS[self.program]['name']} has no edge data")
        if self._vdata is None:
            raise GRError(4, "readEdgeData: no graph bound")
        m = self._n_edges
        out = np.empty(m, self.edge_dtype)
        if m:
            _check(lib().gr_engine_read_edge_data(self._h, out.ctypes.data_as(_P),
                                                  out.dtype.itemsize))
        return out

    def run(self) -> int:
        it = C.c_int32(0)
        _check(lib().gr_engine_run(self._h, C.byref(it)))
        return it.value

    def setIterationCount(self, iteration: int) -> None:
        _check(lib().gr_engine_set_iteration_count(self._h, iteration))

    def readVertexData(self) -> np.ndarray:
        out = np.empty_like(self._vdata)
        _check(lib().gr_engine_read_vertex_data(self._h, out.ctypes.data_as(_P),
                                                out.dtype.itemsize))
        return out

    def resetVertexData(self) -> None:
        """Device-side restore of the vertex data given at setGraph/bindGraph."""
        _check(lib().gr_engine_reset_vertex_data(self._h))

    def setDirection(self, mode: int = 0, alpha: float = 15.0) -> None:
        """0 = automatic push/pull, 1 = push only, 2 = pull only."""
        _check(lib().gr_engine_set_direction(self._h, mode, alpha))

    def setPushShare(self, share: float = 0.6) -> None:
        """Adaptive delta gather: push below this share of changed in-edges."""
        _check(lib().gr_engine_set_push_share(self._h, float(share)))

    def setBlockedGather(self, min_share: float = 0.5) -> None:
        """Gathers over at least min_share * E in-edges run on the graph's
        source-listed row blocks (SSSP, CC); <= 0 turns that off."""
        _check(lib().gr_engine_set_blocked_gather(self._h, float(min_share)))

    def setFuseShare(self, share: float) -> None:
        """SSSP/CC: fuse activation with the next push gather when the changed
        out-edges are at most share * E (0: off; same results)."""
        _check(lib().gr_engine_set_fuse_share(self._h, float(share)))

    def setStreaming(self, max_edges_per_shard: int) -> None:
        """Out-of-core mode for the next setGraph: the edge arrays stay in
        pinned host memory in shards of at most this many edges and are
        streamed per iteration (0: the graph lives in device memory)."""
        _check(lib().gr_engine_set_streaming(self._h, int(max_edges_per_shard)))

    def setVertexOrder(self, order: int = 2) -> None:
        """1: run on the degree-ordered copy of the graph; 0: as given;
        2 (default): the copy when some vertex's in+out degree is at least 32
        times the mean.  Results are in the caller's ids either way; takes
        effect at the next setGraph/bindGraph."""
        _check(lib().gr_engine_set_vertex_order(self._h, order))

    def setDeltaGather(self, mode=1) -> None:
        """SSSP/CC: after an all-active iteration, gather only from sources
        changed in the previous iteration (same results).  0/False: off;
        1/True (default): push under 60 % of the in-edges from changed
        sources, else the full gather; 2: pull-delta whenever valid;
        3: push whenever valid."""
        _check(lib().gr_engine_set_delta_gather(self._h, int(mode)))

    def setGatherTile(self, items_per_thread: int) -> None:
        """Merge-path items per thread of the gather tiles (4, 8 or 16)."""
        _check(lib().gr_engine_set_gather_tile(self._h, items_per_thread))

    def enableTiming(self, on: bool = True) -> None:
        _check(lib().gr_engine_enable_timing(self._h, int(on)))

    def stats(self) -> dict:
        s = _Stats()
        _check(lib().gr_engine_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _Stats._fields_}

    def resetStats(self) -> None:
        _check(lib().gr_engine_reset_stats(self._h))


# ----------------------------------------------------- reference app loops

def _graph_args(graph_or_edges):
    if isinstance(graph_or_edges, Graph):
        return graph_or_edges, None
    n, s, d = graph_or_edges
    return None, (n, _i32(s), _i32(d))


def _bind(eng: GASEngine, graph_or_edges, vd, ed=None):
    g, el = _graph_args(graph_or_edges)
    if g is not None:
        eng.bindGraph(g, vd, ed)
    else:
        n, s, d = el
        eng.setGraph(n, vd, len(s), ed, s, d)


def run_bfs(graph_or_edges, source: int, direction: int = 0):
    """bfs.cu:93-134 loop; returns (depth int32[V], loop trips)."""
    n = graph_or_edges.n_vertices if isinstance(graph_or_edges, Graph) else graph_or_edges[0]
    vd = np.zeros(n, PROGRAMS[BFS]["vertex"])
    vd["depth"] = -1
    eng = GASEngine(BFS)
    eng.setDirection(direction)
    _bind(eng, graph_or_edges, vd)
    eng.setActive(source, source + 1)
    iteration = 0
    eng.setIterationCount(iteration)
    while eng.countActive():
        eng.gather(False)
        eng.apply()
        eng.scatterActivate(False)
        eng.nextIter()
        iteration += 1
        eng.setIterationCount(iteration)
    eng.getResults()
    out = eng.vertexData["depth"].copy()
    eng.close()
    return out, iteration


def run_sssp(graph_or_edges, weights, source: int, direction: int = 0, delta: int = 1):
    """sssp.cu:69-111 loop (all vertices active first); returns (dist, trips)."""
    n = graph_or_edges.n_vertices if isinstance(graph_or_edges, Graph) else graph_or_edges[0]
    vd = np.full(n, SSSP_INF, np.int32)
    vd[source] = 0
    eng = GASEngine(SSSP)
    eng.setDirection(direction)
    eng.setDeltaGather(delta)
    _bind(eng, graph_or_edges, vd, weights)
    eng.setActive(0, n)
    it = 0
    while eng.countActive():
        eng.gather()
        eng.apply()
        eng.scatterActivate()
        eng.nextIter()
        it += 1
    eng.getResults()
    out = eng.vertexData.copy()
    eng.close()
    return out, it


def run_cc(graph_or_edges, direction: int = 0, delta: int = 1):
    """connected_component.cu:66-81 (label = own id, all active, run())."""
    n = graph_or_edges.n_vertices if isinstance(graph_or_edges, Graph) else graph_or_edges[0]
    vd = np.arange(n, dtype=np.int32)
    eng = GASEngine(CC)
    eng.setDirection(direction)
    eng.setDeltaGather(delta)
    _bind(eng, graph_or_edges, vd)
    eng.setActive(0, n)
    it = eng.run()
    eng.getResults()
    out = eng.vertexData.copy()
    eng.close()
    return out, it


def run_pagerank(graph_or_edges, out_degree, direction: int = 0):
    """pagerank.cu:85-101 (rank = 0.15, all active, run()); returns (rank f32, trips)."""
    n = graph_or_edges.n_vertices if isinstance(graph_or_edges, Graph) else graph_or_edges[0]
    vd = np.zeros(n, PROGRAMS[PAGERANK]["vertex"])
    vd["rank"] = np.float32(0.15)
    vd["numOutEdges"] = out_degree
    eng = GASEngine(PAGERANK)
    eng.setDirection(direction)
    _bind(eng, graph_or_edges, vd)
    eng.setActive(0, n)
    it = eng.run()
    eng.getResults()
    out = eng.vertexData["rank"].copy()
    eng.close()
    return out, it


def pagerank_precise(graph: Graph, max_iters: int = 50, tol: float = 1e-7,
                     compensated: bool = True, noalloc: bool = False):
    """All-active PageRank (FP32, compensated sums optional, FP64 L1 residual).

    noalloc=True calls gr_pagerank_precise_noalloc (no device allocation; its
    buffers live in a static 80 MiB device array per GPU).
    Returns (ranks float32[V], info) with info = iterations, residual (last
    step), residuals (per step), ms, gather_ms, gathered_edges."""
    ranks = np.empty(graph.n_vertices, np.float32)
    resid = np.zeros(max(1, max_iters), np.float64)
    r = _PRResult()
    fn = lib().gr_pagerank_precise_noalloc if noalloc else lib().gr_pagerank_precise
    _check(fn(graph.handle, max_iters, tol, int(compensated),
              ranks.ctypes.data_as(C.POINTER(C.c_float)),
              resid.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r)))
    info = {name: getattr(r, name) for name, _ in _PRResult._fields_}
    info["residuals"] = resid[: r.iterations].tolist()
    return ranks, info


def _batch_sources(sources):
    src = _i32(sources)
    if not 1 <= len(src) <= 64:
        raise GRError(1, "batched runs take 1..64 sources")
    return src


MSBFS_STREAM_PULL, MSBFS_PERSISTENT, MSBFS_FILTER = 4, 8, 16


def msbfs(graph: Graph, sources, direction: int = 0, alpha: float = 20.0,
          stream_pull: bool = False, persistent: bool = False, pull_filter: bool = False,
          out: "DeviceArray | None" = None):
    """BFS from up to 64 sources in one batch (one expand + one update launch
    per level; persistent=True: every level in one cooperative launch).

    stream_pull asks for the streaming pull kernel, which the library
    currently refuses (GRError).  Returns (depth int32[len(sources), V], info);
    with `out` (a device array of len(sources) * V int32) the depths stay
    there and None is returned in their place."""
    if stream_pull:
        direction |= MSBFS_STREAM_PULL
    if persistent:
        direction |= MSBFS_PERSISTENT
    if pull_filter:
        direction |= MSBFS_FILTER
    src = _batch_sources(sources)
    n = graph.n_vertices
    nbytes = 4 * n * len(src)
    if out is not None:
        if out.nbytes < nbytes:
            raise GRError(1, "msbfs: output array too small")
        r = _BatchResult()
        _check(lib().gr_msbfs(graph.handle, _p32(src), len(src), out.ptr, direction, alpha,
                              C.byref(r)))
        return None, {name: getattr(r, name) for name, _ in _BatchResult._fields_}
    d = C.c_void_p()
    _check(lib().gr_device_alloc(nbytes, C.byref(d)))
    try:
        r = _BatchResult()
        _check(lib().gr_msbfs(graph.handle, _p32(src), len(src), d, direction, alpha, C.byref(r)))
        out = np.empty((len(src), n), np.int32)
        _check(lib().gr_memcpy_d2h(out.ctypes.data, d, nbytes))
    finally:
        lib().gr_device_free(d)
    return out, {name: getattr(r, name) for name, _ in _BatchResult._fields_}


def mssp(graph: Graph, weights: "DeviceArray", sources, out: "DeviceArray | None" = None):
    """SSSP from up to 64 sources (64 lanes per wave, hipGraph-replayed rounds).

    weights: DeviceArray of int32 per edge in the graph's edge-list order.
    Returns (dist int32[len(sources), V], info); with `out` the distances stay
    in that device array and None is returned in their place."""
    src = _batch_sources(sources)
    n = graph.n_vertices
    nbytes = 4 * n * len(src)
    if out is not None:
        if out.nbytes < nbytes:
            raise GRError(1, "mssp: output array too small")
        r = _BatchResult()
        _check(lib().gr_mssp(graph.handle, weights.ptr, _p32(src), len(src), out.ptr,
                             C.byref(r)))
        return None, {name: getattr(r, name) for name, _ in _BatchResult._fields_}
    d = C.c_void_p()
    _check(lib().gr_device_alloc(nbytes, C.byref(d)))
    try:
        r = _BatchResult()
        _check(lib().gr_mssp(graph.handle, weights.ptr, _p32(src), len(src), d, C.byref(r)))
        out = np.empty((len(src), n), np.int32)
        _check(lib().gr_memcpy_d2h(out.ctypes.data, d, nbytes))
    finally:
        lib().gr_device_free(d)
    return out, {name: getattr(r, name) for name, _ in _BatchResult._fields_}


class Batch:
    """A batch context (gr_batch_create): the state of batched runs on one
    graph, allocated by the first run that needs it and reused by later runs,
    as an engine binds its graph once.  weights (a DeviceArray of int32 per
    edge in edge-list order) is needed for sssp().  The graph must stay open
    while the context is."""

    def __init__(self, graph: Graph, weights: "DeviceArray | None" = None):
        h = _P()
        _check(lib().gr_batch_create(graph.handle, weights.ptr if weights is not None else None,
                                     C.byref(h)))
        self._h = h
        self._graph = graph  # kept alive with the context
        self.n_vertices = graph.n_vertices

    def _run(self, fn, src, out, *args):
        n = self.n_vertices
        nbytes = 4 * n * len(src)
        r = _BatchResult()
        if out is not None:
            if out.nbytes < nbytes:
                raise GRError(1, "batch: output array too small")
            _check(fn(self._h, _p32(src), len(src), out.ptr, *args, C.byref(r)))
            return None, {name: getattr(r, name) for name, _ in _BatchResult._fields_}
        d = C.c_void_p()
        _check(lib().gr_device_alloc(nbytes, C.byref(d)))
        try:
            _check(fn(self._h, _p32(src), len(src), d, *args, C.byref(r)))
            res = np.empty((len(src), n), np.int32)
            _check(lib().gr_memcpy_d2h(res.ctypes.data, d, nbytes))
        finally:
            lib().gr_device_free(d)
        return res, {name: getattr(r, name) for name, _ in _BatchResult._fields_}

    def bfs(self, sources, direction: int = 0, alpha: float = 20.0, persistent: bool = False,
            pull_filter: bool = False, out: "DeviceArray | None" = None):
        """msbfs() on this context; same arguments and results."""
        if persistent:
            direction |= MSBFS_PERSISTENT
        if pull_filter:
            direction |= MSBFS_FILTER
        return self._run(lib().gr_batch_bfs, _batch_sources(sources), out, direction, alpha)

    def sssp(self, sources, out: "DeviceArray | None" = None):
        """mssp() on this context's weights; same results."""
        return self._run(lib().gr_batch_sssp, _batch_sources(sources), out)

    def close(self) -> None:
        if self._h:
            lib().gr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _SHUTDOWN:
            return
        try:
            self.close()
        except Exception:
            pass


def device_empty(nbytes: int) -> DeviceArray:
    """An uninitialised device allocation of nbytes."""
    d = C.c_void_p()
    _check(lib().gr_device_alloc(max(1, nbytes), C.byref(d)))
    return DeviceArray(d, nbytes)


def download(arr: DeviceArray, dtype, shape) -> np.ndarray:
    """Copies a device array to a new host array of the given dtype / shape."""
    out = np.empty(shape, dtype)
    if out.nbytes > arr.nbytes:
        raise GRError(1, "download: host array larger than the device array")
    _check(lib().gr_memcpy_d2h(out.ctypes.data, arr.ptr, out.nbytes))
    return out


def upload_i32(a) -> DeviceArray:
    """Copies a host int32 array into a new DeviceArray."""
    h = _i32(a)
    d = C.c_void_p()
    _check(lib().gr_device_alloc(max(1, h.nbytes), C.byref(d)))
    if h.nbytes:
        _check(lib().gr_memcpy_h2d(d, h.ctypes.data, h.nbytes))
    return DeviceArray(d, h.nbytes)


QUERY_BFS, QUERY_SSSP, QUERY_PPR = 1, 2, 3


class Server:
    """Continuous-batching query server over a device graph (see gr_abi.h)."""

    def __init__(self, graph: Graph, weights: "DeviceArray | None" = None,
                 memory_budget_bytes: int = 0, max_lanes_per_kind: int = 0,
                 ppr_iterations: int = 0):
        cfg = _ServerConfig(memory_budget_bytes, max_lanes_per_kind, ppr_iterations)
        h = _P()
        _check(lib().gr_server_create(graph.handle, weights.ptr if weights else None,
                                      C.byref(cfg), C.byref(h)))
        self._h = h
        self.graph = graph
        self.n = graph.n_vertices
        _LIVE_SERVERS.add(self)

    def close(self):
        if self._h:
            lib().gr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _SHUTDOWN:
            return
        try:
            self.close()
        except Exception:
            pass

    def submit(self, kind: int, vertex: int, keep: bool = True) -> int:
        rid = C.c_int64(0)
        _check(lib().gr_server_submit(self._h, kind, vertex, int(keep), C.byref(rid)))
        return rid.value

    def wait(self, rid: int, kind: int | None = None, want_result: bool = True):
        """Returns (result or None, checksum, latency_ms, steps)."""
        out = None
        ptr = None
        if want_result:
            out = np.empty(self.n, np.float32 if kind == QUERY_PPR else np.int32)
            ptr = out.ctypes.data
        cs = C.c_uint64(0)
        lat = C.c_double(0)
        steps = C.c_int32(0)
        _check(lib().gr_server_wait(self._h, rid, ptr, C.byref(cs), C.byref(lat), C.byref(steps)))
        return out, cs.value, lat.value, steps.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().gr_server_get_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in _ServerStats._fields_}

    def load_test(self, rate_per_s: float, vertices, kinds, keep: bool = False):
        v = _i32(vertices)
        k = _i32(kinds)
        ids = np.empty(len(v), np.int64)
        _check(lib().gr_server_load_test(self._h, rate_per_s, len(v), _p32(v), _p32(k), int(keep),
                                         ids.ctypes.data_as(_I64P)))
        return ids
