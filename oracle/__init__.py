"""CPU oracle of the GAS hot path -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's sequential CPU engine ``GASEngineRef``
(refgas.h:30-230) and of the four app loops, compiled from ``gas_oracle.h`` /
``oracle_capi.cpp`` into ``oracle/build/liboracle.so``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
use it, and only as the checker or the timed CPU baseline; the product never
imports it.

Every driver returns (result array, loop trips, stats) where stats is a dict
with the per-run work counters (iterations, sum #A, sum E_in(A), sum #C,
sum E_out(C)) and the loop time in seconds (timed region of bfs.cu:113-126).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "build" / "liboracle.so"

SSSP_INF = 2**31 - 1 - 100000  # sssp.cu:32-33

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F32P = C.POINTER(C.c_float)
_F64P = C.POINTER(C.c_double)

_SIGS = {
    "oracle_bfs": [C.c_int32, C.c_int64, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_bfs_csr": [C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_sssp": [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_sssp_congestion": [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_sssp_csr": [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_cc": [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, _I64P],
    "oracle_widest": [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, C.c_int32, _I32P, _I64P],
    "oracle_cc_csr": [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, _I64P],
    "oracle_pagerank": [C.c_int32, C.c_int64, _I32P, _I32P, _F32P, _I64P],
    "oracle_pagerank64": [C.c_int32, C.c_int64, _I32P, _I32P, _F64P, _I64P],
    "oracle_pagerank_csr": [C.c_int32, _I32P, _I32P, _I32P, _I32P, _F32P, _I64P],
    "oracle_pagerank_fp64_csr": [C.c_int32, _I32P, _I32P, _I32P, C.c_int32, _F64P, _F64P],
    "oracle_build_csr": [C.c_int32, C.c_int64, _I32P, _I32P, _I32P, _I32P],
}

_lib = None


def lib(path: str | None = None) -> C.CDLL:
    global _lib
    if path is not None:
        L = C.CDLL(path)
        _setup(L)
        return L
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(str(LIB_PATH))
        _setup(_lib)
    return _lib


def _setup(L):
    for name, args in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int if name != "oracle_build_csr" else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t=_I32P):
    return a.ctypes.data_as(t)


def _stats(st):
    return dict(iterations=int(st[0]), active=int(st[1]), in_edges=int(st[2]),
                changed=int(st[3]), out_edges=int(st[4]), seconds=st[5] * 1e-9)


def bfs(n, srcs, dsts, source, L=None):
    L = L or lib()
    s, d = _i32(srcs), _i32(dsts)
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_bfs(n, len(s), _p(s), _p(d), source, _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def bfs_csr(n, off_out, dsts, off_in, srcs, source, L=None):
    L = L or lib()
    a, b, c, d = map(_i32, (off_out, dsts, off_in, srcs))
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_bfs_csr(n, _p(a), _p(b), _p(c), _p(d), source, _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def sssp(n, srcs, dsts, weights, source, L=None):
    L = L or lib()
    s, d, w = _i32(srcs), _i32(dsts), _i32(weights)
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_sssp(n, len(s), _p(s), _p(d), _p(w), source, _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def sssp_congestion(n, srcs, dsts, weights, source, L=None):
    """SSSP on congesting links (a program whose scatter changes edge data):
    returns (dist, trips, stats, weights after the run, edge-list order)."""
    L = L or lib()
    s, d = _i32(srcs), _i32(dsts)
    w = np.array(weights, dtype=np.int32, copy=True)
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_sssp_congestion(n, len(s), _p(s), _p(d), _p(w), source, _p(out),
                                  _p(st, _I64P))
    return out, it, _stats(st), w


def sssp_csr(n, off_out, dsts, off_in, srcs, weights_csc, source, L=None):
    L = L or lib()
    a, b, c, d, w = map(_i32, (off_out, dsts, off_in, srcs, weights_csc))
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_sssp_csr(n, _p(a), _p(b), _p(c), _p(d), _p(w), source, _p(out),
                           _p(st, _I64P))
    return out, it, _stats(st)


def widest(n, srcs, dsts, caps, source, L=None):
    """Widest-path program (a program outside the reference's four): width of
    the best bottleneck path from `source` (INT_MAX there, 0 if unreached)."""
    L = L or lib()
    s, d, c = _i32(srcs), _i32(dsts), _i32(caps)
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_widest(n, len(s), _p(s), _p(d), _p(c), source, _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def cc(n, srcs, dsts, L=None):
    L = L or lib()
    s, d = _i32(srcs), _i32(dsts)
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_cc(n, len(s), _p(s), _p(d), _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def cc_csr(n, off_out, dsts, off_in, srcs, L=None):
    L = L or lib()
    a, b, c, d = map(_i32, (off_out, dsts, off_in, srcs))
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    it = L.oracle_cc_csr(n, _p(a), _p(b), _p(c), _p(d), _p(out), _p(st, _I64P))
    return out, it, _stats(st)


def pagerank(n, srcs, dsts, L=None):
    L = L or lib()
    s, d = _i32(srcs), _i32(dsts)
    out = np.empty(n, np.float32)
    st = np.zeros(6, np.int64)
    it = L.oracle_pagerank(n, len(s), _p(s), _p(d), _p(out, _F32P), _p(st, _I64P))
    return out, it, _stats(st)


def pagerank64(n, srcs, dsts, L=None):
    """pagerank.cu's activation-driven loop in FP64."""
    L = L or lib()
    s, d = _i32(srcs), _i32(dsts)
    out = np.empty(n, np.float64)
    st = np.zeros(6, np.int64)
    it = L.oracle_pagerank64(n, len(s), _p(s), _p(d), _p(out, _F64P), _p(st, _I64P))
    return out, it, _stats(st)


def pagerank_csr(n, off_out, dsts, off_in, srcs, L=None):
    L = L or lib()
    a, b, c, d = map(_i32, (off_out, dsts, off_in, srcs))
    out = np.empty(n, np.float32)
    st = np.zeros(6, np.int64)
    it = L.oracle_pagerank_csr(n, _p(a), _p(b), _p(c), _p(d), _p(out, _F32P), _p(st, _I64P))
    return out, it, _stats(st)


def pagerank_fp64_csr(n, off_out, off_in, srcs, iterations, L=None):
    """All-active FP64 recurrence; returns (rank f64[V], l1 residual per step)."""
    L = L or lib()
    a, c, d = map(_i32, (off_out, off_in, srcs))
    out = np.empty(n, np.float64)
    l1 = np.empty(max(iterations, 1), np.float64)
    L.oracle_pagerank_fp64_csr(n, _p(a), _p(c), _p(d), iterations, _p(out, _F64P),
                               _p(l1, _F64P))
    return out, l1[:iterations]


def build_csr(n, rows, cols, L=None):
    L = L or lib()
    r, c = _i32(rows), _i32(cols)
    off = np.empty(n + 1, np.int32)
    adj = np.empty(len(r), np.int32)
    L.oracle_build_csr(n, len(r), _p(r), _p(c), _p(off), _p(adj))
    return off, adj
