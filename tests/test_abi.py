"""CPU-side checks of the C ABI library: it loads, it exports every function
declared in include/graphreduce_amd/*.h, the host generators are correct,
deterministic and equal an independent numpy statement of their hash
streams, and device calls fail loudly without a GPU."""
import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

import graphreduce_amd as gr

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


def _declared_functions():
    names = set()
    for h in (ROOT / "include" / "graphreduce_amd").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(gr_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) > 30
    L = gr.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python mirror binds every one of them
    unbound = [n for n in sorted(names) if n not in gr._SIGS]
    assert not unbound, f"not bound in graphreduce_amd/__init__.py: {unbound}"


def test_version_and_no_device_is_reported():
    assert b"gfx950" in gr.lib().gr_version()
    n = gr.device_count()
    assert n >= 0


def test_device_calls_fail_loudly_without_gpu():
    if gr.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(gr.GRError):
        gr.Graph.rmat(8, 16, seed=1)
    with pytest.raises(gr.GRError):
        gr.GASEngine(gr.BFS).setGraph(2, np.zeros(2, gr.PROGRAMS[gr.BFS]["vertex"]), 1, None,
                                      np.array([0], np.int32), np.array([1], np.int32))


# ---------------------------------------------------------------- generators

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _base(seed, tag):
    with np.errstate(over="ignore"):
        return _mix(_mix(np.uint64(seed) ^ (np.uint64(tag) * GOLD)) + np.uint64(tag))


def _draw(base, idx):
    with np.errstate(over="ignore"):
        return _mix(base + (np.asarray(idx, np.uint64) + np.uint64(1)) * GOLD)


def _numpy_rmat(scale, ef, seed, a=0.57, b=0.19, c=0.19):
    """Independent numpy statement of the RMAT stream (splitmix64 finaliser,
    two 32-bit uniforms per draw, MSB level first)."""
    m = ef << scale
    s32 = 4294967296.0
    ta, tb, tc = (np.uint64(int(x * s32)) for x in (a, a + b, a + b + c))
    base = _base(seed, 0x524D4154)
    words = (scale + 1) // 2
    k = np.arange(m, dtype=np.uint64)
    src = np.zeros(m, np.uint64)
    dst = np.zeros(m, np.uint64)
    for lvl in range(scale):
        if lvl % 2 == 0:
            w = _draw(base, k * np.uint64(words) + np.uint64(lvl // 2))
        u = (w >> np.uint64(32)) if lvl % 2 else (w & np.uint64(0xFFFFFFFF))
        bit = np.uint64(1 << (scale - 1 - lvl))
        qb = (u >= ta) & (u < tb)
        qc = (u >= tb) & (u < tc)
        qd = u >= tc
        dst |= np.where(qb | qd, bit, np.uint64(0))
        src |= np.where(qc | qd, bit, np.uint64(0))
    keys = np.unique((src << np.uint64(scale)) | dst)
    s = (keys >> np.uint64(scale)).astype(np.int64)
    d = (keys & np.uint64((1 << scale) - 1)).astype(np.int64)
    keep = s != d
    return s[keep].astype(np.int32), d[keep].astype(np.int32)


@pytest.mark.parametrize("scale,seed", [(6, 1), (9, 2), (11, 6)])
def test_host_rmat_equals_numpy_statement(scale, seed):
    s, d = gr.rmat_edges(scale, 16, seed=seed)
    ns, nd = _numpy_rmat(scale, 16, seed)
    assert np.array_equal(s, ns) and np.array_equal(d, nd)


def test_rmat_properties_and_golden_digest():
    digests = json.loads((GOLDEN / "generators.json").read_text())
    for key, want in digests.items():
        kind, scale, seed = key.split(":")
        s, d = gr.rmat_edges(int(scale), 16, seed=int(seed), symmetrize=(kind == "rmat_sym"))
        assert (s != d).all()
        k = s.astype(np.int64) * (1 << int(scale)) + d
        assert (np.diff(k) > 0).all(), "sorted by (src, dst) without duplicates"
        h = hashlib.sha256(s.tobytes() + d.tobytes()).hexdigest()
        assert h == want["sha256"] and len(s) == want["edges"], key
    # RMAT skew: vertex 0 (all bits 0) has the largest expected out-degree
    s, d = gr.rmat_edges(14, 16, seed=3)
    deg = np.bincount(s, minlength=1 << 14)
    assert deg.argmax() == 0
    # fraction of arcs whose src has its top bit set is c + d = 0.24 (before dedup)
    top = (s >> 13) & 1
    assert 0.15 < top.mean() < 0.30


def test_symmetrized_rmat_is_symmetric():
    s, d = gr.rmat_edges(10, 16, seed=4, symmetrize=True)
    fwd = set(zip(s.tolist(), d.tolist()))
    assert all((b, a) in fwd for a, b in fwd)


def test_grid_generator():
    side = 40
    s, d = gr.grid_edges(side, 0.0, seed=5)
    assert len(s) == 4 * side * (side - 1)
    s1, d1 = gr.grid_edges(side, 0.1, seed=5)
    kept = len(s1) / len(s)
    assert 0.85 < kept < 0.95
    pairs = set(zip(s1.tolist(), d1.tolist()))
    assert all((b, a) in pairs for a, b in pairs)
    assert np.all(np.abs(s1 - d1) == 1) or np.all((np.abs(s1 - d1) == 1) | (np.abs(s1 - d1) == side))


def test_hash_weights_equal_numpy_statement():
    w = gr.hash_weights(1000, seed=2)
    base = _base(2, 0x57474854)
    ref = (np.uint64(1) + (_draw(base, np.arange(1000, dtype=np.uint64)) >> np.uint64(58))).astype(np.int32)
    assert np.array_equal(w, ref)
    assert w.min() >= 1 and w.max() <= 64


def _compile_registered(tmp_path, body):
    import shutil
    import subprocess

    src = tmp_path / "reg.cpp"
    src.write_text('#include <climits>\n#include "graphreduce_amd/GASEngineGPU.h"\n' + body)
    cxx = shutil.which("g++") or "g++"
    return subprocess.run([cxx, "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                          capture_output=True, text=True)


def test_register_program_checks_the_device_twin(tmp_path):
    """GR_REGISTER_PROGRAM refuses at compile time a host struct whose layout
    or gatherZero differs from the library's device program."""
    good = """struct SSSP { typedef int VertexData; typedef int EdgeData; typedef int GatherResult;
               static const int gatherZero = INT_MAX - 100000; };
               GR_REGISTER_PROGRAM(SSSP, GR_PROGRAM_SSSP)"""
    assert _compile_registered(tmp_path, good).returncode == 0
    bad_zero = good.replace("INT_MAX - 100000", "INT_MAX")
    r = _compile_registered(tmp_path, bad_zero)
    assert r.returncode != 0 and "gatherZero differs" in r.stderr
    bad_vd = good.replace("typedef int VertexData;", "typedef long VertexData;")
    r = _compile_registered(tmp_path, bad_vd)
    assert r.returncode != 0 and "VertexData differs" in r.stderr
    unregistered = """struct X { typedef int VertexData; typedef int EdgeData; typedef int GatherResult; };
                      graphreduce_amd::GASEngineGPU<X>* make() { return new graphreduce_amd::GASEngineGPU<X>(); }"""
    r = _compile_registered(tmp_path, unregistered)
    assert r.returncode != 0 and "gas_program.hpp" in r.stderr
