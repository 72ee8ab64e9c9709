"""Build recipe of the native library (hipcc, gfx950) and the C++ app drivers.

``build()`` compiles every HIP/C++ source under ``graphreduce_amd/csrc`` to
``graphreduce_amd/lib/libgraphreduce_amd.so`` and the four reference-style app
drivers under ``graphreduce_amd/apps`` to ``graphreduce_amd/bin``.  Objects go
to ``_build/`` at the repository root.  It always recompiles (the round-end
build check calls it on a machine without a GPU; hipcc cross-compiles).

``python -m graphreduce_amd._build --variant NAME -D MACRO=VALUE ...`` builds
the library with compile-time parameters changed (for A/B measurements, e.g.
``-D GR_RELAX_HUB=0``) to ``graphreduce_amd/lib/variants/NAME/``; a process
loads it when ``GRAPHREDUCE_AMD_LIB`` names it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
APPS = PKG / "apps"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
OBJDIR = ROOT / "_build"
LIBNAME = "libgraphreduce_amd.so"

ARCH = os.environ.get("GRAPHREDUCE_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE-correct FP32 division (PageRank's rank / outdeg)
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-ffp-contract=off",
]

SOURCES = [
    "gr_error.cpp",
    "gr_host_graph.cpp",
    "gr_graph.hip",
    "gr_engine.hip",
    "gr_pagerank.hip",
    "gr_batched.hip",
    "gr_serve.hip",
    "gr_graphio.cpp",
]

APP_SOURCES = ["bfs.cpp", "sssp.cpp", "pagerank.cpp", "connected_component.cpp", "mtx2gr.cpp",
               "gr2mtx.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build graphreduce_amd")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build_library(verbose: bool = False, defines: list[str] | None = None,
                  objdir: Path | None = None, libdir: Path | None = None) -> Path:
    cc = hipcc()
    objdir = objdir or OBJDIR
    libdir = libdir or LIBDIR
    objdir.mkdir(parents=True, exist_ok=True)
    libdir.mkdir(parents=True, exist_ok=True)
    extra = [f"-D{d}" for d in (defines or [])]
    objs = []
    cmds = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([cc, *HIP_FLAGS, *extra, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    workers = min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for _ in ex.map(_run, cmds):
            pass
    lib = libdir / LIBNAME
    tmp = libdir / (LIBNAME + ".tmp")
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"])
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


def build_apps(verbose: bool = False) -> list[Path]:
    """The app drivers are plain C++ over the C ABI (no HIP in them)."""
    BINDIR.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    outs = []
    cmds = []
    for src in APP_SOURCES:
        if not (APPS / src).exists():
            continue
        exe = BINDIR / Path(src).stem
        outs.append(exe)
        cmds.append([
            cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{ROOT / 'include' / 'graphreduce_amd'}",
            str(APPS / src), "-o", str(exe),
            f"-L{LIBDIR}", "-lgraphreduce_amd", "-Wl,-rpath,$ORIGIN/../lib",
        ])
    with ThreadPoolExecutor(max_workers=4) as ex:
        for _ in ex.map(_run, cmds):
            pass
    if verbose:
        for o in outs:
            print(f"built {o}")
    return outs


HIP_APP_SOURCES = ["widest_path.hip"]


def build_hip_apps(verbose: bool = False) -> list[Path]:
    """Apps with a vertex program of their own: compiled with hipcc against
    include/graphreduce_amd/gas_program.hpp, which instantiates the engine's
    kernels for the program in the app itself."""
    BINDIR.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    outs, cmds = [], []
    for src in HIP_APP_SOURCES:
        if not (APPS / src).exists():
            continue
        exe = BINDIR / Path(src).stem
        outs.append(exe)
        cmds.append([cc, *HIP_FLAGS, f"-I{ROOT / 'include'}", f"-I{ROOT / 'include' / 'graphreduce_amd'}",
                     "-x", "hip", str(APPS / src), "-o", str(exe), f"-L{LIBDIR}",
                     "-lgraphreduce_amd", "-Wl,-rpath,$ORIGIN/../lib"])
    for c in cmds:
        _run(c)
    if verbose:
        for o in outs:
            print(f"built {o}")
    return outs


def build_oracle(verbose: bool = False) -> Path:
    """CPU oracle (test infrastructure, built here so that tests can run)."""
    _run(["make", "-s", "-B", "-C", str(ROOT / "oracle")])
    lib = ROOT / "oracle" / "build" / "liboracle.so"
    if verbose:
        print(f"built {lib}")
    return lib


def build(verbose: bool = True) -> None:
    build_library(verbose)
    build_apps(verbose)
    build_hip_apps(verbose)
    build_oracle(verbose)


def build_variant(name: str, defines: list[str], verbose: bool = True) -> Path:
    """The library with compile-time parameters changed, for A/B measurements."""
    if not name.replace("_", "").replace("-", "").isalnum():
        raise ValueError(f"variant name {name!r}: letters, digits, '_' and '-' only")
    return build_library(verbose, defines, objdir=OBJDIR / "variants" / name,
                         libdir=LIBDIR / "variants" / name)


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--variant", help="build only the library, with -D flags, as a named variant")
    ap.add_argument("-D", dest="defines", action="append", default=[], metavar="MACRO=VALUE")
    a = ap.parse_args()
    if a.variant:
        build_variant(a.variant, a.defines)
    elif a.defines:
        ap.error("-D needs --variant (the default build uses the sources' values)")
    else:
        build(verbose=True)
