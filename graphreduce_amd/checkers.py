"""Result checkers and timing report of the reference's regression suite.

``check_dist`` restates ``regressions/checkDist.py`` (BFS depths, SSSP
distances, CC labels), ``check_pagerank`` restates
``regressions/checkPageRank.py`` (tolerance histogram) and ``report``
restates ``regressions/report.py`` (CPU vs GPU timing table).  Inputs are
the ``vid value`` files the app drivers write (sorted by vertex id, as the
regression Makefile sorts them, ``regressions/Makefile:66-110``).

Command line::

    python -m graphreduce_amd.checkers dist TEST GOLD
    python -m graphreduce_amd.checkers pagerank TEST GOLD
    python -m graphreduce_amd.checkers report [DIR]

exit status 0 when the files agree, 1 otherwise, as the reference scripts.
"""
from __future__ import annotations

import glob
import math
import os
import sys
from typing import Dict, Iterator, List, Optional, Sequence, Tuple

# PowerGraph's gold files mark an unreachable vertex with 10000000; the apps
# here, like the reference's, write SSSP's gatherZero (checkDist.py:43)
GOLD_UNREACHED = 10000000
TEST_UNREACHED = 2147483647 - 100000

# checkPageRank.py:94-99: share of vertices allowed per |test - gold| bucket
# [0, 1e-4), [1e-4, 1e-3), [1e-3, 1e-2), [1e-2, 0.1), [0.1, 1), [1, 10), >= 10
PAGERANK_TOL_VALS = (1e-4, 1e-3, 1e-2, 1e-1, 1.0, 10.0)
PAGERANK_TOL_ALLOWED = (1.0, 1e-2, 5e-3, 5e-4, 5e-5, 5e-6, 0.0)


def _int_lines(path) -> Iterator[Tuple[int, int]]:
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                yield int(parts[0]), int(parts[1])


def check_dist(test_path, gold_path) -> Tuple[bool, str]:
    """checkDist.py:30-47: walk the sorted gold file; every gold vertex must
    be in the test file with the same value (gold 10000000 matches test
    2147383647); test vertices absent from gold are ignored."""
    test = _int_lines(test_path)
    cur = next(test, None)
    for vid, val in _int_lines(gold_path):
        while cur is not None and cur[0] < vid:
            cur = next(test, None)
        if cur is None or cur[0] != vid:
            return False, f"vertex {vid} is in gold but not in test"
        if cur[1] != val and not (val == GOLD_UNREACHED and cur[1] == TEST_UNREACHED):
            return False, f"{vid} gold={val} ref={cur[1]}"
        cur = next(test, None)
    return True, "No differences found"


def _float_map(path) -> Dict[int, float]:
    out = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                out[int(parts[0])] = float(parts[1])
    return out


def pagerank_histogram(test: Dict[int, float], gold: Dict[int, float],
                       tol_vals: Sequence[float] = PAGERANK_TOL_VALS
                       ) -> Tuple[List[float], List[int]]:
    """checkPageRank.py:40-55: share of the test vertices per difference
    bucket (vertices missing from gold are counted in the total only), and
    the ids missing from gold."""
    counts = [0] * (len(tol_vals) + 1)
    missing = []
    for vid, val in test.items():
        if vid not in gold:
            missing.append(vid)
            continue
        diff = math.fabs(gold[vid] - val)
        pos = len(tol_vals) - 1
        while pos >= 0 and diff < tol_vals[pos]:
            pos -= 1
        counts[pos + 1] += 1
    total = float(len(test)) if test else 1.0
    return [c / total for c in counts], missing


def check_pagerank(test_path, gold_path, tol_vals: Sequence[float] = PAGERANK_TOL_VALS,
                   tol_allowed: Sequence[float] = PAGERANK_TOL_ALLOWED) -> Tuple[bool, str]:
    """checkPageRank.py:58-100: fails when a bucket's share exceeds its limit."""
    shares, missing = pagerank_histogram(_float_map(test_path), _float_map(gold_path), tol_vals)
    notes = [f"vid {v} is in test but not in gold" for v in missing[:10]]
    for share, allowed in zip(shares, tol_allowed):
        if share > allowed:
            notes.append(f"Percentage too high: {allowed} {share}")
            return False, "\n".join(notes)
    return True, "\n".join(notes + ["histogram " + " ".join(f"{s:.3g}" for s in shares)])


def _timing(path) -> float:
    try:
        with open(path) as f:
            return float(f.read())
    except (OSError, ValueError):
        return -1.0


def report(directory: str = ".") -> str:
    """report.py: per algorithm and graph, CPU (`G.A.timing`) and GPU
    (`G.A.timing_gpu`) times and their ratio."""
    cpu = glob.glob(os.path.join(directory, "*.timing"))
    gpu = glob.glob(os.path.join(directory, "*.timing_gpu"))
    names = [os.path.basename(p).split(".") for p in cpu + gpu]
    graphs = sorted({n[0] for n in names})
    algos = sorted({n[1] for n in names})
    lines = []
    for algo in algos:
        lines.append("-" * 51)
        lines.append(f"Algorithm: {algo}")
        lines.append("-" * 51)
        for graph in graphs:
            c = _timing(os.path.join(directory, f"{graph}.{algo}.timing"))
            g = _timing(os.path.join(directory, f"{graph}.{algo}.timing_gpu"))
            ratio = c / g if g else float("inf")
            lines.append("%-20.20s  %8.1f  %8.1f  %8.1fx" % (graph, c, g, ratio))
        lines.append("-" * 51)
    return "\n".join(lines)


def main(argv: Optional[Sequence[str]] = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) == 3 and argv[0] in ("dist", "pagerank"):
        ok, msg = (check_dist if argv[0] == "dist" else check_pagerank)(argv[1], argv[2])
        if msg:
            print(msg)
        return 0 if ok else 1
    if argv and argv[0] == "report" and len(argv) <= 2:
        print(report(argv[1] if len(argv) == 2 else "."))
        return 0
    print("Usage: python -m graphreduce_amd.checkers dist|pagerank TEST GOLD | report [DIR]")
    return 1


if __name__ == "__main__":
    sys.exit(main())
