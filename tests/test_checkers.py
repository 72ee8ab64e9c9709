"""The regression checkers (regressions/checkDist.py, checkPageRank.py,
report.py) on hand-made result files: what they accept and what they reject."""
import subprocess
import sys

import numpy as np
import pytest

from graphreduce_amd import checkers as ck


def _write(path, vids, vals, fmt="%d"):
    with open(path, "w") as f:
        for v, x in zip(vids, vals):
            f.write(f"{v} {fmt % x}\n")
    return path


def test_check_dist_rules(tmp_path):
    vids = list(range(10))
    gold = _write(tmp_path / "gold", vids, [0, 1, 2, 10000000, 4, 5, 6, 7, 8, 9])
    # the unreachable sentinels pair up; extra test vertices are ignored
    test = _write(tmp_path / "test", vids + [10, 11],
                  [0, 1, 2, ck.TEST_UNREACHED, 4, 5, 6, 7, 8, 9, 3, 3])
    assert ck.check_dist(test, gold) == (True, "No differences found")
    bad = _write(tmp_path / "bad", vids, [0, 1, 2, ck.TEST_UNREACHED, 4, 5, 7, 7, 8, 9])
    assert ck.check_dist(bad, gold) == (False, "6 gold=6 ref=7")
    # the sentinel pairing is one way: gold 2147383647 vs test 10000000 differs
    g2 = _write(tmp_path / "g2", [0], [ck.TEST_UNREACHED])
    t2 = _write(tmp_path / "t2", [0], [10000000])
    assert not ck.check_dist(t2, g2)[0]
    short = _write(tmp_path / "short", vids[:5], [0, 1, 2, ck.TEST_UNREACHED, 4])
    assert ck.check_dist(short, gold) == (False, "vertex 5 is in gold but not in test")
    gap = _write(tmp_path / "gap", [0, 1, 2, 3, 5, 6, 7, 8, 9],
                 [0, 1, 2, ck.TEST_UNREACHED, 5, 6, 7, 8, 9])
    assert ck.check_dist(gap, gold) == (False, "vertex 4 is in gold but not in test")


@pytest.mark.parametrize("bucket,count,ok", [
    (1, 10, True), (1, 11, False),      # [1e-4, 1e-3): 1 % allowed
    (2, 5, True), (2, 6, False),        # [1e-3, 1e-2): 0.5 %
    (3, 1, False),                      # [1e-2, 0.1): 0.05 % of 1000 < 1 vertex
    (6, 1, False),                      # >= 10: none allowed
])
def test_check_pagerank_histogram(tmp_path, bucket, count, ok):
    n = 1000
    gold = np.full(n, 1.0)
    test = gold.copy()
    test[:n // 2] += 5e-5  # bucket 0: always allowed
    mid = [None, 5e-4, 5e-3, 5e-2, 0.5, 5.0, 50.0][bucket]
    test[n - count:] = gold[n - count:] + mid
    g = _write(tmp_path / "gold", range(n), gold, "%.9f")
    t = _write(tmp_path / "test", range(n), test, "%.9f")
    assert ck.check_pagerank(t, g)[0] is ok
    shares, _ = ck.pagerank_histogram({i: float(x) for i, x in enumerate(test)},
                                      {i: float(x) for i, x in enumerate(gold)})
    assert abs(shares[bucket] - count / n) < 1e-12 and abs(sum(shares) - 1.0) < 1e-12


def test_check_pagerank_missing_gold_vertices(tmp_path):
    """Test vertices without a gold entry are reported and counted in the
    total only (checkPageRank.py:50-54)."""
    g = _write(tmp_path / "gold", range(100), [1.0] * 100, "%.6f")
    t = _write(tmp_path / "test", range(102), [1.0] * 100 + [7.0, 7.0], "%.6f")
    ok, msg = ck.check_pagerank(t, g)
    assert ok and "vid 100 is in test but not in gold" in msg
    shares, missing = ck.pagerank_histogram({i: 1.0 for i in range(102)},
                                            {i: 1.0 for i in range(100)})
    assert missing == [100, 101] and abs(shares[0] - 100 / 102) < 1e-12


def test_report_and_cli(tmp_path):
    (tmp_path / "g1.sssp.timing").write_text("120.0")
    (tmp_path / "g1.sssp.timing_gpu").write_text("4.0")
    out = ck.report(str(tmp_path))
    assert "Algorithm: sssp" in out and "30.0x" in out
    gold = _write(tmp_path / "gold", [0, 1], [0, 1])
    test = _write(tmp_path / "test", [0, 1], [0, 2])
    r = subprocess.run([sys.executable, "-m", "graphreduce_amd.checkers", "dist", str(test),
                        str(gold)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "1 gold=1 ref=2" in r.stdout
    r = subprocess.run([sys.executable, "-m", "graphreduce_amd.checkers", "dist", str(gold),
                        str(gold)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "No differences found" in r.stdout
