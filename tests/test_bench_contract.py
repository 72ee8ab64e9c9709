"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message on stderr and exit status 2, before any device work,
and nothing is printed on stdout (the JSON line is reserved for results)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", str(n)],
                           capture_output=True, text=True, timeout=120, cwd=ROOT)
        assert p.returncode == 2
        assert p.stdout.strip() == ""
        assert "one GPU" in p.stderr and f"--gpus {n}" in p.stderr
