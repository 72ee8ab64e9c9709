"""Finalizers make no library call once the interpreter is shutting down.

Objects left open at exit are collected during module teardown, possibly
after the HIP runtime's own exit handler has run; a hipFree or stream call
there can fault.  The atexit hook of the package stops live servers (while
the runtime is still up) and then turns every finalizer into a no-op.  Run
in a child process, with the native library replaced by a recorder, so the
hook's global flag does not leak into other tests.
"""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent

SCRIPT = r"""
import gc, sys
sys.path.insert(0, %(root)r)
import graphreduce_amd as gr

calls = []

class Recorder:
    def __getattr__(self, name):
        def fn(*a):
            calls.append(name)
            return 0
        return fn

gr.lib = lambda: Recorder()

def make(cls, attr):
    o = object.__new__(cls)
    setattr(o, attr, 1234)
    return o

# before the hook: a dropped object releases its native handle
g = make(gr.Graph, "_h"); del g; gc.collect()
assert calls == ["gr_graph_destroy"], calls
calls.clear()

keep = [make(gr.Graph, "_h"), make(gr.GASEngine, "_h"), make(gr.DeviceArray, "ptr"),
        make(gr.Batch, "_h")]
srv = make(gr.Server, "_h")
gr._LIVE_SERVERS.add(srv)
gr._at_exit()
# the hook stops the live server while the runtime is up ...
assert calls == ["gr_server_destroy"], calls
calls.clear()
# ... and after it no finalizer reaches the library
del keep, srv
gc.collect()
assert calls == [], calls
print("ok")
"""


def test_no_library_call_after_exit_hook():
    out = subprocess.run([sys.executable, "-c", SCRIPT % {"root": str(ROOT)}],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("ok")
