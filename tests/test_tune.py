"""apply_tuning picks the settings of the right section (no GPU: a stand-in
engine records the setter calls)."""
import json

import graphreduce_amd as gr
from graphreduce_amd import tune


class _Rec:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        return lambda *a: self.calls.append((name,) + a)


def test_apply_tuning_sections(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({
        "programs": {"CC": {"gather_tile": 4, "alpha": 60.0, "fuse_share": 0.0, "push_share": 0.1}},
        "uniform": {"programs": {"CC": {"gather_tile": 8, "alpha": 15.0, "fuse_share": 0.1,
                                        "push_share": 0.6}}},
    }))
    e = _Rec()
    assert tune.apply_tuning(e, gr.CC, p)
    assert e.calls == [("setGatherTile", 4), ("setDirection", 0, 60.0), ("setFuseShare", 0.0),
                       ("setPushShare", 0.1)]
    u = _Rec()
    assert tune.apply_tuning(u, gr.CC, p, uniform=True)
    assert u.calls == [("setGatherTile", 8), ("setDirection", 0, 15.0), ("setFuseShare", 0.1),
                       ("setPushShare", 0.6)]
    # no uniform entry for a program: nothing applied
    assert not tune.apply_tuning(_Rec(), gr.SSSP, p, uniform=True)
    assert not tune.apply_tuning(_Rec(), gr.CC, tmp_path / "missing.json")
