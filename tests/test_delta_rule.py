"""The delta-gather rule, checked on the CPU against the oracle.

For min-monotone programs (SSSP, CC) the engine gathers, after an all-active
iteration, only over in-edges whose source changed in the previous iteration
and substitutes gatherZero for the others.  This numpy model of the
bulk-synchronous loop (refgas.h:128-229) applies exactly that rule; its
values and trip counts must equal the oracle's full-gather run on graphs with
hubs, long paths, duplicates, self-loops and zero weights.  The HIP kernels
are compared with the oracle by the GPU tests."""
import numpy as np
import pytest

import graphreduce_amd as gr
import oracle

SSSP_ZERO = 2**31 - 1 - 100000  # sssp.cu gatherZero
CC_ZERO = 2**31 - 1             # connected_component.cu gatherZero


def _delta_loop(n, s, d, vals, mapped, zero, inert=None):
    """Returns (values, trips) of the loop with the delta rule; `mapped(vals)`
    gives gatherMap per edge-list position.  With `inert`, the first gather
    substitutes gatherZero for the sources whose value is `inert`."""
    vals = vals.astype(np.int64).copy()
    active = np.ones(n, bool)
    changed_prev = None if inert is None else vals != inert
    trips = 0
    while active.any():
        trips += 1
        m = mapped(vals)
        if changed_prev is not None:
            m = np.where(changed_prev[s], m, zero)  # unchanged source -> gatherZero
        g = np.full(n, zero, np.int64)  # every row is folded with gatherZero
        np.minimum.at(g, d, m)
        changed = active & (g < vals)
        vals = np.where(changed, g, vals)
        nxt = np.zeros(n, bool)
        nxt[d[changed[s]]] = True
        changed_prev = changed
        active = nxt
    return vals.astype(np.int32), trips


def _graphs():
    rng = np.random.default_rng(7)
    out = []
    s, d = gr.rmat_edges(12, 16, seed=2)
    out.append(("rmat12", 1 << 12, s, d))
    s, d = gr.rmat_edges(11, 16, seed=4, symmetrize=True)
    out.append(("rmat11_sym", 1 << 11, s, d))
    s, d = gr.grid_edges(40, 0.05, seed=5)
    out.append(("grid40", 1600, s, d))
    n, m = 300, 1500
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    s = np.concatenate([s, s[:50], np.arange(10, dtype=np.int32)])  # duplicates, self-loops
    d = np.concatenate([d, d[:50], np.arange(10, dtype=np.int32)])
    out.append(("multigraph", n, s, d))
    return out


@pytest.mark.parametrize("name,n,s,d", _graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_delta_rule_sssp_matches_oracle(name, n, s, d):
    w = gr.hash_weights(len(s), seed=11, lo=0, span_log2=4)  # includes zero weights
    src = int(np.argmax(np.bincount(s, minlength=n)))
    ref, ref_trips, _ = oracle.sssp(n, s, d, w, src)
    init = np.full(n, SSSP_ZERO, np.int64)
    init[src] = 0
    got, trips = _delta_loop(n, s, d, init, lambda v: v[s] + w, SSSP_ZERO)
    assert np.array_equal(got, ref)
    assert trips == ref_trips


@pytest.mark.parametrize("name,n,s,d", _graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_inert_first_gather_sssp_matches_oracle(name, n, s, d):
    """The first gather reads only sources off gatherZero (weights in
    [0, 100000]: an unreached source maps to at least gatherZero)."""
    w = gr.hash_weights(len(s), seed=12, lo=0, span_log2=4)
    w[::7] = 100000  # the largest weight that keeps gatherZero + w in range
    src = int(np.argmax(np.bincount(s, minlength=n)))
    ref, ref_trips, _ = oracle.sssp(n, s, d, w, src)
    init = np.full(n, SSSP_ZERO, np.int64)
    init[src] = 0
    got, trips = _delta_loop(n, s, d, init, lambda v: v[s] + w, SSSP_ZERO, inert=SSSP_ZERO)
    assert np.array_equal(got, ref)
    assert trips == ref_trips


def test_inert_rule_needs_the_weight_bound():
    """A negative weight lets an unreached source lower its target below
    gatherZero in the full gather; skipping inert sources would miss that."""
    s, d, w = np.array([0, 1]), np.array([1, 2]), np.array([5, -1])
    init = np.array([0, SSSP_ZERO, SSSP_ZERO], np.int64)
    full, _ = _delta_loop(3, s, d, init, lambda v: v[s] + w, SSSP_ZERO)
    skip, _ = _delta_loop(3, s, d, init, lambda v: v[s] + w, SSSP_ZERO, inert=SSSP_ZERO)
    assert full[2] == 4 and skip[2] == 4  # both reach the fixed point ...
    g = np.full(3, SSSP_ZERO, np.int64)  # ... but the first full gather differs:
    np.minimum.at(g, d, init[s] + w)
    assert g[2] == SSSP_ZERO - 1  # vertex 2 changes in trip 1 of the full gather only


@pytest.mark.parametrize("name,n,s,d", _graphs(), ids=lambda x: x if isinstance(x, str) else "")
def test_delta_rule_cc_matches_oracle(name, n, s, d):
    ref, ref_trips, _ = oracle.cc(n, s, d)
    got, trips = _delta_loop(n, s, d, np.arange(n), lambda v: v[s], CC_ZERO)
    assert np.array_equal(got, ref)
    assert trips == ref_trips


def test_delta_rule_needs_the_all_active_start():
    """From a partial start the rule is not exact: vertex 1 never gathered from
    source 2, whose value never changed (the engine therefore keeps the full
    gather after a partial setActive)."""
    s = np.array([3, 0, 2])
    d = np.array([0, 1, 1])
    w = np.array([1, 5, 1])
    vals = np.array([SSSP_ZERO, SSSP_ZERO, 0, 0], np.int64)
    active = np.array([True, False, False, False])
    # two iterations of the delta rule from A0 = {0}
    g = np.full(4, SSSP_ZERO, np.int64)
    np.minimum.at(g, d, vals[s] + w)
    changed = active & (g < vals)
    vals = np.where(changed, g, vals)
    active = np.zeros(4, bool)
    active[d[changed[s]]] = True
    m = np.where(changed[s], vals[s] + w, SSSP_ZERO)
    g = np.full(4, SSSP_ZERO, np.int64)
    np.minimum.at(g, d, m)
    delta_val = np.where(active & (g < vals), g, vals)
    assert delta_val[1] == 6  # the full gather gives 1
