"""The learned per-pair algorithm (uint x uint: merge path or gallop) of the
learned form.  The committed table (emptyheaded_amd/pair_selector.json) was
tabulated from the selector network trained on engine-recorded cycles of
R-MAT scales 16-20; here it is checked for exactness under every choice, and
for what it gains on R-MAT scale 21, which it was not trained on: regret
against the per-pair faster algorithm, next to the reference's ratio rule
(uinteger.hpp:1209-1227), and the wall time of the learned form's
materialized triangle listing against the tuned form (same levels, rule
choice)."""
import sys

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G
from emptyheaded_amd import pairsel as PS
from emptyheaded_amd import plans

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture
def committed_table():
    import json
    with open(PS.PAIR_JSON) as f:
        d = json.load(f)
    yield np.array(d["table"], np.uint8), d
    PS.load_committed(0)


def test_pair_record_layouts_and_bounds(ctx):
    s, d, n = G.rmat_graph(12)
    U = E.Trie(ctx, s, d, E.POLICY_ALL_UINT)
    B = E.Trie(ctx, s, d, E.POLICY_ALL_BITSET)
    i, j, _ = PS.sample_pairs(s, d, 2000, seed=1)
    cu = PS.record(ctx, U, i, j, reps=2)
    assert (cu > 0).all()  # every sampled pair is uint x uint with both sets nonempty
    assert (PS.record(ctx, B, i, j, reps=1) == 0).all()  # bitset pairs are not timed
    with pytest.raises(E.EhxError):
        PS.record(ctx, U, np.array([U.info()["num_heads"] + 1], np.uint32), np.array([0], np.uint32))
    U.close()
    B.close()


def test_pair_tables_never_change_results(ctx, oracle, committed_table):
    """Every table (all merge, all gallop, all rule, the committed one) gives
    the oracle's rows in the learned form's per-tuple levels."""
    table, _ = committed_table
    s, d, n = G.rmat_graph(13)
    rows = oracle.join_materialize(plans.TRIANGLE, [oracle.Trie(s, d)])
    for policy in (E.POLICY_ALL_UINT, E.POLICY_RULE):
        T = E.Trie(ctx, s, d, policy)
        for tab in (np.zeros(576, np.uint8), np.ones(576, np.uint8), np.full(576, 2, np.uint8), table):
            PS.load(0, tab)
            R, st = E.join_materialize(ctx, plans.TRIANGLE, [T], E.FORM_LEARNED, stats=True)
            assert np.array_equal(R, rows)
            assert st["frontier"][2] == len(rows)
        T.close()


def test_pair_selector_gain_held_out(ctx, committed_table):
    table, d = committed_table
    params = np.array(d["params"], np.float32)
    norm = np.array(d["norm"], np.float32)
    s, dd, n = G.rmat_graph(21)
    T = E.Trie(ctx, s, dd, E.POLICY_ALL_UINT)
    X, cyc, rn, fn = PS.dataset(ctx, T, s, dd, 200000, seed=7, reps=3)
    T.close()
    r = PS.evaluate(ctx, params, norm, table, X, cyc, rn, fn)
    say("R-MAT-21 pair regret:", {k: round(v["regret"], 5) for k, v in r.items()}, "pairs", r["rule"]["pairs"])
    assert r["table"]["pairs"] > 100000
    assert r["table"]["regret"] <= 0.01
    assert r["table"]["regret"] < 0.1 * r["rule"]["regret"]
    assert r["table"]["regret"] <= r["always_gallop"]["regret"] + 0.002
    # wall time: the materialized triangle listing, whose last level is one
    # uint x uint intersection per edge; learned form (table) vs tuned (rule)
    s, dd, n = G.rmat_graph(18)
    U = E.Trie(ctx, s, dd, E.POLICY_ALL_UINT)
    out = {}
    for form in (E.FORM_TUNED, E.FORM_LEARNED):
        E.join_materialize(ctx, plans.TRIANGLE, [U], form)
        ts, rows = [], set()
        for _ in range(5):
            R, st = E.join_materialize(ctx, plans.TRIANGLE, [U], form, stats=True)
            ts.append(st["total_ms"])
            rows.add(len(R))
        assert len(rows) == 1
        out[form] = (float(np.median(ts)), rows.pop())
    say("R-MAT-18 triangle listing: tuned (rule) %.3f ms, learned (table) %.3f ms" %
        (out[E.FORM_TUNED][0], out[E.FORM_LEARNED][0]))
    assert out[E.FORM_TUNED][1] == out[E.FORM_LEARNED][1]
    assert out[E.FORM_LEARNED][0] <= out[E.FORM_TUNED][0]
    U.close()
