"""The learned-selector workloads at their full size, on the GPU: R-MAT scale
22 (edge factor 16) triangle COUNT and the 2-hop star and path joins over the
two most frequent predicates of the 50M-triple, 64-predicate Zipf(1.1) table.

Expected values come from tests/golden/synthetic.json, written by the CPU
oracle (triangles) and by numpy closed forms (sum_s d_i(s) d_j(s) and
sum_y in_i(y) out_j(y) for the joins); the inputs are checked against the
SHA-256 recorded there first, so a generator change fails loudly instead of
comparing different graphs.  Every form (plain, tuned, tuned+learned) runs on
density-rule and learned layouts, both triangle attribute orders; counts are
exact integers, so every comparison is bit-exact."""
import hashlib
import sys

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G
from emptyheaded_amd import plans
from emptyheaded_amd import triples as TR
from emptyheaded_amd import tuning as TU

pytestmark = pytest.mark.gpu

FORMS = [("plain", E.FORM_PLAIN), ("tuned", E.FORM_TUNED), ("learned", E.FORM_LEARNED)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def committed_tuning():
    """Committed knobs and selector weights, restored afterwards."""
    saved = E.tune_get()
    applied = TU.apply_committed()
    assert applied["selector"], "no committed selector weights"
    yield applied
    E.tune_set(**saved)


def test_rmat22_triangle_every_form_and_layout(ctx, synthetic_meta, committed_tuning):
    meta = synthetic_meta["rmat22"]
    s, d, n = G.rmat_graph(22)
    assert digest(s, d) == meta["sha256"] and len(s) == meta["edges"]
    expect = meta["triangles"]
    assert expect == 2110480353
    bitsets = {}
    rule_bitsets = None
    for pname, policy in (("rule", E.POLICY_RULE), ("learned", E.POLICY_LEARNED)):
        T = E.Trie(ctx, s, d, policy)
        R = E.Trie(ctx, d, s, policy)
        info = T.info()
        if policy == E.POLICY_RULE:
            rule_bitsets = info["num_bitsets"]
        bitsets[pname] = info["num_staged_bitmaps"]
        if policy == E.POLICY_LEARNED:  # the network, not the rule, chose the staging layouts;
            assert bitsets["learned"] != bitsets["rule"], bitsets  # the rule's stay for the rest
            assert info["num_bitsets"] == rule_bitsets
        for plan_name, plan, tries in (("triangle", plans.TRIANGLE, [T]),
                                       ("triangle_by_middle", plans.TRIANGLE_BY_MIDDLE, [T, R])):
            TU.apply_plan(plan_name)
            for fname, form in FORMS:
                say("R-MAT-22 triangle", pname, plan_name, fname)
                assert E.join_count(ctx, plan, tries, form) == expect, (pname, plan_name, fname)
        T.close()
        R.close()


def test_triple_table_50m_star_and_path(ctx, synthetic_meta, committed_tuning):
    meta = synthetic_meta["triples_50m_top2"]
    s, p, o = TR.triple_table()
    rels = TR.by_predicate(s, p, o, 64)
    del s, p, o
    R = [TR.distinct(*rels[0]), TR.distinct(*rels[1])]
    del rels
    assert digest(R[0][0], R[0][1], R[1][0], R[1][1]) == meta["sha256"]
    # the closed forms again, from these rows
    assert TR.star_count(R[0], R[1]) == meta["star"] and TR.path_count(R[0], R[1]) == meta["path"]
    TU.apply_plan("triangle")
    for pname, policy in (("rule", E.POLICY_RULE), ("all_uint", E.POLICY_ALL_UINT), ("learned", E.POLICY_LEARNED)):
        T = [E.Trie(ctx, *r, policy=policy) for r in R]
        for q, plan in (("star", TR.STAR2), ("path", TR.PATH2)):
            for fname, form in FORMS:
                say("triples", pname, q, fname)
                assert E.join_count(ctx, plan, T, form) == meta[q], (pname, q, fname)
        for x in T:
            x.close()
