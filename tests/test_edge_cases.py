"""Edge cases of the join engine on the GPU, against the oracle: empty and
single-row relations, ids at the top of the uint32 range (the hash staging's
empty-slot marker is 0xFFFFFFFF), duplicate rows, a hub whose staged set and
child lists exceed every LDS staging area, and every form and plan."""
import sys

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G
from emptyheaded_amd import plans

pytestmark = pytest.mark.gpu

FORMS = [E.FORM_PLAIN, E.FORM_TUNED, E.FORM_LEARNED]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def oriented(edges):
    """Undirected edge list -> sorted unique rows with src < dst."""
    e = np.array(sorted({(min(u, v), max(u, v)) for u, v in edges if u != v}), dtype=np.uint64)
    return e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32)


def all_counts(ctx, s, d):
    T = E.Trie(ctx, s, d)
    R = E.Trie(ctx, d, s)
    out = []
    for form in FORMS:
        out.append(E.join_count(ctx, plans.TRIANGLE, [T], form))
        out.append(E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form))
        out.append(E.join_count(ctx, plans.FOUR_CLIQUE, [T], form))
    return out


def test_empty_and_single_row_relations(ctx):
    z = np.zeros(0, np.uint32)
    with pytest.raises(E.EhxError):  # the reference asserts num_rows != 0 (src/Trie.cpp:394)
        E.Trie(ctx, z, z)
    one = np.array([3], np.uint32), np.array([7], np.uint32)
    T = E.Trie(ctx, *one)
    assert E.join_materialize(ctx, plans.TRIANGLE, [T]).shape == (0, 3)
    assert all_counts(ctx, *one) == [0] * 9
    S = E.Trie(ctx, np.array([3, 7], np.uint32), np.array([7, 3], np.uint32))
    assert E.cycle4_count(ctx, S) == 0


def test_query_graph_edge_cases(ctx, oracle):
    """The middle-vertex plan replayed as one graph: no common heads (level 0
    empty, its count known only on the device), one row, and ids at the top of
    the uint32 range, each replayed twice."""
    top = 2**32 - 1
    edges = [(top - i, top - j) for i in range(6) for j in range(i)] + [(1, 2), (2, 3), (1, 3)]
    cases = [(np.array([3], np.uint32), np.array([7], np.uint32)),
             (np.array([1, 2], np.uint32), np.array([5, 6], np.uint32)),
             oriented(edges)]
    saved = E.tune_get()
    try:
        for s, d in cases:
            exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 2)
            T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
            for ug in (0, 1, 1):
                E.tune_set(use_graph=ug)
                for form in (E.FORM_TUNED, E.FORM_LEARNED):
                    assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form) == exp
            T.close()
            R.close()
    finally:
        E.tune_set(**saved)


def test_ids_at_the_top_of_the_range(ctx, oracle):
    """Cliques on the largest uint32 ids: a K6 and a K5 sharing a vertex, plus
    noise; ids 0xFFFFFFFF and 0xFFFFFFFE are in sets that get staged."""
    top = 2**32 - 1
    k6 = [top - i for i in range(6)]
    k5 = [top - 5] + [top - 1000 - i for i in range(4)]
    edges = [(u, v) for c in (k6, k5) for i, u in enumerate(c) for v in c[i + 1:]]
    rng = np.random.default_rng(5)
    edges += [(int(top - rng.integers(0, 3000)), int(top - rng.integers(0, 3000))) for _ in range(4000)]
    s, d = oriented(edges)
    O = oracle.Trie(s, d)
    tri = oracle.join_count(plans.TRIANGLE, [O], 4)
    k4 = oracle.join_count(plans.FOUR_CLIQUE, [O], 4)
    assert tri >= 20 + 10 and k4 >= 15 + 5
    assert all_counts(ctx, s, d) == [tri, tri, k4] * 3
    saved = E.tune_get()
    try:  # every small prefix through the many-per-wavefront bin, 0xFFFFFFFF probed there
        E.tune_set(tiny_work=1 << 20, big_work=1 << 20)
        assert all_counts(ctx, s, d) == [tri, tri, k4] * 3
    finally:
        E.tune_set(**saved)
    ss, sd = G.symmetric(s, d)
    # 4-cycles on ids near the top: relabel-free oracle needs a dense id range, so shift both
    base = int(min(ss.min(), sd.min()))
    exp = oracle.cycle4_count(ss - base, sd - base, int(max(ss.max(), sd.max())) - base + 1, 4)
    assert E.cycle4_count(ctx, E.Trie(ctx, ss, sd)) == exp


def test_duplicate_rows_are_one_tuple(ctx, oracle):
    s, d, n = G.rmat_graph(10)
    rep_s = np.concatenate([s, s, s[::3]])
    rep_d = np.concatenate([d, d, d[::3]])
    T = E.Trie(ctx, rep_s, rep_d)
    assert T.info()["num_rows"] == len(s)
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 4)
    for form in FORMS:
        assert E.join_count(ctx, plans.TRIANGLE, [T], form) == exp


@pytest.mark.parametrize("policy", [E.POLICY_RULE, E.POLICY_ALL_UINT])  # all-bitset: a ~300 GB pool here
def test_hub_beyond_every_staging_area(ctx, oracle, policy):
    """Vertex 0 precedes 60,000 spread-out neighbors (its staged set fits no
    LDS table or bitmap), and a second hub has a 40,000-element child list
    streamed by every lane group; a dense core adds triangles through both."""
    rng = np.random.default_rng(21)
    nb = np.unique(rng.integers(1, 50_000_000, size=60_000))
    edges = [(0, int(v)) for v in nb]
    hub2 = 60_000_000
    edges += [(int(v), hub2) for v in nb[:40_000]]
    core = nb[:400]
    edges += [(int(core[i]), int(core[j])) for i in range(400) for j in rng.choice(400, 8) if i != j]
    s, d = oriented(edges)
    O = oracle.Trie(s, d)
    tri = oracle.join_count(plans.TRIANGLE, [O], 8)
    k4 = oracle.join_count(plans.FOUR_CLIQUE, [O], 8)
    say("hub test triangles", tri, "4-cliques", k4)
    T, R = E.Trie(ctx, s, d, policy), E.Trie(ctx, d, s, policy)
    saved = E.tune_get()
    try:
        for lanes, gap, split, tiny in ((4, 0, 0, 0), (4, 1, 1 << 21, 4096), (64, 1, 4096, 1 << 20)):
            E.tune_set(stream_lanes=lanes, gap_coded=gap, split_work=split, tiny_work=tiny)
            for form in FORMS:
                assert E.join_count(ctx, plans.TRIANGLE, [T], form) == tri
                assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form) == tri
                assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], form) == k4
    finally:
        E.tune_set(**saved)


@pytest.mark.parametrize("lanes", [2, 4, 16, 64])
def test_gap_coded_lists_at_their_limits(ctx, oracle, lanes):
    """Child lists of 16 to 5,000 values (15-value chunks: lengths on, just
    under and just over a chunk boundary), some ending in a 65,535 gap (coded)
    and some in a 65,536 gap (not coded), streamed with and without the coded
    copy; every triangle and 4-clique goes through such a list."""
    rng = np.random.default_rng(31)
    A = np.arange(1, 301)  # the hub's neighbours, pairwise joined through the lists below
    edges = [(0, int(a)) for a in A]
    lengths = [16, 29, 30, 31, 44, 45, 46, 60, 61, 100, 1000, 5000]
    for i, b in enumerate(A[:240]):
        L = lengths[i % len(lengths)]
        mine = A[A > b]
        take = rng.choice(mine, size=min(len(mine), L // 3), replace=False) if len(mine) else np.zeros(0, int)
        rest = L - len(take) - 2
        far = np.unique(rng.integers(1000, 1000 + 60 * max(rest, 1), size=max(rest, 0)))
        top = int(far.max()) if len(far) else 1000
        last_gap = 65535 if i % 3 else 65536
        tail = [top + 7, top + 7 + last_gap]
        for c in list(take) + list(far) + tail:
            edges.append((int(b), int(c)))
    s, d = oriented(edges)
    O = oracle.Trie(s, d)
    tri = oracle.join_count(plans.TRIANGLE, [O], 8)
    k4 = oracle.join_count(plans.FOUR_CLIQUE, [O], 8)
    say("gap-coded lists: triangles", tri, "4-cliques", k4)
    assert tri > 1000 and k4 > 1000
    T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
    saved = E.tune_get()
    try:
        for big_deg in (4, 16):
            for gap in (0, 1):
                E.tune_set(stream_lanes=lanes, big_deg=big_deg, gap_coded=gap, tiny_work=gap * 4096)
                for form in (E.FORM_TUNED, E.FORM_LEARNED):
                    assert E.join_count(ctx, plans.TRIANGLE, [T], form) == tri
                    assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form) == tri
                    assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], form) == k4
    finally:
        E.tune_set(**saved)
