"""Parity of the HIP engine (through the C ABI) with the CPU oracle and the
reference's golden values.  Results are exact integers, so every comparison is
bit-exact."""
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


def random_relation(rng, n_heads, rows, hi):
    src = rng.integers(0, n_heads, size=rows).astype(np.uint32)
    dst = rng.integers(0, hi, size=rows).astype(np.uint32)
    return src, dst


@pytest.mark.parametrize("policy", [E.POLICY_RULE, E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET])
def test_trie_build_matches_oracle(ctx, oracle, policy):
    rng = np.random.default_rng(policy)
    src, dst = random_relation(rng, 500, 20000, 30000)
    dst[:6000] = rng.integers(0, 800, size=6000)          # dense sets -> bitsets
    src = np.concatenate([src, src[:500], [501, 501, 502]]).astype(np.uint32)  # repeated rows, extreme values
    dst = np.concatenate([dst, dst[:500], [2**32 - 1, 2**32 - 2, 0]]).astype(np.uint32)
    say("trie build policy", policy)
    T = E.Trie(ctx, src, dst, policy)
    O = oracle.Trie(src, dst, policy)
    a, b = T.tuples()
    oa, ob = O.tuples()
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    info = T.info()
    assert info["num_rows"] == O.num_rows
    types, cards = T.node_info()
    otypes, ocards = O.node_info(len(types))
    assert np.array_equal(cards, ocards)
    assert np.array_equal(types, otypes)
    heads = np.unique(src)
    nodes = T.lookup(np.concatenate([heads, [10**6, 2**32 - 1]]))
    assert np.array_equal(nodes[: len(heads)], np.arange(len(heads))) and (nodes[len(heads):] == -1).all()


@pytest.mark.parametrize("order", ["sorted", "by_second", "shuffled", "one_swap_first", "one_swap_last"])
def test_trie_build_row_orders(ctx, oracle, order):
    """The build skips its sort for rows already in order and sorts only the
    first column's bits for rows whose second column never decreases; every
    row order gives the oracle's trie.  The swapped cases are sorted rows with
    one adjacent pair exchanged at either end."""
    rng = np.random.default_rng(17)
    src, dst = random_relation(rng, 3000, 200000, 70000)
    dst[:40000] = rng.integers(0, 5000, size=40000)  # dense sets -> bitsets
    src = np.concatenate([src, src[:1000]]).astype(np.uint32)  # repeated rows
    dst = np.concatenate([dst, dst[:1000]]).astype(np.uint32)
    if order == "by_second":
        o = np.lexsort((src, dst))  # second column ascending, first column not
    elif order == "shuffled":
        o = rng.permutation(len(src))
    else:
        o = np.lexsort((dst, src))
        if order == "one_swap_first":
            o[[0, 1]] = o[[1, 0]]
        elif order == "one_swap_last":
            o[[-2, -1]] = o[[-1, -2]]
    src, dst = src[o], dst[o]
    say("trie build row order", order)
    T = E.Trie(ctx, src, dst)
    O = oracle.Trie(src, dst, E.POLICY_RULE)
    a, b = T.tuples()
    oa, ob = O.tuples()
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    types, cards = T.node_info()
    otypes, ocards = O.node_info(len(types))
    assert np.array_equal(cards, ocards) and np.array_equal(types, otypes)


def test_trie_build_large_keys(ctx, oracle):
    rng = np.random.default_rng(3)
    src = rng.integers(0, 2**32, size=30000, dtype=np.uint64).astype(np.uint32)
    dst = rng.integers(0, 2**32, size=30000, dtype=np.uint64).astype(np.uint32)
    src[:10000] = src[0]
    T = E.Trie(ctx, src, dst)
    a, b = T.tuples()
    oa, ob = oracle.Trie(src, dst).tuples()
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    assert T.info()["dense_nav"] == 0


def set_collection(rng, sizes, hi, dense_frac=0.0):
    src, dst = [], []
    for i, n in enumerate(sizes):
        if rng.random() < dense_frac:
            base = int(rng.integers(0, max(1, hi - 64 * n)))
            v = np.unique(base + rng.integers(0, 64 * n + 1, size=n))
        else:
            v = np.unique(rng.integers(0, hi, size=n, dtype=np.int64))
        v = v[v < 2**32]
        src.append(np.full(len(v), i, np.uint32))
        dst.append(v.astype(np.uint32))
    return np.concatenate(src), np.concatenate(dst)


@pytest.mark.parametrize("pa,pb", [(E.POLICY_ALL_UINT, E.POLICY_ALL_UINT), (E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET),
                                   (E.POLICY_ALL_BITSET, E.POLICY_ALL_UINT),
                                   (E.POLICY_ALL_BITSET, E.POLICY_ALL_BITSET), (E.POLICY_RULE, E.POLICY_RULE)])
def test_intersections_all_layout_pairs(ctx, oracle, pa, pb):
    rng = np.random.default_rng(10 * pa + pb)
    sizes = np.concatenate([rng.integers(1, 40, 150), rng.integers(40, 3000, 60), [1, 2, 3, 64, 65, 2000, 5000]])
    src, dst = set_collection(rng, sizes, 200000, dense_frac=0.5)
    say("intersections", pa, pb)
    TA, TB = E.Trie(ctx, src, dst, pa), E.Trie(ctx, src, dst, pb)
    nsets = int(src.max()) + 1
    ka = rng.integers(0, nsets, 3000).astype(np.uint32)
    kb = rng.integers(0, nsets, 3000).astype(np.uint32)
    ka[:5] = [nsets + 5, 0, 1, 2, 3]  # an absent key -> empty
    sets = {i: dst[src == i] for i in range(nsets)}
    exp = [np.intersect1d(sets.get(int(x), np.zeros(0, np.uint32)), sets.get(int(y), np.zeros(0, np.uint32)))
           for x, y in zip(ka, kb)]
    cnt = E.intersect_count(ctx, TA, ka, TB, kb)
    assert np.array_equal(cnt, np.array([len(e) for e in exp], np.uint64))
    offs, vals = E.intersect(ctx, TA, ka, TB, kb)
    for i in range(len(ka)):
        assert np.array_equal(vals[offs[i]:offs[i + 1]], exp[i]), i


def test_gallop_and_merge_boundaries(ctx):
    rng = np.random.default_rng(77)
    rare = np.unique(rng.integers(0, 10**6, 10)).astype(np.uint32)
    freq = [np.unique(rng.integers(0, 10**6, k)).astype(np.uint32) for k in (200, 320, 330, 5000)]
    for f in freq:
        f = np.union1d(f, rare[::2])
        src = np.concatenate([np.zeros(len(rare), np.uint32), np.ones(len(f), np.uint32)])
        dst = np.concatenate([rare, f]).astype(np.uint32)
        T = E.Trie(ctx, src, dst, E.POLICY_ALL_UINT)
        assert E.intersect_count(ctx, T, [0], T, [1])[0] == len(np.intersect1d(rare, f))


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("policy", [E.POLICY_RULE, E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET])
def test_facebook_golden(ctx, facebook, fixtures_meta, form, policy):
    say("facebook form", form, "policy", policy)
    T = E.Trie(ctx, *facebook, policy=policy)
    assert E.join_count(ctx, plans.TRIANGLE, [T], form) == fixtures_meta["facebook_pruned"]["TriangleAgg"]
    assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], form) == fixtures_meta["facebook_pruned"]["FliqueAgg"]


@pytest.mark.parametrize("form", FORMS)
def test_facebook_duplicated_golden(ctx, facebook_dup, fixtures_meta, form):
    T = E.Trie(ctx, *facebook_dup)
    assert E.join_count(ctx, plans.TRIANGLE, [T], form) == fixtures_meta["facebook_duplicated"]["Triangle_rows"]


@pytest.mark.parametrize("scale", [10, 14, 16])
def test_rmat_matches_oracle(ctx, oracle, scale):
    s, d, n = G.rmat_graph(scale)
    say("rmat", scale, len(s))
    O = oracle.Trie(s, d)
    T = E.Trie(ctx, s, d)
    tri = oracle.join_count(plans.TRIANGLE, [O], 8)
    fq = oracle.join_count(plans.FOUR_CLIQUE, [O], 8)
    for form in FORMS:
        assert E.join_count(ctx, plans.TRIANGLE, [T], form) == tri
        assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], form) == fq


def test_er_config_matches_oracle(ctx, oracle):
    s, d, n = G.er_graph()
    O = oracle.Trie(s, d)
    exp = oracle.join_count(plans.TRIANGLE, [O], 8)
    for policy in (E.POLICY_RULE, E.POLICY_ALL_UINT):
        T = E.Trie(ctx, s, d, policy)
        for form in FORMS:
            assert E.join_count(ctx, plans.TRIANGLE, [T], form) == exp


def test_four_cycle_walks_and_two_hop(ctx, oracle):
    s, d, n = G.rmat_graph(11, orient=G.BOTH)
    O, T = oracle.Trie(s, d), E.Trie(ctx, s, d)
    exp = oracle.join_count(plans.FOUR_CYCLE_WALKS, [O], 8)
    for form in FORMS:
        assert E.join_count(ctx, plans.FOUR_CYCLE_WALKS, [T], form) == exp
    rng = np.random.default_rng(2)
    r = rng.integers(0, 3000, size=(20000, 2)).astype(np.uint32)
    t = rng.integers(0, 3000, size=(15000, 2)).astype(np.uint32)
    OR, OT = oracle.Trie(r[:, 0], r[:, 1]), oracle.Trie(t[:, 0], t[:, 1])
    TR, TT = E.Trie(ctx, r[:, 0], r[:, 1]), E.Trie(ctx, t[:, 0], t[:, 1])
    assert E.join_count(ctx, plans.TWO_HOP, [TR, TT]) == oracle.join_count(plans.TWO_HOP, [OR, OT], 4)


def test_range_restriction(ctx, oracle, facebook):
    T = E.Trie(ctx, *facebook)
    O = oracle.Trie(*facebook)
    n0 = oracle.join_level0_size(plans.TRIANGLE, [O])
    for form in FORMS:
        parts = [E.join_count(ctx, plans.TRIANGLE, [T], form, b, min(b + 700, n0)) for b in range(0, n0, 700)]
        assert sum(parts) == 1612010
        assert E.join_count(ctx, plans.TRIANGLE, [T], form, 100, 900) == oracle.join_count(plans.TRIANGLE, [O], 2, 100, 900)


def test_graph_replay_matches_eager(ctx):
    s, d, n = G.rmat_graph(15)
    T = E.Trie(ctx, s, d)
    eager = E.join_count(ctx, plans.TRIANGLE, [T], E.FORM_PLAIN)
    saved = E.tune_get()
    try:
        E.tune_set(use_graph=1)
        for i in range(4):
            assert E.join_count(ctx, plans.TRIANGLE, [T], E.FORM_TUNED) == eager
            assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], E.FORM_PLAIN) >= 0  # eager work between replays
            ctx.synchronize()
        E.tune_set(use_graph=0)
        assert E.join_count(ctx, plans.TRIANGLE, [T], E.FORM_TUNED) == eager
    finally:
        E.tune_set(**saved)


def test_query_graph_replay_matches_eager(ctx, oracle):
    """The middle-vertex plan (level 0 intersects two tries' heads, then the
    tuned level) captured whole in one graph: every replay equals the eager
    count and reports the level-0 count, with eager work between replays, after
    switching to other tries and back (a new capture each time), and after a
    larger query has grown the context's buffers (their addresses change)."""
    s, d, n = G.rmat_graph(14)
    s2, d2, n2 = G.rmat_graph(12)
    T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
    T2, R2 = E.Trie(ctx, s2, d2), E.Trie(ctx, d2, s2)
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    exp2 = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s2, d2)], 8)
    n0 = len(np.intersect1d(np.unique(s), np.unique(d)))
    saved = E.tune_get()
    try:
        E.tune_set(use_graph=0)
        c, st = E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], E.FORM_LEARNED, stats=True)
        assert c == exp and st["frontier"][0] == n0
        E.tune_set(use_graph=1)
        for form in (E.FORM_TUNED, E.FORM_LEARNED):
            for i in range(3):
                c, st = E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form, stats=True)
                assert c == exp and st["frontier"][0] == n0 and st["form_used"] == form
                assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], E.FORM_PLAIN) >= 0  # eager work between replays
                ctx.synchronize()
            assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T2, R2], form) == exp2
            assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], form) == exp
        big_s, big_d, _ = G.rmat_graph(17)
        B, BR = E.Trie(ctx, big_s, big_d), E.Trie(ctx, big_d, big_s)
        assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [B, BR], E.FORM_LEARNED) == \
            E.join_count(ctx, plans.TRIANGLE, [B], E.FORM_PLAIN)
        for i in range(3):
            assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], E.FORM_LEARNED) == exp
        B.close()
        BR.close()
    finally:
        E.tune_set(**saved)
    for t in (T, R, T2, R2):
        t.close()


@pytest.mark.parametrize("lds,big,bitmap_all,lanes,big_work,split,tiny", [
    (2048, 16, 0, 4, 16384, 0, 0), (8192, 256, 1, 64, 256, 1024, 1 << 20), (4096, 64, 1, 16, 4096, 1 << 21, 1024),
    (4096, 4, 0, 8, 64, 64, 64), (2048, 5, 1, 32, 512, 4096, 16384), (4096, 16, 0, 2, 4096, 1 << 21, 256)])
def test_tune_params_never_change_results(ctx, oracle, lds, big, bitmap_all, lanes, big_work, split, tiny):
    """Every knob setting gives the same counts; small big_work and split_work
    send many prefixes to the workgroup bin and split them into parts; tiny_work
    sends the small ones to the many-per-wavefront bin (all of them at 2^20)."""
    s, d, n = G.rmat_graph(13)
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    T = E.Trie(ctx, s, d)
    T10 = E.Trie(ctx, d, s)
    saved = E.tune_get()
    try:
        k4 = E.join_count(ctx, plans.FOUR_CLIQUE, [T], E.FORM_PLAIN)
        for gap in (0, 1):
            E.tune_set(lds_slots=lds, big_deg=big, bitmap_all=bitmap_all, grab=3, waves_per_cu=4, stream_lanes=lanes,
                       gap_coded=gap, big_work=big_work, split_work=split, tiny_work=tiny)
            assert E.join_count(ctx, plans.TRIANGLE, [T], E.FORM_TUNED) == exp
            assert E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T, T10], E.FORM_TUNED) == exp
            assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], E.FORM_TUNED) == k4
    finally:
        E.tune_set(**saved)


def test_rmat20_config_full_size(ctx, oracle):
    """R-MAT scale 20, edge factor 16: plain GPU drop-in vs the CPU oracle."""
    s, d, n = G.rmat_graph(20)
    say("rmat20 edges", len(s))
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 16)
    T = E.Trie(ctx, s, d)
    for form in FORMS:
        assert E.join_count(ctx, plans.TRIANGLE, [T], form) == exp


def test_rmat18_four_clique_config(ctx, oracle, synthetic_meta):
    """4-clique on R-MAT scale 18, edge factor 16: every form, against the oracle."""
    s, d, n = G.rmat_graph(18)
    meta = synthetic_meta["rmat18"]
    assert len(s) == meta["edges"]
    exp = oracle.join_count(plans.FOUR_CLIQUE, [oracle.Trie(s, d)], 16)
    assert exp == meta["four_cliques"]
    for policy in (E.POLICY_RULE, E.POLICY_LEARNED):
        T = E.Trie(ctx, s, d, policy)
        for form in FORMS:
            say("rmat18 4-clique policy", policy, "form", form)
            assert E.join_count(ctx, plans.FOUR_CLIQUE, [T], form) == exp


@pytest.mark.parametrize("form", FORMS)
def test_materialize_facebook_golden(ctx, oracle, facebook, fixtures_meta, form):
    """Materialized Triangle and Flique on facebook: the reference's row counts
    and first rows, and every row equal to the oracle's enumeration."""
    meta = fixtures_meta["facebook_pruned"]
    O = oracle.Trie(*facebook)
    for policy in (E.POLICY_RULE, E.POLICY_ALL_BITSET):
        T = E.Trie(ctx, *facebook, policy=policy)
        for levels, rows_key, row0_key in ((plans.TRIANGLE, "Triangle_rows", "Triangle_row0"),
                                           (plans.FOUR_CLIQUE, "Flique_rows", "Flique_row0")):
            R = E.join_materialize(ctx, levels, [T], form)
            assert len(R) == meta[rows_key] and list(R[0]) == meta[row0_key]
            assert np.array_equal(R, oracle.join_materialize(levels, [O]))


def test_materialize_matches_oracle_and_counts(ctx, oracle):
    s, d, n = G.rmat_graph(12)
    O, T = oracle.Trie(s, d), E.Trie(ctx, s, d)
    for levels in (plans.TRIANGLE, plans.FOUR_CLIQUE):
        R = E.join_materialize(ctx, levels, [T], E.FORM_TUNED)
        assert len(R) == E.join_count(ctx, levels, [T])
        assert np.array_equal(R, oracle.join_materialize(levels, [O]))
    rng = np.random.default_rng(9)
    r = rng.integers(0, 500, size=(3000, 2)).astype(np.uint32)
    t = rng.integers(0, 500, size=(2000, 2)).astype(np.uint32)
    OR, OT = oracle.Trie(r[:, 0], r[:, 1]), oracle.Trie(t[:, 0], t[:, 1])
    TR_, TT = E.Trie(ctx, r[:, 0], r[:, 1]), E.Trie(ctx, t[:, 0], t[:, 1])
    assert np.array_equal(E.join_materialize(ctx, plans.TWO_HOP, [TR_, TT]),
                          oracle.join_materialize(plans.TWO_HOP, [OR, OT]))
    # an empty result: disjoint relations
    a = E.Trie(ctx, np.array([1, 2], np.uint32), np.array([3, 4], np.uint32))
    b = E.Trie(ctx, np.array([7, 8], np.uint32), np.array([9, 9], np.uint32))
    assert E.join_materialize(ctx, plans.TWO_HOP, [a, b]).shape == (0, 3)


@pytest.mark.parametrize("scale", [12, 16])
def test_triangle_by_middle_vertex(ctx, oracle, scale):
    """The triangle ordered by its middle vertex over Edge_0_1 and Edge_1_0:
    the tuned kernel's driver-set shape.  Same count as the (a, b, c) order."""
    s, d, n = G.rmat_graph(scale)
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    assert oracle.join_count(plans.TRIANGLE_BY_MIDDLE, [oracle.Trie(s, d), oracle.Trie(d, s)], 8) == exp
    for policy in (E.POLICY_RULE, E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET):
        T01, T10 = E.Trie(ctx, s, d, policy), E.Trie(ctx, d, s, policy)
        for form in FORMS:
            c, st = E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T01, T10], form, stats=True)
            assert c == exp
            if form != E.FORM_PLAIN:
                assert st["form_used"] == form  # the driver-set shape runs in the tuned kernel


@pytest.mark.parametrize("seed", [1, 2])
def test_trimmed_child_lists_any_orientation(ctx, oracle, seed):
    """The tuned level streams a child list of x from its first value not below
    E(p)'s parent value (link hints) only when S(p)'s first value is not below
    that parent value.  On directed graphs with edges both ways, prefixes of
    both kinds occur; counts equal the oracle's with the hints used by the
    persistent kernel and by the tiny bin, streamed plain and gap-coded."""
    rng = np.random.default_rng(seed)
    n, m = 3000, 60000
    hubs = rng.integers(0, 40, m // 3)  # a few long child lists (gap-coded, streamed by lane groups)
    src = np.concatenate([rng.integers(0, n, m), hubs]).astype(np.uint32)
    dst = np.concatenate([rng.integers(0, n, m), rng.integers(0, n, m // 3)]).astype(np.uint32)
    keep = src != dst
    e = np.unique(np.stack([src[keep], dst[keep]], 1), axis=0)
    s, d = e[:, 0].copy(), e[:, 1].copy()
    exp = oracle.join_count(plans.TRIANGLE_BY_MIDDLE, [oracle.Trie(s, d), oracle.Trie(d, s)], 8)
    assert exp == oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    saved = E.tune_get()
    try:
        for policy in (E.POLICY_RULE, E.POLICY_ALL_UINT):
            T01, T10 = E.Trie(ctx, s, d, policy), E.Trie(ctx, d, s, policy)
            assert T10.link_to(T01)
            for knobs in ({"tiny_work": 0, "gap_coded": 1, "big_deg": 16}, {"tiny_work": 4096, "gap_coded": 1},
                          {"tiny_work": 0, "gap_coded": 0, "big_deg": 8}, {"tiny_work": 256, "gap_coded": 1,
                                                                           "stream_lanes": 4, "big_work": 64}):
                E.tune_set(**knobs)
                for form in (E.FORM_TUNED, E.FORM_LEARNED):
                    c, st = E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T01, T10], form, stats=True)
                    assert st["form_used"] == form
                    assert c == exp, (policy, knobs, form)
            T01.close()
            T10.close()
    finally:
        E.tune_set(**saved)


def dense_block_graph(rng, blocks=24, width=300, p=0.3, extra=3000):
    """Dense communities of `width` consecutive ids (edge probability p) plus
    sparse random edges between them, both directions kept, no loops."""
    src, dst = [], []
    for j in range(blocks):
        a = rng.integers(0, width, int(p * width * width)) + j * width
        b = rng.integers(0, width, int(p * width * width)) + j * width
        src.append(a)
        dst.append(b)
    n = blocks * width
    src.append(rng.integers(0, n, extra))
    dst.append(rng.integers(0, n, extra))
    src, dst = np.concatenate(src), np.concatenate(dst)
    keep = src != dst
    e = np.unique(np.stack([src[keep], dst[keep]], 1).astype(np.uint32), axis=0)
    return e[:, 0].copy(), e[:, 1].copy()


def lists_read_as_words(s, d):
    """Child lists the tuned level reads as bitset words: at least 16 values,
    a range_bitset by the density rule, at least 4 values per 64-bit word."""
    heads, start = np.unique(s, return_index=True)
    n = np.diff(np.append(start, len(s)))
    first, last = d[start].astype(np.int64), d[start + n - 1].astype(np.int64)
    words = (last >> 6) - (first >> 6) + 1
    dense = (n > 2) & (last - first <= 128 * n)
    return int(((n >= 16) & dense & (4 * words <= n)).sum())


@pytest.mark.parametrize("seed", [5, 6])
def test_dense_lists_read_as_words(ctx, oracle, seed):
    """Dense child lists are read as bitset words ANDed with a bitmap staging
    of S(p): lists inside one community, lists reaching into others (their
    words overlap the staged window only in part) and staged windows starting
    at any 32-bit word.  Counts equal the oracle's in both attribute orders,
    both layouts that give words, every stream width, with and without gap
    coding and with every set staged as a bitmap where it fits."""
    rng = np.random.default_rng(seed)
    s, d = dense_block_graph(rng)
    assert lists_read_as_words(s, d) > 1000
    o, r = s < d, s > d
    so, do_ = s[o], d[o]  # oriented half for the (a, b, c) order
    exp = oracle.join_count(plans.TRIANGLE, [oracle.Trie(so, do_)], 8)
    exp_mid = oracle.join_count(plans.TRIANGLE_BY_MIDDLE, [oracle.Trie(s, d), oracle.Trie(d, s)], 8)
    say("dense blocks", len(s), "rows; triangles", exp, exp_mid, "; lists read as words", lists_read_as_words(s, d))
    saved = E.tune_get()
    try:
        for policy in (E.POLICY_RULE, E.POLICY_ALL_BITSET):
            T = E.Trie(ctx, so, do_, policy)
            T01, T10 = E.Trie(ctx, s, d, policy), E.Trie(ctx, d, s, policy)
            for lanes in (2, 4, 64):
                for gap, bitmap_all in ((1, 0), (0, 1)):
                    E.tune_set(stream_lanes=lanes, gap_coded=gap, bitmap_all=bitmap_all, tiny_work=0, big_deg=16)
                    for form in (E.FORM_TUNED, E.FORM_LEARNED):
                        assert E.join_count(ctx, plans.TRIANGLE, [T], form) == exp, (policy, lanes, gap, form)
                        c, st = E.join_count(ctx, plans.TRIANGLE_BY_MIDDLE, [T01, T10], form, stats=True)
                        assert st["form_used"] == form
                        assert c == exp_mid, (policy, lanes, gap, form)
            for t in (T, T01, T10):
                t.close()
    finally:
        E.tune_set(**saved)


# ---- equality selections ------------------------------------------------------

@pytest.mark.parametrize("form", FORMS)
def test_selection_goldens(ctx, facebook_dup, fixtures_meta, form):
    """FliqueSelAgg (x=0) and BarbellSelAgg (p=6) on the symmetric facebook
    graph: the counts travis.py asserts (385-400, 473-489)."""
    s, d = facebook_dup
    T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
    meta = fixtures_meta["facebook_duplicated"]
    assert E.join_count(ctx, plans.flique_sel(0), [T, R], form) == meta["FliqueSelAgg_x0"]
    assert E.join_count(ctx, plans.barbell_sel(6), [T, R], form) == meta["BarbellSelAgg_p6"]


def test_selection_rows(ctx, oracle, facebook_dup, fixtures_meta):
    """FliqueSel(a,b,c,d) rows (travis.py:429-444: 3,759,972) equal the oracle's, in order."""
    s, d = facebook_dup
    levels = plans.flique_sel(0)
    rows = E.join_materialize(ctx, levels, [E.Trie(ctx, s, d), E.Trie(ctx, d, s)], E.FORM_TUNED)
    assert rows.shape == (fixtures_meta["facebook_duplicated"]["FliqueSelAgg_x0"], 4)
    assert np.array_equal(rows, oracle.join_materialize(levels, [oracle.Trie(s, d), oracle.Trie(d, s)]))


def test_selections_vs_oracle(ctx, oracle):
    """Selections on a symmetric R-MAT graph at a hub, at ordinary vertices and
    at an absent value, at level 0, at a later level, as a level's only operand
    (the driver) and as the last level's only operand."""
    s, d, n = G.rmat_graph(12)
    s, d = G.symmetric(s, d)
    T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
    O = [oracle.Trie(s, d), oracle.Trie(d, s)]
    deg = np.bincount(s.astype(np.int64), minlength=n)
    nodes = [int(np.argmax(deg)), int(deg.argsort()[len(deg) // 2]), 7, n + 1000]
    for node in nodes:
        cases = [plans.flique_sel(node),
                 [[(0, -1), (0, -1)], [(0, 0), (0, -1), (1, -2, node)], [(0, 1), (0, 0)]],  # triangles, b ~ node
                 [[(1, -2, node)], [(0, 0)]],                                             # 2-paths from node
                 [[(0, -1)], [(1, -2, node)]]]                                            # heads x N(node)
        # the barbell enumerates (ordered triangles at N(node))^2 results: at the hub
        # of this graph 7.7e12, so only where that is small (vertex 7: 2.8e8)
        if node in (7, n + 1000):
            cases.append(plans.barbell_sel(node))
        for levels in cases:
            exp = oracle.join_count(levels, O, 8)
            for form in FORMS:
                assert E.join_count(ctx, levels, [T, R], form) == exp, (node, levels, form)


# ---- two-bag plans (lollipop, barbell) -----------------------------------------

@pytest.mark.parametrize("form", FORMS)
def test_two_bag_goldens(ctx, facebook_dup, fixtures_meta, form):
    """LollipopAgg = 1,426,911,480 and BarbellAgg = 20,371,831,447,136 on the
    symmetric facebook graph (travis.py:14-70): triangle bag per vertex, then
    the weighted edge fold."""
    from emptyheaded_amd import ghd
    s, d = facebook_dup
    S = E.Trie(ctx, s, d)
    meta = fixtures_meta["facebook_duplicated"]
    for method in ("direct", "oriented", "auto"):
        assert ghd.lollipop_count(ctx, S, form, method) == meta["LollipopAgg"]
        assert ghd.barbell_count(ctx, S, form, method) == meta["BarbellAgg"]


def test_triangle_by_vertex_oriented(ctx, oracle, monkeypatch):
    """Per-vertex triangle counts of the symmetric bag from the three oriented
    passes equal the direct bag's and the oracle's, vertex by vertex, on R-MAT
    graphs; with a loop or a one-way row the automatic choice falls back to
    the direct bag (and still agrees with the oracle)."""
    from emptyheaded_amd import ghd
    for scale in (10, 14):
        s, d, n = G.rmat_graph(scale)
        S, D = G.symmetric(s, d)
        nv = int(max(S.max(), D.max())) + 1
        exp = np.zeros(nv, np.int64)
        exp[np.unique(S)] = oracle.join_count_per_head(plans.TRIANGLE, [oracle.Trie(S, D)], 8)
        T = E.Trie(ctx, S, D)
        for form in (E.FORM_PLAIN, E.FORM_TUNED):
            for method in ("direct", "oriented"):
                got, total = ghd.triangle_by_vertex(ctx, T, form, method)
                assert np.array_equal(got, exp) and total == int(exp.sum()), (scale, form, method)
        lo, rv, cnt = T.orient()
        assert cnt == [len(s), len(s), 0, 0]
        assert lo is not None and rv is not None and lo.num_rows == rv.num_rows == len(s)
        lo.close()
        rv.close()
        lo2, rv2, cnt2 = T.orient(build=False)
        assert lo2 is None and rv2 is None and cnt2 == cnt
        # a loop and a one-way row: the identity does not hold, "auto" goes direct
        # (at any size: the size threshold is lifted here)
        monkeypatch.setattr(ghd, "ORIENT_MIN_ROWS", 0)
        got, _ = ghd.triangle_by_vertex(ctx, T, E.FORM_TUNED, "auto")
        assert np.array_equal(got, exp)
        for extra in ((np.array([5], np.uint32), np.array([5], np.uint32)),
                      (np.array([1], np.uint32), np.array([nv + 3], np.uint32))):
            S2, D2 = np.concatenate([S, extra[0]]), np.concatenate([D, extra[1]])
            nv2 = int(max(S2.max(), D2.max())) + 1
            exp2 = np.zeros(nv2, np.int64)
            exp2[np.unique(S2)] = oracle.join_count_per_head(plans.TRIANGLE, [oracle.Trie(S2, D2)], 8)
            got2, _ = ghd.triangle_by_vertex(ctx, E.Trie(ctx, S2, D2), E.FORM_TUNED, "auto")
            assert np.array_equal(got2, exp2)


def test_orient_rejects_balanced_asymmetric(ctx, oracle, monkeypatch):
    """As many one-way rows x < y as x > y, no loop: the relation is not
    symmetric, so the oriented passes do not apply.  The device's exact check
    (every x < y row needs its reverse) sends "auto" to the direct bag, whose
    per-vertex counts equal the oracle's, and "oriented" raises.  Covered on a
    six-row gadget ({0,1,2} one way up, {4,5,6} one way down: 2 ordered
    triangles, where the oriented identity would give 6) and on a symmetric
    R-MAT graph with one one-way row added in each direction."""
    from emptyheaded_amd import ghd
    monkeypatch.setattr(ghd, "ORIENT_MIN_ROWS", 0)
    gad_s = np.array([0, 1, 0, 5, 6, 6], np.uint32)
    gad_d = np.array([1, 2, 2, 4, 5, 4], np.uint32)
    s, d, n = G.rmat_graph(12)
    S, D = G.symmetric(s, d)
    nv = int(max(S.max(), D.max())) + 1
    big_s = np.concatenate([S, np.array([1, nv + 7], np.uint32)])
    big_d = np.concatenate([D, np.array([nv + 3, 2], np.uint32)])
    for src, dst in ((gad_s, gad_d), (big_s, big_d)):
        T = E.Trie(ctx, src, dst)
        lo, rv, cnt = T.orient()
        assert lo is None and rv is None
        assert cnt[0] == cnt[1] and cnt[2] == 0 and cnt[3] > 0, cnt
        nv2 = int(max(src.max(), dst.max())) + 1
        exp = np.zeros(nv2, np.int64)
        exp[np.unique(src)] = oracle.join_count_per_head(plans.TRIANGLE, [oracle.Trie(src, dst)], 8)
        for form in FORMS:
            got, total = ghd.triangle_by_vertex(ctx, T, form, "auto")
            assert np.array_equal(got, exp) and total == int(exp.sum()), form
            with pytest.raises(ValueError):
                ghd.triangle_by_vertex(ctx, T, form, "oriented")
        if len(src) == 6:
            assert int(exp.sum()) == 2
            assert ghd.lollipop_count(ctx, T, E.FORM_TUNED) == int(exp[src.astype(np.int64)].sum())
        T.close()


def test_weighted_pairs_against_numpy(ctx):
    """The fold over random int64 weights (wrapping mod 2^64), weights shorter
    than the value range (missing values weigh 0), on a graph with hubs."""
    s, d, n = G.rmat_graph(12)
    T = E.Trie(ctx, s, d)
    rng = np.random.default_rng(41)
    for na, nb in ((n, n), (n // 3, n), (n, 17), (0, n)):
        wa = rng.integers(-2**40, 2**40, size=na, dtype=np.int64)
        wb = rng.integers(-2**40, 2**40, size=nb, dtype=np.int64)
        fa = np.zeros(n, np.int64)
        fa[:na] = wa
        fb = np.zeros(n, np.int64)
        fb[:nb] = wb
        with np.errstate(over="ignore"):
            exp = int((fa[s.astype(np.int64)] * fb[d.astype(np.int64)]).sum())
        assert E.weighted_pairs(ctx, T, wa, wb) == exp


# ---- single-source closure (SSSP) ------------------------------------------------

def test_sssp_golden_and_oracle(ctx, oracle, facebook_dup):
    """BFS levels: the SSSP check of travis.py:556-567 on the symmetric facebook
    graph, and the oracle's levels on R-MAT graphs (directed and both
    directions) from a hub, an ordinary vertex, a vertex with no children and a
    value that is not a head, with and without an iteration limit."""
    s, d = facebook_dup
    n = int(max(s.max(), d.max())) + 1
    dist, _ = E.sssp_unweighted(ctx, E.Trie(ctx, s, d), 0, n)
    assert np.array_equal(dist, oracle.sssp_unweighted(s, d, 0, n)[0])
    Tf = E.Trie(ctx, s, d)
    assert np.array_equal(E.sssp_unweighted(ctx, Tf, 0, n, rev=Tf, bottom_up=32)[0], dist)
    visited = np.flatnonzero(dist != E.UNREACHED)
    assert dist[visited[1000]] == 2
    for sym in (False, True):
        s, d, n = G.rmat_graph(16)
        if sym:
            s, d = G.symmetric(s, d)
        T = E.Trie(ctx, s, d)
        deg = np.bincount(s.astype(np.int64), minlength=n)
        starts = [int(np.argmax(deg)), int(np.flatnonzero(deg)[len(np.flatnonzero(deg)) // 2])]
        starts += [int(np.flatnonzero(deg == 0)[0])] if (deg == 0).any() else []
        R = T if sym else E.Trie(ctx, d, s)  # the reversed relation
        for start in starts:
            for it in (-1, 1, 3):
                exp, elv = oracle.sssp_unweighted(s, d, start, n, it)
                # top-down; direction-optimizing at the default switch; bottom-up
                # for every level after the first (n // 2**30 == 0 vertices)
                for rev, bu in ((None, 0), (R, 32), (R, 2 ** 30)):
                    got, lv = E.sssp_unweighted(ctx, T, start, n, it, rev=rev, bottom_up=bu)
                    assert np.array_equal(got, exp), (sym, start, it, bu)
                    assert lv == elv


@pytest.mark.parametrize("form", FORMS)
def test_triangle_projection_golden(ctx, oracle, facebook, fixtures_meta, form):
    """TriangleProj(a,b) (travis.py:282-300) through ehx_join_materialize: the
    reference's 85,658 rows and first row (6,5), every row equal to the
    oracle's, under the density rule and all-bitset layouts."""
    meta = fixtures_meta["facebook_pruned"]
    exp = oracle.join_materialize(plans.TRIANGLE_PROJ, [oracle.Trie(*facebook)])
    for policy in (E.POLICY_RULE, E.POLICY_ALL_BITSET):
        T = E.Trie(ctx, *facebook, policy=policy)
        R = E.join_materialize(ctx, plans.TRIANGLE_PROJ, [T], form)
        assert len(R) == meta["TriangleProj_rows"] and list(R[0]) == meta["TriangleProj_row0"]
        assert np.array_equal(R, exp)
        assert E.join_count(ctx, plans.TRIANGLE_PROJ, [T], form) == len(exp)
        T.close()


def test_batched_node_counts_every_layout_pair(ctx):
    """ehx_intersect_nodes_count (the drop-in's batched build_aggregated_set):
    pairs of small sets (16 lanes a pair), pairs above that bound (a wavefront
    a pair), head sets, empty sets and three-way calls, against numpy, for
    every layout pair."""
    rng = np.random.default_rng(23)
    sizes = np.concatenate([rng.integers(1, 60, 300), rng.integers(200, 600, 60), rng.integers(2000, 9000, 8)])
    src, dst = set_collection(rng, sizes, 400000, dense_frac=0.4)
    nsets = int(src.max()) + 1
    sets = [dst[src == i] for i in range(nsets)]
    for pa, pb in ((E.POLICY_ALL_UINT, E.POLICY_ALL_UINT), (E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET),
                   (E.POLICY_ALL_BITSET, E.POLICY_ALL_BITSET), (E.POLICY_RULE, E.POLICY_RULE)):
        TA, TB = E.Trie(ctx, src, dst, pa), E.Trie(ctx, src, dst, pb)
        heads_a = TA.info()["num_heads"]
        na = rng.integers(0, nsets, 5000).astype(np.uint32)
        nb = rng.integers(0, nsets, 5000).astype(np.uint32)
        na[:3] = heads_a  # the head set
        got = E.intersect_nodes_count(ctx, [TA, TB], [na, nb])
        head = np.unique(src).astype(np.uint32)
        A = [head if x == heads_a else sets[x] for x in na]
        exp = np.array([len(np.intersect1d(a, sets[y])) for a, y in zip(A, nb)], np.uint64)
        assert np.array_equal(got, exp), (pa, pb)
        nc = rng.integers(0, nsets, 5000).astype(np.uint32)
        got3 = E.intersect_nodes_count(ctx, [TA, TB, TA], [na, nb, nc])
        exp3 = np.array([len(np.intersect1d(np.intersect1d(a, sets[y]), sets[z])) for a, y, z in zip(A, nb, nc)],
                        np.uint64)
        assert np.array_equal(got3, exp3), (pa, pb)
        TA.close()
        TB.close()
