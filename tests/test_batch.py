"""Batched uneven workload: ego-networks cut from an R-MAT graph, each with its
own triangle count and 4-cycle count.

CPU tests pin the generator's contract and the 4-cycle oracle (brute force over
4-vertex subsets on tiny graphs; the closed-walk identity through the join
oracle on a larger one).  GPU tests compare one batched launch per query kind
(per-segment counts) with the oracle and with one launch per net."""
import itertools
import json
import os
import sys

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G
from emptyheaded_amd import plans

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORMS = [E.FORM_PLAIN, E.FORM_TUNED, E.FORM_LEARNED]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def brute_cycles4(n, edges):
    adj = np.zeros((n, n), dtype=bool)
    for u, v in edges:
        adj[u, v] = adj[v, u] = True
    c = 0
    for a, b, x, d in itertools.combinations(range(n), 4):
        for p in ((a, b, x, d), (a, b, d, x), (a, x, b, d)):
            if adj[p[0], p[1]] and adj[p[1], p[2]] and adj[p[2], p[3]] and adj[p[3], p[0]]:
                c += 1
    return c


def sym_sorted(edges):
    s = np.array([u for u, v in edges] + [v for u, v in edges], dtype=np.uint32)
    d = np.array([v for u, v in edges] + [u for u, v in edges], dtype=np.uint32)
    o = np.lexsort((d, s))
    return s[o], d[o]


def per_net(values_by_vertex_or_head, keys, base):
    """Sum int64 values into nets by key (vertex id) -> net index."""
    net = np.searchsorted(base, keys, side="right") - 1
    out = np.zeros(len(base) - 1, dtype=np.int64)
    np.add.at(out, net, np.asarray(values_by_vertex_or_head, dtype=np.int64))
    return out


def small_batch(scale=14, nets=64, lo=64, hi=4096):
    s, d, n = G.rmat_graph(scale, orient=G.BOTH)
    return G.egonets(s, d, n, num_nets=nets, min_edges=lo, max_edges=hi, seed=G.EGO_SEED + scale)


# ---- CPU ------------------------------------------------------------------------

def test_cycle4_oracle_brute_force(oracle):
    rng = np.random.default_rng(4)
    for _ in range(40):
        n = int(rng.integers(4, 11))
        p = rng.random()
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        if not edges:
            continue
        s, d = sym_sorted(edges)
        assert oracle.cycle4_count(s, d, n, 2) == brute_cycles4(n, edges)


def test_cycle4_oracle_closed_forms(oracle):
    # K_n has 3 * C(n, 4) four-cycles; the complete bipartite K_{p,q} has C(p,2) * C(q,2)
    for n in (4, 5, 9):
        s, d = sym_sorted([(u, v) for u in range(n) for v in range(u + 1, n)])
        assert oracle.cycle4_count(s, d, n, 1) == 3 * (n * (n - 1) * (n - 2) * (n - 3) // 24)
    p, q = 5, 7
    s, d = sym_sorted([(u, p + v) for u in range(p) for v in range(q)])
    assert oracle.cycle4_count(s, d, p + q, 1) == (p * (p - 1) // 2) * (q * (q - 1) // 2)
    # a 4-cycle per vertex is charged to its largest vertex
    s, d = sym_sorted([(0, 1), (1, 2), (2, 3), (0, 3)])
    tot, pv = oracle.cycle4_count(s, d, 4, 1, per_vertex=True)
    assert tot == 1 and list(pv) == [0, 0, 0, 1]


def test_cycle4_oracle_matches_closed_walk_identity(oracle):
    s, d, n = G.rmat_graph(11, orient=G.BOTH)
    walks = oracle.join_count(plans.FOUR_CYCLE_WALKS, [oracle.Trie(s, d)], 8)
    assert oracle.cycle4_count(s, d, n, 8) == plans.four_cycles(walks, np.bincount(s, minlength=n))


def test_egonets_contract():
    s, d, n = G.rmat_graph(14, orient=G.BOTH)
    es, ed, base, off = G.egonets(s, d, n, num_nets=48, min_edges=64, max_edges=4096, seed=99)
    es2, ed2, base2, off2 = G.egonets(s, d, n, num_nets=48, min_edges=64, max_edges=4096, seed=99)
    assert np.array_equal(es, es2) and np.array_equal(ed, ed2) and np.array_equal(base, base2)  # deterministic
    assert off[0] == 0 and off[-1] == len(es) and base[0] == 0
    m = np.diff(off.astype(np.int64))
    assert m.min() >= 1 and m.max() <= 4096
    assert np.all(es < ed)  # oriented
    order = np.lexsort((ed, es))
    assert np.array_equal(order, np.arange(len(es)))  # sorted rows
    key = es.astype(np.uint64) << np.uint64(32) | ed.astype(np.uint64)
    assert len(np.unique(key)) == len(key)  # no duplicate edges
    for q in range(48):
        a, b = int(off[q]), int(off[q + 1])
        # every edge of net q stays inside its vertex range, and every vertex has an edge
        assert es[a:b].min() >= base[q] and ed[a:b].max() < base[q + 1]
        used = np.union1d(es[a:b], ed[a:b])
        assert len(used) == base[q + 1] - base[q]
    # the nets that reached their target are exact; most do
    assert np.mean(m >= 64) > 0.9


def test_egonet_of_a_clique_is_a_clique_prefix():
    """Inside a clique every joined neighbor links to all vertices already in,
    so a net trimmed to t edges spans the fewest vertices that can hold t."""
    k = 40
    edges = [(u, v) for u in range(k) for v in range(u + 1, k)]
    s, d = sym_sorted(edges)
    es, ed, base, off = G.egonets(s, d, k, num_nets=3, min_edges=100, max_edges=100, seed=5)
    for q in range(3):
        a, b = int(off[q]), int(off[q + 1])
        assert b - a == 100
        nv = int(base[q + 1] - base[q])
        # inside a clique a radius-1 net on nv vertices holds C(nv-1, 2) < 100 <= C(nv, 2) edges
        assert (nv - 1) * (nv - 2) // 2 < 100 <= nv * (nv - 1) // 2


# ---- GPU ------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_segment_counts_match_oracle(ctx, oracle, form):
    es, ed, base, off = small_batch()
    say("small ego batch", len(es), "edges", int(base[-1]), "vertices")
    O = oracle.Trie(es, ed)
    ph = oracle.join_count_per_head(plans.TRIANGLE, [O], 8)
    heads = np.unique(es)
    expect = per_net(ph, heads, base)
    T = E.Trie(ctx, es, ed)
    counts, total = E.join_count_segments(ctx, plans.TRIANGLE, [T], base, form)
    assert total == int(ph.sum())
    assert np.array_equal(counts, expect)
    if form != E.FORM_PLAIN:  # prefixes split into parts add to their segment part by part
        saved = E.tune_get()
        try:
            E.tune_set(big_work=64, split_work=64, tiny_work=0)
            c1, t1 = E.join_count_segments(ctx, plans.TRIANGLE, [T], base, form)
            E.tune_set(big_work=1 << 20, tiny_work=1 << 20)  # every prefix with |S| <= 128 many per wavefront
            c2, t2 = E.join_count_segments(ctx, plans.TRIANGLE, [T], base, form)
        finally:
            E.tune_set(**saved)
        assert t1 == total and np.array_equal(c1, expect)
        assert t2 == total and np.array_equal(c2, expect)
    # 4-clique per net through the same entry point
    ph4 = oracle.join_count_per_head(plans.FOUR_CLIQUE, [O], 8)
    c4, t4 = E.join_count_segments(ctx, plans.FOUR_CLIQUE, [T], base, form)
    assert t4 == int(ph4.sum()) and np.array_equal(c4, per_net(ph4, heads, base))


@pytest.mark.gpu
def test_segment_counts_equal_per_query_launches(ctx):
    es, ed, base, off = small_batch(nets=24)
    T = E.Trie(ctx, es, ed)
    heads = np.unique(es)
    lo = np.searchsorted(heads, base[:-1])
    hi = np.searchsorted(heads, base[1:])
    counts, total = E.join_count_segments(ctx, plans.TRIANGLE, [T], base, E.FORM_LEARNED)
    for q in range(24):
        assert counts[q] == E.join_count(ctx, plans.TRIANGLE, [T], E.FORM_TUNED, int(lo[q]), int(hi[q]))
    # bounds that cover part of the value range, including empty segments
    b = np.array([base[3], base[3], base[5], base[9]], dtype=np.uint32)
    c2, _ = E.join_count_segments(ctx, plans.TRIANGLE, [T], b, E.FORM_PLAIN)
    assert list(c2) == [0, counts[3] + counts[4], counts[5:9].sum()]


@pytest.mark.gpu
def test_segment_bounds_are_checked(ctx):
    s, d, n = G.rmat_graph(8)
    T = E.Trie(ctx, s, d)
    with pytest.raises(E.EhxError):
        E.join_count_segments(ctx, plans.TRIANGLE, [T], np.array([5, 3], dtype=np.uint32))
    with pytest.raises(E.EhxError):
        E.join_count_segments(ctx, plans.TRIANGLE, [T], np.array([5], dtype=np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [8, 12, 15])
def test_cycle4_matches_oracle(ctx, oracle, scale):
    s, d, n = G.rmat_graph(scale, orient=G.BOTH)
    say("cycle4 rmat", scale, len(s))
    exp, pv = oracle.cycle4_count(s, d, n, 8, per_vertex=True)
    T = E.Trie(ctx, s, d)
    got, st = E.cycle4_count(ctx, T, stats=True)
    say("  overflowed vertices", st["frontier"][1])
    assert got == exp
    # per-vertex split through segments of one vertex each (first 200 vertices) and halves
    b = np.arange(0, min(n, 200) + 1, dtype=np.uint32)
    c, _ = E.cycle4_count(ctx, T, bounds=b)
    assert np.array_equal(c, pv[: len(b) - 1])
    half = np.array([0, n // 2, n], dtype=np.uint32)
    c, tot = E.cycle4_count(ctx, T, bounds=half)
    assert tot == exp and c[0] == pv[: n // 2].sum() and c[1] == pv[n // 2:].sum()


@pytest.mark.gpu
@pytest.mark.parametrize("spread", [1, 12, 2000])
def test_cycle4_hub_paths(ctx, oracle, spread):
    """Vertices past the wavefront bin: a dense core plus a wide star.  With
    spread 1 the hub's endpoint window fits the dense LDS counters; with ids
    spread out 12x it does not, and the hub goes through the global table in
    two hash-partitioned passes.  Spread 2000 leaves the ids too sparse for a
    child lookup by value, so the kernels search each neighbor's list instead
    of reading the relation's child pointers into itself."""
    rng = np.random.default_rng(11)
    core = 2000
    edges = set()
    for u in range(core):
        for v in rng.choice(core, size=100, replace=False):
            if u != v:
                edges.add((min(u, int(v)), max(u, int(v))))
    hub = core + 5000
    for v in range(core + 5000):
        edges.add((v, hub))
    for v in range(core, core + 5000):  # the star's leaves touch the core too
        edges.add((int(rng.integers(0, core)), v))
    edges = sorted((u * spread, v * spread) for u, v in edges)
    s, d = sym_sorted(edges)
    n = hub * spread + 1
    exp, pv = oracle.cycle4_count(s, d, n, 8, per_vertex=True)
    T = E.Trie(ctx, s, d, E.POLICY_ALL_UINT)
    got, st = E.cycle4_count(ctx, T, stats=True)
    say("hub test spread", spread, "workgroup-bin vertices", st["frontier"][1])
    assert st["frontier"][1] >= 1
    assert got == exp
    # range restriction by head index: the last head is the hub
    nh = len(np.unique(s))
    assert E.cycle4_count(ctx, T, 0, nh - 1) == exp - pv[hub * spread]


@pytest.mark.gpu
def test_egonet_batch_full_size(ctx, oracle, synthetic_meta):
    """4,096 ego-networks from R-MAT scale 22: batched per-net triangle and
    4-cycle counts equal the oracle's, net by net."""
    s, d, n = G.rmat_graph(22, orient=G.BOTH)
    es, ed, base, off = G.egonets(s, d, n)
    del s, d
    meta = synthetic_meta["egonets_rmat22"]
    assert len(es) == meta["edges"] and int(base[-1]) == meta["vertices"]
    O = oracle.Trie(es, ed)
    heads = np.unique(es)
    tri_exp = per_net(oracle.join_count_per_head(plans.TRIANGLE, [O], 16), heads, base)
    del O
    ss, sd = G.symmetric(es, ed)
    _, pv = oracle.cycle4_count(ss, sd, int(base[-1]), 16, per_vertex=True)
    cyc_exp = per_net(pv, np.arange(int(base[-1])), base)
    assert int(tri_exp.sum()) == meta["triangles"] and int(cyc_exp.sum()) == meta["cycles4"]
    T = E.Trie(ctx, es, ed)
    R = E.Trie(ctx, ed, es)
    S = E.Trie(ctx, ss, sd)
    for form in (E.FORM_PLAIN, E.FORM_LEARNED):
        tri, _ = E.join_count_segments(ctx, plans.TRIANGLE, [T], base, form)
        assert np.array_equal(tri, tri_exp)
    # the middle vertex of a triangle lies in the same net: same per-net counts
    tri, _ = E.join_count_segments(ctx, plans.TRIANGLE_BY_MIDDLE, [T, R], base, E.FORM_LEARNED)
    assert np.array_equal(tri, tri_exp)
    cyc, tot = E.cycle4_count(ctx, S, bounds=base)
    assert np.array_equal(cyc, cyc_exp)
