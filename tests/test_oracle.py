"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: golden constants from
the reference's tests, scipy sparse-matrix algebra, numpy set operations,
brute force on tiny inputs, and closed forms.
"""
import itertools
import os

import numpy as np
import pytest
import scipy.sparse as sp

from emptyheaded_amd import plans

RNG = np.random.default_rng(1234)


def rand_set(n, hi, rng=RNG):
    return np.unique(rng.integers(0, hi, size=n, dtype=np.int64)).astype(np.uint32)


# ---- golden values from the reference's own tests ---------------------------

def test_facebook_triangle_golden(oracle, facebook, fixtures_meta):
    T = oracle.Trie(*facebook)
    assert oracle.join_count(plans.TRIANGLE, [T], 4) == fixtures_meta["facebook_pruned"]["TriangleAgg"]


def test_facebook_four_clique_golden(oracle, facebook, fixtures_meta):
    T = oracle.Trie(*facebook)
    assert oracle.join_count(plans.FOUR_CLIQUE, [T], 4) == fixtures_meta["facebook_pruned"]["FliqueAgg"]


def test_facebook_duplicated_golden(oracle, facebook_dup, fixtures_meta):
    T = oracle.Trie(*facebook_dup)
    assert oracle.join_count(plans.TRIANGLE, [T], 4) == fixtures_meta["facebook_duplicated"]["Triangle_rows"]


def test_fixture_shapes(facebook, facebook_dup, fixtures_meta):
    assert len(facebook[0]) == fixtures_meta["facebook_pruned"]["rows"] == 88234
    assert (facebook[0] > facebook[1]).all()  # every row src > dst
    assert len(facebook_dup[0]) == fixtures_meta["facebook_duplicated"]["rows"] == 176468


# ---- independent algebra -------------------------------------------------------

def sym_adj(src, dst, n=None):
    n = int(max(src.max(), dst.max()) + 1) if n is None else n
    A = sp.coo_matrix((np.ones(len(src)), (src, dst)), shape=(n, n)).tocsr()
    A = ((A + A.T) > 0).astype(np.int64)
    A.setdiag(0)
    A.eliminate_zeros()
    return A


def test_facebook_triangle_matches_scipy(facebook, oracle):
    A = sym_adj(*facebook)
    tri = int((A @ A).multiply(A).sum()) // 6
    assert tri == 1612010


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_triangle_and_clique_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    n = 24
    pairs = rng.integers(0, n, size=(120, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    A = np.zeros((n, n), dtype=np.int64)
    A[pairs[:, 0], pairs[:, 1]] = 1
    A = ((A + A.T) > 0).astype(np.int64)
    iu = np.argwhere(np.triu(A, 1))
    src, dst = iu[:, 0].astype(np.uint32), iu[:, 1].astype(np.uint32)  # src < dst
    T = oracle.Trie(src, dst)
    tri = int(np.trace(A @ A @ A)) // 6
    assert oracle.join_count(plans.TRIANGLE, [T], 2) == tri
    k4 = sum(1 for q in itertools.combinations(range(n), 4)
             if all(A[x, y] for x, y in itertools.combinations(q, 2)))
    assert oracle.join_count(plans.FOUR_CLIQUE, [T], 2) == k4
    # closed 4-walks over the symmetric relation = trace(A^4)
    both = np.argwhere(A)
    S = oracle.Trie(both[:, 0].astype(np.uint32), both[:, 1].astype(np.uint32))
    walks = oracle.join_count(plans.FOUR_CYCLE_WALKS, [S], 2)
    assert walks == int(np.trace(np.linalg.matrix_power(A, 4)))
    c4 = sum(1 for q in itertools.permutations(range(n), 4)
             if A[q[0], q[1]] and A[q[1], q[2]] and A[q[2], q[3]] and A[q[3], q[0]]) // 8
    assert plans.four_cycles(walks, A.sum(1)) == c4


def test_facebook_selection_goldens(oracle, facebook_dup, fixtures_meta):
    """FliqueSelAgg (x=0) and BarbellSelAgg (p=6) on the symmetric facebook
    graph, the constants travis.py asserts (385-400, 473-489)."""
    s, d = facebook_dup
    T, R = oracle.Trie(s, d), oracle.Trie(d, s)
    meta = fixtures_meta["facebook_duplicated"]
    assert oracle.join_count(plans.flique_sel(0), [T, R], 4) == meta["FliqueSelAgg_x0"]
    assert oracle.join_count(plans.barbell_sel(6), [T, R], 4) == meta["BarbellSelAgg_p6"]


@pytest.mark.parametrize("seed", [3, 4])
def test_selections_brute_force(oracle, seed):
    """Selections on a small symmetric graph against enumeration: ordered
    4-cliques with a adjacent to the node, and the barbell count, which
    factors into (ordered triangles at neighbours of p) squared."""
    rng = np.random.default_rng(seed)
    n = 20
    pairs = rng.integers(0, n, size=(110, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    A = np.zeros((n, n), dtype=np.int64)
    A[pairs[:, 0], pairs[:, 1]] = 1
    A = ((A + A.T) > 0).astype(np.int64)
    both = np.argwhere(A)
    s, d = both[:, 0].astype(np.uint32), both[:, 1].astype(np.uint32)
    T, R = oracle.Trie(s, d), oracle.Trie(d, s)
    for node in list(range(0, n, 3)) + [n + 5]:  # n + 5: not a vertex, empty selection
        adj = A[node] if node < n else np.zeros(n, np.int64)
        fs = sum(1 for q in itertools.permutations(range(n), 4)
                 if adj[q[0]] and all(A[x, y] for x, y in itertools.combinations(q, 2)))
        assert oracle.join_count(plans.flique_sel(node), [T, R], 2) == fs
        t6 = sum(1 for a, b, c in itertools.permutations(range(n), 3) if adj[a] and A[a, b] and A[b, c] and A[a, c])
        assert oracle.join_count(plans.barbell_sel(node), [T, R], 2) == t6 * t6


def test_facebook_two_bag_goldens(oracle, facebook_dup, fixtures_meta):
    """LollipopAgg and BarbellAgg (travis.py:14-70) from the oracle's triangle
    counts grouped by first attribute, folded over the edges."""
    s, d = facebook_dup
    T = oracle.Trie(s, d)
    nv = int(max(s.max(), d.max())) + 1
    heads = np.unique(s)
    ph = np.zeros(nv, np.int64)
    ph[heads] = oracle.join_count_per_head(plans.TRIANGLE, [T], 4)
    meta = fixtures_meta["facebook_duplicated"]
    assert int(ph[s].sum()) == meta["LollipopAgg"]
    assert int((ph[s] * ph[d]).sum()) == meta["BarbellAgg"]


def test_sssp_against_scipy_and_golden(oracle, facebook_dup):
    """The closure restated level by level against scipy's unweighted shortest
    paths (directed, random graphs), and the SSSP check travis.py makes
    (556-567): on the symmetric facebook graph from vertex 0, the 1001st
    visited vertex (rows ordered by value) is at distance 2."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import shortest_path
    rng = np.random.default_rng(5)
    for n, m in ((50, 120), (300, 700), (1000, 1500)):
        s = rng.integers(0, n, m)
        d = rng.integers(0, n, m)
        A = sp.csr_matrix((np.ones(m), (s, d)), shape=(n, n))
        for start in (0, n // 2):
            ref = shortest_path(A, directed=True, unweighted=True, indices=start)
            dist, _ = oracle.sssp_unweighted(s, d, start, n)
            exp = np.where(np.isinf(ref), 0xFFFFFFFF, ref).astype(np.uint32)
            assert np.array_equal(dist, exp)
            lim, _ = oracle.sssp_unweighted(s, d, start, n, iterations=2)
            assert np.array_equal(lim, np.where(exp <= 2, exp, 0xFFFFFFFF).astype(np.uint32))
    s, d = facebook_dup
    n = int(max(s.max(), d.max())) + 1
    dist, _ = oracle.sssp_unweighted(s, d, 0, n)
    visited = np.flatnonzero(dist != 0xFFFFFFFF)
    assert dist[visited[1000]] == 2


def test_two_hop_closed_form(oracle):
    rng = np.random.default_rng(7)
    r = np.unique(rng.integers(0, 50, size=(400, 2)), axis=0)
    t = np.unique(rng.integers(0, 50, size=(300, 2)), axis=0)
    R = oracle.Trie(r[:, 0].astype(np.uint32), r[:, 1].astype(np.uint32))
    Tt = oracle.Trie(t[:, 0].astype(np.uint32), t[:, 1].astype(np.uint32))
    indeg_r = np.bincount(r[:, 1], minlength=64)
    outdeg_t = np.bincount(t[:, 0], minlength=64)
    assert oracle.join_count(plans.TWO_HOP, [R, Tt], 2) == int(np.sum(indeg_r * outdeg_t))


def test_relabel_invariance(oracle):
    import emptyheaded_amd.graphs as G
    s, d = G.rmat_samples(11, 16 << 11, 99)
    a1, b1, n1 = G.prepare(s, d, G.KEEP_IDS, G.SRC_LT_DST)
    a2, b2, n2 = G.prepare(s, d, G.DEGREE_ASC, G.SRC_LT_DST)
    a3, b3, n3 = G.prepare(s, d, G.COMPACT, G.SRC_GT_DST)
    c = [oracle.join_count(plans.TRIANGLE, [oracle.Trie(a, b)], 4) for a, b in ((a1, b1), (a2, b2), (a3, b3))]
    assert c[0] == c[1] == c[2] > 0
    A = sym_adj(a1, b1)
    assert c[0] == int((A @ A).multiply(A).sum()) // 6


def test_er_sanity_range(oracle):
    import emptyheaded_amd.graphs as G
    s, d, n = G.er_graph()
    assert 1_040_000 < len(s) < 1_048_576 and (s < d).all()
    tri = oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    # E[triangles] = C(n,3) p^3 with p = m / C(n,2): ~5.5e3 for this config
    p = len(s) / (n * (n - 1) / 2)
    expect = n * (n - 1) * (n - 2) / 6 * p ** 3
    assert abs(tri - expect) < 6 * np.sqrt(expect)


# ---- layout rule and set formats --------------------------------------------------

def test_is_sparse_boundaries(oracle):
    L = oracle.lib()
    # common.hpp:56-67: len <= 2 is always sparse; otherwise range/len > 128
    assert L.orc_is_sparse(2, 0) == 1
    assert L.orc_is_sparse(3, 384) == 0
    assert L.orc_is_sparse(3, 385) == 1
    assert L.orc_is_sparse(1000, 128000) == 0
    assert L.orc_is_sparse(1000, 128001) == 1
    # exact equivalence with the integer test on random inputs
    for n, r in RNG.integers(3, 2**31, size=(2000, 2)):
        assert L.orc_is_sparse(int(n), int(r)) == int(r > 128 * n)


@pytest.mark.parametrize("layout", [0, 1])
def test_set_round_trip_and_find(oracle, layout):
    for n, hi in [(1, 10), (5, 64), (64, 64), (100, 10000), (500, 2**32 - 1), (3, 200)]:
        v = rand_set(n, hi)
        s = oracle.Set(v, layout)
        assert np.array_equal(s.decode(), v)
        for i, x in enumerate(v):
            assert s.find(int(x)) == i
        for x in RNG.integers(0, max(hi, 2), size=50, dtype=np.int64):
            x = int(x) & 0xFFFFFFFF
            exp = int(np.searchsorted(v, x)) if x in set(v.tolist()) else -1
            assert s.find(x) == exp


def test_bitset_byte_format(oracle):
    v = np.array([60, 63, 64, 130], dtype=np.uint32)
    s = oracle.Set(v, 0)
    raw = s.raw()
    # [u64 start_word][W u64 words][W u32 rank], W = word(130) - word(60) + 1 = 3
    assert len(raw) == 8 + 3 * 8 + 3 * 4
    start = np.frombuffer(raw[:8], np.uint64)[0]
    words = np.frombuffer(raw[8:32], np.uint64)
    rank = np.frombuffer(raw[32:44], np.uint32)
    assert start == 0
    assert words[0] == (1 << 60) | (1 << 63) and words[1] == 1 and words[2] == (1 << 2)
    assert list(rank) == [0, 2, 3]


@pytest.mark.parametrize("la,lb", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_intersection_all_layout_pairs(oracle, la, lb):
    cases = [(50, 40, 3000), (5, 500, 3000), (300, 300, 1000), (1, 1000, 5000), (64, 64, 128),
             (40, 2000, 2**20), (33 * 10, 10, 4000)]
    for na, nb, hi in cases:
        a, b = rand_set(na, hi), rand_set(nb, hi)
        A, B = oracle.Set(a, la), oracle.Set(b, lb)
        exp = np.intersect1d(a, b)
        vals, typ, card, rng = oracle.intersect(A, B)
        assert np.array_equal(vals, exp) and card == len(exp)
        assert oracle.intersect_count(A, B) == len(exp)
        if len(exp):
            # result layout follows the inputs: U x U -> U, U x R -> U, R x R -> R
            assert typ == (0 if (la == 0 and lb == 0) else 1)


def test_empty_and_disjoint(oracle):
    e = oracle.Set(np.array([], np.uint32), 1)
    a = oracle.Set(np.array([1, 2, 3], np.uint32), 0)
    vals, typ, card, _ = oracle.intersect(e, a)
    assert card == 0 and typ == 1
    far = oracle.Set(np.array([100000, 100001, 100002], np.uint32), 0)
    assert oracle.intersect_count(a, far) == 0


def test_gallop_rule_boundary(oracle):
    # max/min > 32 gallops (uinteger.hpp:1209-1227); exactly 32 merges
    r = rand_set(10, 10**6)
    while len(r) != 10:
        r = rand_set(10, 10**6)
    f32 = rand_set(400, 10**6)[:320]
    f33 = rand_set(400, 10**6)[:330]
    assert oracle.algorithm_of(oracle.Set(r, 1), oracle.Set(f32, 1)) == 0
    assert oracle.algorithm_of(oracle.Set(r, 1), oracle.Set(f33, 1)) == 1
    for f in (f32, f33):
        assert oracle.intersect_count(oracle.Set(r, 1), oracle.Set(f, 1)) == len(np.intersect1d(r, f))


def test_merge_and_gallop_agree(oracle):
    L = oracle.lib()
    import ctypes as C
    for _ in range(50):
        a, b = rand_set(int(RNG.integers(1, 300)), 5000), rand_set(int(RNG.integers(1, 3000)), 5000)
        pa, pb = a.ctypes.data_as(C.POINTER(C.c_uint32)), b.ctypes.data_as(C.POINTER(C.c_uint32))
        out = np.zeros(len(a) + 1, np.uint32)
        n1 = L.orc_uu_merge(out.ctypes.data_as(C.POINTER(C.c_uint32)), pa, len(a), pb, len(b))
        assert np.array_equal(out[:n1], np.intersect1d(a, b))
        n2 = L.orc_uu_gallop(out.ctypes.data_as(C.POINTER(C.c_uint32)), pa, len(a), pb, len(b))
        assert np.array_equal(out[:n2], np.intersect1d(a, b))


def test_trie_round_trip_and_layouts(oracle):
    rng = np.random.default_rng(5)
    src = rng.integers(0, 300, size=5000).astype(np.uint32)
    dst = rng.integers(0, 20000, size=5000).astype(np.uint32)
    dst[:1000] = rng.integers(0, 300, size=1000)  # some dense child sets
    src = np.concatenate([src, src[:100]])  # repeated rows are dropped
    dst = np.concatenate([dst, dst[:100]])
    T = oracle.Trie(src, dst)
    a, b = T.tuples()
    key = np.unique((src.astype(np.uint64) << 32) | dst)
    assert np.array_equal(a, (key >> 32).astype(np.uint32)) and np.array_equal(b, (key & 0xFFFFFFFF).astype(np.uint32))
    heads = np.unique(src)
    types, cards = T.node_info(len(heads))
    for i, h in enumerate(heads):
        ch = np.unique(dst[src == h])
        assert cards[i] == len(ch)
        span = int(ch[-1]) - int(ch[0]) if len(ch) > 1 else 0
        assert types[i] == (1 if (len(ch) <= 2 or span > 128 * len(ch)) else 0)
        c = T.child(int(h))
        assert c is not None and c.cardinality == len(ch)
    assert T.child(10**6) is None


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_layout_independence(oracle, facebook, policy):
    T = oracle.Trie(*facebook, policy=policy)
    assert oracle.join_count(plans.TRIANGLE, [T], 4) == 1612010


def test_range_restriction_sums(oracle, facebook):
    T = oracle.Trie(*facebook)
    n0 = oracle.join_level0_size(plans.TRIANGLE, [T])
    parts = [oracle.join_count(plans.TRIANGLE, [T], 2, b, min(b + 1000, n0)) for b in range(0, n0, 1000)]
    assert sum(parts) == 1612010
    per = oracle.join_count_per_head(plans.TRIANGLE, [T], 2)
    assert int(per.sum()) == 1612010 and len(per) == n0


def strictly_lexicographic(R):
    """Rows strictly increasing in lexicographic order (sorted and unique)."""
    d = R[1:].astype(np.int64) - R[:-1].astype(np.int64)
    nz = d != 0
    first = np.argmax(nz, axis=1)
    return bool(np.all(nz.any(axis=1)) and np.all(d[np.arange(len(d)), first] > 0))


def test_materialized_golden_rows(oracle, facebook, fixtures_meta):
    """Materialized Triangle and Flique: the reference's row counts and first
    rows (test/graph/travis.py:302-318,342-354), rows sorted and unique, and
    every row a clique of the relation."""
    meta = fixtures_meta["facebook_pruned"]
    T = oracle.Trie(*facebook)
    edges = set(zip(facebook[0].tolist(), facebook[1].tolist()))
    for levels, rows_key, row0_key in ((plans.TRIANGLE, "Triangle_rows", "Triangle_row0"),
                                       (plans.FOUR_CLIQUE, "Flique_rows", "Flique_row0")):
        R = oracle.join_materialize(levels, [T])
        assert len(R) == meta[rows_key]
        assert list(R[0]) == meta[row0_key]
        assert strictly_lexicographic(R[:2_000_000])
        rng = np.random.default_rng(3)
        for i in rng.choice(len(R), 500, replace=False):
            t = R[i]
            for x in range(len(t)):
                for y in range(x + 1, len(t)):
                    assert (t[x], t[y]) in edges


def _inv_degree():
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inv_degree.npz"))
    return d["key"], d["val"]


def pagerank_f64(s, d, w_keys, w_vals, base, add, mul, init, iterations):
    """The rule in float64 sparse algebra (an independent evaluation): R lives
    on the head values (sources); a term exists where z is a head and in W."""
    import scipy.sparse as sp
    s = np.asarray(s, np.int64)
    d = np.asarray(d, np.int64)
    pairs = np.unique(np.stack([s, d], 1), axis=0)
    n = int(max(pairs.max(), int(np.max(w_keys)) if len(w_keys) else 0)) + 1
    A = sp.csr_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    heads = np.unique(pairs[:, 0])
    is_head = np.zeros(n, bool)
    is_head[heads] = True
    w = np.zeros(n)
    w[np.asarray(w_keys, np.int64)] = np.asarray(w_vals, np.float64)
    r = np.where(is_head, float(np.float32(base)), 0.0)
    for _ in range(iterations):
        acc = A @ (init * r * w)
        r = np.where(is_head, add + mul * acc, 0.0)
    return r[heads]


def test_pagerank_oracle_golden(oracle, facebook_dup, fixtures_meta):
    """PageRank over facebook_duplicated with the reference's inv_degree.tsv:
    vertex 0's rank after the base case 1/N and 5 iterations is the value
    travis.py:585-598 checks, to its tolerance 0.0001."""
    meta = fixtures_meta["inv_degree"]
    keys, vals = _inv_degree()
    s, d = facebook_dup
    N = len(np.unique(s))
    assert N == meta["PageRank_N"]
    r = oracle.pagerank(s, d, keys, vals, 1.0 / N, iterations=meta["PageRank_iterations"], nthreads=4)
    assert abs(float(r[0]) - meta["PageRank_row0"]) < meta["PageRank_tolerance"]
    # float64 sparse algebra agrees to float32 accumulation error
    ref = pagerank_f64(s, d, keys, vals, 1.0 / N, 0.15, 0.85, 1.0, 5)
    assert np.allclose(r, ref, rtol=2e-5, atol=0)


def test_pagerank_oracle_against_sparse_algebra(oracle):
    """Directed random graphs (a transposed product or a wrong key set fails),
    W on a subset of the values, several rule constants and iteration counts."""
    rng = np.random.default_rng(11)
    for n, m in ((40, 150), (400, 4000), (2000, 9000)):
        s = rng.integers(0, n, m).astype(np.uint32)
        d = rng.integers(0, n, m).astype(np.uint32)
        wk = np.unique(rng.integers(0, n + 20, n // 2)).astype(np.uint32)
        wv = rng.uniform(0.01, 1.0, len(wk)).astype(np.float32)
        for base, add, mul, init, it in ((1.0 / n, 0.15, 0.85, 1.0, 5), (0.5, 0.1, 0.5, 2.0, 3), (1.0, 0.0, 1.0, 1.0, 0)):
            r = oracle.pagerank(s, d, wk, wv, base, add, mul, init, it, nthreads=3)
            ref = pagerank_f64(s, d, wk, wv, base, add, mul, init, it)
            assert np.allclose(r, ref, rtol=2e-5, atol=1e-30), (n, base, it)


def test_pagerank_oracle_closed_form_complete_graph(oracle):
    """On the complete directed graph K_n with W = 1/(n-1) every rank follows
    r_k = 1 + (r_0 - 1) * 0.85^k."""
    n = 50
    s, d = np.nonzero(~np.eye(n, dtype=bool))
    keys = np.arange(n, dtype=np.uint32)
    vals = np.full(n, 1.0 / (n - 1), np.float32)
    for k in (0, 1, 4, 9):
        r = oracle.pagerank(s.astype(np.uint32), d.astype(np.uint32), keys, vals, 1.0 / n, iterations=k)
        exp = 1 + (1.0 / n - 1) * 0.85 ** k
        assert np.allclose(r, exp, rtol=1e-5)


def test_pagerank_oracle_sequential_fold(oracle):
    """The fold is the generated code's: float32 terms added one by one in
    ascending z.  Terms 1, 2^-24, 2^-24 (ascending z) sum to 1 in float32 and
    would give 1 + 2^-23 in the other order."""
    s = np.array([0, 0, 0, 1, 2, 3], np.uint32)
    d = np.array([1, 2, 3, 0, 0, 0], np.uint32)
    keys = np.array([1, 2, 3], np.uint32)
    vals = np.array([1.0, 2.0 ** -24, 2.0 ** -24], np.float32)
    r = oracle.pagerank(s, d, keys, vals, 1.0, add=0.0, mul=1.0, init=1.0, iterations=1)
    assert r[0] == np.float32(1.0)
    s2 = np.array([0, 0, 0, 1, 2, 3], np.uint32)
    d2 = np.array([3, 2, 1, 0, 0, 0], np.uint32)
    vals2 = np.array([2.0 ** -24, 2.0 ** -24, 1.0], np.float32)  # big term last in z order
    r2 = oracle.pagerank(s2, d2, keys, vals2, 1.0, add=0.0, mul=1.0, init=1.0, iterations=1)
    assert r2[0] == np.float32(1.0 + 2.0 ** -23)


def test_triangles_by_vertex_identity(oracle):
    """The per-vertex counts of the symmetric triangle bag (ordered triangles
    whose first attribute is v) equal twice the triangles {a < b < c} in which
    v is the smallest, middle or largest vertex, counted by the three oriented
    plans segmented by their first attribute -- checked vertex by vertex, and
    against brute force on a small graph."""
    from emptyheaded_amd import graphs as G
    for scale in (8, 11):
        s, d, n = G.rmat_graph(scale)
        S, D = G.symmetric(s, d)
        nv = int(max(S.max(), D.max())) + 1
        sym = oracle.Trie(S, D)
        T = np.zeros(nv, np.int64)
        T[np.unique(S)] = oracle.join_count_per_head(plans.TRIANGLE, [sym], 4)
        O, R = oracle.Trie(s, d), oracle.Trie(d, s)
        tmin = np.zeros(nv, np.int64)
        tmin[np.unique(s)] = oracle.join_count_per_head(plans.TRIANGLE, [O], 4)
        tmid = np.zeros(nv, np.int64)
        tmid[np.intersect1d(np.unique(s), np.unique(d))] = oracle.join_count_per_head(plans.TRIANGLE_BY_MIDDLE, [O, R], 4)
        tmax = np.zeros(nv, np.int64)
        tmax[np.unique(d)] = oracle.join_count_per_head(plans.TRIANGLE_BY_LAST, [O, R], 4)
        assert np.array_equal(T, 2 * (tmin + tmid + tmax))
        if scale == 8:  # brute force: triangles containing each vertex
            A = np.zeros((nv, nv), bool)
            A[S, D] = True
            A3 = (A.astype(np.int64) @ A.astype(np.int64)) * A
            assert np.array_equal(T, A3.sum(1))


def test_triangle_projection_golden(oracle, facebook, fixtures_meta):
    """TriangleProj(a,b) as the reference evaluates it (the existential c is
    dropped by its generator): 85,658 rows with first row (6,5) on
    facebook_pruned (travis.py:282-300) -- exactly the edges whose second
    vertex has a child set, computed here in numpy; the edges that close a
    triangle (scipy: (A A^T) masked by A) are fewer, 84,185."""
    import scipy.sparse as sp
    from emptyheaded_amd import plans
    s, d = facebook
    meta = fixtures_meta["facebook_pruned"]
    R = oracle.join_materialize(plans.TRIANGLE_PROJ, [oracle.Trie(s, d)])
    assert len(R) == meta["TriangleProj_rows"] == 85658 and list(R[0]) == meta["TriangleProj_row0"]
    keep = np.isin(d, np.unique(s))
    exp = np.array(sorted(zip(s[keep].tolist(), d[keep].tolist())), dtype=np.uint32)
    assert np.array_equal(R, exp)
    n = int(max(s.max(), d.max())) + 1
    A = sp.csr_matrix((np.ones(len(s)), (s.astype(np.int64), d.astype(np.int64))), shape=(n, n))
    assert (A.multiply(A @ A.T) > 0).sum() == 84185
