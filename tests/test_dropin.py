"""The generated query code, restated from the code generator's templates,
compiled unchanged against the drop-in C++ API and run on the GPU: the
reference's own golden results (test/graph/travis.py) through the reference's
own call sequence."""
import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import dropin
from emptyheaded_amd import graphs as G
from emptyheaded_amd import plans

pytestmark = pytest.mark.gpu


def test_triangle_agg_golden(tmp_path, facebook, fixtures_meta):
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook)
    v, rows = dropin.run("TriangleAgg", db)
    assert v == fixtures_meta["facebook_pruned"]["TriangleAgg"]
    assert rows == 0  # scalar bag (travis.py:208-209)


def test_flique_agg_golden(tmp_path, facebook, fixtures_meta):
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook)
    v, rows = dropin.run("FliqueAgg", db)
    assert v == fixtures_meta["facebook_pruned"]["FliqueAgg"]
    assert rows == 0


def test_triangle_agg_duplicated(tmp_path, facebook_dup, fixtures_meta):
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook_dup)
    v, _ = dropin.run("TriangleAgg", db)
    assert v == fixtures_meta["facebook_duplicated"]["Triangle_rows"]


@pytest.mark.parametrize("scale", [12, 15])
def test_dropin_matches_oracle(tmp_path, oracle, scale):
    s, d, n = G.rmat_graph(scale)
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, d)
    O = oracle.Trie(s, d)
    assert dropin.run("TriangleAgg", db)[0] == oracle.join_count(plans.TRIANGLE, [O], 8)
    assert dropin.run("FliqueAgg", db)[0] == oracle.join_count(plans.FOUR_CLIQUE, [O], 8)


def test_empty_and_missing_children(tmp_path):
    # a relation whose children never appear as heads: no triangles, no crash
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", np.array([5, 5, 6], np.uint32), np.array([1, 2, 3], np.uint32))
    assert dropin.run("TriangleAgg", db)[0] == 0
    assert dropin.run("FliqueAgg", db)[0] == 0


def test_triangle_materialized_golden(tmp_path, oracle, facebook, fixtures_meta):
    """Triangle(a,b,c): build_set / allocate_next / foreach_builder / set_level
    produce the output trie; Trie::foreach reads the reference's 1,612,010 rows
    with first row (6,5,2) (test/graph/travis.py:206-220), every row equal to
    the oracle's enumeration."""
    meta = fixtures_meta["facebook_pruned"]
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook)
    R, _, num_rows, got = dropin.rows("Triangle", db, meta["Triangle_rows"] + 1)
    assert num_rows == got == meta["Triangle_rows"]
    assert list(R[0]) == meta["Triangle_row0"]
    assert np.array_equal(R, dropin_oracle_rows(oracle, facebook, plans.TRIANGLE))


def dropin_oracle_rows(oracle, graph, levels):
    return oracle.join_materialize(levels, [oracle.Trie(*graph)])


@pytest.mark.parametrize("scale", [10, 13])
def test_triangle_materialized_rmat(tmp_path, oracle, scale):
    s, d, n = G.rmat_graph(scale)
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, d)
    exp = dropin_oracle_rows(oracle, (s, d), plans.TRIANGLE)
    R, _, num_rows, got = dropin.rows("Triangle", db, len(exp) + 1)
    assert num_rows == got == len(exp)
    assert np.array_equal(R, exp)


def test_triangle_counts_annotated(tmp_path, oracle, facebook, fixtures_meta):
    """TriangleCounts(a;z) z<-COUNT(b,c): build_set + allocate_annotation +
    set_annotation give one row per head value of Edge with its triangle count
    (the oracle's per-head counts; they sum to the reference's 1,612,010);
    num_rows is what the generated code folds: one per (a,b) iteration."""
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook)
    src, dst = facebook
    heads = np.unique(src)
    R, ann, num_rows, got = dropin.rows("TriangleCounts", db, len(heads) + 1)
    assert got == len(heads)
    assert np.array_equal(R[:, 0], heads)
    per_head = oracle.join_count_per_head(plans.TRIANGLE, [oracle.Trie(src, dst)])
    assert np.array_equal(ann.astype(np.int64), per_head)
    assert int(ann.sum()) == fixtures_meta["facebook_pruned"]["TriangleAgg"]
    assert num_rows == int(np.isin(dst, heads).sum())


def test_materialized_empty_children(tmp_path):
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", np.array([5, 5, 6], np.uint32), np.array([1, 2, 3], np.uint32))
    R, _, num_rows, got = dropin.rows("Triangle", db, 4)
    assert got == 0 and num_rows == 0
    R, ann, num_rows, got = dropin.rows("TriangleCounts", db, 4)
    assert list(R[:, 0]) == [5, 6] and list(ann) == [0, 0] and num_rows == 0


def write_reference_db(oracle, db, src, dst, **kw):
    """A database folder as the reference's loader lays it out: the relation's
    saved trie image and the uint32_t encoding (size_t count, then each key's
    value, src/Encoding.cpp:81-104; here the value of key k is 3k + 1)."""
    import os
    oracle.write_refdb(src, dst, os.path.join(db, "relations", "Edge", "Edge_0_1"), **kw)
    enc = os.path.join(db, "encodings", "uint32_t")
    os.makedirs(enc, exist_ok=True)
    n = int(max(src.max(), dst.max())) + 1
    with open(os.path.join(enc, "encoding.bin"), "wb") as f:
        f.write(np.uint64(n).tobytes())
        f.write((np.arange(n, dtype=np.uint32) * 3 + 1).tobytes())


@pytest.mark.parametrize("mmapped", [False, True])
def test_reference_format_database(tmp_path, oracle, facebook, fixtures_meta, mmapped):
    """The generated code's own Trie::load and Encoding::from_binary open a
    database in the reference's on-disk format: TriangleAgg and FliqueAgg
    goldens, and Triangle's rows (first row (6,5,2))."""
    meta = fixtures_meta["facebook_pruned"]
    db = str(tmp_path / "db")
    write_reference_db(oracle, db, *facebook, nbuffers=4, mmapped=mmapped)
    assert dropin.run("TriangleAgg", db)[0] == meta["TriangleAgg"]
    assert dropin.run("FliqueAgg", db)[0] == meta["FliqueAgg"]
    R, _, num_rows, got = dropin.rows("Triangle", db, meta["Triangle_rows"] + 1)
    assert got == num_rows == meta["Triangle_rows"] and list(R[0]) == meta["Triangle_row0"]


def test_trie_load_reference_device(ctx, tmp_path, oracle):
    """ehx_trie_load_reference builds the device trie of a saved relation:
    its tuples equal the relation's."""
    s, d, n = G.rmat_graph(12)
    p = str(tmp_path / "Edge_0_1")
    T = oracle.write_refdb(s, d, p, nbuffers=3, policy=2)
    D = E.Trie.load_reference(ctx, p)
    a, b = D.tuples()
    ea, eb = T.tuples()
    assert np.array_equal(a, ea) and np.array_equal(b, eb)
    D.close()


def test_encoded_store_database(ctx, tmp_path, facebook, fixtures_meta):
    """A database that holds only the relation's encoded column store
    (relations/Edge/encoded.bin): the trie of each ordering is built on the
    device from it, through the C ABI and through the generated code's
    Trie::load of relations/Edge/Edge_0_1."""
    import os
    from test_refdb import write_encoded
    s, d = facebook
    db = str(tmp_path / "db")
    rel = os.path.join(db, "relations", "Edge")
    write_encoded(rel, [s, d])
    os.makedirs(os.path.join(db, "encodings", "uint32_t"), exist_ok=True)
    for (i, j), (a, b) in (((0, 1), (s, d)), ((1, 0), (d, s))):
        T = E.Trie.from_encoded(ctx, rel, i, j)
        R = E.Trie(ctx, a, b)
        ta, tb = T.tuples()
        ra, rb = R.tuples()
        assert np.array_equal(ta, ra) and np.array_equal(tb, rb)
    assert dropin.run("TriangleAgg", db)[0] == fixtures_meta["facebook_pruned"]["TriangleAgg"]


def test_annotated_relations_pagerank(ctx, tmp_path, facebook_dup, fixtures_meta):
    """Annotated input tries through the drop-in API (PageRankSteps.cpp: the
    reference's PageRank rule in the generated code's idiom, hand-written):
    InvDegree as a unary float relation read from its encoded column store,
    each step's PageRank output read back through get_annotation by the next
    step.  The drop-in is built without fp contraction, so its ranks equal the
    oracle's with an unfused finish bit for bit, and the device's plain form
    (fused finish, as the reference's build) to one float ulp; both meet the
    reference's golden (travis.py:585-598)."""
    import os
    from test_refdb import write_encoded
    meta = fixtures_meta["inv_degree"]
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inv_degree.npz"))
    keys, vals = d["key"], d["val"]
    s, t = facebook_dup
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, t)
    write_encoded(os.path.join(db, "relations", "InvDegree"), [keys], vals, width=4)
    heads = np.unique(s)
    R, ann, num_rows, got = dropin.rows("PageRankSteps", db, len(heads) + 1)
    assert got == len(heads) and np.array_equal(R[:, 0], heads)
    ranks = ann.astype(np.float32)
    assert abs(float(ranks[0]) - meta["PageRank_row0"]) < meta["PageRank_tolerance"]
    from oracle import oracle as O
    unfused = O.pagerank(s, t, keys, vals, 1.0 / len(heads), iterations=5, nthreads=4, fused_finish=False)
    assert np.array_equal(ranks.view(np.uint32), unfused.view(np.uint32))
    T = E.Trie(ctx, s, t)
    exp, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=5, form=E.FORM_PLAIN)
    ulps = np.abs(ranks.view(np.int32).astype(np.int64) - exp.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1


def test_annotated_binary_relation(tmp_path):
    """An annotated binary relation (WeightedDegree.cpp, hand-written in the
    generated code's idiom): per x, the float32 sum of its edges' annotations
    in ascending z through TrieIterator::get_annotation(1, z), bit for bit;
    duplicate rows keep the last annotation in file order."""
    import os
    from test_refdb import write_encoded
    rng = np.random.default_rng(9)
    s = rng.integers(0, 3000, 40000).astype(np.uint32)
    d = rng.integers(0, 3000, 40000).astype(np.uint32)
    w = rng.uniform(-1, 1, 40000).astype(np.float32)
    db = str(tmp_path / "db")
    write_encoded(os.path.join(db, "relations", "Edge"), [s, d], w, width=4)
    order = np.lexsort((d, s))
    last = {}
    for i in order:  # stable: the last of equal rows in file order wins
        last[(int(s[i]), int(d[i]))] = w[i]
    heads = np.unique(s)
    exp = np.zeros(len(heads), np.float32)
    keys = sorted(last)
    hi = {int(h): k for k, h in enumerate(heads)}
    for (x, z) in keys:  # ascending (x, z)
        exp[hi[x]] = np.float32(exp[hi[x]] + np.float32(last[(x, z)]))
    R, ann, num_rows, got = dropin.rows("WeightedDegree", db, len(heads) + 1)
    assert got == len(heads) and np.array_equal(R[:, 0], heads)
    assert np.array_equal(ann.astype(np.float32).view(np.uint32), exp.view(np.uint32))


def test_annotated_reference_images(tmp_path, oracle, facebook_dup):
    """The same annotated programs over relations saved in the reference's
    on-disk format with their annotations (InvDegree as a unary float image,
    weighted edges as a binary float image): ranks equal the encoded-store run,
    edge sums the sequential float32 fold."""
    import os
    from test_refdb import write_encoded
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inv_degree.npz"))
    keys, vals = d["key"], d["val"]
    s, t = facebook_dup
    db1, db2 = str(tmp_path / "enc"), str(tmp_path / "img")
    for db in (db1, db2):
        dropin.write_relation(db, "Edge", s, t)
    write_encoded(os.path.join(db1, "relations", "InvDegree"), [keys], vals, width=4)
    oracle.write_refdb(keys, None, os.path.join(db2, "relations", "InvDegree", "InvDegree_0"), nbuffers=2,
                       annos=vals, unary=True)
    heads = np.unique(s)
    _, a1, _, _ = dropin.rows("PageRankSteps", db1, len(heads) + 1)
    _, a2, _, _ = dropin.rows("PageRankSteps", db2, len(heads) + 1)
    assert np.array_equal(a1.astype(np.float32).view(np.uint32), a2.astype(np.float32).view(np.uint32))
    # weighted edges as a saved image
    rng = np.random.default_rng(10)
    T = oracle.Trie(s, t)
    ra, rb = T.tuples()
    w = rng.uniform(-1, 1, len(ra)).astype(np.float32)
    db3 = str(tmp_path / "w")
    oracle.write_refdb(s, t, os.path.join(db3, "relations", "Edge", "Edge_0_1"), nbuffers=4, annos=w)
    hx = np.unique(ra)
    exp = np.zeros(len(hx), np.float32)
    idx = np.searchsorted(hx, ra)
    for i in range(len(ra)):  # ascending (x, z)
        exp[idx[i]] = np.float32(exp[idx[i]] + w[i])
    R, ann, _, got = dropin.rows("WeightedDegree", db3, len(hx) + 1)
    assert got == len(hx) and np.array_equal(R[:, 0], hx)
    assert np.array_equal(ann.astype(np.float32).view(np.uint32), exp.view(np.uint32))


def test_selection_golden(tmp_path, facebook_dup, fixtures_meta):
    """FliqueSelAgg (x=0) through the drop-in's selection calls (hand-written in
    the generated code's idiom): the reference's 3,759,972 (travis.py:385-400),
    and the same after every vertex is relabelled through a permutation that a
    reference-format encoding.bin records, so value_to_key.at(0) must map the
    constant to its key."""
    import os
    s, t = facebook_dup
    exp = fixtures_meta["facebook_duplicated"]["FliqueSelAgg_x0"]
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, t)
    dropin.write_relation(db, "Edge", t, s, ordering="1_0")
    assert dropin.run("FliqueSelAgg", db)[0] == exp
    n = int(max(s.max(), t.max())) + 1
    perm = np.random.default_rng(12).permutation(n).astype(np.uint32)  # key of value v
    db2 = str(tmp_path / "db2")
    dropin.write_relation(db2, "Edge", perm[s], perm[t])
    dropin.write_relation(db2, "Edge", perm[t], perm[s], ordering="1_0")
    key_to_value = np.empty(n, np.uint32)
    key_to_value[perm] = np.arange(n, dtype=np.uint32)
    with open(os.path.join(db2, "encodings", "uint32_t", "encoding.bin"), "wb") as f:
        f.write(np.uint64(n).tobytes())
        f.write(key_to_value.tobytes())
    assert dropin.run("FliqueSelAgg", db2)[0] == exp


def test_sssp_generated_program(tmp_path, oracle, facebook_dup):
    """The generated SSSP program (travis.py:556-567: a MIN recursion from the
    selected source 0, compiled to tc::unweighted_single_source) through the
    drop-in: the reference's check (the output's row 1000, rows in value order,
    has annotation 2) and every row against the oracle's BFS levels (the start
    keeps annotation 0, every visited value is a row, num_rows counts them);
    the *[i=2] form stops after two levels.  On the symmetric facebook graph
    (direction-optimizing on the device) and on a directed random graph."""
    rng = np.random.default_rng(12)
    rs = rng.integers(0, 6000, 30000).astype(np.uint32)
    rd = rng.integers(0, 6000, 30000).astype(np.uint32)
    rs[:3], rd[:3] = 0, [5, 17, 4000]
    for name, (s, d) in (("facebook", facebook_dup), ("directed", (rs, rd))):
        db = str(tmp_path / name)
        dropin.write_relation(db, "Edge", s, d)
        n = int(max(s.max(), d.max())) + 1
        exp, _ = oracle.sssp_unweighted(s, d, 0, n)
        for query, limit in (("SSSP", None), ("SSSPIter2", 2)):
            keep = exp != E.UNREACHED if limit is None else exp <= limit
            reached = np.flatnonzero(keep)
            rows, ann, num_rows, got = dropin.rows(query, db, len(reached) + 16)
            assert num_rows == got == len(reached), (name, query)
            assert np.array_equal(rows[:, 0], reached), (name, query)
            assert np.array_equal(ann, exp[reached].astype(np.float64)), (name, query)
            v, nr = dropin.run(query, db)
            assert nr == num_rows
        if name == "facebook":  # the reference's own check
            _, full_ann, _, _ = dropin.rows("SSSP", db, n)
            assert full_ann[1000] == 2


# ---- multi-bag plans (LollipopAgg / BarbellAgg and their materialized forms) -----

def _sym_sets(s, d):
    n = int(max(s.max(), d.max())) + 1
    N = [set() for _ in range(n)]
    for a, b in zip(s.tolist(), d.tolist()):
        N[a].add(b)
    return N


def _lollipop_rows(N):
    """Lollipop(a,b,c,x) by brute force, as (a, x, b, c) tuples sorted: the
    output trie's level order."""
    out = []
    for a in range(len(N)):
        for b in N[a]:
            for c in N[b] & N[a]:
                for x in N[a]:
                    out.append((a, x, b, c))
    return np.array(sorted(out), dtype=np.uint32).reshape(-1, 4)


def _barbell_rows(N):
    """Barbell(a,b,c,x,y,z) by brute force, as (a, x, y, z, b, c) sorted."""
    tri = [[(b, c) for b in N[v] for c in (N[b] & N[v])] for v in range(len(N))]
    out = []
    for a in range(len(N)):
        for x in N[a]:
            for (y, z) in tri[x]:
                for (b, c) in tri[a]:
                    out.append((a, x, y, z, b, c))
    return np.array(sorted(out), dtype=np.uint32).reshape(-1, 6)


def reference_dataframe(trie_rows, ordering):
    """The reference's DataFrame of an output trie (codegen/trie/Trie.scala:
    219-230): column i holds the trie's level ordering[i], where ordering is
    the output relation's attribute permutation (QueryPlan.scala:156-170)."""
    return trie_rows[:, list(ordering)]


def test_multibag_generated_aggregates_golden(tmp_path, facebook_dup, fixtures_meta):
    """The generated LollipopAgg and BarbellAgg programs (a triangle bag
    annotated by its COUNT per vertex, then the root bag over Edge(a,x) folding
    those annotations) give the reference's 1,426,911,480 and
    20,371,831,447,136 on the symmetric facebook graph (travis.py:14-70)."""
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook_dup)
    meta = fixtures_meta["facebook_duplicated"]
    assert dropin.run("LollipopAgg", db) == (meta["LollipopAgg"], 0)
    assert dropin.run("BarbellAgg", db) == (meta["BarbellAgg"], 0)


@pytest.mark.parametrize("scale", [9, 12])
def test_multibag_generated_aggregates_oracle(tmp_path, oracle, scale):
    """The same programs on symmetric R-MAT graphs against the oracle's
    per-vertex triangle counts folded over the edges in numpy."""
    s, d, n = G.rmat_graph(scale)
    s, d = G.symmetric(s, d)
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, d)
    nv = int(max(s.max(), d.max())) + 1
    T = np.zeros(nv, np.int64)
    T[np.unique(s)] = oracle.join_count_per_head(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    s64, d64 = s.astype(np.int64), d.astype(np.int64)
    with np.errstate(over="ignore"):
        assert dropin.run("LollipopAgg", db)[0] == int(T[s64].sum())
        assert dropin.run("BarbellAgg", db)[0] == int((T[s64] * T[d64]).sum())


def test_multibag_materialized_simple_golden(tmp_path, fixtures_meta):
    """The generated Lollipop and Barbell programs (triangle bag, root bag,
    top-down pass reading the three-level triangle trie) on the reference's
    simple.tsv: 28 and 56 rows, every row equal to brute force in the output
    trie's level order, and the reference's DataFrame rows travis.py checks:
    df.iloc[27] = (5,3,4,4) and df.iloc[55] = (5,3,4,3,4,5)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "simple.npz"))
    s, d = z["src"], z["dst"]
    meta = fixtures_meta["simple"]
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", s, d)
    N = _sym_sets(s, d)
    rows, _, nr, got = dropin.rows("Lollipop", db, 1000)
    assert nr == got == meta["Lollipop_rows"] == 28
    assert np.array_equal(rows, _lollipop_rows(N))
    assert tuple(reference_dataframe(rows, (0, 3, 1, 2))[27]) == tuple(meta["Lollipop_row27"])
    rows, _, nr, got = dropin.rows("Barbell", db, 1000)
    assert nr == got == meta["Barbell_rows"] == 56
    assert np.array_equal(rows, _barbell_rows(N))
    assert tuple(reference_dataframe(rows, (0, 3, 4, 5, 1, 2))[55]) == tuple(meta["Barbell_row55"])
    assert dropin.run("Lollipop", db)[1] == 28 and dropin.run("Barbell", db)[1] == 56


def test_multibag_materialized_random(tmp_path):
    """Materialized Lollipop and Barbell on random symmetric graphs with hubs
    and isolated values, every row against brute force."""
    rng = np.random.default_rng(31)
    for n, m in ((40, 120), (90, 300)):
        a = rng.integers(0, n, m)
        b = (a + rng.integers(1, 6, m)) % n  # short-range edges make triangles
        a = np.concatenate([a, np.zeros(12, np.int64)])
        b = np.concatenate([b, rng.integers(1, n, 12)])  # vertex 0 a hub
        keep = a != b
        s = np.concatenate([a[keep], b[keep]]).astype(np.uint32)
        d = np.concatenate([b[keep], a[keep]]).astype(np.uint32)
        db = str(tmp_path / ("g%d" % n))
        dropin.write_relation(db, "Edge", s, d)
        N = _sym_sets(s, d)
        for q, brute in (("Lollipop", _lollipop_rows), ("Barbell", _barbell_rows)):
            exp = brute(N)
            rows, _, nr, got = dropin.rows(q, db, len(exp) + 10)
            assert nr == got == len(exp), (q, n)
            assert np.array_equal(rows, exp), (q, n)


def test_triangle_projection_generated(tmp_path, oracle, facebook, fixtures_meta):
    """The generated TriangleProj program (a child bag that keeps (a, b) with c
    dropped, then the root over Edge(a,b)) through the drop-in: the
    reference's 85,658 rows, first row (6,5), every row equal to the oracle's
    one-bag plan (travis.py:282-300)."""
    meta = fixtures_meta["facebook_pruned"]
    db = str(tmp_path / "db")
    dropin.write_relation(db, "Edge", *facebook)
    rows, _, nr, got = dropin.rows("TriangleProj", db, 100000)
    assert nr == got == meta["TriangleProj_rows"] and list(rows[0]) == meta["TriangleProj_row0"]
    assert np.array_equal(rows, oracle.join_materialize(plans.TRIANGLE_PROJ, [oracle.Trie(*facebook)]))
