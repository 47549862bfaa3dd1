"""The generated query code, restated from the code generator's templates,
compiled unchanged against the drop-in C++ API and run on the GPU: the
reference's own golden results (test/graph/travis.py) through the reference's
own call sequence."""
import numpy as np
import pytest

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
