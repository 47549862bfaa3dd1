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
