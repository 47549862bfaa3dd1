"""Two-bag plans, evaluated as the reference's GHD planner splits them.

LollipopAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x)
BarbellAgg(;w)  :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x),Edge(x,y),Edge(y,z),Edge(x,z)
(test/graph/travis.py:14-70) have a triangle bag hanging off the edge Edge(a,x).
The generated code builds each lower bag as an intermediate trie annotated by
its COUNT(*) grouped by the shared attribute, then folds the top bag over those
annotations (query_compiler/src/main/scala/codegen/query/EHGenerator.scala:
399-434, 711-726).  Here the triangle bag runs once as a batched query with one
segment per vertex (T(v) = ordered triangles whose first attribute is v), and
the top bag is the weighted pair sum over the edge trie:
  LollipopAgg = sum over Edge(a,x) of T(a),  BarbellAgg = sum over Edge(a,x) of T(a) T(x).
"""
import numpy as np

from . import plans


def triangle_by_vertex(ctx, sym, form):
    """T(v) for every value v of the symmetric edge trie `sym` (int64 array
    indexed by value), and the total (the ordered triangle count)."""
    import emptyheaded_amd as E
    nv = int(sym.info()["max_value"]) + 1
    bounds = np.arange(nv + 1, dtype=np.uint32)
    counts, total = E.join_count_segments(ctx, plans.TRIANGLE, [sym], bounds, form)
    return np.asarray(counts, dtype=np.int64), total


def lollipop_count(ctx, sym, form):
    import emptyheaded_amd as E
    T, _ = triangle_by_vertex(ctx, sym, form)
    return E.weighted_pairs(ctx, sym, T, np.ones(len(T), dtype=np.int64))


def barbell_count(ctx, sym, form):
    import emptyheaded_amd as E
    T, _ = triangle_by_vertex(ctx, sym, form)
    return E.weighted_pairs(ctx, sym, T, T)
