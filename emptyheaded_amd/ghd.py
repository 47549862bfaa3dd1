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


ORIENT_MIN_ROWS = 1 << 20


def triangle_by_vertex(ctx, sym, form, method="auto"):
    """T(v) for every value v of the symmetric edge trie `sym` (int64 array
    indexed by value), and the total (the ordered triangle count).

    method "direct" runs the bag as written: the triangle plan over `sym` with
    one segment per vertex.  "oriented" uses that for a symmetric relation
    without loops T(v) = 2 (t_min(v) + t_mid(v) + t_max(v)), the triangles
    {a < b < c} in which v is the smallest, middle or largest vertex: three
    per-segment passes over the relation's oriented half (rows x < y, built on
    the device) and its reverse (plans.TRIANGLE, TRIANGLE_BY_MIDDLE and
    TRIANGLE_BY_LAST, segmented by their first attribute), each finding every
    triangle once; it raises ValueError for a relation that is not symmetric
    or has a loop.  The device checks symmetry exactly: every row x < y must
    have its reverse row, and the rows x > y must be as many.  "auto" takes
    "oriented" for symmetric loop-free relations of at least ORIENT_MIN_ROWS
    rows (below that, orienting and building two tries costs more than it
    saves), else "direct".  The knobs in force when called are restored."""
    import emptyheaded_amd as E
    if method not in ("auto", "direct", "oriented"):
        raise ValueError("method must be 'auto', 'direct' or 'oriented', not %r" % (method,))
    nv = int(sym.info()["max_value"]) + 1
    bounds = np.arange(nv + 1, dtype=np.uint32)
    if method == "auto" and sym.num_rows < ORIENT_MIN_ROWS:
        method = "direct"
    if method != "direct":
        lower, rev, (n_lt, n_gt, n_eq, n_miss) = sym.orient(getattr(sym, "policy", E.POLICY_RULE))
        symmetric = n_lt == n_gt and n_eq == 0 and n_miss == 0
        if method == "oriented" and not symmetric:
            raise ValueError("oriented triangle passes need a symmetric relation without loops "
                             "(rows x<y %d, x>y %d, x==y %d, x<y rows without their reverse %d)" %
                             (n_lt, n_gt, n_eq, n_miss))
        if symmetric:
            if lower is None:  # no rows at all
                return np.zeros(nv, dtype=np.int64), 0
            from . import tuning as TU
            saved = E.tune_get()
            t = np.zeros(nv, dtype=np.int64)
            try:
                # each pass under its plan's tuned knobs
                for knobs, levels, tries in (("triangle_by_middle", plans.TRIANGLE_BY_MIDDLE, [lower, rev]),
                                             ("triangle_by_middle", plans.TRIANGLE_BY_LAST, [lower, rev]),
                                             ("triangle", plans.TRIANGLE, [lower])):
                    E.tune_set(**TU.plan_knobs(knobs))
                    c, _ = E.join_count_segments(ctx, levels, tries, bounds, form)
                    t += np.asarray(c, dtype=np.int64)
            finally:
                E.tune_set(**saved)
                lower.close()
                rev.close()
            T = 2 * t
            return T, int(T.sum())
    counts, total = E.join_count_segments(ctx, plans.TRIANGLE, [sym], bounds, form)
    return np.asarray(counts, dtype=np.int64), total


def lollipop_count(ctx, sym, form, method="auto"):
    import emptyheaded_amd as E
    T, _ = triangle_by_vertex(ctx, sym, form, method)
    return E.weighted_pairs(ctx, sym, T, np.ones(len(T), dtype=np.int64))


def barbell_count(ctx, sym, form, method="auto"):
    import emptyheaded_amd as E
    T, _ = triangle_by_vertex(ctx, sym, form, method)
    return E.weighted_pairs(ctx, sym, T, T)
