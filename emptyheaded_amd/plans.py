"""Join plans as the EmptyHeaded code generator emits them.

A plan lists, per attribute in attribute order, the operand sets intersected
at that level: (rel, parent) with parent = -1 for the head set of relation
`rel` and parent = j for the child set of attribute j's value, or
(rel, -2, value) for the child set of a constant (an equality selection).  The operand
order follows the accessor order of the generated code
(query_compiler/src/main/scala/codegen/query/EHGenerator.scala:273-838); the
attribute order is the sorted attribute names (parsers/Datalog.scala:150-188).
"""

# TriangleAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),z:long<-[COUNT(*)].
# Iterators Edge_a_b, Edge_b_c, Edge_a_c all read trie Edge_0_1 (relation 0):
#   a: head(Edge_a_b) ∩ head(Edge_a_c)
#   b: child(Edge_a_b, a) ∩ head(Edge_b_c)
#   c: child(Edge_b_c, b) ∩ child(Edge_a_c, a)
TRIANGLE = [[(0, -1), (0, -1)],
            [(0, 0), (0, -1)],
            [(0, 1), (0, 0)]]

# The same TriangleAgg with attribute order (b, a, c): relation 0 = Edge_0_1,
# relation 1 = Edge_1_0 (the reverse ordering, which the reference builds for
# every binary relation: all attribute permutations,
# query_compiler/src/main/scala/codegen/create/createDB.scala:232-233).
#   b: head(Edge_0_1) ∩ head(Edge_1_0)
#   a: child(Edge_1_0, b) ∩ head(Edge_0_1)
#   c: child(Edge_0_1, a) ∩ child(Edge_0_1, b)
# Every triangle a < b < c is found once, at its middle vertex b.  With the
# ascending-degree orientation the last level streams sum_a d+(a)^2 values
# instead of sum_b d-(b) d+(b) for the (a, b, c) order.
TRIANGLE_BY_MIDDLE = [[(0, -1), (1, -1)],
                      [(1, 0), (0, -1)],
                      [(0, 1), (0, 0)]]

# The same triangles found at their largest vertex c, attribute order (c, a, b)
# over relation 0 = Edge_0_1 and relation 1 = Edge_1_0:
#   c: head(Edge_1_0)
#   a: child(Edge_1_0, c)
#   b: child(Edge_0_1, a) ∩ child(Edge_1_0, c)
# The last level streams the out-lists of the a's (sum_a d+(a)^2 values, as the
# middle-vertex order does); per-segment counts by c are the triangles whose
# largest vertex is c.
TRIANGLE_BY_LAST = [[(1, -1)],
                    [(1, 0)],
                    [(0, 1), (1, 0)]]

# TriangleProj(a,b) :- Edge(a,b),Edge(b,c),Edge(a,c) (test/graph/travis.py:282-300),
# as the reference evaluates it.  The planner puts Edge(a,b) in the root bag and
# Edge(b,c),Edge(a,c) in a child bag that keeps (a, b) (solver/GHDSolver.scala,
# GHDNode.scala:121-147); the code generator then drops every attribute that is
# neither materialized, aggregated nor selected (codegen/query/QueryPlan.scala:
# 244), so the child bag never intersects on c: it is head(Edge) x head(Edge),
# and the root keeps the edges (a, b) whose b has a child set.  One bag with the
# same result:
#   a: head(Edge_0_1)
#   b: child(Edge_0_1, a) ∩ head(Edge_0_1)
# (85,658 rows on facebook_pruned, the reference's golden; the edges that close
# a triangle, the projection as written, are 84,185.)
TRIANGLE_PROJ = [[(0, -1)],
                 [(0, 0), (0, -1)]]

# FliqueAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,d),Edge(b,d),Edge(c,d),COUNT(*).
#   a: heads of Edge_a_b, Edge_a_c, Edge_a_d
#   b: child(Edge_a_b, a) ∩ head(Edge_b_c) ∩ head(Edge_b_d)
#   c: child(Edge_b_c, b) ∩ child(Edge_a_c, a) ∩ head(Edge_c_d)
#   d: child(Edge_a_d, a) ∩ child(Edge_b_d, b) ∩ child(Edge_c_d, c)
FOUR_CLIQUE = [[(0, -1), (0, -1), (0, -1)],
               [(0, 0), (0, -1), (0, -1)],
               [(0, 1), (0, 0), (0, -1)],
               [(0, 0), (0, 1), (0, 2)]]

HEAD, CONST = -1, -2


# Equality selections.  The generated code puts a selected attribute first and,
# when it has one accessor, replaces its level by one lookup:
# Iterators_<acc>.get_next_block(selection_x_0), so the next attribute reads
# the selected value's child set (EHGenerator.scala:657-701).  Here that child
# set is an operand (rel, CONST, value); relation 1 is the reverse ordering
# Edge_1_0 (the accessor Edge(a,x) ordered (x, a)).

def flique_sel(node):
    """FliqueSelAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,d),Edge(b,d),
    Edge(c,d),Edge(a,x),x=node,z:long<-[COUNT(*)] (test/graph/travis.py:385-400):
    the 4-clique plan with a restricted to child(Edge_1_0, node)."""
    return [[(0, HEAD), (0, HEAD), (0, HEAD), (1, CONST, node)],
            [(0, 0), (0, HEAD), (0, HEAD)],
            [(0, 1), (0, 0), (0, HEAD)],
            [(0, 0), (0, 1), (0, 2)]]


def barbell_sel(node):
    """BarbellSelAgg(;w) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,p),Edge(p,x),
    Edge(x,y),Edge(y,z),Edge(x,z),p=node,w:long<-[COUNT(*)] (test/graph/
    travis.py:473-489), as one bag over (a, b, c, x, y, z): a in
    child(Edge_1_0, node), x in child(Edge_0_1, node)."""
    return [[(0, HEAD), (0, HEAD), (1, CONST, node)],
            [(0, 0), (0, HEAD)],
            [(0, 1), (0, 0)],
            [(0, HEAD), (0, HEAD), (0, CONST, node)],
            [(0, 3), (0, HEAD)],
            [(0, 4), (0, 3)]]


# Closed 4-walks over a symmetric relation S (both edge directions):
#   Q(;z) :- S(a,b),S(b,c),S(c,d),S(a,d),COUNT(*)
# counts homomorphic images of the 4-cycle; the number of 4-cycles of the
# undirected graph is (Q - 2*sum(deg^2) + 2*m) / 8 (four_cycles below).
FOUR_CYCLE_WALKS = [[(0, -1), (0, -1)],
                    [(0, 0), (0, -1)],
                    [(0, 1), (0, -1)],
                    [(0, 2), (0, 0)]]

# 2-hop path through two relations R(x,y), T(y,z):  Q(;z) :- R(x,y),T(y,z),COUNT(*)
# with relation 0 = R keyed by x and relation 1 = T keyed by y.
TWO_HOP = [[(0, -1)],
           [(0, 0), (1, -1)],
           [(1, 1)]]


def four_cycles(closed_walks, degrees):
    """Number of undirected 4-cycles from the closed 4-walk count."""
    deg = [int(d) for d in degrees]
    m2 = sum(deg)  # 2m
    s2 = sum(d * d for d in deg)
    v = closed_walks - 2 * s2 + m2
    assert v % 8 == 0, "closed-walk count inconsistent with degrees"
    return v // 8
