"""Child pointers between tries and their trim hints (ehx_trie_link_hints).

hint[e] = how many values of row e's child set in the target trie lie below
row e's first-column value.  numpy gives it exactly: with the target's rows as
sorted keys (x << 32) | v, it is searchsorted((x << 32) | y) -
searchsorted(x << 32) for row e = (y, x).  Both routes are checked: one search
per row, and one sort of the rows by value when the target looks like the
source's reverse ordering, whose candidates are checked against the target and
searched where they fail.
"""
import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G

SEARCHED, TRANSPOSED = 2, 3


def say(*a):
    print(*a, flush=True)


def expected_hints(src_rows, dst_rows):
    """src_rows, dst_rows: (first, second) column arrays, any order, repeats
    allowed; the hints come in the source trie's row order."""
    ys, xs = (c.astype(np.uint64) for c in rows_of(*src_rows))
    dk = np.unique((np.asarray(dst_rows[0], dtype=np.uint64) << 32) | np.asarray(dst_rows[1], dtype=np.uint64))
    return (np.searchsorted(dk, (xs << 32) | ys) - np.searchsorted(dk, xs << 32)).astype(np.uint32)


def rows_of(s, d):
    """Distinct rows sorted by first column, then second (a trie's row order)."""
    k = np.unique((np.asarray(s, dtype=np.uint64) << 32) | np.asarray(d, dtype=np.uint64))
    return (k >> 32).astype(np.uint32), (k & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("route", ["sort", "search"])
def test_reverse_ordering_hints(ctx, route):
    s, d, n = G.rmat_graph(15)
    say("link rmat15 oriented rows", len(s), "route", route)
    T = E.Trie(ctx, s, d)
    R = E.Trie(ctx, d, s)
    h, state = R.link_hints(T, route)
    assert state == (TRANSPOSED if route == "sort" else SEARCHED)
    assert np.array_equal(h, expected_hints((d, s), (s, d)))
    assert h.any()
    # a relation linked to itself in ascending orientation keeps no hints
    h2, state2 = T.link_hints(T, route)
    assert state2 == 1 and not h2.any()


@pytest.mark.gpu
def test_symmetric_relation_linked_to_itself(ctx):
    s, d, n = G.rmat_graph(14, orient=G.BOTH)
    say("link rmat14 symmetric rows", len(s))
    T = E.Trie(ctx, s, d)
    h, state = T.link_hints(T, "sort")
    assert state == TRANSPOSED
    assert np.array_equal(h, expected_hints((s, d), (s, d)))


@pytest.mark.gpu
@pytest.mark.parametrize("drop", [0.001, 0.3])
def test_reverse_ordering_with_rows_removed(ctx, drop):
    """Every kept row still finds its parent value in its child set, so the
    sampled test takes the sort, but positions shift: the hints that fail the
    check against the target are searched."""
    s, d, n = G.rmat_graph(15)
    rng = np.random.default_rng(5)
    keep = rng.random(len(s)) >= drop
    say("link rows removed", drop, "kept", int(keep.sum()), "of", len(s))
    T = E.Trie(ctx, s, d)
    R = E.Trie(ctx, d[keep], s[keep])
    h, state = R.link_hints(T, "sort")
    assert state == TRANSPOSED
    assert np.array_equal(h, expected_hints((d[keep], s[keep]), (s, d)))


@pytest.mark.gpu
@pytest.mark.parametrize("route", ["sort", "auto"])
def test_rows_foreign_to_the_target(ctx, route):
    """Rows whose parent value is absent from their child set: whichever route
    the sampled test picks, every hint equals the exact count."""
    s, d, n = G.rmat_graph(15)
    rng = np.random.default_rng(9)
    k = 5000
    fs = rng.integers(0, n, size=k).astype(np.uint32)
    fd = rng.integers(0, n, size=k).astype(np.uint32)
    rs, rd = rows_of(np.concatenate([d, fs]), np.concatenate([s, fd]))
    T = E.Trie(ctx, s, d)
    R = E.Trie(ctx, rs, rd)
    h, state = R.link_hints(T, route)
    say("link foreign rows: route", route, "state", state)
    assert state in (SEARCHED, TRANSPOSED)
    assert np.array_equal(h, expected_hints((rs, rd), (s, d)))


@pytest.mark.gpu
def test_small_relation_is_searched(ctx):
    s, d, n = G.rmat_graph(10)
    assert len(s) < 65536
    T = E.Trie(ctx, s, d)
    R = E.Trie(ctx, d, s)
    h, state = R.link_hints(T)
    assert state == SEARCHED
    assert np.array_equal(h, expected_hints((d, s), (s, d)))


@pytest.mark.gpu
def test_link_route_is_checked(ctx):
    s, d, n = G.rmat_graph(10)
    T = E.Trie(ctx, s, d)
    with pytest.raises(ValueError):
        T.link_hints(T, "fastest")


def test_expected_hints_by_brute_force():
    """The numpy formula against a direct count on a small random pair."""
    rng = np.random.default_rng(3)
    a = rows_of(rng.integers(0, 40, 300), rng.integers(0, 40, 300))
    b = rows_of(rng.integers(0, 40, 300), rng.integers(0, 40, 300))
    child = {}
    for x, v in zip(*b):
        child.setdefault(int(x), []).append(int(v))
    want = [sum(1 for v in child.get(int(x), []) if v < int(y)) for y, x in zip(*a)]
    assert np.array_equal(expected_hints(a, b), np.array(want, dtype=np.uint32))
