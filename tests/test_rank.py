"""Recursive annotated SUM (the reference's PageRank rule) on the GPU against
the oracle: the plain form bit-identical to the generated code's float32 fold,
the tuned form within the rounding of a different summation order."""
import os

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


def inv_degree():
    d = np.load(os.path.join(HERE, "golden", "inv_degree.npz"))
    return d["key"], d["val"]


def tuned_close(a, b, iterations):
    # Both orders round each of a head's d float32 additions once: |a - b| is at
    # most 2 d 2^-24 times the sum of the terms; through the recursion the error
    # of the terms adds in.  rtol 1e-4 holds for the graphs here (degrees up to
    # a few thousand, 5 iterations, terms of one sign).
    return np.allclose(a, b, rtol=1e-4 if iterations else 0, atol=0)


def test_pagerank_facebook_golden(ctx, oracle, facebook_dup, fixtures_meta):
    meta = fixtures_meta["inv_degree"]
    keys, vals = inv_degree()
    s, d = facebook_dup
    N = meta["PageRank_N"]
    T = E.Trie(ctx, s, d)
    exp = oracle.pagerank(s, d, keys, vals, 1.0 / N, iterations=5, nthreads=4)
    for form in (E.FORM_PLAIN, E.FORM_TUNED):
        say("facebook pagerank form", form)
        r, ms = E.pagerank(ctx, T, keys, vals, 1.0 / N, iterations=5, form=form)
        assert abs(float(r[0]) - meta["PageRank_row0"]) < meta["PageRank_tolerance"]
        if form == E.FORM_PLAIN:
            assert np.array_equal(r.view(np.uint32), exp.view(np.uint32))
        else:
            assert tuned_close(r, exp, 5)
            r2, _ = E.pagerank(ctx, T, keys, vals, 1.0 / N, iterations=5, form=form)
            assert np.array_equal(r.view(np.uint32), r2.view(np.uint32))  # deterministic


@pytest.mark.parametrize("scale", [10, 14, 16])
def test_pagerank_rmat_symmetric(ctx, oracle, scale):
    """Both directions of R-MAT edges (hubs of more than 512 children take the
    wavefront kernel at scale 16), W = 1/out-degree; iteration counts 0, 1, 5."""
    s, d, n = G.rmat_graph(scale)
    S = np.concatenate([s, d])
    D = np.concatenate([d, s])
    heads, deg = np.unique(S, return_counts=True)
    keys = heads.astype(np.uint32)
    vals = (1.0 / deg).astype(np.float32)
    T = E.Trie(ctx, S, D)
    if scale == 16:
        assert deg.max() > 512
    for it in (0, 1, 5):
        exp = oracle.pagerank(S, D, keys, vals, 1.0 / len(heads), iterations=it, nthreads=8)
        say("rmat", scale, "iterations", it)
        r, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=it, form=E.FORM_PLAIN)
        assert np.array_equal(r.view(np.uint32), exp.view(np.uint32))
        r, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=it, form=E.FORM_TUNED)
        assert tuned_close(r, exp, it)


def test_pagerank_sparse_ids_and_partial_w(ctx, oracle):
    """Ids spread over a wide domain (child lookup by binary search, no dense
    array), W on part of the values and with keys past the largest value, a
    directed graph, other rule constants."""
    rng = np.random.default_rng(3)
    n, m = 3000, 40000
    s = (rng.integers(0, n, m) * 977).astype(np.uint32)
    d = (rng.integers(0, n, m) * 977).astype(np.uint32)
    wk = np.unique(np.concatenate([rng.integers(0, n, n // 3) * 977, [n * 977 + 5, 2 ** 31 + 7]])).astype(np.uint32)
    wv = rng.uniform(0.0, 2.0, len(wk)).astype(np.float32)
    T = E.Trie(ctx, s, d)
    for base, add, mul, init, it in ((0.25, 0.1, 0.7, 2.0, 4), (1.0, 0.0, 1.0, 1.0, 1)):
        exp = oracle.pagerank(s, d, wk, wv, base, add, mul, init, it, nthreads=8)
        r, _ = E.pagerank(ctx, T, wk, wv, base, add, mul, init, it, form=E.FORM_PLAIN)
        assert np.array_equal(r.view(np.uint32), exp.view(np.uint32))
        r, _ = E.pagerank(ctx, T, wk, wv, base, add, mul, init, it, form=E.FORM_TUNED)
        assert tuned_close(r, exp, it)


def test_pagerank_edge_cases(ctx, oracle):
    s = np.array([0, 0, 1, 2], np.uint32)
    d = np.array([1, 2, 2, 0], np.uint32)
    T = E.Trie(ctx, s, d)
    # no annotations: every term is absent, every rank is `add`
    r, _ = E.pagerank(ctx, T, np.zeros(0, np.uint32), np.zeros(0, np.float32), 0.5, iterations=2,
                      form=E.FORM_PLAIN)
    assert np.array_equal(r, np.full(3, np.float32(0.15)))
    # unsorted or repeated keys are rejected before any launch
    for bad in (np.array([2, 1], np.uint32), np.array([1, 1], np.uint32)):
        with pytest.raises(E.EhxError):
            E.pagerank(ctx, T, bad, np.ones(2, np.float32), 0.5)
    # the fold order of the generated code, bit for bit (see the oracle's test)
    s = np.array([0, 0, 0, 1, 2, 3], np.uint32)
    d = np.array([1, 2, 3, 0, 0, 0], np.uint32)
    T = E.Trie(ctx, s, d)
    keys = np.array([1, 2, 3], np.uint32)
    vals = np.array([1.0, 2.0 ** -24, 2.0 ** -24], np.float32)
    r, _ = E.pagerank(ctx, T, keys, vals, 1.0, add=0.0, mul=1.0, init=1.0, iterations=1, form=E.FORM_PLAIN)
    assert r[0] == np.float32(1.0)


def test_pagerank_hubs_across_chunks(ctx, oracle):
    """Hubs of 600, 5,000 and 20,000 children (one, two and five chunks of
    4,096 in the tuned form) beside small heads, both edge directions."""
    rng = np.random.default_rng(8)
    n = 30000
    parts = [(np.full(k, h), rng.choice(n, k, replace=False)) for h, k in ((7, 600), (11, 5000), (29999, 20000))]
    s = np.concatenate([p[0] for p in parts] + [rng.integers(0, n, 60000)]).astype(np.uint32)
    d = np.concatenate([p[1] for p in parts] + [rng.integers(0, n, 60000)]).astype(np.uint32)
    S, D = np.concatenate([s, d]), np.concatenate([d, s])
    heads, deg = np.unique(S, return_counts=True)
    keys, vals = heads.astype(np.uint32), (1.0 / deg).astype(np.float32)
    T = E.Trie(ctx, S, D)
    exp = oracle.pagerank(S, D, keys, vals, 1.0 / len(heads), iterations=5, nthreads=8)
    r, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=5, form=E.FORM_PLAIN)
    assert np.array_equal(r.view(np.uint32), exp.view(np.uint32))
    r, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=5, form=E.FORM_TUNED)
    assert tuned_close(r, exp, 5)
    r2, _ = E.pagerank(ctx, T, keys, vals, 1.0 / len(heads), iterations=5, form=E.FORM_TUNED)
    assert np.array_equal(r.view(np.uint32), r2.view(np.uint32))
