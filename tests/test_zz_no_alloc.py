"""The entry point that allocates nothing (ehx_intersect_count_batch_noalloc)
against ehx_intersect_count_batch, through the C ABI: equal counts on two
calls in a row (the second smaller, so stale counts of the first would show),
on every layout pair, absent keys included; the size limit is an error."""
import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G

pytestmark = pytest.mark.gpu


def test_noalloc_counts_equal_allocating_entry(ctx):
    s, d, n = G.rmat_graph(14)
    rng = np.random.default_rng(99)
    tries = {p: E.Trie(ctx, s, d, p) for p in (E.POLICY_RULE, E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET)}
    heads = np.unique(s)
    for pa, pb, sizes in ((E.POLICY_RULE, E.POLICY_RULE, (300_000, 1_000)),
                          (E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET, (E.NOALLOC_MAX_PAIRS, 77)),
                          (E.POLICY_ALL_BITSET, E.POLICY_ALL_BITSET, (5_000, 4_999))):
        for k in sizes:  # two calls in a row, the second smaller
            ka = rng.choice(heads, k).astype(np.uint32)
            kb = rng.choice(heads, k).astype(np.uint32)
            ka[:3] = [n + 5, heads[0], 2**32 - 1]  # absent keys count 0
            want = E.intersect_count(ctx, tries[pa], ka, tries[pb], kb)
            got = E.intersect_count_noalloc(ctx, tries[pa], ka, tries[pb], kb)
            assert np.array_equal(got, want), (pa, pb, k)
            assert got[0] == 0 and got[2] == 0
    T = tries[E.POLICY_RULE]
    assert len(E.intersect_count_noalloc(ctx, T, [], T, [])) == 0
    with pytest.raises(E.EhxError):
        big = np.zeros(E.NOALLOC_MAX_PAIRS + 1, np.uint32)
        E.intersect_count_noalloc(ctx, T, big, T, big)
