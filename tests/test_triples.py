"""Triple-table joins (2-hop star and path over two predicates) under every
layout policy, including the learned selector: results are bit-exact and equal
the closed forms sum_s d_i(s) d_j(s) and sum_y in_i(y) out_j(y)."""
import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import triples as TR

PAIRS = [(0, 1), (0, 5), (3, 2), (1, 1)]


def small_table():
    s, p, o = TR.triple_table(n=2_000_000, preds=16, scale=12, seed=TR.TRIPLE_SEED + 1)
    return [TR.distinct(*r) for r in TR.by_predicate(s, p, o, 16)]


def test_zipf_generator():
    z = TR.zipf(400_000, 64, 1.1, 7)
    assert np.array_equal(z, TR.zipf(400_000, 64, 1.1, 7))
    f = np.bincount(z, minlength=64)
    assert z.max() < 64 and f.sum() == 400_000
    # frequencies follow (r+1)^-1.1 within sampling noise
    expect = np.arange(1, 65, dtype=np.float64) ** -1.1
    expect *= 400_000 / expect.sum()
    assert np.all(np.abs(f - expect) < 6 * np.sqrt(expect) + 5)
    assert np.all(TR.zipf(1000, 1, 1.1, 3) == 0)


def test_oracle_joins_match_closed_forms(oracle):
    rels = small_table()
    for i, j in PAIRS:
        Ti, Tj = oracle.Trie(*rels[i]), oracle.Trie(*rels[j])
        assert oracle.join_count(TR.STAR2, [Ti, Tj], 4) == TR.star_count(rels[i], rels[j])
        assert oracle.join_count(TR.PATH2, [Ti, Tj], 4) == TR.path_count(rels[i], rels[j])


def test_bitset_pool_words():
    src = np.array([0, 0, 0, 5, 5], dtype=np.uint32)
    dst = np.array([1, 63, 64, 700, 701], dtype=np.uint32)
    # sets {1,63,64}: words 0..1 = 2; {700,701}: 1; head {0,5}: 1
    assert TR.bitset_pool_words(src, dst) == 4


@pytest.mark.gpu
@pytest.mark.parametrize("policy", [E.POLICY_RULE, E.POLICY_ALL_UINT, E.POLICY_ALL_BITSET, E.POLICY_LEARNED])
def test_triple_joins_every_policy(ctx, policy):
    from emptyheaded_amd import tuning
    tuning.apply_committed()
    rels = small_table()
    tries = {}
    for i, j in PAIRS:
        for k in (i, j):
            if k not in tries:
                tries[k] = E.Trie(ctx, *rels[k], policy=policy)
        star, path = TR.star_count(rels[i], rels[j]), TR.path_count(rels[i], rels[j])
        for form in (E.FORM_PLAIN, E.FORM_LEARNED):
            assert E.join_count(ctx, TR.STAR2, [tries[i], tries[j]], form) == star
            assert E.join_count(ctx, TR.PATH2, [tries[i], tries[j]], form) == path
