"""Synthetic triple table (subject, predicate, object) and the 2-hop joins the
learned layout selector is evaluated on.

The table holds `n` uint32 triples: predicates follow Zipf(s) over `preds`
ranks (rank 0 the most frequent); (subject, object) pairs are R-MAT samples
over 2^scale ids, so subjects and objects have skewed degrees.  Each predicate
is one binary relation P_k(subject, object) with set semantics (duplicates
dropped by the trie build).  Queries (COUNT(*), single bag):

  star  Q(;c) :- P_i(s,o1), P_j(s,o2)      = sum_s d_i(s) * d_j(s)
  path  Q(;c) :- P_i(x,y), P_j(y,z)        = sum_y in_i(y) * out_j(y)
"""
import numpy as np

from . import check, lib, ptr_u32
from .graphs import SEED_BASE, rmat_samples

TRIPLE_SEED = SEED_BASE + 0x50

# star: attributes (s, o1, o2); relation 0 = P_i, 1 = P_j, both keyed by subject
STAR2 = [[(0, -1), (1, -1)],
         [(0, 0)],
         [(1, 0)]]
# path: attributes (x, y, z); relation 0 = P_i keyed by x, 1 = P_j keyed by y
PATH2 = [[(0, -1)],
         [(0, 0), (1, -1)],
         [(1, 1)]]


def zipf(n, k, s, seed):
    out = np.zeros(n, dtype=np.uint32)
    check(lib().ehx_gen_zipf(n, k, s, seed, ptr_u32(out)))
    return out


def triple_table(n=50_000_000, preds=64, zipf_s=1.1, scale=22, seed=TRIPLE_SEED):
    """Returns (subject, predicate, object) uint32 columns."""
    s, o = rmat_samples(scale, n, seed)
    p = zipf(n, preds, zipf_s, seed ^ 0x5A5A5A5A)
    return s, p, o


def by_predicate(s, p, o, preds):
    """Per-predicate (subject, object) columns, in table order."""
    order = np.argsort(p, kind="stable")
    bounds = np.zeros(preds + 1, dtype=np.int64)
    np.cumsum(np.bincount(p, minlength=preds), out=bounds[1:])
    ss, oo = s[order], o[order]
    return [(ss[bounds[k]:bounds[k + 1]], oo[bounds[k]:bounds[k + 1]]) for k in range(preds)]


def distinct(src, dst):
    """Distinct rows, sorted."""
    key = np.unique((src.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64))
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def star_count(rel_i, rel_j):
    """sum over subjects of d_i(s) * d_j(s) (relations as distinct rows)."""
    si, sj = rel_i[0], rel_j[0]
    n = int(max(si.max(initial=0), sj.max(initial=0))) + 1
    return int(np.dot(np.bincount(si, minlength=n).astype(np.int64), np.bincount(sj, minlength=n).astype(np.int64)))


def path_count(rel_i, rel_j):
    """sum over y of in_i(y) * out_j(y) (relations as distinct rows)."""
    yi, sj = rel_i[1], rel_j[0]
    n = int(max(yi.max(initial=0), sj.max(initial=0))) + 1
    return int(np.dot(np.bincount(yi, minlength=n).astype(np.int64), np.bincount(sj, minlength=n).astype(np.int64)))


def bitset_pool_words(src, dst):
    """Words a range_bitset layout of every child set (and the head) would take:
    (last >> 6) - (first >> 6) + 1 per set (distinct sorted rows)."""
    if len(src) == 0:
        return 0
    starts = np.flatnonzero(np.r_[True, src[1:] != src[:-1]])
    ends = np.r_[starts[1:], len(src)] - 1
    words = (dst[ends] >> 6).astype(np.int64) - (dst[starts] >> 6).astype(np.int64) + 1
    heads = src[starts]
    return int(words.sum() + (int(heads[-1]) >> 6) - (int(heads[0]) >> 6) + 1)
