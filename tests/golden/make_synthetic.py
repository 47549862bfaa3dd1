"""Writes tests/golden/synthetic.json: the oracle's counts on the synthetic
workloads the benchmarks run (R-MAT and ego-network batch generators with
their committed seeds).  These pin the generators and guard regressions; the
oracle itself is pinned by the reference's golden counts and by brute force
(tests/test_oracle.py, tests/test_batch.py).

    python tests/golden/make_synthetic.py [--only-triples]
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from emptyheaded_amd import graphs as G  # noqa: E402
from emptyheaded_amd import plans  # noqa: E402
from oracle import oracle as O  # noqa: E402

THREADS = os.cpu_count() or 8


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def triples_entry():
    """The triple table's two most frequent predicates (distinct rows) and the
    closed forms of their 2-hop star and path joins (no join engine involved)."""
    from emptyheaded_amd import triples as TR
    s, p, o = TR.triple_table()
    rels = TR.by_predicate(s, p, o, 64)
    del s, p, o
    R = [TR.distinct(*rels[0]), TR.distinct(*rels[1])]
    return {"triples": 50_000_000, "predicates": 64, "zipf_s": 1.1, "rows": [len(R[0][0]), len(R[1][0])],
            "sha256": digest(R[0][0], R[0][1], R[1][0], R[1][1]),
            "star": TR.star_count(R[0], R[1]), "path": TR.path_count(R[0], R[1])}


def main():
    if "--only-triples" in sys.argv:  # refresh one entry of the existing file
        with open(os.path.join(HERE, "synthetic.json")) as f:
            meta = json.load(f)
        meta["triples_50m_top2"] = triples_entry()
        with open(os.path.join(HERE, "synthetic.json"), "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print(json.dumps(meta["triples_50m_top2"], indent=1))
        return
    meta = {}
    s, d, n = G.er_graph()
    meta["er_65536"] = {"vertices": n, "edges": len(s), "triangles": O.join_count(plans.TRIANGLE, [O.Trie(s, d)], THREADS)}
    for scale in (18, 20, 22):
        s, d, n = G.rmat_graph(scale)
        T = O.Trie(s, d)
        e = {"vertices": n, "edges": len(s), "sha256": digest(s, d),
             "triangles": O.join_count(plans.TRIANGLE, [T], THREADS)}
        if scale == 18:
            e["four_cliques"] = O.join_count(plans.FOUR_CLIQUE, [T], THREADS)
        meta["rmat%d" % scale] = e
        del T
    s, d, n = G.rmat_graph(22, orient=G.BOTH)
    es, ed, base, off = G.egonets(s, d, n)
    del s, d
    O_T = O.Trie(es, ed)
    tri = O.join_count(plans.TRIANGLE, [O_T], THREADS)
    del O_T
    ss, sd = G.symmetric(es, ed)
    cyc = O.cycle4_count(ss, sd, int(base[-1]), THREADS)
    meta["egonets_rmat22"] = {"nets": len(base) - 1, "edges": len(es), "vertices": int(base[-1]),
                              "sha256": digest(es, ed, base, off), "triangles": tri, "cycles4": cyc,
                              "seed": G.EGO_SEED, "min_edges": 64, "max_edges": 65536}
    meta["triples_50m_top2"] = triples_entry()
    with open(os.path.join(HERE, "synthetic.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
