"""Writes the committed golden fixtures for the generic-join parity tests.

Inputs are data files held by the reference's own tests (SNAP ego-Facebook,
a public dataset): test/graph/data/facebook_pruned.tsv (88,234 edges, every
row src > dst) used by test/graph/travis.py:197-358.  The expected results are
the constants those tests assert:
  TriangleAgg  = 1,612,010   (travis.py:197-213)
  FliqueAgg    = 30,004,668  (travis.py:238-254)
  Triangle     = 1,612,010 rows, first row (6,5,2) (travis.py:302-318)
  Flique       = 30,004,668 rows, first row (9,8,7,0) (travis.py:342-354)
  TriangleProj = 85,658 rows, first row (6,5)  (travis.py:282-300)
and, for the graph with both edge directions (facebook_duplicated.tsv),
9,672,060 triangle tuples (docs/notebooks/Getting Started.ipynb, cell 22),
and the selections test_duplicated runs (travis.py:647-693):
  FliqueSelAgg, x=0  = 3,759,972  (travis.py:385-400; FliqueSel rows, :429-444)
  BarbellSelAgg, p=6 = 26,936,100 (travis.py:473-489; BarbellSel rows, :519-530)
  LollipopAgg        = 1,426,911,480      (travis.py:14-30)
  BarbellAgg         = 20,371,831,447,136 (travis.py:55-70)
and the materialized multi-bag queries over test/graph/data/simple.tsv (two
triangles joined by an edge, both directions), read back through the
reference's DataFrame (its columns the output trie's levels permuted by the
relation's ordering, codegen/trie/Trie.scala:219-230):
  Lollipop(a,b,c,x)     = 28 rows, df.iloc[27] = (5,3,4,4)      (travis.py:157-172)
  Barbell(a,b,c,x,y,z)  = 56 rows, df.iloc[55] = (5,3,4,3,4,5)  (travis.py:109-125,
                          checked there under check_big_out)
the LUBM relations of test/rdf/travis.py:245-300 (ten string-typed relations
from test/rdf/data/, dictionary-encoded into one string encoding) with the row
counts its rules assert: lubm1 4, lubm2 0, lubm4 14, lubm6 5,916, lubm7 59,
lubm8 5,916, lubm12 125 (travis.py:8-223), stored as the sorted dictionary
and each relation's key columns (tests/golden/lubm.npz);
and PageRank over facebook_duplicated with test/graph/data/inv_degree.tsv
(4,039 float32 inverse degrees): the rank of the first vertex after the base
case and 5 iterations is 15.227079960463206, checked to 0.0001
(travis.py:585-598; the relation is loaded as float32, travis.py:659-667).

Run once in a container that holds the reference tree:
    python tests/golden/make_fixtures.py /root/reference
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_tsv(path):
    raw = open(path, "rb").read()
    arr = np.loadtxt(path, dtype=np.int64, delimiter="\t", ndmin=2)
    return raw, arr


def main(ref_root):
    src = os.path.join(ref_root, "test", "graph", "data")
    raw, fb = load_tsv(os.path.join(src, "facebook_pruned.tsv"))
    assert fb.shape == (88234, 2) and fb.min() >= 0 and fb.max() < 2**32
    np.savez_compressed(os.path.join(HERE, "facebook_pruned.npz"),
                        src=fb[:, 0].astype(np.uint32), dst=fb[:, 1].astype(np.uint32))
    rawd, dup = load_tsv(os.path.join(src, "facebook_duplicated.tsv"))
    rawt, simple = load_tsv(os.path.join(src, "simple.tsv"))
    np.savez_compressed(os.path.join(HERE, "simple.npz"),
                        src=simple[:, 0].astype(np.uint32), dst=simple[:, 1].astype(np.uint32))
    # facebook_duplicated holds both directions of every edge, under the
    # original SNAP vertex ids (facebook_pruned is a relabelled copy).
    np.savez_compressed(os.path.join(HERE, "facebook_duplicated.npz"),
                        src=dup[:, 0].astype(np.uint32), dst=dup[:, 1].astype(np.uint32))
    rawi = open(os.path.join(src, "inv_degree.tsv"), "rb").read()
    inv = np.loadtxt(os.path.join(src, "inv_degree.tsv"), dtype=np.float64, delimiter="\t", ndmin=2)
    assert inv.shape == (4039, 2)
    np.savez_compressed(os.path.join(HERE, "inv_degree.npz"), key=inv[:, 0].astype(np.uint32),
                        val=inv[:, 1].astype(np.float32))
    # LUBM: one string encoding over the ten relations
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from emptyheaded_amd import rdf
    rdir = os.path.join(ref_root, "test", "rdf", "data")
    rels = {n: rdf.read_tsv_strings(os.path.join(rdir, f)) for n, f in rdf.LUBM_RELATIONS}
    enc, coded = rdf.encode_database(rels)
    np.savez_compressed(os.path.join(HERE, "lubm.npz"),
                        strings=np.frombuffer("\n".join(enc.key_to_value).encode("utf-8"), dtype=np.uint8),
                        **{n: np.stack(coded[n], 1).astype(np.uint32) for n, _ in rdf.LUBM_RELATIONS})
    lubm_meta = {
        "source": "test/rdf/data/*.tsv (relations of test/rdf/travis.py:245-300)",
        "sha256": {f: hashlib.sha256(open(os.path.join(rdir, f), "rb").read()).hexdigest()
                   for _, f in rdf.LUBM_RELATIONS},
        "rows": {n: len(rels[n][0]) for n, _ in rdf.LUBM_RELATIONS},
        "distinct_strings": enc.num_distinct,
        "rows_expected": {"lubm1": 4, "lubm2": 0, "lubm4": 14, "lubm6": 5916, "lubm7": 59, "lubm8": 5916,
                          "lubm12": 125},
        "cite": "test/rdf/travis.py:8-223",
    }
    pairs = {tuple(r) for r in dup.tolist()}
    dup_is_sym = all((y, x) in pairs for (x, y) in pairs)
    meta = {
        "facebook_pruned": {
            "source": "test/graph/data/facebook_pruned.tsv",
            "sha256": hashlib.sha256(raw).hexdigest(),
            "rows": int(fb.shape[0]),
            "TriangleAgg": 1612010,
            "FliqueAgg": 30004668,
            "Triangle_rows": 1612010,
            "Triangle_row0": [6, 5, 2],
            "Flique_rows": 30004668,
            "Flique_row0": [9, 8, 7, 0],
            "TriangleProj_rows": 85658,
            "TriangleProj_row0": [6, 5],
            "cite": "test/graph/travis.py:197-213,238-254,282-300,302-318,342-354",
        },
        "facebook_duplicated": {
            "source": "test/graph/data/facebook_duplicated.tsv",
            "sha256": hashlib.sha256(rawd).hexdigest(),
            "rows": int(dup.shape[0]),
            "symmetric": bool(dup_is_sym),
            "Triangle_rows": 9672060,
            "FliqueSelAgg_x0": 3759972,
            "BarbellSelAgg_p6": 26936100,
            "LollipopAgg": 1426911480,
            "BarbellAgg": 20371831447136,
            "cite": "docs/notebooks/Getting Started.ipynb cell 22; test/graph/travis.py:14-30,55-70,385-400,429-444,473-489,519-530",
        },
        "inv_degree": {
            "source": "test/graph/data/inv_degree.tsv",
            "sha256": hashlib.sha256(rawi).hexdigest(),
            "rows": int(inv.shape[0]),
            "PageRank_N": int(len(np.unique(dup[:, 0]))),
            "PageRank_iterations": 5,
            "PageRank_row0": 15.227079960463206,
            "PageRank_tolerance": 0.0001,
            "cite": "test/graph/travis.py:585-598,659-667",
        },
        "lubm": lubm_meta,
        "simple": {
            "source": "test/graph/data/simple.tsv",
            "sha256": hashlib.sha256(rawt).hexdigest(),
            "rows": int(simple.shape[0]),
            "Lollipop_rows": 28,
            "Lollipop_row27": [5, 3, 4, 4],
            "Barbell_rows": 56,
            "Barbell_row55": [5, 3, 4, 3, 4, 5],
            "cite": "test/graph/travis.py:109-125,157-172",
        },
    }
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
