"""Runs the restated generated queries (emptyheaded_amd/engine/queries) through
the C++ drop-in API, from a database folder laid out as the generated code
expects: <db>/relations/<R>/<R>_0_1/ehx_rows.bin."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN_LIB = os.path.join(HERE, "engine", "libehx_dropin.so")
QUERIES = ("TriangleAgg", "FliqueAgg", "Triangle", "TriangleCounts", "PageRankSteps", "WeightedDegree",
           "FliqueSelAgg", "SSSP", "SSSPIter2", "LollipopAgg", "BarbellAgg", "Lollipop", "Barbell", "TriangleProj", "lubm1", "lubm6")
# Materializing queries: output columns and whether rows carry an annotation.
MATERIALIZED = {"Triangle": (3, False), "TriangleCounts": (1, True), "PageRankSteps": (1, True),
                "WeightedDegree": (1, True), "SSSP": (1, True), "SSSPIter2": (1, True),
                "Lollipop": (4, False), "Barbell": (6, False), "TriangleProj": (2, False), "lubm1": (1, False),
                "lubm6": (1, False)}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(DROPIN_LIB):
            raise RuntimeError("libehx_dropin.so is not built: run `python -m emptyheaded_amd.build`")
        from . import lib as ehx_lib
        ehx_lib()  # libehx.so first
        L = C.CDLL(DROPIN_LIB)
        for q in QUERIES:
            f = getattr(L, "c_run_" + q)
            f.argtypes = [C.c_char_p, C.POINTER(C.c_ulong)]
            f.restype = C.c_long
        for q, (nc, annotated) in MATERIALIZED.items():
            f = getattr(L, "c_rows_" + q)
            f.argtypes = ([C.c_char_p, C.POINTER(C.c_uint32)] + ([C.POINTER(C.c_double)] if annotated else []) +
                          [C.c_ulong, C.POINTER(C.c_ulong)])
            f.restype = C.c_long
        _lib = L
    return _lib


def write_relation(db, name, src, dst, ordering="0_1"):
    """Stores a binary relation for Trie<A,M>::load (rows, then both columns)."""
    d = os.path.join(db, "relations", name, "%s_%s" % (name, ordering))
    os.makedirs(d, exist_ok=True)
    s = np.ascontiguousarray(src, np.uint32)
    t = np.ascontiguousarray(dst, np.uint32)
    with open(os.path.join(d, "ehx_rows.bin"), "wb") as f:
        f.write(np.uint64(len(s)).tobytes())
        f.write(s.tobytes())
        f.write(t.tobytes())
    os.makedirs(os.path.join(db, "encodings", "uint32_t"), exist_ok=True)
    n = int(max(s.max(), t.max())) + 1 if len(s) else 0
    with open(os.path.join(db, "encodings", "uint32_t", "ehx_encoding.bin"), "wb") as f:
        f.write(np.uint64(n).tobytes())
    return d


def write_string_database(db, encoding, relations):
    """A database of string-typed relations as the generated code reads it:
    relations/<R>/<R>_0_1 and <R>_1_0 (key rows) for every relation, and the
    type's encoding in the reference's layout under encodings/string/.
    relations: {name: (key col0, key col1)}; encoding: rdf.StringEncoding."""
    for name, (a, b) in relations.items():
        write_relation(db, name, a, b, "0_1")
        write_relation(db, name, b, a, "1_0")
    encoding.to_binary(os.path.join(db, "encodings", "string"))


def run(query, db):
    """Returns (annotation, num_rows) of the query's output bag."""
    rows = C.c_ulong(0)
    v = getattr(lib(), "c_run_" + query)(db.encode(), C.byref(rows))
    return v, rows.value


def rows(query, db, cap):
    """Runs a materializing query and reads its output trie back through
    Trie::foreach: returns (rows [r, ncols] uint32, annotations or None,
    num_rows, rows read).  At most `cap` rows are copied."""
    nc, annotated = MATERIALIZED[query]
    out = np.zeros((max(1, cap), nc), np.uint32)
    ann = np.zeros(max(1, cap), np.float64) if annotated else None
    num_rows = C.c_ulong(0)
    args = [db.encode(), out.ctypes.data_as(C.POINTER(C.c_uint32))]
    if annotated:
        args.append(ann.ctypes.data_as(C.POINTER(C.c_double)))
    args += [cap, C.byref(num_rows)]
    got = getattr(lib(), "c_rows_" + query)(*args)
    k = min(got, cap)
    return out[:k], (ann[:k] if annotated else None), num_rows.value, got
