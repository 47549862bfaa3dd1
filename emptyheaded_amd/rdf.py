"""String-typed relations, as the reference's RDF tests use them (LUBM,
test/rdf/travis.py), on the join engine.

Dictionary encoding (storage_engine/src/Encoding.cpp:21-48): the values of one
attribute type across every relation of the database form one encoding, and
each distinct value's key is its rank in sorted order (a std::set of the
values, so std::string order: bytewise).  StringEncoding writes and reads the
reference's on-disk form of Encoding<std::string> (Encoding.cpp:54-77,
106-118: a size_t count, then per key a size_t length and the bytes).

The LUBM queries are single rules whose non-output attributes are all
equality selections on constants; each becomes a plan over the relations'
two orderings (Edge_0_1-style and the reverse _1_0 for a selected first
column), with the selected constants as CONST operands: the child set of the
constant's key, as the generated selection path reads it
(query_compiler/src/main/scala/codegen/query/EHGenerator.scala:657-701).
"""
import os

import numpy as np

from .plans import CONST, HEAD

# The relations of test/rdf/travis.py:245-300, in its order, with their files
# (undegraduateDegreeFrom is the name the test gives undergraduateDegreeFrom.tsv).
LUBM_RELATIONS = [("takesCourse", "takesCourse.tsv"), ("memberOf", "memberOf.tsv"),
                  ("subOrganizationOf", "subOrganizationOf.tsv"),
                  ("undegraduateDegreeFrom", "undergraduateDegreeFrom.tsv"), ("rdftype", "rdftype.tsv"),
                  ("worksFor", "worksFor.tsv"), ("name", "name.tsv"), ("emailAddress", "emailAddress.tsv"),
                  ("telephone", "telephone.tsv"), ("teacherOf", "teacherOf.tsv")]

UB = "http://www.lehigh.edu/~zhp2/2004/0401/univ-bench.owl#"


class StringEncoding:
    """Encoding<std::string>: keys are ranks of the sorted distinct values."""

    def __init__(self, values):
        self.key_to_value = sorted(set(values), key=lambda s: s.encode("utf-8"))
        self._keys = {v: k for k, v in enumerate(self.key_to_value)}

    @property
    def num_distinct(self):
        return len(self.key_to_value)

    def value_to_key(self, v):
        return self._keys[v]  # KeyError for an absent value, as value_to_key.at throws

    def encode(self, column):
        return np.fromiter((self._keys[v] for v in column), dtype=np.uint32, count=len(column))

    def to_binary(self, folder):
        """<folder>/encoding.bin in the reference's string layout."""
        os.makedirs(folder, exist_ok=True)
        with open(os.path.join(folder, "encoding.bin"), "wb") as f:
            f.write(np.uint64(self.num_distinct).tobytes())
            for v in self.key_to_value:
                b = v.encode("utf-8")
                f.write(np.uint64(len(b)).tobytes())
                f.write(b)

    @classmethod
    def from_binary(cls, folder):
        with open(os.path.join(folder, "encoding.bin"), "rb") as f:
            raw = f.read()
        n = int(np.frombuffer(raw[:8], np.uint64)[0])
        vals, p = [], 8
        for _ in range(n):
            ln = int(np.frombuffer(raw[p:p + 8], np.uint64)[0])
            vals.append(raw[p + 8:p + 8 + ln].decode("utf-8"))
            p += 8 + ln
        if p != len(raw):
            raise ValueError("encoding.bin: %d trailing bytes" % (len(raw) - p))
        e = cls.__new__(cls)
        e.key_to_value = vals
        e._keys = {v: k for k, v in enumerate(vals)}
        return e


def read_tsv_strings(path):
    """(first column, second column) of a two-column tab-separated file."""
    a, b = [], []
    with open(path, encoding="utf-8") as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            x, y = line.split("\t")
            a.append(x)
            b.append(y)
    return a, b


def encode_database(relations):
    """relations: {name: (col0 strings, col1 strings)} -> (StringEncoding,
    {name: (key col0, key col1)}), one encoding for the string type."""
    enc = StringEncoding(v for a, b in relations.values() for col in (a, b) for v in col)
    return enc, {n: (enc.encode(a), enc.encode(b)) for n, (a, b) in relations.items()}


def lubm_queries(enc):
    """The LUBM rules of test/rdf/travis.py with their expected row counts:
    {name: (tries, levels, head, expected)}.  `tries` lists (relation,
    ordering) pairs, "0_1" or "1_0"; `levels` is the plan over them (attribute
    order in `head`, the output attributes); selections are CONST operands."""
    k = enc.value_to_key
    d0 = "http://www.Department0.University0.edu"
    return {
        # lubm1(a) :- b='...GraduateCourse0', c=ub:GraduateStudent, takesCourse(a,b),rdftype(a,c).  (:8-20)
        "lubm1": ([("takesCourse", "1_0"), ("rdftype", "1_0")],
                  [[(0, CONST, k(d0 + "/GraduateCourse0")), (1, CONST, k(UB + "GraduateStudent"))]],
                  ("a",), 4),
        # lubm2(a,b,c) :- x=ub:GraduateStudent, y=ub:Department, z=ub:University, memberOf(a,b),
        #   subOrganizationOf(b,c),undegraduateDegreeFrom(a,c),rdftype(a,x),rdftype(b,y),rdftype(c,z).  (:38-51)
        "lubm2": ([("memberOf", "0_1"), ("undegraduateDegreeFrom", "0_1"), ("rdftype", "1_0"),
                   ("subOrganizationOf", "0_1")],
                  [[(0, HEAD), (1, HEAD), (2, CONST, k(UB + "GraduateStudent"))],
                   [(0, 0), (3, HEAD), (2, CONST, k(UB + "Department"))],
                   [(3, 1), (1, 0), (2, CONST, k(UB + "University"))]],
                  ("a", "b", "c"), 0),
        # lubm4(a,b,c,d) :- e=Department0, f=ub:AssociateProfessor, worksFor(a,e),name(a,b),
        #   emailAddress(a,d),telephone(a,c),rdftype(a,f).  (:76-88)
        "lubm4": ([("worksFor", "1_0"), ("name", "0_1"), ("emailAddress", "0_1"), ("telephone", "0_1"),
                   ("rdftype", "1_0")],
                  [[(0, CONST, k(d0)), (1, HEAD), (2, HEAD), (3, HEAD), (4, CONST, k(UB + "AssociateProfessor"))],
                   [(1, 0)],
                   [(3, 0)],
                   [(2, 0)]],
                  ("a", "b", "c", "d"), 14),
        # lubm6(a) :- b=ub:UndergraduateStudent, rdftype(a,b).  (:111-122)
        "lubm6": ([("rdftype", "1_0")],
                  [[(0, CONST, k(UB + "UndergraduateStudent"))]],
                  ("a",), 5916),
        # lubm7(a,b) :- c='.../AssociateProfessor0', d=ub:Course, e=ub:UndergraduateStudent,
        #   teacherOf(c,b),takesCourse(a,b),rdftype(b,d),rdftype(a,e).  (:139-152); order (b, a)
        "lubm7": ([("teacherOf", "0_1"), ("rdftype", "1_0"), ("takesCourse", "1_0")],
                  [[(0, CONST, k(d0 + "/AssociateProfessor0")), (1, CONST, k(UB + "Course")), (2, HEAD)],
                   [(2, 0), (1, CONST, k(UB + "UndergraduateStudent"))]],
                  ("b", "a"), 59),
        # lubm8(a,b,c) :- d=ub:UndergraduateStudent, e=University0, f=ub:Department, memberOf(a,b),
        #   emailAddress(a,c),rdftype(a,d),subOrganizationOf(b,e),rdftype(b,f).  (:174-187)
        "lubm8": ([("memberOf", "0_1"), ("emailAddress", "0_1"), ("rdftype", "1_0"), ("subOrganizationOf", "1_0")],
                  [[(0, HEAD), (1, HEAD), (2, CONST, k(UB + "UndergraduateStudent"))],
                   [(0, 0), (3, CONST, k("http://www.University0.edu")), (2, CONST, k(UB + "Department"))],
                   [(1, 0)]],
                  ("a", "b", "c"), 5916),
        # lubm12(a,b) :- c=ub:FullProfessor, d=University0, e=ub:Department, worksFor(a,b),
        #   rdftype(a,c),subOrganizationOf(b,d),rdftype(b,e).  (:210-223)
        "lubm12": ([("worksFor", "0_1"), ("rdftype", "1_0"), ("subOrganizationOf", "1_0")],
                   [[(0, HEAD), (1, CONST, k(UB + "FullProfessor"))],
                    [(0, 0), (2, CONST, k("http://www.University0.edu")), (1, CONST, k(UB + "Department"))]],
                   ("a", "b"), 125),
    }


def relation_columns(encoded, name, ordering):
    """(first, second) key columns of relation `name` under `ordering`."""
    a, b = encoded[name]
    return (a, b) if ordering == "0_1" else (b, a)


def load_fixture(path):
    """(StringEncoding, {name: (key col0, key col1)}) from a file holding the
    sorted dictionary ("strings", newline-joined) and each relation's key
    columns as an (n, 2) array."""
    z = np.load(path)
    strings = bytes(z["strings"]).decode("utf-8").split("\n")
    enc = StringEncoding(strings)
    assert enc.key_to_value == strings, "dictionary not in encoding order"
    return enc, {n: (z[n][:, 0].copy(), z[n][:, 1].copy()) for n, _ in LUBM_RELATIONS}
