"""String-typed relations: the reference's LUBM tests (test/rdf/travis.py:8-223)
over its own data (test/rdf/data, ten relations dictionary-encoded into one
string encoding, committed as tests/golden/lubm.npz).  Every rule's row count
is the reference's assertion; rows are sets, compared exactly."""
import os

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import rdf

HERE = os.path.dirname(os.path.abspath(__file__))
FORMS = [E.FORM_PLAIN, E.FORM_TUNED, E.FORM_LEARNED]


@pytest.fixture(scope="module")
def lubm():
    return rdf.load_fixture(os.path.join(HERE, "golden", "lubm.npz"))


def test_string_encoding_rules_and_file_format(tmp_path, lubm, fixtures_meta):
    """Keys are ranks in bytewise order over every value of the type
    (Encoding.cpp:21-48); the binary file is the reference's string layout
    (a size_t count, then per key a size_t length and the bytes,
    Encoding.cpp:106-118) and reads back unchanged."""
    enc, coded = lubm
    meta = fixtures_meta["lubm"]
    assert enc.num_distinct == meta["distinct_strings"] == 26407
    assert all(len(coded[n][0]) == meta["rows"][n] for n in meta["rows"])
    b = [v.encode() for v in enc.key_to_value]
    assert all(x < y for x, y in zip(b, b[1:]))
    enc.to_binary(str(tmp_path / "string"))
    raw = open(tmp_path / "string" / "encoding.bin", "rb").read()
    assert np.frombuffer(raw[:8], np.uint64)[0] == enc.num_distinct
    first = enc.key_to_value[0].encode()
    assert np.frombuffer(raw[8:16], np.uint64)[0] == len(first) and raw[16:16 + len(first)] == first
    back = rdf.StringEncoding.from_binary(str(tmp_path / "string"))
    assert back.key_to_value == enc.key_to_value
    small = rdf.StringEncoding(["b", "a", "ab", "B", "b"])
    assert small.key_to_value == ["B", "a", "ab", "b"] and small.value_to_key("ab") == 2
    with pytest.raises(KeyError):
        small.value_to_key("absent")


def test_lubm_oracle_goldens(oracle, lubm, fixtures_meta):
    enc, coded = lubm
    exp = fixtures_meta["lubm"]["rows_expected"]
    for q, (tries, levels, head, rows) in rdf.lubm_queries(enc).items():
        T = [oracle.Trie(*rdf.relation_columns(coded, n, o)) for n, o in tries]
        R = oracle.join_materialize(levels, T)
        assert len(R) == rows == exp[q], q
        assert len({tuple(r) for r in R.tolist()}) == len(R)


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS)
def test_lubm_goldens_every_form(ctx, oracle, lubm, fixtures_meta, form):
    """The seven LUBM rules through ehx_join_materialize and ehx_join_count:
    the reference's row counts and every row equal to the oracle's, under the
    density rule and all-uint layouts."""
    enc, coded = lubm
    exp = fixtures_meta["lubm"]["rows_expected"]
    cache = {}
    for policy in (E.POLICY_RULE, E.POLICY_ALL_UINT):
        for q, (tries, levels, head, rows) in rdf.lubm_queries(enc).items():
            T = []
            for n, o in tries:
                if (n, o, policy) not in cache:
                    cache[(n, o, policy)] = E.Trie(ctx, *rdf.relation_columns(coded, n, o), policy=policy)
                T.append(cache[(n, o, policy)])
            R = E.join_materialize(ctx, levels, T, form)
            assert len(R) == rows == exp[q], (q, form, policy)
            O = [oracle.Trie(*rdf.relation_columns(coded, n, o)) for n, o in tries]
            assert np.array_equal(R.reshape(-1, len(head)), oracle.join_materialize(levels, O).reshape(-1, len(head)))
            assert E.join_count(ctx, levels, T, form) == rows
    for t in cache.values():
        t.close()


@pytest.mark.gpu
def test_lubm_generated_programs_dropin(tmp_path, oracle, lubm, fixtures_meta):
    """The generated lubm1 and lubm6 programs through the drop-in: their
    selection constants go through Encoding<string> read from the reference's
    string layout, and the output rows (keys) equal the oracle's."""
    from emptyheaded_amd import dropin
    enc, coded = lubm
    db = str(tmp_path / "db")
    dropin.write_string_database(db, enc, coded)
    exp = fixtures_meta["lubm"]["rows_expected"]
    queries = rdf.lubm_queries(enc)
    for q in ("lubm1", "lubm6"):
        tries, levels, head, rows = queries[q]
        got, _, nr, n = dropin.rows(q, db, 10000)
        assert nr == n == exp[q], q
        O = [oracle.Trie(*rdf.relation_columns(coded, nm, o)) for nm, o in tries]
        assert np.array_equal(got[:, 0], oracle.join_materialize(levels, O)[:, 0])
        assert dropin.run(q, db)[1] == exp[q]
