"""The C-ABI library loads and exports every symbol the headers declare; host
logic that needs no GPU (generators, graph clean-up, argument checks)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import emptyheaded_amd as E
from emptyheaded_amd import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ehx_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    L = E.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and the binding knows each of them
    assert names <= set(E.SIGNATURES), sorted(names - set(E.SIGNATURES))


def test_version_and_errors_without_gpu():
    L = E.lib()
    assert L.ehx_version() >= 1
    h = C.c_void_p()
    if E.device_count() == 0:
        rc = L.ehx_ctx_create(0, C.byref(h))
        assert rc == -2 and b"GPU" in L.ehx_last_error()
    assert L.ehx_ctx_create(0, None) == -1
    assert L.ehx_trie_build(None, None, None, 0, 0, 0, None) == -1


def test_tune_params_validation():
    p = E.tune_get()
    with pytest.raises(E.EhxError):
        E.tune_set(lds_slots=3000)
    assert E.tune_get() == p
    E.tune_set(**p)


def test_generators_deterministic():
    a = G.rmat_samples(10, 5000, 42)
    b = G.rmat_samples(10, 5000, 42)
    c = G.rmat_samples(10, 5000, 43)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    assert a[0].max() < 1024 and a[1].max() < 1024
    e = G.er_samples(1000, 10000, 3)
    assert e[0].max() < 1000 and e[1].max() < 1000


def test_rmat_quadrant_frequencies():
    s, d = G.rmat_samples(1, 200000, 11)  # one level: quadrant probabilities directly
    q = np.bincount(s * 2 + d, minlength=4) / 200000
    assert np.allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_prepare_semantics():
    src = np.array([5, 1, 1, 2, 9, 9, 3], np.uint32)
    dst = np.array([1, 5, 2, 2, 3, 3, 9], np.uint32)
    a, b, n = G.prepare(src, dst, G.KEEP_IDS, G.SRC_LT_DST)
    assert list(zip(a, b)) == [(1, 2), (1, 5), (3, 9)] and n == 10
    a, b, n = G.prepare(src, dst, G.KEEP_IDS, G.BOTH)
    assert len(a) == 6
    a, b, n = G.prepare(src, dst, G.COMPACT, G.SRC_LT_DST)
    assert n == 5 and list(zip(a, b)) == [(0, 1), (0, 3), (2, 4)]
    # ascending degree: deg 1 -> {2,5,3,9}, deg 2 -> {1}; ties by id
    a, b, n = G.prepare(src, dst, G.DEGREE_ASC, G.SRC_LT_DST)
    assert n == 5
    lab = {2: 0, 3: 1, 5: 2, 9: 3, 1: 4}
    exp = sorted(tuple(sorted((lab[x], lab[y]))) for x, y in [(1, 2), (1, 5), (3, 9)])
    assert list(zip(a.tolist(), b.tolist())) == exp


def test_degree_orientation_bounds_outdegree():
    s, d, n = G.rmat_graph(14)
    dout = np.bincount(s, minlength=n)
    deg = G.degrees_undirected(s, d, n)
    assert (s < d).all()
    assert dout.max() <= np.sqrt(2 * len(s)) + 1
    assert deg.max() > 4 * dout.max()


def test_plan_encoding():
    p = E.make_plan(E.plans.TRIANGLE)
    assert p.num_attrs == 3 and list(p.num_ops)[:3] == [2, 2, 2]
    assert (p.ops[2][0].rel, p.ops[2][0].parent) == (0, 1)
    with pytest.raises(ValueError):
        E.make_plan([])


def test_tune_params_are_checked():
    saved = E.tune_get()
    try:
        for bad in ({"stream_lanes": 1}, {"stream_lanes": 12}, {"stream_lanes": 128}, {"lds_slots": 1000}):
            with pytest.raises(E.EhxError):
                E.tune_set(**bad)
        assert E.tune_set(stream_lanes=4)["stream_lanes"] == 4
    finally:
        E.tune_set(**saved)


def test_dropin_library_exports_every_query():
    """The drop-in runtime with the restated generated queries loads without a
    GPU and exports each query's c_run_<id> (and the row readers of the
    materializing ones)."""
    from emptyheaded_amd import dropin
    L = dropin.lib()
    for q in dropin.QUERIES:
        assert hasattr(L, "c_run_" + q)
    for q in dropin.MATERIALIZED:
        assert hasattr(L, "c_rows_" + q)
