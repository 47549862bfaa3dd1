"""ctypes binding of the CPU oracle (oracle/libehx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libehx_oracle.so")

RANGE_BITSET, UINTEGER = 0, 1
POLICY_RULE, POLICY_ALL_UINT, POLICY_ALL_BITSET = 0, 1, 2
MAX_ATTRS, MAX_OPS = 8, 8


class OrcSet(C.Structure):
    _fields_ = [("cardinality", C.c_uint32), ("range", C.c_uint32),
                ("number_of_bytes", C.c_uint64), ("type", C.c_uint8),
                ("data", C.c_void_p)]


class OrcOp(C.Structure):
    _fields_ = [("rel", C.c_int32), ("parent", C.c_int32), ("value", C.c_uint32)]


class OrcPlan(C.Structure):
    _fields_ = [("num_attrs", C.c_int32), ("num_ops", C.c_int32 * MAX_ATTRS),
                ("ops", (OrcOp * MAX_OPS) * MAX_ATTRS)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None
_lib_info = {"path": LIB_PATH, "march": "x86-64-v3"}


def use_native(outdir):
    """Compiles the oracle for the host CPU (gcc -O3 -march=native, otherwise
    the Makefile's flags) into outdir and makes lib() return that build, for
    the CPU-baseline timing.  Call before anything else loads the oracle.
    Returns the build description, or None when the compile fails (lib() then
    stays the committed-flags build)."""
    global _lib, _lib_info
    if _lib is not None:
        raise RuntimeError("the oracle is already loaded")
    path = os.path.join(outdir, "libehx_oracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fPIC", "-std=gnu11", "-pthread", "-shared",
           "-o", path, os.path.join(HERE, "ehx_oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None
    _lib = _configure(C.CDLL(path))
    _lib_info = {"path": path, "march": "native", "command": " ".join(cmd[:8])}
    return dict(_lib_info)


def lib_info():
    return dict(_lib_info)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def _configure(L):
    u32p, u64 = C.POINTER(C.c_uint32), C.c_uint64
    L.orc_is_sparse.argtypes = [u64, u64]
    L.orc_is_sparse.restype = C.c_int
    L.orc_layout_of.argtypes = [u32p, u64]
    L.orc_layout_of.restype = C.c_int
    L.orc_bytes_needed.argtypes = [u64, u32p, C.c_int]
    L.orc_bytes_needed.restype = u64
    L.orc_set_build.argtypes = [C.c_void_p, u32p, u64, C.c_int, C.POINTER(OrcSet)]
    L.orc_set_build.restype = u64
    L.orc_set_find.argtypes = [C.POINTER(OrcSet), C.c_uint32]
    L.orc_set_find.restype = C.c_int64
    L.orc_set_decode.argtypes = [C.POINTER(OrcSet), u32p]
    L.orc_set_decode.restype = u64
    L.orc_intersect.argtypes = [C.POINTER(OrcSet), C.c_void_p, C.POINTER(OrcSet), C.POINTER(OrcSet)]
    L.orc_intersect.restype = u64
    L.orc_intersect_count.argtypes = [C.POINTER(OrcSet), C.POINTER(OrcSet)]
    L.orc_intersect_count.restype = u64
    L.orc_algorithm_of.argtypes = [C.POINTER(OrcSet), C.POINTER(OrcSet)]
    L.orc_algorithm_of.restype = C.c_int
    L.orc_uu_merge.argtypes = [u32p, u32p, u64, u32p, u64]
    L.orc_uu_merge.restype = u64
    L.orc_uu_gallop.argtypes = [u32p, u32p, u64, u32p, u64]
    L.orc_uu_gallop.restype = u64
    L.orc_trie_build.argtypes = [u32p, u32p, u64, C.c_int]
    L.orc_trie_build.restype = C.c_void_p
    L.orc_trie_free.argtypes = [C.c_void_p]
    L.orc_trie_num_rows.argtypes = [C.c_void_p]
    L.orc_trie_num_rows.restype = u64
    L.orc_trie_head.argtypes = [C.c_void_p, C.POINTER(OrcSet)]
    L.orc_trie_child.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(OrcSet)]
    L.orc_trie_child.restype = C.c_int
    L.orc_trie_tuples.argtypes = [C.c_void_p, u32p, u32p]
    L.orc_trie_tuples.restype = u64
    L.orc_trie_node_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), u32p]
    L.orc_join_count.argtypes = [C.POINTER(OrcPlan), C.POINTER(C.c_void_p), C.c_int, u64, u64]
    L.orc_join_count.restype = C.c_int64
    L.orc_join_count_sampled.argtypes = [C.POINTER(OrcPlan), C.POINTER(C.c_void_p), C.c_int, u64]
    L.orc_join_count_sampled.restype = C.c_int64
    L.orc_join_level0_size.argtypes = [C.POINTER(OrcPlan), C.POINTER(C.c_void_p)]
    L.orc_join_level0_size.restype = u64
    L.orc_join_count_per_head.argtypes = [C.POINTER(OrcPlan), C.POINTER(C.c_void_p), C.c_int,
                                          C.POINTER(C.c_int64)]
    L.orc_join_count_per_head.restype = C.c_int
    L.orc_join_materialize.argtypes = [C.POINTER(OrcPlan), C.POINTER(C.c_void_p), u32p, u64]
    L.orc_join_materialize.restype = u64
    L.orc_cycle4_count.argtypes = [C.POINTER(C.c_uint64), u32p, C.c_uint32, C.c_int, C.POINTER(C.c_int64)]
    L.orc_cycle4_count.restype = C.c_int64
    L.orc_pagerank.argtypes = [C.c_void_p, u32p, C.POINTER(C.c_float), u64, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.orc_pagerank.restype = C.c_int
    L.orc_write_refdb.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
    L.orc_write_refdb.restype = C.c_int
    L.orc_write_refdb_annotated.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                            C.c_int]
    L.orc_write_refdb_annotated.restype = C.c_int
    L.orc_set_pin_threads.argtypes = [C.c_int]
    L.orc_set_pin_threads.restype = None
    return L


def set_pin_threads(on):
    lib().orc_set_pin_threads(1 if on else 0)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


class Set:
    """A set in the reference's byte layout (uinteger or range_bitset)."""

    def __init__(self, values, layout=None):
        v, p = _u32(np.asarray(values, dtype=np.uint32))
        L = lib()
        if layout is None:
            layout = L.orc_layout_of(p, len(v))
        self.buf = np.zeros(int(L.orc_bytes_needed(len(v), p, layout)) + 16, dtype=np.uint8)
        self.s = OrcSet()
        L.orc_set_build(self.buf.ctypes.data, p, len(v), layout, C.byref(self.s))
        self.values = v

    @property
    def layout(self):
        return self.s.type

    def find(self, key):
        return lib().orc_set_find(C.byref(self.s), key)

    def decode(self):
        out = np.zeros(max(1, self.s.cardinality), dtype=np.uint32)
        n = lib().orc_set_decode(C.byref(self.s), out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out[:n]

    def raw(self):
        return bytes(self.buf[: self.s.number_of_bytes])


def intersect(a, b):
    """Materialized intersection; returns (values, result_layout, cardinality, range)."""
    L = lib()
    cap = max(a.s.number_of_bytes, b.s.number_of_bytes) + 64
    cbuf = np.zeros(int(cap), dtype=np.uint8)
    c = OrcSet()
    L.orc_intersect(C.byref(c), cbuf.ctypes.data, C.byref(a.s), C.byref(b.s))
    out = np.zeros(max(1, c.cardinality), dtype=np.uint32)
    n = L.orc_set_decode(C.byref(c), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out[:n], c.type, c.cardinality, c.range


def intersect_count(a, b):
    return lib().orc_intersect_count(C.byref(a.s), C.byref(b.s))


def algorithm_of(a, b):
    return lib().orc_algorithm_of(C.byref(a.s), C.byref(b.s))


class Trie:
    """Two-level trie over a binary relation (src, dst)."""

    def __init__(self, src, dst, policy=POLICY_RULE):
        s, ps = _u32(src)
        d, pd = _u32(dst)
        assert len(s) == len(d)
        self.h = C.c_void_p(lib().orc_trie_build(ps, pd, len(s), policy))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_trie_free(self.h)
            self.h = None

    @property
    def num_rows(self):
        return lib().orc_trie_num_rows(self.h)

    def tuples(self):
        n = self.num_rows
        a = np.zeros(max(1, n), dtype=np.uint32)
        b = np.zeros(max(1, n), dtype=np.uint32)
        k = lib().orc_trie_tuples(self.h, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  b.ctypes.data_as(C.POINTER(C.c_uint32)))
        return a[:k], b[:k]

    def head(self):
        s = OrcSet()
        lib().orc_trie_head(self.h, C.byref(s))
        return s

    def child(self, key):
        s = OrcSet()
        ok = lib().orc_trie_child(self.h, key, C.byref(s))
        return s if ok else None

    def node_info(self, nheads):
        t = np.zeros(max(1, nheads), dtype=np.uint8)
        c = np.zeros(max(1, nheads), dtype=np.uint32)
        lib().orc_trie_node_info(self.h, t.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 c.ctypes.data_as(C.POINTER(C.c_uint32)))
        return t[:nheads], c[:nheads]


def make_plan(levels):
    """levels: list (one per attribute) of lists of (rel, parent) pairs, or
    (rel, -2, value) triples for the child set of a constant (a selection)."""
    p = OrcPlan()
    p.num_attrs = len(levels)
    for k, ops in enumerate(levels):
        p.num_ops[k] = len(ops)
        for i, op in enumerate(ops):
            p.ops[k][i].rel = op[0]
            p.ops[k][i].parent = op[1]
            p.ops[k][i].value = op[2] if len(op) > 2 else 0
    return p


def _tries_arg(tries):
    arr = (C.c_void_p * len(tries))(*[t.h.value for t in tries])
    return arr


def join_count(levels, tries, nthreads=1, begin=0, end=2**64 - 1):
    p = make_plan(levels)
    return lib().orc_join_count(C.byref(p), _tries_arg(tries), nthreads, begin, end)


def join_count_sampled(levels, tries, nthreads, stride):
    p = make_plan(levels)
    return lib().orc_join_count_sampled(C.byref(p), _tries_arg(tries), nthreads, stride)


def join_level0_size(levels, tries):
    p = make_plan(levels)
    return lib().orc_join_level0_size(C.byref(p), _tries_arg(tries))


def join_count_per_head(levels, tries, nthreads=1):
    p = make_plan(levels)
    n = join_level0_size(levels, tries)
    out = np.zeros(max(1, n), dtype=np.int64)
    rc = lib().orc_join_count_per_head(C.byref(p), _tries_arg(tries), nthreads,
                                       out.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    return out[:n]


def cycle4_count(src, dst, num_vertices, nthreads=1, per_vertex=False):
    """4-cycles of the undirected graph whose edges (src, dst) hold both
    directions, rows sorted.  With per_vertex, also the cycles charged to each
    vertex (the cycle's largest vertex)."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    off = np.zeros(num_vertices + 1, dtype=np.uint64)
    np.cumsum(np.bincount(src, minlength=num_vertices), out=off[1:])
    pv = np.zeros(max(1, num_vertices), dtype=np.int64)
    tot = lib().orc_cycle4_count(off.ctypes.data_as(C.POINTER(C.c_uint64)), dst.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 num_vertices, nthreads,
                                 pv.ctypes.data_as(C.POINTER(C.c_int64)) if per_vertex else None)
    return (tot, pv[:num_vertices]) if per_vertex else tot


def join_materialize(levels, tries):
    """Every result tuple in lexicographic order: (rows, num_attrs) uint32."""
    p = make_plan(levels)
    n = lib().orc_join_materialize(C.byref(p), _tries_arg(tries), None, 0)
    out = np.zeros((len(levels), max(n, 1)), dtype=np.uint32)
    lib().orc_join_materialize(C.byref(p), _tries_arg(tries), out.ctypes.data_as(C.POINTER(C.c_uint32)), max(n, 1))
    return out[:, :n].T.copy()


def sssp_unweighted(src, dst, start, n, iterations=-1):
    """Unweighted single-source closure, level by level as
    tc::unweighted_single_source does (storage_engine/include/
    TransitiveClosure.hpp:12-123): the start is visited at level 0; each level
    visits the unvisited children of the frontier.  Returns (dist, levels) with
    dist[v] = level of first visit (0xFFFFFFFF never), for v < n."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src[src < n] + 1, 1)
    off = np.cumsum(off)
    dist = np.full(n, 0xFFFFFFFF, dtype=np.uint32)
    dist[start] = 0
    frontier = np.array([start], dtype=np.int64)
    level = 0
    while len(frontier) and (iterations < 0 or level < iterations):
        lo, hi = off[frontier], off[frontier + 1]
        cnt = hi - lo
        idx = np.repeat(lo - np.cumsum(cnt) + cnt, cnt) + np.arange(int(cnt.sum()))  # the frontier's child rows
        kids = np.unique(dst[idx]) if len(idx) else np.zeros(0, np.int64)
        kids = kids[(kids < n)]
        kids = kids[dist[kids] == 0xFFFFFFFF]
        level += 1
        dist[kids] = level
        frontier = kids
    return dist, level


def pagerank(src, dst, w_keys, w_vals, base, add=0.15, mul=0.85, init=1.0, iterations=5, nthreads=1,
             policy=POLICY_RULE, trie=None, fused_finish=True):
    """The reference's recursive PageRank rule over Edge(src,dst) with the
    unary float relation W, folded per head in ascending order with float32
    accumulation as the generated code does (orc_pagerank): float32 ranks per
    head value of Edge, in head order."""
    T = trie if trie is not None else Trie(src, dst, policy)
    nh = int(T.head().cardinality)
    k = np.ascontiguousarray(w_keys, dtype=np.uint32)
    v = np.ascontiguousarray(w_vals, dtype=np.float32)
    out = np.zeros(max(1, nh), dtype=np.float32)
    rc = lib().orc_pagerank(T.h, k.ctypes.data_as(C.POINTER(C.c_uint32)), v.ctypes.data_as(C.POINTER(C.c_float)),
                            len(k), C.c_double(base), C.c_double(add), C.c_double(mul), C.c_double(init),
                            int(iterations), int(nthreads), 1 if fused_finish else 0,
                            out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0
    return out[:nh]


def write_refdb(src, dst, path, nbuffers=4, mmapped=False, policy=POLICY_RULE, annos=None, unary=False):
    """Saves the relation (src, dst) as the reference engine does
    (orc_write_refdb): <path>/ram/ or <path>/mmap/ with trieinfo.bin and the
    buffer images.  unary: the relation of src alone (dst ignored).  annos:
    one value per distinct row in lexicographic row order (its dtype's width is
    sizeof(A))."""
    if unary:
        dst = src
    T = Trie(src, dst, policy)
    if annos is None and not unary:
        rc = lib().orc_write_refdb(T.h, path.encode(), int(nbuffers), 1 if mmapped else 0)
    else:
        a = None if annos is None else np.ascontiguousarray(annos)
        rc = lib().orc_write_refdb_annotated(T.h, path.encode(), int(nbuffers), 1 if mmapped else 0,
                                             1 if unary else 2, None if a is None else a.ctypes.data,
                                             8 if a is None else a.dtype.itemsize)
    assert rc == 0
    return T
