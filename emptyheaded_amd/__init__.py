"""emptyheaded_amd -- MI355X-native generic-join engine (libehx) for EmptyHeaded's
storage-engine hot path.

This module is a thin ctypes binding of the C ABI in include/ehx.h and
include/ehx_ext.h.  The library must be built (python -m emptyheaded_amd.build);
importing the module without it raises, and nothing here falls back to a CPU
implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libehx.so")

OK = 0
RANGE_BITSET, UINTEGER = 0, 1
POLICY_RULE, POLICY_ALL_UINT, POLICY_ALL_BITSET, POLICY_LEARNED = 0, 1, 2, 3
FORM_PLAIN, FORM_TUNED, FORM_LEARNED = 0, 1, 2
MAX_ATTRS, MAX_OPS = 8, 8
ALGORITHMS = ("merge", "gallop", "probe", "and", "empty")


class EhxError(RuntimeError):
    pass


class Op(C.Structure):
    _fields_ = [("rel", C.c_int32), ("parent", C.c_int32), ("value", C.c_uint32)]


class Plan(C.Structure):
    _fields_ = [("num_attrs", C.c_int32), ("num_ops", C.c_int32 * MAX_ATTRS),
                ("ops", (Op * MAX_OPS) * MAX_ATTRS)]


class TrieMeta(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_heads", C.c_uint64), ("num_bitsets", C.c_uint64),
                ("bitset_words", C.c_uint64), ("bitset_elems", C.c_uint64), ("max_card", C.c_uint32),
                ("head_layout", C.c_uint32), ("dense_nav", C.c_uint32), ("max_value", C.c_uint32),
                ("build_ms", C.c_double), ("num_staged_bitmaps", C.c_uint64)]


class JoinStats(C.Structure):
    _fields_ = [("level_ms", C.c_double * MAX_ATTRS), ("frontier", C.c_uint64 * MAX_ATTRS),
                ("pairs_by_alg", C.c_uint64 * 5), ("total_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("form_used", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {"level_ms": list(self.level_ms), "frontier": list(self.frontier),
                "pairs_by_alg": dict(zip(ALGORITHMS, list(self.pairs_by_alg))),
                "total_ms": self.total_ms, "main_kernel_ms": self.main_kernel_ms, "form_used": self.form_used}


class TuneParams(C.Structure):
    _fields_ = [("lds_slots", C.c_int32), ("big_deg", C.c_int32), ("waves_per_cu", C.c_int32),
                ("grab", C.c_int32), ("bitmap_all", C.c_int32), ("use_graph", C.c_int32), ("big_work", C.c_int32),
                ("stream_lanes", C.c_int32), ("gap_coded", C.c_int32), ("split_work", C.c_int32),
                ("tiny_work", C.c_int32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# Every symbol the headers declare, with ctypes signatures.
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
SIGNATURES = {
    "ehx_version": ([], C.c_int),
    "ehx_last_error": ([], C.c_char_p),
    "ehx_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "ehx_ctx_create": ([C.c_int, C.POINTER(_vp)], C.c_int),
    "ehx_ctx_destroy": ([_vp], C.c_int),
    "ehx_ctx_synchronize": ([_vp], C.c_int),
    "ehx_trie_build": ([_vp, _vp, _vp, C.c_uint64, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
    "ehx_trie_destroy": ([_vp], C.c_int),
    "ehx_trie_info": ([_vp, C.POINTER(TrieMeta)], C.c_int),
    "ehx_trie_export": ([_vp, _vp, _u32p, _u32p], C.c_int),
    "ehx_trie_node_info": ([_vp, _vp, C.POINTER(C.c_uint8), _u32p], C.c_int),
    "ehx_trie_lookup": ([_vp, _vp, _u32p, C.c_uint64, C.POINTER(C.c_int64)], C.c_int),
    "ehx_trie_device_arrays": ([_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)], C.c_int),
    "ehx_intersect_count_batch": ([_vp, _vp, _u32p, _vp, _u32p, C.c_uint64, _u64p], C.c_int),
    "ehx_intersect_count_batch_noalloc": ([_vp, _vp, _u32p, _vp, _u32p, C.c_uint64, _u64p], C.c_int),
    "ehx_intersect_batch": ([_vp, _vp, _u32p, _vp, _u32p, C.c_uint64, _u64p, _u32p, C.c_uint64, _u64p],
                            C.c_int),
    "ehx_intersect_nodes_count": ([_vp, C.c_int, C.POINTER(_vp), C.POINTER(_u32p), C.c_uint64, _u64p], C.c_int),
    "ehx_intersect_nodes": ([_vp, C.c_int, C.POINTER(_vp), C.POINTER(_u32p), C.c_uint64, _u64p, _u32p, C.c_uint64,
                             _u64p], C.c_int),
    "ehx_trie_host_arrays": ([_vp, _vp, _u32p, _u32p], C.c_int),
    "ehx_join_count": ([_vp, C.POINTER(Plan), C.POINTER(_vp), C.c_int, C.POINTER(C.c_int64),
                        C.POINTER(JoinStats)], C.c_int),
    "ehx_join_count_range": ([_vp, C.POINTER(Plan), C.POINTER(_vp), C.c_int, C.c_uint64, C.c_uint64,
                              C.POINTER(C.c_int64), C.POINTER(JoinStats)], C.c_int),
    "ehx_join_count_segments": ([_vp, C.POINTER(Plan), C.POINTER(_vp), C.c_int, _u32p, C.c_uint64,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(JoinStats)], C.c_int),
    "ehx_join_materialize": ([_vp, C.POINTER(Plan), C.POINTER(_vp), C.c_int, C.POINTER(_u32p), C.c_uint64, _u64p,
                              C.POINTER(JoinStats)], C.c_int),
    "ehx_cycle4_count": ([_vp, _vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_int64), C.POINTER(JoinStats)], C.c_int),
    "ehx_weighted_pairs": ([_vp, _vp, C.POINTER(C.c_int64), C.c_uint64, C.POINTER(C.c_int64), C.c_uint64,
                            C.POINTER(C.c_int64)], C.c_int),
    "ehx_sssp_unweighted": ([_vp, _vp, C.c_uint32, C.c_int32, _u32p, C.c_uint64, _u32p], C.c_int),
    "ehx_sssp_unweighted_dir": ([_vp, _vp, _vp, C.c_uint32, C.c_int32, _u32p, C.c_uint64, _u32p, C.c_int32],
                                C.c_int),
    "ehx_refdb_open": ([C.c_char_p, C.c_int, C.POINTER(_vp)], C.c_int),
    "ehx_refdb_info": ([_vp, _u32p, _u64p, C.POINTER(C.c_int32)], C.c_int),
    "ehx_refdb_columns": ([_vp, C.POINTER(_u32p), _u64p], C.c_int),
    "ehx_refdb_close": ([_vp], C.c_int),
    "ehx_trie_orient": ([_vp, _vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), _u64p], C.c_int),
    "ehx_trie_link": ([_vp, _vp, _vp, C.POINTER(C.c_int)], C.c_int),
    "ehx_trie_link_hints": ([_vp, _vp, _vp, C.c_int32, C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_int)],
                            C.c_int),
    "ehx_trie_load_reference": ([_vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
    "ehx_encoded_open": ([C.c_char_p, C.POINTER(_vp)], C.c_int),
    "ehx_trie_build_encoded": ([_vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
    "ehx_pagerank": ([_vp, _vp, _u32p, C.POINTER(C.c_float), C.c_uint64, C.c_void_p, C.c_int,
                      C.POINTER(C.c_float), C.POINTER(C.c_double)], C.c_int),
    "ehx_cycle4_count_segments": ([_vp, _vp, _u32p, C.c_uint64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(JoinStats)], C.c_int),
    "ehx_gen_egonets": ([_u32p, _u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                         C.POINTER(_u32p), C.POINTER(_u32p), _u64p, _u32p, _u64p], C.c_int),
    "ehx_gen_er": ([C.c_uint64, C.c_uint64, C.c_uint64, _u32p, _u32p], C.c_int),
    "ehx_gen_zipf": ([C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, _u32p], C.c_int),
    "ehx_gen_rmat": ([C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64, _u32p, _u32p],
                     C.c_int),
    "ehx_graph_prepare": ([_u32p, _u32p, C.c_uint64, C.c_int, C.c_int, C.POINTER(_u32p), C.POINTER(_u32p),
                           _u64p, _u32p], C.c_int),
    "ehx_free": ([_vp], None),
    "ehx_tune_get": ([C.POINTER(TuneParams)], C.c_int),
    "ehx_tune_set": ([C.POINTER(TuneParams)], C.c_int),
    "ehx_selector_num_params": ([], C.c_int),
    "ehx_selector_load": ([C.c_int, C.POINTER(C.c_uint16), C.c_uint64, C.POINTER(C.c_float), C.c_float],
                          C.c_int),
    "ehx_selector_clear": ([], C.c_int),
    "ehx_pair_record": ([_vp, _vp, _vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint64, C.c_int,
                         C.POINTER(C.c_uint64)], C.c_int),
    "ehx_pair_table_load": ([C.c_int, C.POINTER(C.c_uint8)], C.c_int),
    "ehx_trie_set_features": ([_vp, _vp, C.POINTER(C.c_float)], C.c_int),
    "ehx_copy_calibrate": ([_vp, C.c_uint64, C.POINTER(C.c_double)], C.c_int),
    "ehx_host_alloc": ([C.c_uint64, C.POINTER(_vp)], C.c_int),
    "ehx_host_free": ([_vp], C.c_int),
    "ehx_join_record": ([_vp, C.POINTER(Plan), C.POINTER(_vp), C.c_int, _u64p, C.c_uint64, _u64p,
                         C.POINTER(C.c_int64)], C.c_int),
    "ehx_mlp_create": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint64,
                        C.POINTER(_vp)], C.c_int),
    "ehx_mlp_destroy": ([_vp], C.c_int),
    "ehx_mlp_grad": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_float)], C.c_int),
    "ehx_mlp_predict": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint64,
                         C.POINTER(C.c_float)], C.c_int),
}

_lib = None


def lib():
    """Load libehx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EhxError("libehx.so is not built: run `python -m emptyheaded_amd.build` (%s)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().ehx_last_error()
        raise EhxError("ehx error %d: %s" % (rc, msg.decode() if msg else ""))


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def ptr_u32(a):
    return a.ctypes.data_as(_u32p)


def device_count():
    n = C.c_int(0)
    check(lib().ehx_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU: a HIP stream and grow-only scratch (ehx_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().ehx_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            check(lib().ehx_ctx_destroy(self.h))
            self.h = None

    def synchronize(self):
        check(lib().ehx_ctx_synchronize(self.h))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Trie:
    """Device trie over a binary relation (src, dst) -- Trie<A,M> of the reference."""

    def __init__(self, ctx, src, dst, policy=POLICY_RULE):
        s, d = u32(src), u32(dst)
        if s.shape != d.shape:
            raise ValueError("columns differ in length")
        h = _vp()
        check(lib().ehx_trie_build(ctx.h, s.ctypes.data, d.ctypes.data, len(s), policy, 0, C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.policy = policy

    def link_to(self, target):
        """Builds (once) the child pointers from this trie's values into
        `target`'s child sets; True when the link exists."""
        ok = C.c_int(0)
        check(lib().ehx_trie_link(self.ctx.h, self.h, target.h, C.byref(ok)))
        return bool(ok.value)

    def link_hints(self, target, route="auto"):
        """(hints, state): the link's trim hints (uint32 per row: how many values
        of the row's child set in `target` lie below the row's first column)
        and how they were found (0 no link, 1 none kept, 2 searched, 3 from the
        transposition sort).  route ("auto", "search" or "sort") applies when
        the link is built by this call."""
        routes = {"auto": 0, "search": 1, "sort": 2}
        if route not in routes:
            raise ValueError("route must be 'auto', 'search' or 'sort', not %r" % (route,))
        n = int(self.num_rows)
        out = np.zeros(n, dtype=np.uint32)
        st = C.c_int(0)
        check(lib().ehx_trie_link_hints(self.ctx.h, self.h, target.h, routes[route],
                                        out.ctypes.data_as(C.POINTER(C.c_uint32)), n, C.byref(st)))
        return out, st.value

    def orient(self, policy=POLICY_RULE, build=True):
        """(lower, reverse, counts): the rows x < y as a device trie and its
        reverse ordering (both None unless `build` and the relation is
        symmetric without loops), and the numbers of rows with x < y, x > y,
        x == y, and of rows x < y whose reverse row is absent."""
        lo, rv = _vp(), _vp()
        cnt = np.zeros(4, dtype=np.uint64)
        check(lib().ehx_trie_orient(self.ctx.h, self.h, policy, C.byref(lo) if build else None,
                                    C.byref(rv) if build else None, cnt.ctypes.data_as(_u64p)))
        out = []
        for h in (lo, rv):
            t = None
            if h.value:
                t = Trie.__new__(Trie)
                t.h, t.ctx, t.policy = h, self.ctx, policy
            out.append(t)
        return out[0], out[1], [int(x) for x in cnt]

    @classmethod
    def from_encoded(cls, ctx, path, first=0, second=1, policy=POLICY_RULE):
        """The trie of ordering (first, second) of a relation's encoded column
        store (<path>/encoded.bin), built on the device."""
        h = _vp()
        check(lib().ehx_trie_build_encoded(ctx.h, path.encode(), first, second, policy, C.byref(h)))
        t = cls.__new__(cls)
        t.h, t.ctx, t.policy = h, ctx, policy
        return t

    @classmethod
    def load_reference(cls, ctx, path, anno_bytes=8, policy=POLICY_RULE):
        """A relation the reference engine saved to disk (ram/ or mmap/ image),
        as a device trie: Trie<A,M>::load."""
        h = _vp()
        check(lib().ehx_trie_load_reference(ctx.h, path.encode(), anno_bytes, policy, C.byref(h)))
        t = cls.__new__(cls)
        t.h, t.ctx, t.policy = h, ctx, policy
        return t

    def close(self):
        if getattr(self, "h", None):
            check(lib().ehx_trie_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        m = TrieMeta()
        check(lib().ehx_trie_info(self.h, C.byref(m)))
        return {f: getattr(m, f) for f, _ in TrieMeta._fields_}

    @property
    def num_rows(self):
        return self.info()["num_rows"]

    def tuples(self):
        n = self.num_rows
        a = np.zeros(n, dtype=np.uint32)
        b = np.zeros(n, dtype=np.uint32)
        check(lib().ehx_trie_export(self.ctx.h, self.h, ptr_u32(a), ptr_u32(b)))
        return a, b

    def node_info(self):
        nh = self.info()["num_heads"]
        t = np.zeros(nh, dtype=np.uint8)
        c = np.zeros(nh, dtype=np.uint32)
        check(lib().ehx_trie_node_info(self.ctx.h, self.h, t.ctypes.data_as(C.POINTER(C.c_uint8)), ptr_u32(c)))
        return t, c

    def lookup(self, keys):
        k = u32(keys)
        out = np.zeros(len(k), dtype=np.int64)
        check(lib().ehx_trie_lookup(self.ctx.h, self.h, ptr_u32(k), len(k),
                                    out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out


def intersect_count(ctx, ta, keys_a, tb, keys_b):
    ka, kb = u32(keys_a), u32(keys_b)
    out = np.zeros(len(ka), dtype=np.uint64)
    check(lib().ehx_intersect_count_batch(ctx.h, ta.h, ptr_u32(ka), tb.h, ptr_u32(kb), len(ka),
                                          out.ctypes.data_as(_u64p)))
    return out


def intersect_nodes_count(ctx, tries, nodes):
    """|set(tries[0], nodes[0][i]) ∩ ... | for every call i: node == num_heads
    names a trie's head set (TrieBuilder::build_aggregated_set over blocks)."""
    ns = [u32(x) for x in nodes]
    n = len(ns[0])
    out = np.zeros(n, dtype=np.uint64)
    tarr = (_vp * len(tries))(*[t.h.value for t in tries])
    narr = (_u32p * len(ns))(*[ptr_u32(x) for x in ns])
    check(lib().ehx_intersect_nodes_count(ctx.h, len(tries), tarr, narr, n, out.ctypes.data_as(_u64p)))
    return out


NOALLOC_MAX_PAIRS = 1 << 20


def intersect_count_noalloc(ctx, ta, keys_a, tb, keys_b):
    """intersect_count through the entry point that allocates nothing (at most
    NOALLOC_MAX_PAIRS pairs a call; calls on one GPU must not overlap)."""
    ka, kb = u32(keys_a), u32(keys_b)
    out = np.zeros(len(ka), dtype=np.uint64)
    check(lib().ehx_intersect_count_batch_noalloc(ctx.h, ta.h, ptr_u32(ka), tb.h, ptr_u32(kb), len(ka),
                                                  out.ctypes.data_as(_u64p)))
    return out


def intersect(ctx, ta, keys_a, tb, keys_b):
    """Materialized intersections: (offsets[n+1], values)."""
    ka, kb = u32(keys_a), u32(keys_b)
    n = len(ka)
    offs = np.zeros(n + 1, dtype=np.uint64)
    total = C.c_uint64(0)
    cap = int(intersect_count(ctx, ta, ka, tb, kb).sum()) if n else 0
    vals = np.zeros(max(cap, 1), dtype=np.uint32)
    check(lib().ehx_intersect_batch(ctx.h, ta.h, ptr_u32(ka), tb.h, ptr_u32(kb), n, offs.ctypes.data_as(_u64p),
                                    ptr_u32(vals), cap, C.byref(total)))
    return offs, vals[: total.value]


def make_plan(levels):
    """levels: one list per attribute of (rel, parent) operands; parent -1 = head
    set; (rel, -2, value) = the child set of the constant value (a selection)."""
    if not 1 <= len(levels) <= MAX_ATTRS:
        raise ValueError("1..%d attributes" % MAX_ATTRS)
    p = Plan()
    p.num_attrs = len(levels)
    for k, ops in enumerate(levels):
        p.num_ops[k] = len(ops)
        for i, op in enumerate(ops):
            p.ops[k][i].rel = op[0]
            p.ops[k][i].parent = op[1]
            p.ops[k][i].value = op[2] if len(op) > 2 else 0
    return p


def join_count(ctx, levels, tries, form=FORM_PLAIN, begin=None, end=None, stats=False):
    p = make_plan(levels) if not isinstance(levels, Plan) else levels
    arr = (_vp * len(tries))(*[t.h.value for t in tries])
    res = C.c_int64(0)
    st = JoinStats()
    if begin is None and end is None:
        check(lib().ehx_join_count(ctx.h, C.byref(p), arr, form, C.byref(res), C.byref(st)))
    else:
        check(lib().ehx_join_count_range(ctx.h, C.byref(p), arr, form, begin or 0,
                                         2**64 - 1 if end is None else end, C.byref(res), C.byref(st)))
    return (res.value, st.as_dict()) if stats else res.value


def join_count_segments(ctx, levels, tries, bounds, form=FORM_PLAIN, stats=False):
    """Per-segment COUNT(*): counts[s] for level-0 values in [bounds[s], bounds[s+1])."""
    p = make_plan(levels) if not isinstance(levels, Plan) else levels
    arr = (_vp * len(tries))(*[t.h.value for t in tries])
    b = np.ascontiguousarray(bounds, dtype=np.uint32)
    counts = np.zeros(max(len(b) - 1, 1), dtype=np.int64)
    res = C.c_int64(0)
    st = JoinStats()
    check(lib().ehx_join_count_segments(ctx.h, C.byref(p), arr, form, ptr_u32(b), len(b) - 1,
                                        counts.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(res), C.byref(st)))
    counts = counts[: len(b) - 1]
    return (counts, res.value, st.as_dict()) if stats else (counts, res.value)


def join_materialize(ctx, levels, tries, form=FORM_PLAIN, stats=False):
    """Every result tuple in lexicographic order: an (rows, num_attrs) uint32
    array (and the second pass's stats when `stats`)."""
    p = make_plan(levels) if not isinstance(levels, Plan) else levels
    arr = (_vp * len(tries))(*[t.h.value for t in tries])
    rows = C.c_uint64(0)
    cols = (_u32p * p.num_attrs)()
    rc = lib().ehx_join_materialize(ctx.h, C.byref(p), arr, form, cols, 0, C.byref(rows), None)
    if rc not in (0, -3):
        check(rc)
    n = rows.value
    out = np.zeros((p.num_attrs, max(n, 1)), dtype=np.uint32)
    for j in range(p.num_attrs):
        cols[j] = ptr_u32(out[j])
    st = JoinStats()
    check(lib().ehx_join_materialize(ctx.h, C.byref(p), arr, form, cols, n, C.byref(rows), C.byref(st)))
    res = out[:, :n].T.copy()
    return (res, st.as_dict()) if stats else res


def cycle4_count(ctx, sym, begin=None, end=None, bounds=None, stats=False):
    """4-cycles of the undirected graph held in both directions by trie `sym`;
    with bounds, also the per-segment counts (by each cycle's largest vertex)."""
    res = C.c_int64(0)
    st = JoinStats()
    if bounds is None:
        check(lib().ehx_cycle4_count(ctx.h, sym.h, begin or 0, 2**64 - 1 if end is None else end, C.byref(res),
                                     C.byref(st)))
        return (res.value, st.as_dict()) if stats else res.value
    b = np.ascontiguousarray(bounds, dtype=np.uint32)
    counts = np.zeros(max(len(b) - 1, 1), dtype=np.int64)
    check(lib().ehx_cycle4_count_segments(ctx.h, sym.h, ptr_u32(b), len(b) - 1,
                                          counts.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(res), C.byref(st)))
    counts = counts[: len(b) - 1]
    return (counts, res.value, st.as_dict()) if stats else (counts, res.value)


def weighted_pairs(ctx, trie, wa, wb):
    """Sum over the tuples (a, x) of trie of wa[a] * wb[x] (int64 arrays indexed
    by value; values past an array's end weigh 0), modulo 2^64."""
    a = np.ascontiguousarray(wa, dtype=np.int64)
    b = np.ascontiguousarray(wb, dtype=np.int64)
    res = C.c_int64(0)
    check(lib().ehx_weighted_pairs(ctx.h, trie.h, a.ctypes.data_as(C.POINTER(C.c_int64)), len(a),
                                   b.ctypes.data_as(C.POINTER(C.c_int64)), len(b), C.byref(res)))
    return res.value


UNREACHED = 0xFFFFFFFF


def sssp_unweighted(ctx, trie, start, n, iterations=-1, rev=None, bottom_up=0):
    """BFS levels from `start` over the trie's child sets, for values 0..n-1
    (UNREACHED where never reached), and the number of levels run.  With `rev`
    (the reversed relation's trie; the trie itself when symmetric), levels whose
    frontier holds at least n / bottom_up vertices are found bottom-up."""
    dist = np.empty(n, dtype=np.uint32)
    lv = C.c_uint32(0)
    if rev is None:
        check(lib().ehx_sssp_unweighted(ctx.h, trie.h, start, iterations, ptr_u32(dist), n, C.byref(lv)))
    else:
        check(lib().ehx_sssp_unweighted_dir(ctx.h, trie.h, rev.h, start, iterations, ptr_u32(dist), n, C.byref(lv),
                                            bottom_up))
    return dist, lv.value


def refdb_read(path, anno_bytes=8, encoded=False):
    """Rows (columns of uint32, lexicographic) and annotations (raw 64-bit
    values, or None) of a relation the reference engine saved to disk; with
    encoded=True, the columns of its encoded column store (file order)."""
    h = _vp()
    if encoded:
        check(lib().ehx_encoded_open(path.encode(), C.byref(h)))
    else:
        check(lib().ehx_refdb_open(path.encode(), anno_bytes, C.byref(h)))
    try:
        nc, rows, ann = C.c_uint32(0), C.c_uint64(0), C.c_int32(0)
        check(lib().ehx_refdb_info(h, C.byref(nc), C.byref(rows), C.byref(ann)))
        cols = [np.zeros(max(1, rows.value), dtype=np.uint32) for _ in range(nc.value)]
        ptrs = (_u32p * max(1, nc.value))(*[ptr_u32(c) for c in cols])
        annos = np.zeros(max(1, rows.value), dtype=np.uint64) if ann.value else None
        check(lib().ehx_refdb_columns(h, ptrs, annos.ctypes.data_as(_u64p) if annos is not None else None))
        return [c[:rows.value] for c in cols], (annos[:rows.value] if annos is not None else None)
    finally:
        lib().ehx_refdb_close(h)


class RankRule(C.Structure):
    """ehx_rank_rule: R(x;y) <- base; R(x;y)*[i=iterations] <- add + mul*SUM(z;init)."""
    _fields_ = [("base", C.c_double), ("add", C.c_double), ("mul", C.c_double), ("init", C.c_double),
                ("iterations", C.c_int32), ("pad", C.c_int32)]


def pagerank(ctx, trie, w_keys, w_vals, base, add=0.15, mul=0.85, init=1.0, iterations=5, form=FORM_TUNED):
    """The reference's recursive PageRank rule over the trie's child sets with
    the unary float relation W (sorted distinct keys): per-head float32 ranks
    (head order) and the device time of the iterations in ms."""
    k = np.ascontiguousarray(w_keys, dtype=np.uint32)
    v = np.ascontiguousarray(w_vals, dtype=np.float32)
    assert len(k) == len(v)
    nh = trie.info()["num_heads"]
    out = np.empty(max(1, nh), dtype=np.float32)
    rule = RankRule(float(base), float(add), float(mul), float(init), int(iterations), 0)
    ms = C.c_double(0.0)
    check(lib().ehx_pagerank(ctx.h, trie.h, ptr_u32(k), v.ctypes.data_as(C.POINTER(C.c_float)), len(k),
                             C.cast(C.pointer(rule), C.c_void_p), form, out.ctypes.data_as(C.POINTER(C.c_float)),
                             C.byref(ms)))
    return out[:nh], ms.value


def tune_get():
    t = TuneParams()
    check(lib().ehx_tune_get(C.byref(t)))
    return t.as_dict()


def tune_set(**kw):
    t = TuneParams()
    check(lib().ehx_tune_get(C.byref(t)))
    for k, v in kw.items():
        setattr(t, k, int(v))
    check(lib().ehx_tune_set(C.byref(t)))
    return t.as_dict()


from . import graphs, plans  # noqa: E402,F401
