"""Synthetic inputs (host): Erdos-Renyi and R-MAT edge samples and the
undirected clean-up each workload applies.  Generation is counter-based, so a
seed names the same graph on every machine."""
import ctypes as C

import numpy as np

from . import check, lib, ptr_u32

RMAT_ABC = (0.57, 0.19, 0.19)
# relabel modes of ehx_graph_prepare
KEEP_IDS, DEGREE_ASC, COMPACT = 0, 1, 2
# orientation modes
BOTH, SRC_LT_DST, SRC_GT_DST = 0, 1, 2

SEED_BASE = 0xE4E40001


def er_samples(num_vertices, samples, seed):
    s = np.zeros(samples, dtype=np.uint32)
    d = np.zeros(samples, dtype=np.uint32)
    check(lib().ehx_gen_er(num_vertices, samples, seed, ptr_u32(s), ptr_u32(d)))
    return s, d


def rmat_samples(scale, samples, seed, abc=RMAT_ABC):
    s = np.zeros(samples, dtype=np.uint32)
    d = np.zeros(samples, dtype=np.uint32)
    a, b, c = abc
    check(lib().ehx_gen_rmat(scale, samples, a, b, c, seed, ptr_u32(s), ptr_u32(d)))
    return s, d


def prepare(src, dst, relabel=DEGREE_ASC, orient=SRC_LT_DST):
    """Drop loops, symmetrize, de-duplicate, relabel, orient.  Returns
    (src, dst, num_vertices) with rows sorted."""
    s, d = np.ascontiguousarray(src, np.uint32), np.ascontiguousarray(dst, np.uint32)
    os_, od_ = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    m, nv = C.c_uint64(0), C.c_uint32(0)
    check(lib().ehx_graph_prepare(ptr_u32(s), ptr_u32(d), len(s), relabel, orient, C.byref(os_), C.byref(od_),
                                  C.byref(m), C.byref(nv)))
    try:
        n = m.value
        a = np.ctypeslib.as_array(os_, shape=(max(n, 1),))[:n].copy()
        b = np.ctypeslib.as_array(od_, shape=(max(n, 1),))[:n].copy()
    finally:
        lib().ehx_free(C.cast(os_, C.c_void_p))
        lib().ehx_free(C.cast(od_, C.c_void_p))
    return a, b, nv.value


def er_graph(num_vertices=65536, samples=1 << 20, seed=SEED_BASE + 0):
    """Erdos-Renyi workload: uniform pairs, loops dropped, de-duplicated,
    oriented src < dst, ids kept."""
    s, d = er_samples(num_vertices, samples, seed)
    return prepare(s, d, KEEP_IDS, SRC_LT_DST)


def rmat_graph(scale, edge_factor=16, seed=None, relabel=DEGREE_ASC, orient=SRC_LT_DST):
    """R-MAT workload: 2^scale vertices, edge_factor * 2^scale samples,
    symmetrized, relabelled by ascending degree (isolated vertices dropped),
    oriented src < dst (low degree -> high degree)."""
    if seed is None:
        seed = SEED_BASE + scale
    s, d = rmat_samples(scale, edge_factor << scale, seed)
    return prepare(s, d, relabel, orient)


def degrees_undirected(src, dst, num_vertices):
    """Degrees of the undirected graph given one orientation of each edge."""
    return np.bincount(src, minlength=num_vertices) + np.bincount(dst, minlength=num_vertices)


def algorithmic_bytes(src, dst, num_vertices):
    """Operand bytes of the triangle query's last level: every oriented edge
    (a, b) streams both child sets, 4 bytes per element:
    4 * sum_v dout(v) * (dout(v) + din(v))."""
    dout = np.bincount(src, minlength=num_vertices).astype(np.int64)
    din = np.bincount(dst, minlength=num_vertices).astype(np.int64)
    return int(4 * np.sum(dout * (dout + din)))


EGO_SEED = SEED_BASE + 0x40  # ego-network batch drawn from R-MAT scale 22


def egonets(src_sym, dst_sym, num_vertices, num_nets=4096, min_edges=64, max_edges=65536, seed=EGO_SEED):
    """Batch of ego-networks cut from an undirected graph held in both
    directions (see ehx_gen_egonets).  Returns (src, dst, net_base, net_off):
    oriented src < dst edges of all nets, net q owning vertex ids
    [net_base[q], net_base[q+1]) and edges [net_off[q], net_off[q+1])."""
    s, d = np.ascontiguousarray(src_sym, np.uint32), np.ascontiguousarray(dst_sym, np.uint32)
    os_, od_ = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    m = C.c_uint64(0)
    base = np.zeros(num_nets + 1, dtype=np.uint32)
    off = np.zeros(num_nets + 1, dtype=np.uint64)
    check(lib().ehx_gen_egonets(ptr_u32(s), ptr_u32(d), len(s), num_vertices, num_nets, min_edges, max_edges, seed,
                                C.byref(os_), C.byref(od_), C.byref(m), ptr_u32(base),
                                off.ctypes.data_as(C.POINTER(C.c_uint64))))
    try:
        n = m.value
        a = np.ctypeslib.as_array(os_, shape=(max(n, 1),))[:n].copy()
        b = np.ctypeslib.as_array(od_, shape=(max(n, 1),))[:n].copy()
    finally:
        lib().ehx_free(C.cast(os_, C.c_void_p))
        lib().ehx_free(C.cast(od_, C.c_void_p))
    return a, b, base, off


def symmetric(src, dst):
    """Both directions of an oriented edge list, rows sorted (ids kept)."""
    s = np.concatenate([src, dst]).astype(np.uint32)
    d = np.concatenate([dst, src]).astype(np.uint32)
    o = np.lexsort((d, s))
    return s[o], d[o]
