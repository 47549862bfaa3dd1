"""Benchmark: triangle COUNT on R-MAT scale 22 (edge factor 16) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 22] [--no-cpu]

A step is one triangle-count query over the device-resident tries (the
reference's QUERY TIME: trie build and loading excluded; the reference's loader
builds both orderings of every relation, Edge_0_1 and Edge_1_0).  `value` is
the tuned+learned form running the triangle with its middle vertex first; the
tuned+learned form in the reference's attribute order, the tuned form on
density-rule layouts and the plain GPU drop-in (the generated code's order) are
timed beside it; the CPU oracle (a port of the reference engine) is timed on a
bounded sample on the host cores.  Also measured: 4-clique on R-MAT-18, the
ego-network batch and the triple-table joins.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "R-MAT-22 triangle count edges/s + % HBM peak, tuned+learned vs plain, 1 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_tuning(E, selector_path=None):
    """Autotuned knobs and learned selector weights committed with the package."""
    from emptyheaded_amd import tuning
    return tuning.apply_committed(selector_path=selector_path)


def time_form(E, ctx, T, plan, form, steps, warmup, tries=None):
    tries = tries or [T]
    for _ in range(warmup):
        E.join_count(ctx, plan, tries, form)
    ctx.synchronize()
    kernel_ms, total = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        total, st = E.join_count(ctx, plan, tries, form, stats=True)
        kernel_ms.append(st["main_kernel_ms"])
    ctx.synchronize()
    t1 = time.perf_counter()
    return total, (t1 - t0) * 1e3 / steps, sum(kernel_ms) / len(kernel_ms), st


def bench_small_graphs(E, ctx, steps, cpu=True):
    """Erdos-Renyi (65,536 vertices, 2^20 samples; uint layouts) and R-MAT-20
    triangle counts: the plain GPU form and the tuned+learned form next to the
    CPU oracle (16 threads), all on the same graph."""
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans
    from emptyheaded_amd import tuning as TU
    out = {}
    for name, (s, d, n), pol in (("er_65536", G.er_graph(), E.POLICY_ALL_UINT), ("rmat20", G.rmat_graph(20), None)):
        T = E.Trie(ctx, s, d, pol if pol is not None else E.POLICY_RULE)
        TL = E.Trie(ctx, s, d, pol if pol is not None else
                    (E.POLICY_LEARNED if _selector_loaded(E) else E.POLICY_RULE))
        TU.apply_plan("triangle")
        pc, pms, _, _ = time_form(E, ctx, T, plans.TRIANGLE, E.FORM_PLAIN, steps, 1)
        lc, lms, _, _ = time_form(E, ctx, TL, plans.TRIANGLE, E.FORM_LEARNED, steps, 1)
        assert pc == lc
        r = {"oriented_edges": len(s), "triangles": pc, "plain_ms": pms, "learned_ms": lms,
             "layouts": "all uint" if pol is not None else "density rule (plain), learned (tuned+learned)"}
        if cpu:
            from oracle import oracle as O
            OT = O.Trie(s, d)
            t = time.perf_counter()
            oc = O.join_count(plans.TRIANGLE, [OT], min(16, os.cpu_count() or 1))
            r["cpu_oracle_ms"] = (time.perf_counter() - t) * 1e3
            r["cpu_threads"] = min(16, os.cpu_count() or 1)
            assert oc == pc
        out[name] = r
        log("%s: %d triangles; plain %.3f ms, tuned+learned %.3f ms%s" %
            (name, pc, pms, lms, (", CPU oracle %.1f ms" % r["cpu_oracle_ms"]) if cpu else ""))
        T.close()
        TL.close()
    return out


def bench_selections(E, ctx, steps, cpu=True):
    """Equality selections on the symmetric facebook graph the reference's
    tests use (FliqueSelAgg x=0, BarbellSelAgg p=6; expected counts from the
    committed fixtures), every GPU form next to the CPU oracle."""
    import numpy as np
    from emptyheaded_amd import plans
    z = np.load(os.path.join(ROOT, "tests", "golden", "facebook_duplicated.npz"))
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        meta = json.load(f)["facebook_duplicated"]
    s, d = z["src"], z["dst"]
    T, R = E.Trie(ctx, s, d), E.Trie(ctx, d, s)
    out = {"workload": "facebook_duplicated (176,468 rows, both directions): FliqueSelAgg x=0, BarbellSelAgg p=6"}
    for name, levels, key in (("flique_sel_x0", plans.flique_sel(0), "FliqueSelAgg_x0"),
                              ("barbell_sel_p6", plans.barbell_sel(6), "BarbellSelAgg_p6")):
        r = {"expected": meta[key]}
        for fname, form in (("plain", E.FORM_PLAIN), ("tuned", E.FORM_TUNED), ("learned", E.FORM_LEARNED)):
            c, ms, _, _ = time_form(E, ctx, T, levels, form, steps, 1, tries=[T, R])
            assert c == meta[key], "%s %s: %d" % (name, fname, c)
            r[fname + "_ms"] = ms
        if cpu:
            from oracle import oracle as O
            t = time.perf_counter()
            assert O.join_count(levels, [O.Trie(s, d), O.Trie(d, s)], min(16, os.cpu_count() or 1)) == meta[key]
            r["cpu_oracle_ms_including_trie_build"] = (time.perf_counter() - t) * 1e3
        out[name] = r
        log("%s: %d; plain %.3f ms, tuned %.3f ms, learned %.3f ms" % (name, meta[key], r["plain_ms"], r["tuned_ms"],
                                                                     r["learned_ms"]))
    T.close()
    R.close()
    return out


def bench_two_bag(E, ctx, steps, cpu=True):
    """LollipopAgg and BarbellAgg (triangle bag per vertex, then the weighted
    edge fold) on the symmetric facebook graph, checked against the committed
    reference constants, and on R-MAT-20 with both edge directions, checked
    against the CPU oracle's per-vertex triangle counts folded in numpy."""
    import numpy as np
    from emptyheaded_amd import ghd, plans
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import tuning as TU
    z = np.load(os.path.join(ROOT, "tests", "golden", "facebook_duplicated.npz"))
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        meta = json.load(f)["facebook_duplicated"]
    rs, rd, _ = G.rmat_graph(20)
    rs, rd = G.symmetric(rs, rd)
    out = {}
    TU.apply_plan("triangle")
    for name, (s, d) in (("facebook_duplicated", (z["src"], z["dst"])), ("rmat20_symmetric", (rs, rd))):
        S = E.Trie(ctx, s, d)
        r = {"rows": len(s)}
        # "learned": the triangle bag from three oriented passes (orienting the relation on the
        # device is timed with it); "learned_direct" / "plain": the bag as written
        for fname, form, method in (("plain", E.FORM_PLAIN, "direct"), ("learned_direct", E.FORM_LEARNED, "direct"),
                                    ("learned", E.FORM_LEARNED, "auto")):
            for q, fn in (("lollipop", ghd.lollipop_count), ("barbell", ghd.barbell_count)):
                fn(ctx, S, form, method)
                ctx.synchronize()
                t = time.perf_counter()
                for _ in range(steps):
                    v = fn(ctx, S, form, method)
                ctx.synchronize()
                assert q not in r or r[q] == v
                r[q] = v
                r["%s_%s_ms" % (q, fname)] = (time.perf_counter() - t) * 1e3 / steps
        if name == "facebook_duplicated":
            assert r["lollipop"] == meta["LollipopAgg"] and r["barbell"] == meta["BarbellAgg"]
        if cpu:
            from oracle import oracle as O
            t = time.perf_counter()
            OT = O.Trie(s, d)
            nv = int(max(s.max(), d.max())) + 1
            ph = np.zeros(nv, np.int64)
            ph[np.unique(s)] = O.join_count_per_head(plans.TRIANGLE, [OT], min(16, os.cpu_count() or 1))
            s64, d64 = s.astype(np.int64), d.astype(np.int64)
            lol, bar = int(ph[s64].sum()), int((ph[s64] * ph[d64]).sum())
            r["cpu_oracle_ms_including_trie_build"] = (time.perf_counter() - t) * 1e3
            assert lol == r["lollipop"] and bar == r["barbell"]
        out[name] = r
        log("%s: lollipop %d, barbell %d; learned %.3f / %.3f ms (direct bag %.3f / %.3f), plain %.3f / %.3f ms" %
            (name, r["lollipop"], r["barbell"], r["lollipop_learned_ms"], r["barbell_learned_ms"],
             r["lollipop_learned_direct_ms"], r["barbell_learned_direct_ms"], r["lollipop_plain_ms"],
             r["barbell_plain_ms"]))
        S.close()
    return out


def bench_sssp(E, ctx, steps, cpu=True, scale=22):
    """Unweighted single-source closure (BFS levels) on R-MAT scale 22 with both
    edge directions (128 M rows) from the highest-degree vertex: traversed
    edges per second (the rows of every reached vertex's child set) of the
    direction-optimizing form, the top-down form's time, and the oracle's
    level-by-level restatement on the host."""
    import numpy as np
    from emptyheaded_amd import graphs as G
    s, d, n = G.rmat_graph(scale)
    s, d = G.symmetric(s, d)
    T = E.Trie(ctx, s, d)
    deg = np.bincount(s.astype(np.int64), minlength=n)
    start = int(np.argmax(deg))
    E.sssp_unweighted(ctx, T, start, n)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        dist_td, _ = E.sssp_unweighted(ctx, T, start, n)
    td_ms = (time.perf_counter() - t) * 1e3 / steps
    # direction-optimizing: bottom-up levels while the frontier holds >= n/32 vertices
    E.sssp_unweighted(ctx, T, start, n, rev=T, bottom_up=32)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        dist, levels = E.sssp_unweighted(ctx, T, start, n, rev=T, bottom_up=32)
    ms = (time.perf_counter() - t) * 1e3 / steps
    assert np.array_equal(dist, dist_td)
    reached = dist != E.UNREACHED
    traversed = int(deg[reached].sum())
    out = {"workload": "BFS levels from the highest-degree vertex, R-MAT scale %d both directions" % scale,
           "rows": len(s), "reached": int(reached.sum()), "levels": levels, "ms": ms,
           "traversed_edges_per_s": traversed / (ms * 1e-3), "top_down_ms": td_ms,
           "direction": "bottom-up while the frontier holds >= 1/32 of the vertices"}
    if cpu:
        from oracle import oracle as O
        t = time.perf_counter()
        od, _ = O.sssp_unweighted(s, d, start, n)
        out["cpu_oracle_ms"] = (time.perf_counter() - t) * 1e3
        assert np.array_equal(od, dist)
    log("sssp: %d reached in %d levels, %.3f ms (%.3g traversed edges/s)" % (out["reached"], levels, ms,
                                                                          out["traversed_edges_per_s"]))
    T.close()
    return out


def bench_pagerank(E, ctx, steps, cpu=True, scale=22, iterations=5):
    """The reference's PageRank rule (base 1/N, 5 iterations of
    0.15 + 0.85*SUM(z;1.0) over Edge(x,z),PageRank(z),InvDegree(z)) on R-MAT
    scale 22 with both edge directions, InvDegree = 1/out-degree as float32:
    device time of the iterations for the plain (reference-order) and tuned
    folds, rows*iterations per second, and the HBM rate on algorithmic bytes
    (per iteration 16 B per value of the domain for the term pass, 8 B per row
    and 12 B per head for the fold).  The oracle times one iteration on the
    host cores (trie built beforehand) and checks the plain result bit for bit."""
    import numpy as np
    from emptyheaded_amd import graphs as G
    s, d, n = G.rmat_graph(scale)
    s, d = G.symmetric(s, d)
    T = E.Trie(ctx, s, d)
    heads, deg = np.unique(s, return_counts=True)
    keys, vals = heads.astype(np.uint32), (1.0 / deg).astype(np.float32)
    base = 1.0 / len(heads)
    rows, nh, V = len(s), len(heads), int(max(s.max(), d.max())) + 1
    alg_bytes = iterations * (16.0 * V + 8.0 * rows + 12.0 * nh)
    out = {"workload": "PageRank, R-MAT scale %d both directions, %d iterations" % (scale, iterations),
           "rows": rows, "heads": nh, "iterations": iterations, "algorithmic_bytes": alg_bytes}
    res = {}
    for name, form in (("plain", E.FORM_PLAIN), ("tuned", E.FORM_TUNED)):
        E.pagerank(ctx, T, keys, vals, base, iterations=iterations, form=form)
        times = []
        for _ in range(steps):
            r, ms = E.pagerank(ctx, T, keys, vals, base, iterations=iterations, form=form)
            times.append(ms)
        ms = float(np.median(times))
        res[name] = r
        out[name + "_ms"] = ms
        out[name + "_rows_per_s"] = rows * iterations / (ms * 1e-3)
        out[name + "_hbm_gbs"] = alg_bytes / (ms * 1e-3) / 1e9
    out["tuned_vs_plain_max_rel_diff"] = float(np.max(np.abs(res["tuned"] - res["plain"]) / res["plain"]))
    if cpu:
        from oracle import oracle as O
        OT = O.Trie(s, d)
        r1, _ = E.pagerank(ctx, T, keys, vals, base, iterations=1, form=E.FORM_PLAIN)
        t = time.perf_counter()
        o1 = O.pagerank(s, d, keys, vals, base, iterations=1, nthreads=16, trie=OT)
        out["cpu_oracle_ms_per_iteration"] = (time.perf_counter() - t) * 1e3
        out["cpu_oracle_threads"] = 16
        out["plain_bit_identical_to_oracle"] = bool(np.array_equal(r1.view(np.uint32), o1.view(np.uint32)))
        del OT
    log("pagerank: plain %.3f ms, tuned %.3f ms for %d iterations (%.3g rows/s, %.0f GB/s)" % (
        out["plain_ms"], out["tuned_ms"], iterations, out["tuned_rows_per_s"], out["tuned_hbm_gbs"]))
    T.close()
    return out


def bench_four_clique(E, ctx, steps, scale=18):
    """4-clique COUNT on R-MAT scale 18: plain vs tuned vs tuned+learned."""
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans
    s, d, n = G.rmat_graph(scale)
    out = {"workload": "FliqueAgg COUNT(*) on R-MAT scale %d, edge factor 16, pruned src<dst" % scale,
           "oriented_edges": len(s)}
    from emptyheaded_amd import tuning as TU
    TU.apply_plan("four_clique")
    T_rule = E.Trie(ctx, s, d, E.POLICY_RULE)
    T_learn = E.Trie(ctx, s, d, E.POLICY_LEARNED if _selector_loaded(E) else E.POLICY_RULE)
    counts = set()
    for name, T, form in (("plain", T_rule, E.FORM_PLAIN), ("tuned", T_rule, E.FORM_TUNED),
                          ("learned", T_learn, E.FORM_LEARNED)):
        c, ms, kms, st = time_form(E, ctx, T, plans.FOUR_CLIQUE, form, steps, 1)
        counts.add(c)
        out[name + "_ms"] = ms
        out[name + "_last_level_ms"] = kms
        log("4-clique R-MAT-%d %s: %d, %.3f ms" % (scale, name, c, ms))
    assert len(counts) == 1, "4-clique forms disagree"
    out["four_cliques"] = counts.pop()
    out["speedup_learned_vs_plain"] = out["plain_ms"] / out["learned_ms"]
    TU.apply_plan("triangle")
    T_rule.close()
    T_learn.close()
    return out


def _selector_loaded(E):
    from emptyheaded_amd import tuning
    return tuning.selector_active()


def bench_egonets(E, ctx, steps):
    """4,096 ego-networks from R-MAT scale 22, triangle and 4-cycle count per
    net: one batched launch per query kind vs one launch per net."""
    import numpy as np
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans
    t = time.perf_counter()
    s, d, n = G.rmat_graph(22, orient=G.BOTH)
    es, ed, base, off = G.egonets(s, d, n)
    del s, d
    ss, sd = G.symmetric(es, ed)
    nets = len(base) - 1
    log("ego batch: %d nets, %d edges, %d vertices (%.1fs)" % (nets, len(es), int(base[-1]), time.perf_counter() - t))
    from emptyheaded_amd import tuning as TU
    pol = E.POLICY_LEARNED if _selector_loaded(E) else E.POLICY_RULE
    T = E.Trie(ctx, es, ed, pol)
    S = E.Trie(ctx, ss, sd)
    # the reference attribute order: per net it is faster here than the middle-vertex order,
    # whose level 0 joins two head sets
    TRI, TT = plans.TRIANGLE, [T]
    TU.apply_plan("triangle")
    saved = E.tune_get()
    # batched: per-segment counts, the tuned triangle level replayed from a hipGraph
    E.tune_set(use_graph=1)
    tri, _ = E.join_count_segments(ctx, TRI, TT, base, E.FORM_LEARNED)
    cyc, _ = E.cycle4_count(ctx, S, bounds=base)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        E.join_count_segments(ctx, TRI, TT, base, E.FORM_LEARNED)
    t1 = time.perf_counter()
    for _ in range(steps):
        E.cycle4_count(ctx, S, bounds=base)
    t2 = time.perf_counter()
    E.tune_set(**saved)
    batched_tri, batched_cyc = (t1 - t0) * 1e3 / steps, (t2 - t1) * 1e3 / steps
    # one launch sequence per net (eager: every net has its own range)
    E.tune_set(use_graph=0)
    heads = np.unique(es)
    lo, hi = np.searchsorted(heads, base[:-1]), np.searchsorted(heads, base[1:])
    pq_tri = np.zeros(nets, dtype=np.int64)
    pq_cyc = np.zeros(nets, dtype=np.int64)
    t0 = time.perf_counter()
    for q in range(nets):
        pq_tri[q] = E.join_count(ctx, TRI, TT, E.FORM_LEARNED, int(lo[q]), int(hi[q]))
    t1 = time.perf_counter()
    for q in range(nets):
        pq_cyc[q] = E.cycle4_count(ctx, S, int(base[q]), int(base[q + 1]))
    t2 = time.perf_counter()
    E.tune_set(**saved)
    assert np.array_equal(pq_tri, tri) and np.array_equal(pq_cyc, cyc), "batched and per-net counts disagree"
    out = {"workload": "4096 ego-networks (radius 1, trimmed to log-uniform 64..65536 edges) cut from R-MAT "
                       "scale 22; triangle COUNT and 4-cycle COUNT per net",
           "nets": nets, "edges": int(len(es)), "vertices": int(base[-1]),
           "triangles": int(tri.sum()), "cycles4": int(cyc.sum()),
           "batched_triangle_ms": batched_tri, "batched_cycle4_ms": batched_cyc,
           "batched_ms": batched_tri + batched_cyc,
           "per_query_triangle_ms": (t1 - t0) * 1e3, "per_query_cycle4_ms": (t2 - t1) * 1e3,
           "per_query_ms": (t2 - t0) * 1e3}
    out["speedup_batched_vs_per_query"] = out["per_query_ms"] / out["batched_ms"]
    log("ego batch: batched %.2f ms (tri %.2f, 4-cycle %.2f); per-net launches %.1f ms" %
        (out["batched_ms"], batched_tri, batched_cyc, out["per_query_ms"]))
    TU.apply_plan("triangle")
    for x in (T, S):
        x.close()
    return out


def bench_triples(E, ctx, steps, pool_budget_words=8 << 30):
    """2-hop star and path joins over the two most frequent predicates of a
    50M-triple table (64 Zipf(1.1) predicates), under each layout policy; the
    learned selector's regret is taken against the best policy per query."""
    from emptyheaded_amd import triples as TR
    t = time.perf_counter()
    s, p, o = TR.triple_table()
    rels = TR.by_predicate(s, p, o, 64)
    del s, p, o
    R = [TR.distinct(*rels[0]), TR.distinct(*rels[1])]
    del rels
    exp = {"star": TR.star_count(R[0], R[1]), "path": TR.path_count(R[0], R[1])}
    log("triples: P0 %d rows, P1 %d rows (%.1fs)" % (len(R[0][0]), len(R[1][0]), time.perf_counter() - t))
    out = {"workload": "COUNT(*) of P0(s,o1),P1(s,o2) (star) and P0(x,y),P1(y,z) (path); P0, P1 = the two most "
                       "frequent of 64 Zipf(1.1) predicates in 50M R-MAT-22 (s,o) triples",
           "rows": [len(R[0][0]), len(R[1][0])], "star": exp["star"], "path": exp["path"], "policies": {}}
    policies = [("rule", E.POLICY_RULE), ("all_uint", E.POLICY_ALL_UINT), ("all_bitset", E.POLICY_ALL_BITSET)]
    if _selector_loaded(E):
        policies.append(("learned", E.POLICY_LEARNED))
    for name, pol in policies:
        if pol == E.POLICY_ALL_BITSET:
            words = sum(TR.bitset_pool_words(*r) for r in R)
            if words > pool_budget_words:
                out["policies"][name] = {"infeasible": "bitset pool of %.1f GB" % (words * 8 / 1e9)}
                continue
        T = [E.Trie(ctx, *r, policy=pol) for r in R]
        res = {"build_ms": sum(x.info()["build_ms"] for x in T)}
        for q, plan in (("star", TR.STAR2), ("path", TR.PATH2)):
            c, ms, _, st = time_form(E, ctx, T[0], plan, E.FORM_LEARNED if pol == E.POLICY_LEARNED else E.FORM_PLAIN,
                                     steps, 1, tries=T)
            assert c == exp[q], (name, q, c, exp[q])
            res[q + "_ms"] = ms
            res[q + "_level_ms"] = [round(x, 4) for x in st["level_ms"][:3]]
        out["policies"][name] = res
        log("triples %s: star %.3f ms, path %.3f ms" % (name, res["star_ms"], res["path_ms"]))
        for x in T:
            x.close()
    for q in ("star", "path"):
        times = {k: v[q + "_ms"] for k, v in out["policies"].items() if q + "_ms" in v}
        best = min(times.values())
        out[q + "_best_policy"] = min(times, key=times.get)
        for k, v in times.items():
            out["policies"][k][q + "_regret"] = v / best - 1.0
    return out


def bench_dropin(src, dst, expect, reps=2):
    """The generated TriangleAgg, compiled unchanged against the C++ drop-in API
    (record/replay passes, every intersection batched on the GPU): wall time of
    c_run_TriangleAgg on a DB folder holding the relation (load and trie build
    included, as the generated code does them)."""
    import shutil
    import tempfile
    from emptyheaded_amd import dropin
    db = tempfile.mkdtemp(prefix="ehx_db_")
    try:
        dropin.write_relation(db, "Edge", src, dst)
        times = []
        timer_log = os.path.join(db, "timers.txt")
        sys.stdout.flush()
        saved = os.dup(1)
        fd = os.open(timer_log, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        os.dup2(fd, 1)  # the generated code prints its timers to stdout; keep the JSON line alone there
        try:
            for _ in range(reps):
                t = time.perf_counter()
                v, rows = dropin.run("TriangleAgg", db)
                times.append((time.perf_counter() - t) * 1e3)
                assert v == expect, (v, expect)
        finally:
            sys.stdout.flush()
            os.dup2(saved, 1)
            os.close(saved)
            os.close(fd)
        with open(timer_log) as f:
            timers = f.read()
    finally:
        shutil.rmtree(db, ignore_errors=True)
    query = [float(l.split(":")[1].split()[0]) * 1e3 for l in timers.splitlines() if l.startswith("Time[QUERY TIME]")]
    log("C++ drop-in TriangleAgg: wall %s ms, QUERY TIME %s ms" %
        (", ".join("%.1f" % x for x in times), ", ".join("%.1f" % x for x in query)))
    return {"workload": "generated TriangleAgg (engine/queries/TriangleAgg.cpp) compiled unchanged against the C++ "
                        "drop-in API, same R-MAT graph", "query_ms": min(query) if query else None,
            "wall_ms": min(times), "wall_includes": "Trie::load, trie build, query"}


def cpu_baseline(src, dst, m, expect, budget_s=20.0):
    from emptyheaded_amd import plans
    from oracle import oracle as O
    cores = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    OT = O.Trie(src, dst)
    build_s = time.perf_counter() - t
    # stratified sample: every stride-th chunk of 128 level-0 values
    stride = 64
    while True:
        t = time.perf_counter()
        O.join_count_sampled(plans.TRIANGLE, [OT], cores, stride)
        dt = time.perf_counter() - t
        if dt * 4 > budget_s or stride == 1:
            break
        stride = max(1, stride // 4)
    full = None
    if stride == 1:
        full = O.join_count(plans.TRIANGLE, [OT], cores)
        assert full == expect, (full, expect)
    est_full_s = dt * stride
    return {"value": m / est_full_s, "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": ("full workload" if stride == 1 else
                       "every %d-th chunk of 128 level-0 vertices (1/%d of heads), time scaled by %d" %
                       (stride, stride, stride)),
            "sample_seconds": dt, "trie_build_seconds": build_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--selector", default=None, help="selector weights (default: the committed ones)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 4-clique R-MAT-18, ego-network batch and triple-table measurements")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline form (for profiling its kernel alone)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one copy kernel of this many bytes (traffic-counter reference)")
    ap.add_argument("--config", default=None, choices=["pagerank", "sssp", "two_bag", "selections"],
                    help="run only this secondary workload and print its JSON (for profiling it alone)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the join's multi-GPU form (replicated trie, first attribute "
              "sharded, one 8-byte all-reduce) is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans

    if args.config:
        ctx = E.Context(0)
        fn = {"pagerank": bench_pagerank, "sssp": bench_sssp, "two_bag": bench_two_bag,
              "selections": bench_selections}[args.config]
        print(json.dumps(fn(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu)))
        ctx.close()
        return

    t = time.perf_counter()
    src, dst, nv = G.rmat_graph(args.scale, args.edge_factor)
    m = len(src)
    log("graph: scale %d, %d vertices, %d oriented edges (%.1fs)" % (args.scale, nv, m, time.perf_counter() - t))
    balg = G.algorithmic_bytes(src, dst, nv)
    tuning = load_tuning(E, args.selector)
    log("tuning:", json.dumps(tuning))

    from emptyheaded_amd import tuning as TU
    ctx = E.Context(0)
    learned_pol = E.POLICY_LEARNED if tuning.get("selector") else E.POLICY_RULE
    T_rule = E.Trie(ctx, src, dst, E.POLICY_RULE)
    ctx.synchronize()
    # wall time of the headline's two tries from host arrays (PCIe copy included)
    t_build = time.perf_counter()
    T_learn = E.Trie(ctx, src, dst, learned_pol)
    # the reverse ordering Edge_1_0, built for every relation as the reference's loader does
    R_learn = E.Trie(ctx, dst, src, learned_pol)
    ctx.synchronize()
    build_wall_ms = (time.perf_counter() - t_build) * 1e3
    R_rule = E.Trie(ctx, dst, src, E.POLICY_RULE)
    info = T_learn.info()
    log("tries built: rule %.2f ms, learned %.2f ms (+ reverse %.2f / %.2f ms); bitsets %d (learned) vs %d (rule)" %
        (T_rule.info()["build_ms"], info["build_ms"], R_rule.info()["build_ms"], R_learn.info()["build_ms"],
         info["num_bitsets"], T_rule.info()["num_bitsets"]))

    if args.copy_calib:
        import ctypes as C
        cms = C.c_double(0)
        E.check(E.lib().ehx_copy_calibrate(ctx.h, args.copy_calib, C.byref(cms)))
        log("copy kernel: %d bytes in %.3f ms (%.1f GB/s read+write)" %
            (args.copy_calib, cms.value, 2 * args.copy_calib / cms.value / 1e6))

    MID, REF = "triangle_by_middle", "triangle"
    PLAN = {MID: (plans.TRIANGLE_BY_MIDDLE, [T_learn, R_learn], [T_rule, R_rule]),
            REF: (plans.TRIANGLE, [T_learn], [T_rule])}
    res = {}

    def run(key, plan_name, which, form, steps, warmup):
        TU.apply_plan(plan_name)
        plan, learned_tries, rule_tries = PLAN[plan_name]
        tries = learned_tries if which == "learned" else rule_tries
        r = time_form(E, ctx, tries[0], plan, form, steps, warmup, tries=tries)
        res[key] = r
        log("%-28s %d triangles, %.3f ms/step, last level %.3f ms" % (key, r[0], r[1], r[2]))
        return r

    # headline: tuned+learned, attribute order by middle vertex
    count, ms, kms, st = run("learned_by_middle", MID, "learned", E.FORM_LEARNED, args.steps, args.warmup)
    if not args.headline_only:
        run("learned_reference_order", REF, "learned", E.FORM_LEARNED, args.steps, args.warmup)
        run("tuned_rule_by_middle", MID, "rule", E.FORM_TUNED, args.steps, args.warmup)
        run("tuned_rule_reference_order", REF, "rule", E.FORM_TUNED, args.steps, args.warmup)
    if not args.no_plain and not args.headline_only:
        run("plain_reference_order", REF, "rule", E.FORM_PLAIN, max(2, args.steps // 2), 1)
        run("plain_by_middle", MID, "rule", E.FORM_PLAIN, max(2, args.steps // 2), 1)
    for k, r in res.items():
        assert r[0] == count, "%s disagrees: %d vs %d" % (k, r[0], count)
    # the headline query with its tuned level replayed from a captured hipGraph
    gms, graph_used = None, False
    if not args.headline_only:
        TU.apply_plan(MID)
        E.tune_set(use_graph=1)
        gc, gms, _, _ = time_form(E, ctx, T_learn, plans.TRIANGLE_BY_MIDDLE, E.FORM_LEARNED, args.steps, args.warmup,
                                  tries=[T_learn, R_learn])
        graph_used = E.tune_get()["use_graph"] == 1
        TU.apply_plan(MID)
        assert gc == count, "graph replay disagrees"
        log("learned_by_middle, hipGraph replay: %.3f ms/step (graph %s)" % (gms, "used" if graph_used else "refused"))

    TU.apply_plan(MID)
    headline_knobs = E.tune_get()
    achieved = balg / (kms * 1e-3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_k_tuned_rmat22.json")
    if os.path.exists(tf) and args.scale == 22:
        with open(tf) as f:
            tj = json.load(f)
        if tj.get("plan") == MID:
            traffic = tj.get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": m / (ms * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "TriangleAgg COUNT(*) on R-MAT scale %d, edge factor %d, (a,b,c,d)=(0.57,0.19,0.19,0.05), "
                        "symmetrized, ascending-degree relabel, pruned src<dst" % (args.scale, args.edge_factor),
            "scale": args.scale, "edge_factor": args.edge_factor, "vertices": nv, "oriented_edges": m,
            "triangles": count, "form": "tuned+learned",
            "plan": "attribute order (b, a, c) over Edge_0_1 and Edge_1_0 (middle vertex first)",
        },
        "results_per_s": count / (ms * 1e-3),
        "algorithmic_bytes": balg,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": "k_tuned",
                     "kernel_ms": kms,
                     "traffic_frac": (traffic / (kms * 1e-3) / HBM_PEAK) if traffic else None},
        "trie_build_wall_ms": build_wall_ms,
        "edges_per_s_including_build": m / ((ms + build_wall_ms) * 1e-3),
        "ms_per_step_by_form": {k: r[1] for k, r in res.items()},
        "last_level_ms_by_form": {k: r[2] for k, r in res.items()},
        "graph_replay_ms_per_step": gms if graph_used else None,
        "tuning": tuning,
        "knobs_of_headline_plan": headline_knobs,
    }
    if "plain_reference_order" in res:
        out["plain_ms_per_step"] = res["plain_reference_order"][1]
        out["plain_edges_per_s"] = m / (out["plain_ms_per_step"] * 1e-3)
        out["speedup_vs_plain"] = out["plain_ms_per_step"] / ms
        out["speedup_vs_plain_same_plan"] = res["plain_by_middle"][1] / ms
        out["speedup_learned_vs_plain_reference_order"] = (res["plain_reference_order"][1] /
                                                         res["learned_reference_order"][1])
        out["plain_pairs_by_alg"] = res["plain_reference_order"][3]["pairs_by_alg"]
    if not args.no_configs and not args.headline_only:
        out["configs"] = {"small_graphs": bench_small_graphs(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu),
                          "rmat18_four_clique": bench_four_clique(E, ctx, max(2, args.steps // 2)),
                          "egonet_batch": bench_egonets(E, ctx, max(2, args.steps // 2)),
                          "triple_star_path": bench_triples(E, ctx, max(2, args.steps // 2)),
                          "selections": bench_selections(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu),
                          "two_bag": bench_two_bag(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu),
                          "sssp": bench_sssp(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu),
                          "pagerank": bench_pagerank(E, ctx, max(2, args.steps // 2), cpu=not args.no_cpu),
                          "dropin_cpp_triangle": bench_dropin(src, dst, count)}
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(src, dst, m, count)
    else:
        out["cpu_baseline"] = None
    ctx.synchronize()
    for T in (T_rule, T_learn, R_rule, R_learn):
        T.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
