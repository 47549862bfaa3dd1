"""Autotuner for the tuned form's knobs (LDS staging size, wave/workgroup bin
cut, long-child-set cutoff, grid size, queue grab, bitmap rule, hipGraph).

Coordinate descent over a small grid on training workloads (R-MAT graphs
smaller than the evaluation graph), timing the tuned last level with HIP
events; every configuration's count is checked against the first one (knobs
never change results).  Writes a timing DB:

    python -m emptyheaded_amd.autotune --out gpurun_out/tuning.json
"""
import argparse
import itertools
import json
import sys
import time

GRID = {
    "lds_slots": [2048, 4096, 8192],
    "big_deg": [16, 32, 64, 128],
    "waves_per_cu": [8, 16, 24, 32, 48],
    "grab": [1, 2, 4, 8],
    "big_work": [1024, 4096, 16384, 65536, 262144],
    "stream_lanes": [2, 4, 8, 16, 32, 64],
    "gap_coded": [0, 1],
    "split_work": [0, 1 << 19, 1 << 20, 1 << 21, 1 << 22],
    "tiny_work": [0, 64, 128, 256, 1024, 4096, 16384],
    "bitmap_all": [0, 1],
    "use_graph": [0, 1],
}

# knobs whose best values depend on each other
JOINT = [("stream_lanes", "gap_coded"), ("lds_slots", "big_deg")]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measure(E, ctx, tries, plan, form, reps):
    E.join_count(ctx, plan, tries, form)  # warm (and capture, in graph mode)
    ctx.synchronize()
    best, total = None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        total, st = E.join_count(ctx, plan, tries, form, stats=True)
        dt = (time.perf_counter() - t0) * 1e3
        best = dt if best is None else min(best, dt)
    return total, best


def tune(E, ctx, workloads, passes=2, reps=3, grid=GRID, form=None):
    """workloads: list of (name, tries, plan).  Objective: sum of best-of-reps
    wall ms per query over the workloads."""
    form = E.FORM_TUNED if form is None else form
    cur = E.tune_get()
    ref = {}
    history = []

    def evaluate(params):
        E.tune_set(**params)
        tot = 0.0
        for name, tries, plan in workloads:
            cnt, ms = measure(E, ctx, tries, plan, form, reps)
            if name in ref and ref[name] != cnt:
                raise AssertionError("knobs changed the result of %s: %d vs %d" % (name, cnt, ref[name]))
            ref.setdefault(name, cnt)
            tot += ms
        history.append({"params": dict(params), "ms": tot})
        log("  %.3f ms  %s" % (tot, json.dumps(params)))
        return tot

    best_ms = evaluate(cur)
    for _ in range(passes):
        improved = False
        for knob, values in grid.items():
            for v in values:
                if cur[knob] == v:
                    continue
                trial = dict(cur)
                trial[knob] = v
                ms = evaluate(trial)
                if ms < best_ms * 0.99:
                    best_ms, cur, improved = ms, trial, True
        # coupled knobs, moved together (one at a time can stall between them)
        for group in JOINT:
            if not all(k in grid for k in group):
                continue
            for vals in itertools.product(*(grid[k] for k in group)):
                if all(cur[k] == v for k, v in zip(group, vals)):
                    continue
                trial = dict(cur)
                trial.update(zip(group, vals))
                ms = evaluate(trial)
                if ms < best_ms * 0.99:
                    best_ms, cur, improved = ms, trial, True
        if not improved:
            break
    E.tune_set(**cur)
    return cur, best_ms, history, ref


PLANS = ("mixed", "triangle", "triangle_by_middle", "four_clique")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out/tuning.json")
    ap.add_argument("--plan", default="mixed", choices=PLANS,
                    help="mixed: triangle and 4-clique together (the global knobs); otherwise one plan's knobs")
    ap.add_argument("--scales", default="20")
    ap.add_argument("--clique-scale", type=int, default=17)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--layouts", default="learned", choices=("learned", "rule"),
                    help="tries built with the committed selector (and timed in the learned form) or the density rule")
    ap.add_argument("--fixed", default="bitmap_all=0",
                    help="knobs held at a value, e.g. bitmap_all=0 (the learned layouts decide staging)")
    args = ap.parse_args()
    fixed = dict((kv.split("=")[0], int(kv.split("=")[1])) for kv in args.fixed.split(",") if kv)
    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans, tuning
    applied = tuning.apply_committed()  # start from the committed knobs and selector
    if args.plan != "mixed":
        tuning.apply_plan(args.plan)  # and from the plan's own committed knobs
    E.tune_set(**fixed)
    learned = args.layouts == "learned" and applied.get("selector")
    pol = E.POLICY_LEARNED if learned else E.POLICY_RULE
    form = E.FORM_LEARNED if learned else E.FORM_TUNED
    grid = {k: v for k, v in GRID.items() if k not in fixed}
    ctx = E.Context(0)
    workloads, keep = [], []
    if args.plan in ("mixed", "triangle", "triangle_by_middle"):
        for s in [int(x) for x in args.scales.split(",")]:
            src, dst, nv = G.rmat_graph(s)
            T = E.Trie(ctx, src, dst, pol)
            keep.append(T)
            if args.plan == "triangle_by_middle":
                T10 = E.Trie(ctx, dst, src, pol)
                keep.append(T10)
                workloads.append(("triangle_by_middle_rmat%d" % s, [T, T10], plans.TRIANGLE_BY_MIDDLE))
            else:
                workloads.append(("triangle_rmat%d" % s, [T], plans.TRIANGLE))
    if args.plan in ("mixed", "four_clique") and args.clique_scale:
        src, dst, nv = G.rmat_graph(args.clique_scale)
        T = E.Trie(ctx, src, dst, pol)
        keep.append(T)
        workloads.append(("clique4_rmat%d" % args.clique_scale, [T], plans.FOUR_CLIQUE))
    best, ms, hist, ref = tune(E, ctx, workloads, passes=args.passes, grid=grid, form=form)
    out = {"plan": args.plan, "best": best, "objective_ms": ms, "workloads": [w[0] for w in workloads],
           "counts": ref, "fixed": fixed, "layouts": "learned" if learned else "rule", "history": hist}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log("best", json.dumps(best), "%.3f ms" % ms)
    print(json.dumps({"plan": args.plan, "best": best, "objective_ms": ms}))


if __name__ == "__main__":
    main()
