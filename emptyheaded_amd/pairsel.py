"""Learned per-pair algorithm choice for uint x uint intersections: merge path
or galloping the smaller set into the larger.

The reference picks by a fixed ratio: gallop when the larger set holds more
than 32 times the smaller one's values, merge otherwise
(storage_engine/include/trie/set/ops/intersection/uinteger.hpp:1209-1227).
Here the selector network decides, for the learned form's levels that run one
intersection per tuple (every level the persistent kernel does not run:
materialized levels, plain last levels).

Data: pairs of child sets that a triangle level intersects, N+(a) and N+(b)
for an edge (a, b) of an R-MAT graph in uint layout; the engine records the
shader cycles of counting each pair by merge path and by galloping
(ehx_pair_record, the fewest of `reps` runs, both counts checked equal).
Features (8 per pair, r the smaller set, f the larger): log2 |r|, log2 |f|,
log2 |f|/|r|, log2 of each value range, log2 of each density, and the share
of r's value range that f's range covers.

Training: the layout selector's 8 -> 32 -> 32 -> 2 network and its GPU
forward/backward kernels (selector.train), output 0 learning log2(merge
cycles) - log2(gallop cycles) and output 1 zero, each pair weighted by its
cycle gap (the regret a wrong choice costs); the network picks merge when
output 0 < output 1.

Deployment: the device reads the choice from a table over (log2 |r|,
log2 |f|/|r|) bins (ehx_pair_table_load), filled with the network's decision
for the median features of the bin's training pairs; bins without training
pairs keep the reference's rule.  Regret: extra cycles over the per-pair
faster algorithm, reported for the rule, the table and the network itself.

    python -m emptyheaded_amd.pairsel --out gpurun_out/pair_selector.json
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

from . import check, lib
from . import selector as SEL

HERE = os.path.dirname(os.path.abspath(__file__))
PAIR_JSON = os.path.join(HERE, "pair_selector.json")
SMALL_BINS = RATIO_BINS = 24  # EHX_PAIR_SMALL_BINS, EHX_PAIR_RATIO_BINS
MERGE, GALLOP, RULE = 0, 1, 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def set_table(src, dst):
    """Per head of the relation (src, dst) sorted by src: heads, cardinality,
    first and last value of its child set."""
    o = np.lexsort((dst, src))
    s, d = src[o], dst[o]
    heads, start, card = np.unique(s, return_index=True, return_counts=True)
    first = d[start]
    last = d[start + card - 1]
    return heads, card.astype(np.int64), first.astype(np.int64), last.astype(np.int64)


def sample_pairs(src, dst, n, seed=0):
    """Node pairs (i, j) of the trie of (src, dst): N+(a), N+(b) for up to n
    edges (a, b) whose both ends are heads."""
    heads, card, first, last = set_table(src, dst)
    is_head = np.isin(dst, heads)
    a, b = src[is_head], dst[is_head]
    rng = np.random.default_rng(seed)
    if len(a) > n:
        pick = rng.choice(len(a), n, replace=False)
        a, b = a[pick], b[pick]
    i = np.searchsorted(heads, a).astype(np.uint32)
    j = np.searchsorted(heads, b).astype(np.uint32)
    return i, j, (heads, card, first, last)


def features(i, j, table):
    _, card, first, last = table
    ni, nj = card[i], card[j]
    small_i = ni <= nj
    rn = np.where(small_i, ni, nj).astype(np.float64)
    fn = np.where(small_i, nj, ni).astype(np.float64)
    r0 = np.where(small_i, first[i], first[j]).astype(np.float64)
    r1 = np.where(small_i, last[i], last[j]).astype(np.float64)
    f0 = np.where(small_i, first[j], first[i]).astype(np.float64)
    f1 = np.where(small_i, last[j], last[i]).astype(np.float64)
    rr, fr = r1 - r0 + 1, f1 - f0 + 1
    cover = np.clip(np.minimum(r1, f1) - np.maximum(r0, f0) + 1, 0, None) / rr
    X = np.stack([np.log2(rn + 1), np.log2(fn + 1), np.log2(fn / rn), np.log2(rr + 1), np.log2(fr + 1),
                  np.log2(rn / rr), np.log2(fn / fr), cover], 1)
    return X.astype(np.float32), rn.astype(np.int64), fn.astype(np.int64)


def record(ctx, T, i, j, reps=3):
    """cycles (n, 2): merge path, gallop (0: not a uint x uint pair)."""
    i = np.ascontiguousarray(i, np.uint32)
    j = np.ascontiguousarray(j, np.uint32)
    cyc = np.zeros(2 * max(len(i), 1), np.uint64)
    u32 = C.POINTER(C.c_uint32)
    check(lib().ehx_pair_record(ctx.h, T.h, T.h, i.ctypes.data_as(u32), j.ctypes.data_as(u32), len(i), reps,
                                cyc.ctypes.data_as(C.POINTER(C.c_uint64))))
    return cyc[: 2 * len(i)].reshape(-1, 2).astype(np.float64)


def dataset(ctx, T, src, dst, n, seed=0, reps=3):
    i, j, table = sample_pairs(src, dst, n, seed)
    X, rn, fn = features(i, j, table)
    cyc = record(ctx, T, i, j, reps)
    ok = (cyc[:, 0] > 0) & (cyc[:, 1] > 0)
    return X[ok], cyc[ok], rn[ok], fn[ok]


def targets(cyc):
    """Output 0: log2(merge) - log2(gallop); output 1: 0; weights |gap|."""
    Y = np.stack([np.log2(cyc[:, 0]) - np.log2(cyc[:, 1]), np.zeros(len(cyc))], 1).astype(np.float32)
    w = np.abs(cyc[:, 0] - cyc[:, 1])
    w = w / max(w.mean(), 1.0)
    return Y, np.stack([w, w], 1).astype(np.float32)


def bins(rn, fn):
    """The device's bins (pair_gallop): floor(log2 |r|), floor(log2 (|f| div |r|))."""
    sb = np.minimum(np.floor(np.log2(np.maximum(rn, 1))).astype(np.int64), SMALL_BINS - 1)
    rb = np.minimum(np.floor(np.log2(np.maximum(fn // np.maximum(rn, 1), 1))).astype(np.int64), RATIO_BINS - 1)
    return sb, rb


def rule_gallop(rn, fn):
    return fn > 32 * rn


def net_gallop(ctx, params, norm, X):
    pred = SEL.predict(ctx, params, norm, X)
    return pred[:, 0] >= pred[:, 1]


def make_table(ctx, params, norm, X, rn, fn, min_pairs=8):
    sb, rb = bins(rn, fn)
    key = sb * RATIO_BINS + rb
    table = np.full(SMALL_BINS * RATIO_BINS, RULE, np.uint8)
    reps, cells = [], []
    for k in np.unique(key):
        sel = key == k
        if sel.sum() < min_pairs:
            continue
        reps.append(np.median(X[sel], axis=0))
        cells.append(k)
    if cells:
        g = net_gallop(ctx, params, norm, np.array(reps, np.float32))
        table[np.array(cells)] = np.where(g, GALLOP, MERGE)
    return table


def table_gallop(table, rn, fn):
    sb, rb = bins(rn, fn)
    c = table[sb * RATIO_BINS + rb]
    return np.where(c == RULE, rule_gallop(rn, fn), c == GALLOP)


def regret(cyc, gallop):
    best = np.minimum(cyc[:, 0], cyc[:, 1]).sum()
    chosen = np.where(gallop, cyc[:, 1], cyc[:, 0]).sum()
    return {"pairs": int(len(cyc)), "best_cycles": float(best), "chosen_cycles": float(chosen),
            "regret": float((chosen - best) / max(best, 1.0))}


def evaluate(ctx, params, norm, table, X, cyc, rn, fn):
    return {"rule": regret(cyc, rule_gallop(rn, fn)), "table": regret(cyc, table_gallop(table, rn, fn)),
            "network": regret(cyc, net_gallop(ctx, params, norm, X)),
            "always_merge": regret(cyc, np.zeros(len(cyc), bool)),
            "always_gallop": regret(cyc, np.ones(len(cyc), bool))}


def load(device, table):
    t = np.ascontiguousarray(table, np.uint8)
    assert t.size == SMALL_BINS * RATIO_BINS
    check(lib().ehx_pair_table_load(device, t.ctypes.data_as(C.POINTER(C.c_uint8))))


def unload(device=0):
    check(lib().ehx_pair_table_load(device, None))


def load_committed(device=0, path=PAIR_JSON):
    with open(path) as f:
        d = json.load(f)
    load(device, np.array(d["table"], np.uint8))
    return {"trained_on": d.get("trained_on"), "regret": d.get("regret")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out/pair_selector.json")
    ap.add_argument("--train-scales", default="16,17,18,19,20")
    ap.add_argument("--eval-scale", type=int, default=21)
    ap.add_argument("--pairs", type=int, default=200000, help="sampled pairs per graph")
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    ctx = E.Context(0)
    Xs, Cs, Rs, Fs = [], [], [], []
    for sc in [int(x) for x in args.train_scales.split(",")]:
        s, d, _ = G.rmat_graph(sc)
        T = E.Trie(ctx, s, d, E.POLICY_ALL_UINT)
        X, cyc, rn, fn = dataset(ctx, T, s, d, args.pairs, seed=sc, reps=args.reps)
        T.close()
        log("R-MAT-%d: %d pairs, merge faster on %.1f%%" % (sc, len(X), 100 * np.mean(cyc[:, 0] < cyc[:, 1])))
        Xs.append(X), Cs.append(cyc), Rs.append(rn), Fs.append(fn)
    X, cyc, rn, fn = (np.concatenate(a) for a in (Xs, Cs, Rs, Fs))
    Y, M = targets(cyc)
    params, norm, hist = SEL.train(ctx, X, Y, M, steps=args.steps)
    table = make_table(ctx, params, norm, X, rn, fn)
    train_regret = evaluate(ctx, params, norm, table, X, cyc, rn, fn)
    s, d, _ = G.rmat_graph(args.eval_scale)
    T = E.Trie(ctx, s, d, E.POLICY_ALL_UINT)
    Xe, ce, re_, fe = dataset(ctx, T, s, d, args.pairs, seed=args.eval_scale, reps=args.reps)
    T.close()
    held = evaluate(ctx, params, norm, table, Xe, ce, re_, fe)
    for name, r in (("train", train_regret), ("R-MAT-%d" % args.eval_scale, held)):
        log("%s regret: %s" % (name, {k: round(v["regret"], 5) for k, v in r.items()}))
    out = {"params": params.tolist(), "norm": norm.tolist(), "table": table.tolist(),
           "trained_on": {"scales": args.train_scales, "pairs": int(len(X)), "steps": args.steps, "reps": args.reps,
                          "target": "log2 merge cycles - log2 gallop cycles, weighted by the cycle gap",
                          "final_loss": hist[-1][1]},
           "regret": {"train": train_regret, "eval_scale": args.eval_scale, "held_out": held},
           "table_cells": {"merge": int((table == MERGE).sum()), "gallop": int((table == GALLOP).sum()),
                           "rule": int((table == RULE).sum())}}
    with open(args.out, "w") as f:
        json.dump(out, f)
    print(json.dumps({"regret_held_out": {k: v["regret"] for k, v in held.items()}, "cells": out["table_cells"]}))
    ctx.close()


if __name__ == "__main__":
    main()
