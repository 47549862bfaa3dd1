"""Learned layout selector: training data, training, evaluation, loading.

Data: the engine records, per prefix of the tuned triangle level, the shader
cycles spent with the prefix's set staged as an LDS bitmap and as an LDS hash
table (ehx_join_record, force_mode 0 and 1).  For the triangle query the
prefix is a head vertex a and its staged set is N(a), so each sample is one
child set of the trie: features from ehx_trie_set_features (the same device
code the trie build evaluates), targets log2(cycles) per staging, masked where
the forced staging did not fit.

Training runs on the GPU: ehx_mlp_grad computes the loss and gradient with
the project's own forward/backward kernels; the Adam update of the FP32
master weights runs on the host (1,410 parameters).  The density rule stays
the fallback (no weights loaded, or a set too wide to stage as a bitmap) and
the judge of regret: per prefix, the oracle-best staging is the faster of the
two recorded runs, and regret is the extra cycles a chooser spends over it.

    python -m emptyheaded_amd.selector --out gpurun_out/selector.json
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import check, lib, make_plan, plans

HERE = os.path.dirname(os.path.abspath(__file__))
SELECTOR_JSON = os.path.join(HERE, "selector.json")
IN, H, OUT = 8, 32, 2
NPARAM = H * IN + H + H * H + H + OUT * H + OUT
MODE_SHIFT = 62


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _sigs():
    return lib()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def features(ctx, T):
    L = _sigs()
    nh = T.info()["num_heads"]
    out = np.zeros((nh, IN), np.float32)
    check(L.ehx_trie_set_features(ctx.h, T.h, _fp(out)))
    return out


def record(ctx, T, force_mode, plan=plans.TRIANGLE, tries=None, nprefix=None):
    """Cycles per prefix and the staging actually used (-1: no work)."""
    L = _sigs()
    p = make_plan(plan)
    tries = tries or [T]
    nh = T.info()["num_heads"] if nprefix is None else nprefix
    cyc = np.zeros(max(nh, 1), np.uint64)
    n = C.c_uint64(0)
    res = C.c_int64(0)
    arr = (C.c_void_p * len(tries))(*[t.h.value for t in tries])
    check(L.ehx_join_record(ctx.h, C.byref(p), arr, force_mode, cyc.ctypes.data_as(C.POINTER(C.c_uint64)), nh,
                            C.byref(n), C.byref(res)))
    assert n.value == nh, (n.value, nh)
    cyc = cyc[:nh]
    mode = np.where(cyc == 0, -1, (cyc >> np.uint64(MODE_SHIFT)).astype(np.int64))
    cycles = (cyc & np.uint64((1 << MODE_SHIFT) - 1)).astype(np.float64)
    return cycles, mode, res.value


def dataset(ctx, T, reps=2, plan="triangle", src=None, dst=None):
    """X, Y (log2 cycles: bitmap, hash), M (masks), plus raw cycles.

    plan "triangle": the prefix is a head vertex a of T and the staged set is
    N+(a).  plan "triangle_by_middle": the prefix is a middle vertex b (a head of
    both T and its reverse, built here from src, dst) and the staged set is
    N+(b), again a set of T, so every sample is one set of T either way."""
    X = features(ctx, T)
    if plan == "triangle_by_middle":
        import emptyheaded_amd as E
        R = E.Trie(ctx, dst, src)
        heads = np.unique(src)
        mids = np.intersect1d(heads, np.unique(dst))
        X = X[np.searchsorted(heads, mids)]
        rec = dict(plan=plans.TRIANGLE_BY_MIDDLE, tries=[T, R], nprefix=len(mids))
    else:
        rec = {}
    c0s, c1s = [], []
    for _ in range(reps):
        c0, m0, r0 = record(ctx, T, 0, **rec)
        c1, m1, r1 = record(ctx, T, 1, **rec)
        assert r0 == r1, "staging changed the count"
        c0s.append(c0)
        c1s.append(c1)
    c0 = np.min(c0s, axis=0)
    c1 = np.min(c1s, axis=0)
    v0 = (m0 == 0) & (c0 > 0)
    v1 = (m1 == 1) & (c1 > 0)
    Y = np.stack([np.log2(c0 + 1), np.log2(c1 + 1)], 1).astype(np.float32)
    M = np.stack([v0, v1], 1).astype(np.float32)
    if plan == "triangle_by_middle":
        R.close()
    return X, Y, M, c0, c1, r0


def difference_targets(Y, M, c0, c1):
    """Targets for the decision itself: output 0 learns log2(bitmap cost) -
    log2(hash cost), output 1 learns 0, on prefixes where both stagings were
    recorded, weighted by |bitmap cost - hash cost| (the regret a wrong choice
    costs).  The selector's rule 'output 0 < output 1' then picks the cheaper."""
    both = (M[:, 0] > 0) & (M[:, 1] > 0)
    Yd = np.stack([Y[:, 0] - Y[:, 1], np.zeros(len(Y), np.float32)], 1).astype(np.float32)
    gap = np.abs(c0 - c1)
    w = np.where(both, gap, 0.0)
    w = w / max(w[w > 0].mean(), 1.0) if (w > 0).any() else w
    return Yd, np.stack([w, w], 1).astype(np.float32)


def cost_weights(M, c0, c1):
    """Per-sample loss weights proportional to the prefix's cost: regret is
    measured in cycles, so expensive prefixes must be fitted best."""
    big = np.where(M[:, 0] > 0, c0, 0.0)
    big = np.where((M[:, 1] > 0) & ((M[:, 0] == 0) | (c1 < big)), c1, big)
    w = big / max(big[big > 0].mean(), 1.0) if (big > 0).any() else np.ones(len(M))
    return (M * w[:, None]).astype(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((IN, H), (H, H), (H, OUT)):
        p[o:o + fan_in * fan_out] = rng.normal(0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


class GpuMlp:
    def __init__(self, ctx, X, Y, M):
        L = _sigs()
        self.L = L
        self.X, self.Y, self.M = (np.ascontiguousarray(a, np.float32) for a in (X, Y, M))
        h = C.c_void_p()
        check(L.ehx_mlp_create(ctx.h, _fp(self.X), _fp(self.Y), _fp(self.M), len(self.X), C.byref(h)))
        self.h = h

    def grad(self, params, norm):
        g = np.zeros(NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        n = np.ascontiguousarray(norm, np.float32)
        check(self.L.ehx_mlp_grad(self.h, _fp(p), _fp(n), _fp(g), _fp(loss)))
        return float(loss[0]), g

    def close(self):
        if self.h:
            check(self.L.ehx_mlp_destroy(self.h))
            self.h = None


def predict(ctx, params, norm, X):
    L = _sigs()
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros((len(X), OUT), np.float32)
    p = np.ascontiguousarray(params, np.float32)
    n = np.ascontiguousarray(norm, np.float32)
    check(L.ehx_mlp_predict(ctx.h, _fp(p), _fp(n), _fp(X), len(X), _fp(out)))
    return out


def make_norm(X):
    mu = X.mean(0)
    sd = X.std(0)
    sd[sd < 1e-6] = 1.0
    return np.concatenate([mu, 1.0 / sd]).astype(np.float32)


def train(ctx, X, Y, M, steps=600, lr=3e-3, seed=0, log_every=100):
    norm = make_norm(X)
    p = init_params(seed)
    # the output bias starts at the mean target so training begins near the data
    for k in range(OUT):
        sel = M[:, k] > 0
        if sel.any():
            p[NPARAM - OUT + k] = float(Y[sel, k].mean())
    mlp = GpuMlp(ctx, X, Y, M)
    m1 = np.zeros(NPARAM, np.float64)
    m2 = np.zeros(NPARAM, np.float64)
    b1, b2, eps = 0.9, 0.999, 1e-8
    hist = []
    master = p.astype(np.float64)
    try:
        for t in range(1, steps + 1):
            loss, g = mlp.grad(master.astype(np.float32), norm)
            g = g.astype(np.float64)
            m1 = b1 * m1 + (1 - b1) * g
            m2 = b2 * m2 + (1 - b2) * g * g
            step = lr * (m1 / (1 - b1 ** t)) / (np.sqrt(m2 / (1 - b2 ** t)) + eps)
            master -= step
            if t == 1 or t % log_every == 0 or t == steps:
                hist.append((t, loss))
                log("  step %4d  loss %.4f" % (t, loss))
    finally:
        mlp.close()
    return master.astype(np.float32), norm, hist


def regret(c0, c1, M, choose_bitmap):
    """Extra cycles over the per-prefix best staging, on prefixes where both
    stagings were recorded."""
    both = (M[:, 0] > 0) & (M[:, 1] > 0)
    best = np.minimum(c0, c1)[both]
    chosen = np.where(choose_bitmap[both], c0[both], c1[both])
    return {"prefixes": int(both.sum()), "best_cycles": float(best.sum()), "chosen_cycles": float(chosen.sum()),
            "regret": float((chosen.sum() - best.sum()) / max(best.sum(), 1.0))}


def rule_choice(ctx, T, rows=None):
    types, cards = T.node_info()
    types = types if rows is None else types[rows]
    return types == 0  # bitset layout -> staged as a bitmap


def learned_choice(ctx, params, norm, X, max_words):
    pred = predict(ctx, params, norm, X)
    words64 = np.exp2(X[:, 4])
    n = np.exp2(X[:, 0]) - 1
    ok = (words64 <= max_words + 0.5) & (n > 0.5)
    return ok & (pred[:, 0] < pred[:, 1]), pred


def params_fp16_bits(params):
    return np.asarray(params, np.float32).astype(np.float16).view(np.uint16)


def load(device, params, norm, max_words):
    L = lib()
    bits = np.ascontiguousarray(params_fp16_bits(params))
    n = np.ascontiguousarray(norm, np.float32)
    check(L.ehx_selector_load(device, bits.ctypes.data_as(C.POINTER(C.c_uint16)), len(bits), _fp(n),
                              float(max_words)))


def load_committed(device=0, path=SELECTOR_JSON):
    with open(path) as f:
        d = json.load(f)
    load(device, np.array(d["params"], np.float32), np.array(d["norm"], np.float32), d["max_words"])
    return {"trained_on": d.get("trained_on"), "regret": d.get("regret")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out/selector.json")
    ap.add_argument("--train-scales", default="16,17,18,19,20")
    ap.add_argument("--eval-scale", type=int, default=22)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--weighting", choices=("uniform", "cost"), default="cost")
    ap.add_argument("--target", choices=("costs", "diff"), default="diff")
    ap.add_argument("--plans", default="triangle_by_middle",
                    help="comma list of plans whose recorded costs train the selector "
                         "(triangle, triangle_by_middle); regret is reported for each")
    args = ap.parse_args()
    train_plans = [x for x in args.plans.split(",") if x]
    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import tuning
    tuning.apply_committed()
    tp = E.tune_get()
    max_words = tp["lds_slots"] // 2  # bitmap capacity of a workgroup, in 64-bit words
    ctx = E.Context(0)
    Xs, Ys, Ms = [], [], []
    t0 = time.perf_counter()
    for s in [int(x) for x in args.train_scales.split(",")]:
        src, dst, nv = G.rmat_graph(s)
        T = E.Trie(ctx, src, dst)
        for plan in train_plans:
            X, Y, M, c0, c1, cnt = dataset(ctx, T, plan=plan, src=src, dst=dst)
            log("scale %d %s: %d sets, %d with both stagings" %
                (s, plan, len(X), int(((M[:, 0] > 0) & (M[:, 1] > 0)).sum())))
            if args.target == "diff":
                Y, M = difference_targets(Y, M, c0, c1)
            elif args.weighting == "cost":
                M = cost_weights(M, c0, c1)
            keep = M.sum(1) > 0
            Xs.append(X[keep])
            Ys.append(Y[keep])
            Ms.append(M[keep])
        T.close()
    X, Y, M = np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)
    log("training on %d samples (%.1fs of recording)" % (len(X), time.perf_counter() - t0))
    params, norm, hist = train(ctx, X, Y, M, steps=args.steps)
    # evaluation graph
    src, dst, nv = G.rmat_graph(args.eval_scale)
    T = E.Trie(ctx, src, dst)
    regrets = {"eval_scale": args.eval_scale}
    for plan in ("triangle_by_middle", "triangle"):
        Xe, Ye, Me, c0, c1, cnt = dataset(ctx, T, plan=plan, src=src, dst=dst)
        rows = None
        if plan == "triangle_by_middle":
            heads = np.unique(src)
            rows = np.searchsorted(heads, np.intersect1d(heads, np.unique(dst)))
        rule = rule_choice(ctx, T, rows)
        learned, pred = learned_choice(ctx, params, norm, Xe, max_words)
        regrets[plan] = {"rule": regret(c0, c1, Me, rule), "learned": regret(c0, c1, Me, learned),
                         "bitmap_when_fits": regret(c0, c1, Me, Me[:, 0] > 0),
                         "bitmap_sets": {"rule": int(rule.sum()), "learned": int(learned.sum())}}
    out = {
        "params": params.tolist(), "norm": norm.tolist(), "max_words": max_words,
        "trained_on": {"scales": args.train_scales, "plans": train_plans, "samples": int(len(X)),
                       "steps": args.steps, "weighting": args.weighting, "target": args.target,
                       "final_loss": hist[-1][1]},
        "loss_history": hist,
        "regret": regrets,
    }
    log(json.dumps(out["regret"], indent=1))
    with open(args.out, "w") as f:
        json.dump(out, f)
    print(json.dumps(out["regret"]))


if __name__ == "__main__":
    main()
