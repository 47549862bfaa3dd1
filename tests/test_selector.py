"""Learned selector: the CPU oracle's gradient is pinned by finite
differences, the GPU forward/backward kernels match the oracle, training on
engine-recorded costs reduces the loss, and layouts chosen by the network
never change results."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(2.0, 3.0, size=(n, MO.IN))
    y = rng.normal(10.0, 2.0, size=(n, MO.OUT))
    m = (rng.random((n, MO.OUT)) < 0.8).astype(np.float64)
    m[0] = 1.0
    params = rng.normal(0, 0.4, MO.NPARAM)
    norm = np.concatenate([x.mean(0), 1.0 / x.std(0)])
    return params, norm, x, y, m


def test_oracle_gradient_matches_finite_differences():
    params, norm, x, y, m = rand_problem(40, 1)
    _, g = MO.loss_and_grad(params, norm, x, y, m, half=False)
    gn = MO.numeric_grad(params, norm, x, y, m)
    assert np.allclose(g, gn, rtol=1e-5, atol=1e-7)


def test_oracle_dimensions():
    assert MO.NPARAM == 32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2 == 1410
    import emptyheaded_amd as E
    assert E.lib().ehx_selector_num_params() == MO.NPARAM


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(ctx):
    from emptyheaded_amd import selector as S
    params, norm, x, y, m = rand_problem(5000, 2)
    mlp = S.GpuMlp(ctx, x, y, m)
    try:
        loss, g = mlp.grad(params.astype(np.float32), norm.astype(np.float32))
    finally:
        mlp.close()
    # the oracle sees the same FP32 inputs and FP16-rounded weights
    lo, go = MO.loss_and_grad(params.astype(np.float32), norm.astype(np.float32), x.astype(np.float32),
                              y.astype(np.float32), m.astype(np.float32), half=True)
    # FP32 accumulation on the device vs float64 here
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.allclose(g, go, rtol=2e-3, atol=2e-4 * np.abs(go).max())


@pytest.mark.gpu
def test_gpu_predict_matches_oracle(ctx):
    from emptyheaded_amd import selector as S
    params, norm, x, y, m = rand_problem(3000, 3)
    out = S.predict(ctx, params.astype(np.float32), norm.astype(np.float32), x.astype(np.float32))
    ref, _ = MO.forward(params.astype(np.float32), norm.astype(np.float32), x.astype(np.float32), half=True)
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.gpu
def test_training_on_recorded_costs_and_learned_layouts_exact(ctx, oracle):
    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans
    from emptyheaded_amd import selector as S
    s, d, n = G.rmat_graph(14)
    T = E.Trie(ctx, s, d)
    X, Y, M, c0, c1, cnt = S.dataset(ctx, T, reps=1)
    assert cnt == oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    assert (M.sum(0) > 0).all()
    keep = M.sum(1) > 0
    params, norm, hist = S.train(ctx, X[keep], Y[keep], M[keep], steps=150, log_every=50)
    assert hist[-1][1] < hist[0][1]
    try:
        S.load(0, params, norm, 2048)
        TL = E.Trie(ctx, s, d, E.POLICY_LEARNED)
        for form in (E.FORM_PLAIN, E.FORM_TUNED, E.FORM_LEARNED):
            assert E.join_count(ctx, plans.TRIANGLE, [TL], form) == cnt
        assert E.join_count(ctx, plans.FOUR_CLIQUE, [TL], E.FORM_LEARNED) == \
            E.join_count(ctx, plans.FOUR_CLIQUE, [T], E.FORM_PLAIN)
    finally:
        E.lib().ehx_selector_clear()


@pytest.mark.gpu
def test_recording_the_middle_vertex_plan(ctx, oracle):
    """Recorded staging costs for the triangle ordered by its middle vertex: one
    sample per middle vertex, its staged set a set of Edge_0_1; the recorded
    runs count every triangle."""
    import emptyheaded_amd as E
    from emptyheaded_amd import graphs as G
    from emptyheaded_amd import plans
    from emptyheaded_amd import selector as S
    s, d, n = G.rmat_graph(13)
    T = E.Trie(ctx, s, d)
    X, Y, M, c0, c1, cnt = S.dataset(ctx, T, reps=1, plan="triangle_by_middle", src=s, dst=d)
    assert cnt == oracle.join_count(plans.TRIANGLE, [oracle.Trie(s, d)], 8)
    mids = np.intersect1d(np.unique(s), np.unique(d))
    assert X.shape == (len(mids), S.IN) and len(c0) == len(mids)
    assert (M.sum(0) > 0).all()
    # features of middle vertex b are those of b's set in Edge_0_1
    allx = S.features(ctx, T)
    assert np.array_equal(X, allx[np.searchsorted(np.unique(s), mids)])
