"""The cost model's MLP: CPU oracle pinned by finite differences; GPU kernels against the oracle."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def rand_problem(n, seed):
    r = np.random.default_rng(seed)
    w = r.normal(0, 0.3, M.NPARAM)
    z = r.normal(0, 1, (n, M.NIN))
    t = np.stack([r.normal(8, 2, n), r.normal(0, 1, n)], 1)
    return w, z, t


def test_oracle_gradient_matches_finite_differences():
    w, z, t = rand_problem(37, 1)
    loss, g = M.loss_grad(w, z, t, 0.3)
    assert abs(loss - M.loss_only(w, z, t, 0.3)) < 1e-12
    r = np.random.default_rng(2)
    for p in r.choice(M.NPARAM, 60, replace=False):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (M.loss_only(wp, z, t, 0.3) - M.loss_only(wm, z, t, 0.3)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 + 1e-5 * abs(fd), p


def test_makespan_longest_first_beats_random():
    from canu_amd import learned
    r = np.random.default_rng(0)
    costs = r.lognormal(8, 2, 5000)
    rand = learned.makespan(costs, r.permutation(5000), 64)
    lpt = learned.makespan(costs, np.argsort(-costs), 64)
    assert lpt <= rand
    assert lpt >= costs.sum() / 64 - 1e-6 and lpt >= costs.max()


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    from canu_amd import oic
    eng = oic.Engine()
    for n, seed in ((1, 3), (64, 4), (1000, 5)):
        w, z, t = rand_problem(n, seed)
        loss, g = eng.model_grad(w.astype(np.float32), z.astype(np.float32), t.astype(np.float32), 0.3)
        # oracle on the float32-rounded inputs
        ol, og = M.loss_grad(w.astype(np.float32).astype(np.float64), z.astype(np.float32).astype(np.float64),
                             t.astype(np.float32).astype(np.float64), 0.3)
        assert abs(loss - ol) <= 1e-4 * abs(ol) + 1e-5
        scale = np.abs(og).max()
        assert np.max(np.abs(g - og)) <= 2e-4 * scale, (n, np.max(np.abs(g - og)), scale)


@pytest.mark.gpu
def test_gpu_training_fits_and_predicts():
    from canu_amd import oic
    r = np.random.default_rng(7)
    n = 20000
    x = r.normal(0, 1, (n, 12)).astype(np.float32) * 2 + 3
    y0 = 1.5 * x[:, 0] - 0.7 * x[:, 3] + np.maximum(x[:, 5], 0) + 0.1 * r.normal(0, 1, n)
    y1 = 0.5 * x[:, 2]
    t = np.stack([y0, y1], 1).astype(np.float32)
    eng = oic.Engine()
    loss = eng.train_model(x, t, epochs=30, batch=1024, lr=3e-3)
    pred = eng.predict(x)
    assert loss < 0.05, loss
    assert np.corrcoef(pred[:, 0], y0)[0, 1] > 0.99
    w, mu, sd = eng.model()
    # oracle forward on the same weights
    yo, _ = M.forward(w.astype(np.float64), (x - mu) / sd)
    assert np.max(np.abs(yo - pred)) < 1e-3 * (1 + np.abs(yo).max())


@pytest.mark.gpu
def test_gpu_learned_model_beats_static_binning():
    """What the trained network gains over the method without it: the static length-based cost
    estimate (nodes x shorter read), on a C3-generator instance the model was not trained on.
    Ranking quality of predicted DP cells, and the makespan of longest-first orders over 4096
    workers evaluated on the true cells."""
    from canu_amd import learned, oic
    ftr, ttr = learned.collect_traces("c3", 3_000_000, seed=4, max_traces=600_000)
    fte, tte = learned.collect_traces("c3", 2_000_000, seed=9, max_traces=300_000)
    assert len(ftr) > 50_000 and len(fte) > 20_000
    eng = oic.Engine()
    eng.train_model(ftr, ttr, epochs=12)
    pred = eng.predict(fte)[:, 0]
    err = learned.prediction_error(tte, pred, fte)
    ms = learned.ordering_gain(fte, tte, pred, workers=4096, chunk=80_000)
    print(f"\nheld-out: {err}\nmakespans (cells): {ms}")
    assert err["learned_spearman"] > err["static_spearman"] + 0.02
    # With 4096 workers a chunk's makespan is usually its longest job, which both longest-first
    # orders reach: the learned order must stay within 1% of the ideal one.  The strict gain in
    # measured DP time is test_zy_learned_gain's.
    assert ms["learned"] <= ms["static"]
    assert ms["learned"] <= 1.01 * ms["ideal"]
    assert ms["learned"] < ms["plain"]
