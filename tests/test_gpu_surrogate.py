"""GPU checks of the surrogate ensemble kernels (csrc/mlp.hip) through the C ABI.

Gradients and forward passes of the BF16-MFMA kernels are compared with the
CPU oracle (oracle/mlp_oracle.py) evaluated on the same parameters with the
same BF16 rounding points; sums differ only by FP32-vs-FP64 accumulation and
rare BF16 tie flips, so the bars are relative norms of 5e-3 per tensor.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _ctx():
    from tardiis_and_claritas_amd._lib import Context
    return Context(0)


def _flush(msg):
    print(msg, flush=True)


def _tensors(shape, flat):
    """[(name, array)] per member and layer for error reports."""
    out = []
    for m, layers in enumerate(shape.unpack(flat)):
        for l, (W, b) in enumerate(layers):
            out += [(f'm{m} W{l}', W), (f'm{m} b{l}', b)]
    return out


def test_params_roundtrip(gpu_available):
    _flush('surrogate params roundtrip')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=2, hidden=64, n_hidden_layers=2, seed=3)
    p = ens.get_params()
    assert p.size == 2 * mo.Shape(2, 8, 64, 3, 2).per_member
    q = np.random.default_rng(0).normal(0, 0.1, p.size).astype(np.float32)
    ens.set_params(q)
    np.testing.assert_array_equal(ens.get_params(), q)


@pytest.mark.parametrize('n', [100, 256])
def test_gradients_match_oracle(gpu_available, n):
    _flush(f'surrogate gradient check n={n}')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=3, hidden=256, n_hidden_layers=3, seed=11)
    shape = mo.Shape(3, 8, 256, 3, 3)
    g = np.random.default_rng(n)
    x = g.normal(size=(n, 8)).astype(np.float32)
    y = g.normal(size=(n, 3)).astype(np.float32)
    loss, grad = ens.grad(x, y)
    flat = ens.get_params()
    loss_o, grad_o = mo.ensemble_loss_grad(shape, flat, x.astype(np.float64), y.astype(np.float64), bf16=True)
    np.testing.assert_allclose(loss, loss_o, rtol=1e-3)
    for (name, a), (_, b) in zip(_tensors(shape, grad), _tensors(shape, grad_o)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 5e-3, f'{name}: relative gradient error {err:.2e}'


def test_weighted_gradients_match_oracle(gpu_available):
    _flush('surrogate weighted gradient check n=192')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=2, hidden=256, n_hidden_layers=3, seed=13)
    shape = mo.Shape(2, 8, 256, 3, 3)
    g = np.random.default_rng(17)
    x = g.normal(size=(192, 8)).astype(np.float32)
    y = g.normal(size=(192, 3)).astype(np.float32)
    w = g.uniform(0.0, 4.0, size=(192, 3)).astype(np.float32)
    w[::7] = 0.0                                    # rows that carry no weight at all
    loss, grad = ens.grad(x, y, w)
    loss_o, grad_o = mo.ensemble_loss_grad(shape, ens.get_params(), x.astype(np.float64), y.astype(np.float64),
                                           bf16=True, w=w.astype(np.float64))
    np.testing.assert_allclose(loss, loss_o, rtol=1e-3)
    for (name, a), (_, b) in zip(_tensors(shape, grad), _tensors(shape, grad_o)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 5e-3, f'{name}: relative gradient error {err:.2e}'
    # unit weights give the unweighted gradient exactly (same kernels, same order)
    l1, g1 = ens.grad(x, y, np.ones_like(w))
    l0, g0 = ens.grad(x, y)
    np.testing.assert_array_equal(g1, g0)


def test_forward_matches_oracle(gpu_available):
    _flush('surrogate forward check')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=5, hidden=256, n_hidden_layers=3, seed=5)
    shape = mo.Shape(5, 8, 256, 3, 3)
    x = np.random.default_rng(9).normal(size=(1000, 8)).astype(np.float32)
    mean, std = ens.predict(x)
    mean_o, std_o = mo.ensemble_predict(shape, ens.get_params(), x.astype(np.float64), bf16=True)
    scale = np.abs(mean_o).max()
    assert np.abs(mean - mean_o).max() <= 2e-3 * scale
    assert np.abs(std - std_o).max() <= 2e-3 * scale


def test_training_fits_smooth_function(gpu_available):
    _flush('surrogate training on a smooth function')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=5, hidden=256, n_hidden_layers=3, seed=1)
    g = np.random.default_rng(2)
    x = g.uniform(-1, 1, size=(16384, 8)).astype(np.float32)

    def f(x):
        return np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.exp(0.5 * x[:, 3]) - x[:, 4] ** 2,
                         np.tanh(x[:, 5] + x[:, 6] - x[:, 7])], axis=1).astype(np.float32)

    y = f(x)
    first = ens.train(x, y, batch=512, epochs=1, lr=2e-3, seed=0)
    ens.train(x, y, batch=512, epochs=20, lr=1e-3, seed=1)
    last = ens.train(x, y, batch=512, epochs=10, lr=2e-4, seed=2)
    assert np.all(last < 0.02 * first), (first, last)
    xt = g.uniform(-1, 1, size=(2048, 8)).astype(np.float32)
    mean, std = ens.predict(xt)
    rmse = np.sqrt(np.mean((mean - f(xt)) ** 2))
    assert rmse < 0.1 * np.std(f(xt)), rmse
    # spread of independently initialised members is small where data is dense
    assert np.median(std) < 0.1 * np.std(f(xt))


def test_inference_timing(gpu_available):
    _flush('surrogate inference timing')
    from tardiis_and_claritas_amd.surrogate import MlpEnsemble
    ens = MlpEnsemble(_ctx(), n_members=5, hidden=256, n_hidden_layers=3, seed=0)
    ms = ens.time_infer(65536, 5)
    assert 0.0 < ms < 1000.0


def test_surrogate_gain_over_monte_carlo(gpu_available):
    """What the trained network gains over batched Monte Carlo, the method
    without it.  A 5-member ensemble is trained on the GPU on the committed
    config-3 training fixture (8192 configurations, labels from batched Monte
    Carlo at ~150 expected 90-degree window counts) and scored on the 256
    spot-check configurations (labels at ~4000 counts, 0.3 % noise):

    - at equal cost (the Monte Carlo photons one surrogate query's time buys,
      from the live inference rate and the live batched Monte Carlo rate on
      the same configurations) Monte Carlo's NTU error is many times the
      network's;
    - Monte Carlo needs more than 1000x the network's time per query to
      match the network's own rms error;
    - the members' spread ranks the errors: the 10 % most uncertain queries
      err more than the rest, which is what the Monte Carlo fallback uses."""
    _flush('surrogate gain over Monte Carlo (config-3 fixture)')
    from pathlib import Path
    from tardiis_and_claritas_amd import surrogate as sg
    from tardiis_and_claritas_amd.batched import BatchedRunner
    d = np.load(Path(__file__).parent / 'golden' / 'surrogate_config3.npz')
    keys = sg.OPTICAL + sg.TRANSPORT
    P_tr = {k: d[f'train_{k}'] for k in keys}
    P_sp = {k: d[f'spot_{k}'] for k in keys}
    S_tr, n_tr, S_sp, n_sp = d['train_S'], d['train_hist'], d['spot_S'], d['spot_hist']
    runner = BatchedRunner(0)
    sur = sg.Surrogate(runner.ctx, seed=7)
    keep = np.all(S_tr > 0, axis=1)
    sur.fit(sg.features(P_tr)[keep], S_tr[keep], epochs=120, batch=256, lr=1e-3, seed=7, n_hist=n_tr[keep])
    pred, spread = sur.predict(sg.features(P_sp))
    ok = np.all(S_sp > 0, axis=1)
    err = sg.ntu_relative_error(pred[ok], S_sp[ok])
    q_s = sur.ens.time_infer(65536, 10) * 1e-3 / 65536
    # live batched Monte Carlo rate on the spot configurations at ~1e6 histories each
    specs_n = np.full(len(S_sp), 1_000_000, dtype=np.int64)
    sg.monte_carlo(P_sp, specs_n, 11, runner=runner)
    t0 = __import__('time').perf_counter()
    sg.monte_carlo(P_sp, specs_n, 12, runner=runner)
    rate = specs_n.sum() / (__import__('time').perf_counter() - t0)
    n_eq = max(int(q_s * rate), 1)
    S_mc = sg.monte_carlo(P_sp, n_eq, 13, runner=runner)
    mc_ok = np.all(S_mc[ok][:, :2] > 0, axis=1)
    err_mc = np.ones(int(ok.sum()))
    err_mc[mc_ok] = sg.ntu_relative_error(S_mc[ok][mc_ok], S_sp[ok][mc_ok])
    m = sg.set_metrics(pred, S_sp, n_sp)
    t_match = np.median(sg.mc_matched_photons(pred[ok], m['surrogate_rms_est'])) / rate
    print(f'network: median {np.median(err):.4f} p90 {np.percentile(err, 90):.4f} rms {m["surrogate_rms_est"]:.4f}; '
          f'{1 / q_s:.3g} queries/s; Monte Carlo {rate:.3g} histories/s, {n_eq} histories per query at equal '
          f'cost: median {np.median(err_mc):.3f}; matched-error Monte Carlo {t_match / q_s:.3g}x the query time',
          flush=True)
    assert np.median(err) < 0.05
    assert np.median(err_mc) > 5 * np.median(err)
    assert t_match / q_s > 1000
    top = spread[ok, 0] >= np.quantile(spread[ok, 0], 0.9)
    assert err[top].mean() > err[~top].mean()


def test_committed_weights_meet_one_percent(gpu_available):
    """The committed ensemble on the committed spot-check sets: the network's
    own rms error on the NTU reading S90/S0 (the labels' Monte Carlo variance
    subtracted from the mean squared log error) is within 1 %, and so is the
    median error against the labels."""
    _flush('committed surrogate weights on the spot-check sets')
    from pathlib import Path
    from tardiis_and_claritas_amd import surrogate as sg
    d = Path(sg.__file__).resolve().parent
    sur = sg.Surrogate.load(_ctx(), d / 'surrogate_weights.npz')
    for name, bar in (('surrogate_spotcheck.npz', 0.01), ('surrogate_heldout.npz', 0.01)):
        P, S, n = sg.load_spot(d / name)
        pred, spread = sur.predict(sg.features(P))
        m = sg.set_metrics(pred, S, n)
        _flush(f'{name}: median {m["ntu_error_median"]:.4f} rms vs labels {m["rms_vs_labels"]:.4f} '
               f'label rms {m["label_rms"]:.4f} network rms {m["surrogate_rms_est"]:.4f}')
        assert m['n'] == 1024 and m['n_without_counts'] == 0
        assert m['surrogate_rms_est'] <= bar
        assert m['ntu_error_median'] <= bar
        assert np.all(np.isfinite(spread)) and np.all(spread >= 0.0)


def test_query_falls_back_to_monte_carlo(gpu_available):
    """Surrogate.query answers from the network where the ensemble agrees and
    by batched Monte Carlo (histories sized for 1 % on S90/S0) where it does
    not: unflagged answers equal predict(), flagged ones are Monte Carlo
    estimates that agree with the spot-check labels within their combined
    counting error."""
    _flush('surrogate query with Monte Carlo fallback')
    from pathlib import Path
    from tardiis_and_claritas_amd import surrogate as sg
    from tardiis_and_claritas_amd.batched import BatchedRunner
    d = Path(sg.__file__).resolve().parent
    runner = BatchedRunner(0)
    sur = sg.Surrogate.load(runner.ctx, d / 'surrogate_weights.npz')
    P, S_ref, n_ref = sg.load_spot(d / 'surrogate_spotcheck.npz')
    sel = np.arange(64)
    P = {k: np.asarray(v)[sel] for k, v in P.items()}
    S_ref, n_ref = S_ref[sel], n_ref[sel]
    pred, spread = sur.predict(sg.features(P))
    thr = float(np.sort(spread[:, 0])[-5])          # the four most uncertain queries fall back
    S, mc = sur.query(P, threshold=thr, seed=5, runner=runner)
    assert int(mc.sum()) == 4
    np.testing.assert_array_equal(S[~mc], pred[~mc])
    n_mc = np.clip(sg.mc_matched_photons(pred[mc], 0.01), 200_000, 2_000_000_000)
    # relative standard error of the ratio S90/S0 from both Monte Carlo runs
    se = np.sqrt(sg.label_variance(S_ref[mc], n_ref[mc])[:, 0] + sg.label_variance(S[mc], n_mc)[:, 0])
    d_log = np.log((S[mc, 0] / S[mc, 1]) / (S_ref[mc, 0] / S_ref[mc, 1]))
    _flush(f'fallback ratio errors {np.round(d_log, 4)} against standard errors {np.round(se, 4)}')
    assert np.all(np.abs(d_log) < 4.0 * se)
    assert sg.default_fallback_threshold() > 0.0
