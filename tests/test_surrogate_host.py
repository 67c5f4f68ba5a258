"""Host-side pieces of the surrogate pipeline (no GPU): label variances and
weights, photon budgets, the NTU error and the matched Monte Carlo cost."""
import numpy as np

from tardiis_and_claritas_amd import surrogate as sg


def test_label_variance_matches_poisson_simulation():
    # delta-method variances of the log targets against simulated Poisson counts
    g = np.random.default_rng(0)
    mean = np.array([400.0, 90000.0, 900.0])         # expected counts: S90, S0, S135
    n_hist = 1e7
    c = g.poisson(mean, size=(200000, 3)).astype(np.float64)
    Y = sg.Surrogate.targets(c / n_hist)
    var_sim = Y.var(0)
    var = sg.label_variance((mean / n_hist)[None, :], [n_hist])[0]
    np.testing.assert_allclose(var_sim, var, rtol=0.03)


def test_label_weights_scale_and_floor():
    S = np.array([[1e-5, 1e-2, 1e-5], [4e-5, 1e-2, 4e-5], [1e-3, 1e-1, 1e-3]])
    n = np.array([1e7, 1e7, 1e9])
    w = sg.label_weights(S, n, rel_floor=0.0)
    np.testing.assert_allclose(w.mean(0), 1.0, rtol=1e-6)
    v = sg.label_variance(S, n)
    np.testing.assert_allclose(w[0] / w[1], v[1] / v[0], rtol=1e-6)
    # a floor bounds the ratio of the largest to the smallest weight
    wf = sg.label_weights(S, n, rel_floor=0.01)
    assert wf[:, 0].max() / wf[:, 0].min() < w[:, 0].max() / w[:, 0].min()
    assert np.all(wf[:, 0].max() / wf[:, 0].min() <= (v[:, 0].max() + 1e-4) / 1e-4)


def test_ntu_error_and_matched_cost():
    S_ref = np.array([[2e-5, 1e-2, 3e-5]])
    S_pred = S_ref * np.array([[1.01, 1.0, 1.0]])
    np.testing.assert_allclose(sg.ntu_relative_error(S_pred, S_ref), [0.01], rtol=1e-12)
    # equal scaling of S90 and S0 leaves the NTU reading unchanged
    np.testing.assert_allclose(sg.ntu_relative_error(S_ref * 3.0, S_ref), [0.0], atol=1e-15)
    # histories for a 1 % relative error: 1/c90 + 1/c0 = 1e-4
    n = sg.mc_matched_photons(S_ref)[0]
    np.testing.assert_allclose(1.0 / (2e-5 * n) + 1.0 / (1e-2 * n), 1e-4, rtol=1e-12)


def test_photon_budget_clips():
    b = sg.photon_budget(np.array([1.0, 1000.0, 1e7]), counts=100.0)
    assert b[0] == 200_000_000 and b[2] == 200_000
    assert b[1] == int(100.0 / (1e-7 * 1000.0))


def test_adaptive_budget_targets_counts_from_pilot():
    # pilot yields 1e-3, 1e-6 and none (half a count of 1e6 -> 5e-7), and one
    # above the counts/lo ratio: budgets counts/S90, clipped to [lo, hi]
    b = sg.adaptive_budget(np.array([1e-3, 1e-6, 0.0, 0.5]), 1_000_000, 20000.0, lo=200_000, hi=400_000_000)
    assert b[0] == 20_000_000
    assert b[1] == 400_000_000            # 2e10 clipped to hi
    assert b[2] == 400_000_000            # 4e10 clipped to hi
    assert b[3] == 200_000                # 4e4 clipped to lo
    # the expected 90 deg count of an unclipped budget is the target
    s90 = np.array([3e-4, 2e-5])
    bb = sg.adaptive_budget(s90, 1_000_000, 5000.0)
    np.testing.assert_allclose(bb * s90, 5000.0, rtol=1e-6)


def test_ratio_budget_sets_the_ntu_label_variance():
    # with the 0 deg yield as well, the budget makes the binomial relative
    # variance of S90/S0, 1/(n S90) + 1/(n S0), equal to 1/counts
    s90, s0 = np.array([3e-4, 2e-5, 1e-3]), np.array([0.4, 1e-5, 2e-3])
    b = sg.adaptive_budget(s90, 1_000_000, 1000.0, s0_pilot=s0)
    np.testing.assert_allclose(1.0 / (b * s90) + 1.0 / (b * s0), 1e-3, rtol=1e-5)
    # an extinct transmitted beam (no pilot count) gets half a count of 1e6
    b0 = sg.adaptive_budget(np.array([1e-4]), 1_000_000, 100.0, hi=10**12, s0_pilot=np.array([0.0]))
    assert b0[0] == int(100.0 * (1e4 + 2e6))
    # the ratio budget is never below the 90 deg one
    assert np.all(b >= sg.adaptive_budget(s90, 1_000_000, 1000.0))


def test_set_metrics_recovers_network_error_under_label_noise():
    # predictions off by a known 2 % rms in S90/S0, labels with Poisson noise of
    # a comparable size: the label variance is subtracted from mean d^2
    g = np.random.default_rng(1)
    n = 40000
    true = np.stack([g.uniform(1e-5, 1e-4, n), g.uniform(1e-2, 5e-2, n), g.uniform(1e-5, 1e-4, n)], axis=1)
    n_hist = np.full(n, 5e6)
    labels = g.poisson(true * n_hist[:, None]) / n_hist[:, None]
    pred = true.copy()
    pred[:, 0] *= np.exp(g.normal(0.0, 0.02, n))
    m = sg.set_metrics(pred, labels, n_hist)
    assert m['n'] == n
    assert abs(m['surrogate_rms_est'] - 0.02) < 0.002
    assert m['rms_vs_labels'] > m['surrogate_rms_est']
    # exact predictions: the estimate is at the sampling noise of the subtraction
    assert sg.set_metrics(true, labels, n_hist)['surrogate_rms_est'] < 0.006


def test_spread_calibration_bins_and_matched_speedup():
    g = np.random.default_rng(2)
    n = 20000
    true = np.stack([np.full(n, 5e-5), np.full(n, 2e-2), np.full(n, 5e-5)], axis=1)
    n_hist = np.full(n, 1e12)                      # labels practically exact
    spread = np.zeros((n, 3))
    spread[:, 0] = g.uniform(0.0, 0.1, n)
    pred = true.copy()
    pred[:, 0] *= np.exp(g.normal(0.0, 1.0, n) * spread[:, 0])   # error grows with the spread
    rows = sg.spread_calibration(pred, spread, true, n_hist)
    assert len(rows) == 10 and sum(r['n'] for r in rows) == n
    est = [r['rms_est'] for r in rows]
    assert est[-1] > 5 * est[0]
    mid = [0.5 * (r['spread_lo'] + r['spread_hi']) for r in rows]
    np.testing.assert_allclose(est[2:], np.asarray(mid[2:]) * 1.0, rtol=0.15)
    s = sg.matched_speedup(true, 0.02, 1e-8, 1e9)
    # 1 % needs 4x the histories of 2 %
    np.testing.assert_allclose(s['speedup_at_1pct'] / s['speedup_at_matched_rms'], 4.0, rtol=1e-9)
    assert not s['meets_1pct'] and sg.matched_speedup(true, 0.009, 1e-8, 1e9)['meets_1pct']


def test_committed_spot_check_set():
    # the spot-check set the bench scores: 1,024 configurations of the
    # config-3 space, each with its own optical draw, whose S90/S0 labels
    # carry under 1 % relative Monte Carlo error
    from pathlib import Path
    P, S, n = sg.load_spot(Path(sg.__file__).resolve().parent / 'surrogate_spotcheck.npz')
    assert len(S) == 1024 and np.all(S > 0)
    assert len({tuple(P[k][i] for k in sg.OPTICAL) for i in range(len(S))}) == 1024
    v = sg.label_variance(S, n)[:, 0]
    assert np.sqrt(np.mean(v)) < 0.01
    assert np.all((P['conc_mg_l'] >= 10.0) & (P['conc_mg_l'] <= 10000.0))
    assert np.all((P['median_um'] >= 2.0) & (P['median_um'] <= 60.0))
    assert set(np.unique(P['wavelength_nm'])) <= {530.0, 860.0}
