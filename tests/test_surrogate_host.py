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
