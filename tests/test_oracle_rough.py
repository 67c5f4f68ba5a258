"""Pins of the oracle's rough-particle physics (no GPU).

The Smith microsurface walk (CLARITAS.py:1792-2430) and the spheroid wave
channel are restated in oracle/oracle.c.  Each piece is checked against an
independent statement of the mathematics:

- inverse error function: mpmath.erfinv;
- B(x) = 1 - x sqrt(pi) erfcx(x): mpmath at 50 digits (the x >= 5 branch is
  the reference's 7-term asymptotic series, so its bar is the series'
  truncation error);
- Smith Lambda: the textbook Beckmann closed form
  (erf(a) - 1)/2 + exp(-a^2)/(2 a sqrt(pi)), a = cot(theta)/alpha, and the
  signed relation Lambda(-w) = -1 - Lambda(w);
- height step: escape iff U > 1 - C0^Lambda, next height from the inverse
  C' = C0 (1 - U)^(-1/Lambda), in mpmath;
- visible-slope sampler: the sampled slope inverts the visible-slope CDF
  (mpmath) in both the ordinary and the deep-backside branch;
- whole walks: energy conservation for a lossless interface (every walk
  escapes), an index-matched interface is invisible (direction unchanged,
  always transmitted), the alpha -> 0 limit reflects with the smooth
  Fresnel probability, and reflections are symmetric about the plane of
  incidence.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

mp = pytest.importorskip('mpmath')
mp.mp.dps = 50


def test_erfinv_matches_mpmath():
    ys = np.concatenate([np.linspace(-0.999, 0.999, 401), [1e-300, -1e-12, 0.5, 0.9999999, -0.99999999999,
                                                          1 - 2.0 ** -32, -(1 - 2.0 ** -33)]])
    got = po.erfinv(ys)
    for y, g in zip(ys, got):
        ref = float(mp.erfinv(mp.mpf(y)))
        assert abs(g - ref) <= 4e-16 * max(abs(ref), 1e-300), (y, g, ref)


def _B_ref(x):
    x = mp.mpf(x)
    return float(1 - x * mp.sqrt(mp.pi) * mp.exp(x * x) * mp.erfc(x))


def test_beckmann_B():
    xs = np.concatenate([np.linspace(0.0, 4.99, 200), np.linspace(5.0, 60.0, 120)])
    got = po.beckmann_B(xs)
    for x, g in zip(xs, got):
        ref = _B_ref(x)
        # the asymptotic branch truncates after 7 terms: next term 2027025/256 x^-16
        tol = 1e-12 if x < 5.0 else 1.1 * (2027025.0 / 256.0) * x ** -16 / ref + 1e-14
        assert abs(g - ref) <= tol * ref, (x, g, ref)
    assert po.beckmann_B([-1.0])[0] == -1.0


def _lambda_ref(mu, alpha):
    mu, alpha = mp.mpf(mu), mp.mpf(alpha)
    a = abs(mu) / (alpha * mp.sqrt(1 - mu * mu))
    lam = -mp.erfc(a) / 2 + mp.exp(-a * a) / (2 * a * mp.sqrt(mp.pi))    # (erf(a) - 1)/2 + ...
    return float(lam) if mu > 0 else float(-1 - lam)


@pytest.mark.parametrize('alpha', [0.05, 0.4663, 1.3])
def test_smith_lambda_closed_form(alpha):
    alpha = float(np.float32(alpha))
    for th in np.linspace(0.01, math.pi - 0.01, 61):
        if abs(th - math.pi / 2) < 1e-6:
            continue
        w = np.array([math.sin(th), 0.0, math.cos(th)], dtype=np.float32)
        mu = float(w[2])              # the walk takes w as given (unit to float rounding)
        got = po.smith_lambda(w, (0.0, 0.0, 1.0), alpha)
        ref = _lambda_ref(mu, alpha)
        a = abs(mu) / (alpha * math.sqrt(1 - mu * mu))
        tol = 1e-12
        if a >= 5.0:      # B(a) from the 7-term asymptotic series
            tol += 1.1 * (2027025.0 / 256.0) * a ** -16 / _B_ref(a)
        mag = abs(ref) if mu > 0 else abs(ref + 1.0)
        assert abs(got - ref) <= tol * mag + 1e-300, (th, got, ref)
    # signed relation and vertical limits
    w = (0.3, 0.2, 0.9)
    assert abs(po.smith_lambda(tuple(-v for v in w), (0, 0, 1), alpha) + 1 + po.smith_lambda(w, (0, 0, 1), alpha)) < 1e-15
    assert po.smith_lambda((0, 0, 1), (0, 0, 1), alpha) == 0.0
    assert po.smith_lambda((0, 0, -1), (0, 0, 1), alpha) == -1.0


def test_smith_height_step():
    alpha = float(np.float32(0.4663))
    g = np.random.default_rng(3)
    for _ in range(200):
        th = g.uniform(0.05, math.pi - 0.05)
        w = np.array([math.sin(th), 0.0, math.cos(th)], dtype=np.float32)
        h = float(np.float32(g.uniform(-0.99, 0.99)))
        U = g.uniform(1e-9, 1 - 1e-9)
        st, nh, fc = po.smith_height(U, w, (0, 0, 1), alpha, h)
        lam = mp.mpf(po.smith_lambda(w, (0, 0, 1), alpha))
        C0 = (mp.mpf(h) + 1) / 2
        if w[2] > 0:
            p_hit = 1 - C0 ** lam
            if U > p_hit:
                assert st == 2, (th, h, U)
                continue
        assert st == 1 and fc == 0
        Cn = C0 * (1 - mp.mpf(U)) ** (-1 / lam)
        ref = float(2 * Cn - 1)
        assert abs(nh - ref) <= 2e-7 * max(1.0, abs(ref)), (th, h, U, nh, ref)
        if w[2] < 0:
            assert nh <= h + 1e-6          # moving down the surface
        else:
            assert nh >= h - 1e-6


def _vis_cdf(s, si):
    s, si = mp.mpf(s), mp.mpf(si)
    c = 1 / (2 * mp.sqrt(mp.pi))
    q = c * mp.exp(-si * si) + si / 2 * mp.erfc(-si)
    return (c * mp.exp(-s * s) + si / 2 * mp.erfc(-s)) / q


@pytest.mark.parametrize('si', [3.0, 0.7, 0.0, -0.6, -1.0, -1.5, -4.0, -12.0])
def test_visible_slope_sampler_inverts_cdf(si):
    g = np.random.default_rng(7)
    u1 = g.uniform(1e-6, 1 - 1e-6, 40)
    u2 = g.uniform(1e-6, 1 - 1e-6, 40)
    sx, sy, ok = po.beckmann_slopes(np.full(40, si), u1, u2)
    assert np.all(ok == 1)
    for a, b, u, v in zip(sx, sy, u1, u2):
        assert a <= si + 1e-12
        F = _vis_cdf(a, si)
        # bisection to 2^-48 (ordinary) / 2^-52 (tail) of the bracket; compare in u
        dF = abs(float(F) - u)
        slope = float(mp.diff(lambda s: _vis_cdf(s, si), a))
        assert dF <= 1e-9 * max(1.0, slope) + 1e-12, (si, a, u, float(F))
        assert abs(b - float(mp.erfinv(2 * mp.mpf(v) - 1))) <= 1e-14 * max(1.0, abs(b))


def _fresnel_unpolarised(ci, n1, n2):
    s2 = (n1 / n2) ** 2 * (1 - ci * ci)
    if s2 >= 1:
        return 1.0
    ct = math.sqrt(1 - s2)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (rs * rs + rp * rp)


def _rays(n, th, seed):
    g = np.random.default_rng(seed)
    b = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), (n, 1))
    d = np.tile(np.array([math.sin(th), 0.0, -math.cos(th)], dtype=np.float32), (n, 1))
    # rotate every (b, d) pair by a random rotation: the walk is frame independent
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    return (np.einsum('nij,nj->ni', R, b).astype(np.float32), np.einsum('nij,nj->ni', R, d).astype(np.float32), R)


@pytest.mark.parametrize('th', [0.1, 0.8, 1.4])
def test_walk_lossless_always_escapes(th):
    b, d, _ = _rays(20000, th, 1)
    for n_top, n_bot in ((1.33, 1.55), (1.55, 1.33)):
        out, st, top, fc = po.micro_walk(11, 0.4663, b, d, n_top, 0.0, n_bot, 0.0)
        assert np.all(st == 2), (np.unique(st), np.unique(fc))
        assert np.all((top == 0) | (top == 1))
        np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=2e-6)
        # left on the side it says
        side = np.einsum('ni,ni->n', out, b)
        assert np.all(side[top == 1] > 0) and np.all(side[top == 0] < 0)


def test_walk_index_matched_interface_is_invisible():
    b, d, _ = _rays(5000, 0.9, 2)
    out, st, top, fc = po.micro_walk(5, 0.4663, b, d, 1.4, 0.0, 1.4, 0.0)
    assert np.all(st == 2) and np.all(top == 0)
    np.testing.assert_allclose(out, d, atol=3e-6)


@pytest.mark.parametrize('th', [0.3, 1.0])
def test_walk_smooth_limit_is_fresnel(th):
    n = 40000
    b, d, _ = _rays(n, th, 3)
    out, st, top, fc = po.micro_walk(17, 1e-4, b, d, 1.33, 0.0, 1.55, 0.0)
    assert np.all(st == 2)
    R = _fresnel_unpolarised(math.cos(th), 1.33, 1.55)
    frac = top.mean()
    assert abs(frac - R) <= 4 * math.sqrt(R * (1 - R) / n) + 1e-4, (frac, R)
    # reflected rays leave in the mirror direction
    mirror = d - 2 * np.einsum('ni,ni->n', d, b)[:, None] * b
    cos = np.einsum('ni,ni->n', out[top == 1], mirror[top == 1])
    assert np.all(cos > 1 - 1e-6)


def test_walk_reflection_symmetric_about_plane_of_incidence():
    n = 60000
    th = 0.7
    b = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), (n, 1))
    d = np.tile(np.array([math.sin(th), 0.0, -math.cos(th)], dtype=np.float32), (n, 1))
    out, st, top, fc = po.micro_walk(23, 0.4663, b, d, 1.33, 0.0, 1.55, 0.0)
    assert np.all(st == 2)
    for sel in (top == 1, top == 0):
        y = out[sel, 1]
        assert abs(y.mean()) <= 4 * y.std() / math.sqrt(sel.sum())


def _rough_model(**over):
    from tardiis_and_claritas_amd import ClaritasForwardModel, workloads
    kw, mat, conc, n, seed = workloads.config1_rough(**over)
    return ClaritasForwardModel(**kw), mat, conc, seed


@pytest.mark.parametrize('over', [dict(), dict(aspect=1.0), dict(orientation='local_tangential_axial_vmf', kappa=4.0),
                                  dict(event_model='hybrid_mie_excess')])
def test_rough_transport_closure_and_determinism(oracle_lib, over):
    m, mat, conc, seed = _rough_model(**over)
    om = oracle_lib.OracleModel(m, mat, conc)
    N = 20000
    a, _ = om.run(N, seed, 0, threads=1)
    b, _ = om.run(N, seed, 0, threads=4)
    for k in ('det', 'status', 'counters', 'ic_hist', 'failures'):
        assert np.array_equal(a[k], b[k]), k
    assert a['status'].sum() == N
    ci = __import__('tardiis_and_claritas_amd.abi', fromlist=['C_IDX']).C_IDX
    c = a['counters']
    # particle events happened, most wave events are morphology (coherent fraction ~0.02)
    assert c[ci['sum_ic']] > 0.2 * N
    assert c[ci['sum_wic']] - c[ci['sum_wcoh']] > 10 * c[ci['sum_wcoh']]
    # failures (numerical walk failures, watchdogs, spheroid geometry) stay rare
    assert a['failures'].sum() <= 1e-3 * N, a['failures']


def test_rough_transport_offset_additivity(oracle_lib):
    m, mat, conc, seed = _rough_model()
    om = oracle_lib.OracleModel(m, mat, conc)
    x, _ = om.run(6000, seed, 0)
    y, _ = om.run(6000, seed, 6000)
    z, _ = om.run(12000, seed, 0)
    assert np.array_equal(x['status'] + y['status'], z['status'])
    assert np.array_equal(x['counters'] + y['counters'], z['counters'])
