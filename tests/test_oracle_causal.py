"""Pins of the oracle's causal path provenance (orc_run_causal) and of the
host reduction into the reference's detector-path / causal / H170
diagnostics (no GPU); CLARITAS.py:2497-2599 (per-encounter record),
4629-5103 (per-history state), 8017-8480 and 8967-9440 (reductions).

- smooth-sphere ray optics fixes the record of a finite-particle turn: an
  external reflection at incidence i turns the ray by 180 - 2i and leaves at
  i; a transmission turns it by 2(i - t) (sin t = n_w sin i / n_p) with an
  internal path of 2 r cos t; a path with k internal reflections turns it by
  2(i - t) + k(180 - 2t) (folded into [0, 180]) with k + 1 chords, and the
  exit angle equals i; every smooth walk has microfacet order 1;
- per-history identities: events = geometric + wave, the deflection
  classes sum to the events, a one-event history has equal first, last,
  largest and cumulative turns (squared turn = cumulative square) and one
  site, first backscatter is a first turn beyond 90 deg, the plain-Mie law
  has only coherent events (no interface angles or paths);
- the reduction: every detected history has a record (the record score
  equals the tallies' hardware symmetry score), every exclusive partition
  recovers its parent (the reference's own closure checks), and the
  vectorised tables equal a per-record loop on a sample.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tardiis_and_claritas_amd import causal, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel


def _run(kw, mat, conc, n=20_000, seed=11):
    m = ClaritasForwardModel(**kw)
    om = po.OracleModel(m, mat, conc)
    T, rec = om.run_causal(n, seed)
    nb = len(m.resolve_material(mat).diameter_m)
    return m, T, rec, nb


def _geometric_cell():
    """Config-1 cell, geometric-optics law, smooth 40 um spheres at 5 g/L."""
    kw, _, _, _, _ = workloads.config1('geometric')
    kw.update(particle_surface_rms_slope_deg=0.0, wave_surface_rms_height_m=0.0)
    return kw, workloads.monodisperse(40e-6, 1.46, 0.0), 5.0


def test_smooth_sphere_turn_records_obey_ray_optics():
    kw, mat, conc = _geometric_cell()
    m, T, rec, nb = _run(kw, mat, conc, n=60_000)
    r = rec[rec['ic'] > 0]
    assert len(r) > 500
    nw, npart = float(m.n_water), 1.46
    inc = np.radians(r['max_incidence'].astype(np.float64))
    t = np.arcsin(np.clip(np.sin(inc) * nw / npart, -1, 1))
    k = r['max_macro_bounces'].astype(np.float64)
    typ, defl = r['max_type'], r['max_deflection'].astype(np.float64)
    assert set(np.unique(typ)) <= {1, 2, 3}
    assert np.all(r['max_micro_order'] == 1)
    # exit angle equals the incidence angle for all three (to FP32 acos)
    np.testing.assert_allclose(r['max_exit'], r['max_incidence'], atol=0.05)
    ext = typ == 1
    assert ext.sum() > 10
    np.testing.assert_allclose(defl[ext], 180.0 - 2.0 * np.degrees(inc[ext]), atol=0.05)
    assert np.all(r['max_macro_bounces'][ext] == 0) and np.all(r['max_internal_path'][ext] == 0.0)
    tr = typ == 2
    assert tr.sum() > 100 and np.all(k[tr] == 0)
    np.testing.assert_allclose(defl[tr], np.degrees(2 * (inc[tr] - t[tr])), atol=0.05)
    np.testing.assert_allclose(r['max_path_over_diameter'][tr], np.cos(t[tr]), atol=2e-4)
    ir = typ == 3
    assert ir.sum() > 3 and np.all(k[ir] >= 1)
    D = np.degrees(2 * (inc[ir] - t[ir])) + k[ir] * (180.0 - 2 * np.degrees(t[ir]))
    folded = np.degrees(np.arccos(np.cos(np.radians(D))))
    np.testing.assert_allclose(defl[ir], folded, atol=0.1)
    np.testing.assert_allclose(r['max_path_over_diameter'][ir], (k[ir] + 1) * np.cos(t[ir]), rtol=1e-3)
    np.testing.assert_allclose(r['max_internal_path'], r['max_path_over_diameter'] * 40e-6, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize('law', ['plain_mie', 'hybrid_mie_exact_rate'])
def test_history_identities(law):
    kw, mat, conc, _, _ = workloads.config1(law)
    m, T, rec, nb = _run(kw, mat, conc, n=30_000)
    ic = rec['ic']
    assert np.all(ic == rec['gic'] + rec['wic'])
    assert np.all(rec['small_deflections'] + rec['medium_deflections'] + rec['large_deflections'] == ic)
    z = rec[ic == 0]
    assert len(z) > 0
    for k in ('cumulative_deflection', 'max_deflection', 'mean_r', 'min_r', 'water_path_first', 'max_incidence',
              'first_deflection', 'water_path_after_last'):
        assert np.all(z[k] == -1.0), k
    assert np.all(z['max_bin'] == -1) and np.all(z['first_type'] == 0) and np.all(z['turn_x'] == 0.0)
    s = rec[ic == 1]
    assert len(s) > 50
    for a, b in (('first_deflection', 'last_deflection'), ('first_deflection', 'max_deflection'),
                 ('first_deflection', 'cumulative_deflection'), ('first_x', 'last_x'), ('first_z', 'turn_z'),
                 ('min_r', 'max_r'), ('mean_z', 'first_z'), ('water_path_first', 'water_path_last')):
        np.testing.assert_array_equal(s[a], s[b], err_msg=f'{a} vs {b}')
    np.testing.assert_allclose(s['cumulative_deflection_sq'], s['max_deflection'].astype(float) ** 2, rtol=1e-6)
    np.testing.assert_array_equal(s['first_type'], s['max_type'])
    np.testing.assert_allclose(s['mean_r'], np.hypot(s['first_x'], s['first_y']), rtol=1e-6, atol=1e-9)
    p = rec[ic > 0]
    np.testing.assert_array_equal(p['first_backscatter'], (p['first_deflection'] > 90.0).astype(np.int32))
    assert np.all(p['min_r'] <= p['mean_r'] * (1 + 1e-6)) and np.all(p['mean_r'] <= p['max_r'] * (1 + 1e-6))
    assert np.all(p['water_path_first'] <= p['water_path_last']) and np.all(p['water_path_after_last'] >= 0.0)
    assert np.all((p['max_order'] >= 1) & (p['max_order'] <= p['ic']))
    assert np.all(p['max_deflection'] >= p['last_deflection']) and np.all(p['max_deflection'] >= p['first_deflection'])
    if law == 'plain_mie':
        assert np.all(p['max_type'] == 4) and np.all(p['max_incidence'] == -1.0) and np.all(p['internal_path'] == 0)
    else:
        assert set(np.unique(p['max_type'])) <= {1, 2, 3, 4} and np.any(p['max_type'] != 4)
        g = p[(p['ic'] == 1) & (p['gic'] == 1)]
        assert np.all(g['first_geometric_bin'] == g['first_bin']) and np.all(g['first_wave_bin'] == -1)


def test_reduction_closure_and_record_completeness():
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d')
    m, T, rec, nb = _run(kw, mat, conc, n=40_000)
    r, z = causal.spatial_edges(m)
    P = causal.PathTally(nb, True, r, z)
    P.add(rec[: len(rec) // 2])
    P.add(rec[len(rec) // 2:])
    det = T['det'].sum(axis=0)
    hw = 0.5 * (det[0] + det[1])
    d = causal.diagnostics_dict(P, hw, 'closure', {}, 16384)  # raises on a failed partition
    np.testing.assert_array_equal(P.get('hardware', (18,)), hw)
    np.testing.assert_array_equal(d['any_particle_scores'] + d['ballistic_scores'], hw)
    np.testing.assert_allclose(d['interaction_order_partition_closure_error_scores'], 0.0, atol=1e-12)
    assert d['H170_last_event_angle_histogram_scores_by_event_code'][0].sum() == pytest.approx(
        d['H170_last_event_angle_histogram_scores_by_event_code'][1:].sum())
    # a planted mistake in one record breaks a closure
    bad = rec.copy()
    i = int(np.flatnonzero(bad['ic'] > 0)[0])
    bad['max_order'][i] = 0
    P2 = causal.PathTally(nb, True, r, z)
    P2.add(bad)
    with pytest.raises(RuntimeError, match='closure'):
        causal.diagnostics_dict(P2, hw, 'closure', {}, 16384)


def test_tables_match_a_per_record_loop():
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate')
    m, T, rec, nb = _run(kw, mat, conc, n=8_000)
    r, z = causal.spatial_edges(m)
    P = causal.PathTally(nb, True, r, z)
    P.add(rec)
    order = np.zeros((6, 18))
    ball = np.zeros(18)
    first_map = np.zeros((5, 5))
    for q in rec:
        for ch in (int(q['native_id']), int(q['mirror_id'])):
            if not 0 <= ch < 18:
                continue
            order[min(int(q['ic']), 5), ch] += 0.5
            if q['ic'] == 0:
                ball[ch] += 0.5
            if ch == 17 and q['last_type'] > 0:
                fr = math.sqrt(max(float(q['first_x']) ** 2 + float(q['first_y']) ** 2, 0.0))
                ri = [j for j in range(5) if r[j] <= fr < r[j + 1]]
                zi = [j for j in range(5) if z[j] <= float(q['first_z']) < z[j + 1]]
                if ri and zi:
                    first_map[ri[0], zi[0]] += 0.5
    np.testing.assert_allclose(P.get('exact_particle_order', (6, 18)), order)
    np.testing.assert_allclose(P.get('path:ballistic', (18,)), ball)
    np.testing.assert_allclose(P.get('h170_first_map', (25,)).reshape(5, 5), first_map, atol=1e-6)


def test_causal_run_keeps_the_tallies():
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate')
    m = ClaritasForwardModel(**kw)
    om = po.OracleModel(m, mat, conc)
    Ta, _ = om.run_causal(10_000, 3)
    Tb, _ = om.run(10_000, 3, 0)
    for k in ('det', 'status', 'counters'):
        np.testing.assert_array_equal(Ta[k], Tb[k])


@pytest.mark.parametrize('case', ['plain', 'hybrid', 'rough'])
def test_native_binning_equals_numpy(case):
    """clr_causal_bin (one pass in the library) equals the numpy reduction on
    the same records, table by table (only summation order differs)."""
    if case == 'plain':
        kw, mat, conc, _, _ = workloads.config1('plain_mie')
    elif case == 'hybrid':
        kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d')
    else:
        kw, mat, conc, _, _ = workloads.config1_rough()
    m, T, rec, nb = _run(kw, mat, conc, n=20_000)
    r, z = causal.spatial_edges(m)
    for sym in (True, False):
        A = causal.PathTally(nb, sym, r, z, engine='native')
        B = causal.PathTally(nb, sym, r, z, engine='numpy')
        A.add(rec[:5000]); A.add(rec[5000:])
        B.add(rec[:5000]); B.add(rec[5000:])
        assert set(A.t) == set(B.t), set(A.t) ^ set(B.t)
        for k in B.t:
            assert A.t[k].shape == B.t[k].shape, k
            np.testing.assert_allclose(A.t[k], B.t[k], rtol=1e-12, atol=1e-12, err_msg=k)


def test_conditional_sampler_counts():
    """All ten conditional_sampler_counts (CLARITAS.py:9119-9125): every
    acceptance is counted once by channel and once by outcome, so the two
    families have equal sums; a run's counts are the sum of its halves'
    (they are per-history); the unconditional laws count nothing; the
    closure's own counts follow its probe outcomes."""
    kw, mat, conc, n, seed = workloads.config1('hybrid_mie_exact_rate')
    m = ClaritasForwardModel(**kw)
    om = po.OracleModel(m, mat, conc)
    T, _ = om.run(8000, seed, 0)
    c = causal.conditional_counts(T['counters'])
    assert list(c) == list(causal.CONDITIONAL_COUNT_NAMES)
    by_channel = c['accepted_geometric'] + c['accepted_coherent'] + c['accepted_morphology']
    assert by_channel == c['accepted_absorptions'] + c['accepted_scatterings'] > 1000
    # rejections and inaccessible placements are trials that were not accepted
    assert c['candidate_trials'] >= by_channel + c['outcome_rejections'] + c['inaccessible_trials']
    A, _ = om.run(3000, seed, 0)
    B, _ = om.run(5000, seed, 3000)
    cab = causal.conditional_counts(A['counters'] + B['counters'])
    assert cab == c
    # the spheroid-morphology channel appears with a coherent fraction below 1
    kw2, mat2, conc2, _, seed2 = workloads.config1_rough()
    m2 = ClaritasForwardModel(**kw2)
    c2 = causal.conditional_counts(po.OracleModel(m2, mat2, conc2).run(4000, seed2, 0)[0]['counters'])
    assert c2['accepted_morphology'] > 0 and c2['accepted_geometric'] > 0
    assert (c2['accepted_geometric'] + c2['accepted_coherent'] + c2['accepted_morphology'] ==
            c2['accepted_absorptions'] + c2['accepted_scatterings'])
    kw3, mat3, conc3, _, seed3 = workloads.config1('plain_mie')
    m3 = ClaritasForwardModel(**kw3)
    c3 = causal.conditional_counts(po.OracleModel(m3, mat3, conc3).run(2000, seed3, 0)[0]['counters'])
    assert all(v == 0 for v in c3.values())
