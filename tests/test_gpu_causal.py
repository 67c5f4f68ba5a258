"""GPU parity of the causal path provenance (clr_run_causal) against the
oracle (orc_run_causal), history by history, and of the
detector_particle_path_diagnostics of simulate.

Bars: on the same seeded histories >= 99.9 % of the oracle's detected
histories carry a device record with identical integer fields (ids,
counts, event types, bins, orders, classes); their float fields agree as
FP32 on both sides allows (different contraction): turn angles as cosines
within 1e-5 (acos is ill-conditioned near 0 deg), cumulative turns within
0.05 deg per event, sites within 5e-8 m, water paths within 5e-7 m,
particle paths within 2e-4 relative, for >= 99.9 % of the histories and
within 100 times that for all.  Ranges of one run add
up to the whole run exactly; the CAUSAL and LIGHT variants record the same
histories; simulate's diagnostics pass the reference's closure checks and
keep the production tallies.
"""
import sys

import numpy as np
import pytest

from tardiis_and_claritas_amd import abi, causal, light, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel

pytestmark = pytest.mark.gpu

INT_FIELDS = [k for k in abi.CAUSAL_INT_FIELDS]
ANGLES = ('first_deflection', 'last_deflection', 'cumulative_deflection', 'max_deflection', 'max_incidence',
          'max_exit')


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _device_records(m, n, seed, offset=0, ranges=None, light_acc=None):
    recs, total = [], None
    import ctypes as C
    ctx = m.ctx
    ranges = ranges or [(0, n)]
    for r0, rn in ranges:
        buf = np.zeros(rn, dtype=abi.causal_dtype())
        co = abi.CausalOut()
        co.records, co.capacity, co.range_start, co.range_count = buf.ctypes.data, rn, r0, rn
        t = abi.Tallies()
        if light_acc is not None:
            out, arrs = abi.light_out(*light_acc.shape())
            ctx.check(ctx.lib.clr_run_causal(ctx.h, n, seed, offset, m._launch(), C.byref(light_acc.struct),
                                             C.byref(t), C.byref(out), C.byref(co)), 'clr_run_causal')
            light_acc.add(arrs, abi.light_scalars(out))
        else:
            ctx.check(ctx.lib.clr_run_causal(ctx.h, n, seed, offset, m._launch(), None, C.byref(t), None,
                                             C.byref(co)), 'clr_run_causal')
        assert co.count <= rn
        recs.append(buf[:co.count])
        total = causal.sum_tallies(total, abi.tallies_to_numpy(t))
    r = np.concatenate(recs)
    return total, r[np.argsort(r['history'], kind='stable')]


def _compare(g, o, min_frac=0.999):
    common, gi, oi = np.intersect1d(g['history'], o['history'], return_indices=True)
    G, O = g[gi], o[oi]
    same = np.ones(len(common), dtype=bool)
    for k in INT_FIELDS:
        same &= G[k] == O[k]
    frac = same.sum() / max(len(o), 1)
    assert frac >= min_frac, (frac, len(g), len(o),
                              {k: int((G[k] != O[k]).sum()) for k in INT_FIELDS if (G[k] != O[k]).any()})
    G, O = G[same], O[same]
    ic = G['ic'].astype(np.float64)
    bad = {}
    for k in abi.CAUSAL_FLOAT_FIELDS:
        a, b = G[k].astype(np.float64), O[k].astype(np.float64)
        if k in ANGLES and k != 'cumulative_deflection':
            # FP32 acos of a unit-vector dot product: compared as cosines
            # (within 1e-5: ulp differences of the directions, grown over a
            # rough-surface walk), sentinels exactly
            err = np.where((a < 0) != (b < 0), np.inf, np.abs(np.cos(np.radians(a)) - np.cos(np.radians(b))))
            tol = np.full_like(a, 1e-5)
        elif k in ('cumulative_deflection', 'cumulative_deflection_sq'):
            err = np.abs(a - b)  # 0.05 deg per event (and its square's share)
            tol = (0.05 if k == 'cumulative_deflection' else 0.05 * 2 * 180.0) * np.maximum(ic, 1) + 1e-9
        elif k.startswith('water_path'):
            err, tol = np.abs(a - b), 5e-7 + 1e-5 * np.abs(b)  # FP32 sums of flights (and differences)
        elif k in ('internal_path', 'internal_opl', 'max_internal_path', 'max_path_over_diameter'):
            err, tol = np.abs(a - b), 1e-12 + 2e-4 * np.abs(b)  # FP32 chords (spheroid quadratics)
        else:
            err, tol = np.abs(a - b), 5e-8 + 1e-5 * np.abs(b)  # sites: ~10 FP32 ulp at the cell's 5 cm
        # the tight bar for >= 99.9 % of the histories (rounding differences
        # grow chaotically along a few rough-surface paths), 100x for all
        nb = int((err > tol).sum())
        if nb > 1e-3 * len(a) or np.any(err > 100.0 * tol):
            j = int(np.argmax(err - tol))
            bad[k] = (nb, float(a[j]), float(b[j]))
    assert not bad, bad
    return len(common)


def assert_per_event(a, b, ic, per_event, name):
    """|a - b| within per_event for each of the history's events."""
    bad = np.abs(a - b) > per_event * np.maximum(ic, 1) + 1e-9
    assert not bad.any(), (name, a[bad][:5], b[bad][:5], ic[bad][:5])


CASES = {
    'plain_mie': lambda: workloads.config1('plain_mie'),
    'hybrid': lambda: workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d'),
    'rough_spheroid': lambda: workloads.config1_rough('hybrid_mie_exact_rate'),
}


@pytest.mark.parametrize('case', list(CASES))
def test_records_match_oracle_history_by_history(gpu_available, oracle_lib, case):
    _flush(f'causal records vs oracle: {case}')
    kw, mat, conc, _, _ = CASES[case]()
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 40_000
    Tg, g = _device_records(m, N, 77)
    To, o = oracle_lib.OracleModel(m, mat, conc).run_causal(N, 77)
    assert len(o) > 1000
    n = _compare(g, o)
    _flush(f'  {case}: {len(g)} device / {len(o)} oracle records, {n} common')
    assert abs(len(g) - len(o)) <= 1e-3 * len(o) + 2
    assert (g['ic'] > 0).sum() > 50


def test_ranges_add_up_to_the_run(gpu_available):
    _flush('causal ranges vs one run')
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate')
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 30_000
    Ta, ra = _device_records(m, N, 5)
    Tb, rb = _device_records(m, N, 5, ranges=[(0, 7_001), (7_001, 15_000), (22_001, 7_999)])
    for k in ('det', 'group_histories', 'status', 'photodiode'):
        np.testing.assert_array_equal(Ta[k], Tb[k], err_msg=k)
    np.testing.assert_array_equal(ra['history'], rb['history'])
    for k in INT_FIELDS:
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
    # the production kernel traces the same histories
    Tp, _, _ = m.run_tallies(N, 5, 0)
    assert np.abs(Ta['det'] - Tp['det']).sum() <= 1e-3 * N
    assert int(Ta['det'][:, :2].sum()) == int((ra['native_id'] >= 0).sum() + (ra['mirror_id'] >= 0).sum())


def test_light_variant_records_the_same_histories(gpu_available):
    _flush('causal records: LIGHT vs CAUSAL variant')
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate')
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 20_000
    cfgd = {'enabled': True, 'nx': 8, 'ny': 8, 'nz': 8, 'representative_rays': 4, 'max_segments_per_ray': 8}
    lacc = light.LightAccumulator(m, cfgd, N)
    _, rl = _device_records(m, N, 9, light_acc=lacc)
    _, rc = _device_records(m, N, 9)
    _compare(rl, rc)
    assert lacc.scalars['error_count'] == 0 and lacc.scalars['segment_count'] > 0


@pytest.mark.parametrize('builder', ['config1_hybrid', 'config1_rough'])
def test_simulate_path_diagnostics(gpu_available, builder):
    _flush(f'simulate path diagnostics {builder}')
    kw, mat, conc, n, seed = (workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d')
                              if builder == 'config1_hybrid' else workloads.config1_rough())
    m = ClaritasForwardModel(**kw)
    r = m.simulate(mat, conc, n_rays=200_000, seed=seed)
    d = r.detector_particle_path_diagnostics
    assert d is not None
    causal.check_closure(d)
    hw = d['hardware_symmetry_scores']
    np.testing.assert_array_equal(d['any_particle_scores'] + d['ballistic_scores'], hw)
    assert d['any_particle_scores'].sum() > 0
    seen = hw > 0
    assert np.all(np.isfinite(d['mean_particle_interactions_by_channel'][seen]))
    assert np.all((d['any_particle_fraction'] >= 0) & (d['any_particle_fraction'] <= 1))
    # mean interactions >= fraction of histories with any interaction
    assert np.all(d['mean_particle_interactions_by_channel'][seen] >= d['any_particle_fraction'][seen] - 1e-12)
    assert d['conditional_max_trials'] == 16384 and d['conditional_sampler_counts']['candidate_trials'] >= 0
    # result fields of the unaggregated PSD, the morphology share and the vortex geometry
    assert r.aggregation_model == 'none' and r.aggregation_steps == 0 and r.aggregation_projected_area_ratio == 1.0
    assert 0.0 < r.aggregation_largest_bin_mass_fraction <= 1.0
    assert 0.0 <= r.morphology_fraction_of_wave_interactions <= 1.0
    assert r.mean_morphology_wave_interactions == pytest.approx(
        r.mean_wave_interactions - r.mean_coherent_wave_interactions, abs=1e-12)
    assert r.ray_data is None and r.heatmap_xy is None
    r2 = m.simulate(mat, conc, n_rays=200_000, seed=seed, path_diagnostics=False)
    assert r2.detector_particle_path_diagnostics is None
    assert np.abs(r.hardware_native_direct_hits - r2.hardware_native_direct_hits).sum() <= 200


def test_adaptive_simulate_with_all_diagnostics(gpu_available):
    """Adaptive stopping (batches of statistics_batch_rays) with the path
    diagnostics and the heat maps: every batch's records and maps are
    accumulated, and the closures hold over the whole run."""
    _flush('adaptive simulate with path diagnostics and heat maps')
    kw, mat, conc, n, seed = workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d')
    m = ClaritasForwardModel(statistics_batch_rays=50_000, **kw)
    r = m.simulate(mat, conc, n_rays=50_000, seed=seed, target_detector_score=1e12, max_rays=150_000,
                   heatmap_size=32)
    assert r.n_rays == 150_000 and r.stop_reason == 'max_rays_before_target'  # CLARITAS.py:8872-8873
    d = r.detector_particle_path_diagnostics
    causal.check_closure(d)
    np.testing.assert_allclose(d['hardware_symmetry_scores'], r.hardware_symmetry_scores)
    assert r.heatmap_xy.shape == (32, 32) and r.heatmap_xy.sum() > 0
    assert r.production_failure_diagnostics['count_closure_ok']
