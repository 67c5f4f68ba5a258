"""GPU parity of the reflecting collimator bores, ring-face cavity and SFH213
photodiode weights (CLARITAS.py:1263-1374, 4283-4376) against the oracle.

The device draws the collimator and scoring random numbers from each
history's own stream in the oracle's order, so histories agree one for one:
>= 99.9 % identical status, detector ids and draw counts; hit, stray,
face-stray and photodiode tallies within 0.2 % (+ a few counts); source
collimator counters within 0.2 %.  A larger run agrees within 4 combined
standard errors per channel, and simulate() reports consistent fields.
"""
import sys

import numpy as np
import pytest

from tardiis_and_claritas_amd import abi, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel

pytestmark = pytest.mark.gpu

ALL_ON = dict(source_bore_surface_model='smooth_fresnel_nylon', detector_bore_surface_model='smooth_fresnel_nylon',
              ring_inner_face_surface_model='smooth_fresnel_nylon', ring_inner_face_refractive_index=4.0,
              bore_refractive_index=2.0, photodiode_response_model='sfh213_datasheet_fit')


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _default_cell(law, **over):
    kw = dict(particle_event_model=law, free_surface_model='flat', sediment_transport_model='uniform',
              particle_surface_rms_slope_deg=0.0, wave_surface_rms_height_m=0.0, mie_phase_grid_size=1800)
    kw.update(over)
    return kw, workloads.monodisperse(10e-6, 1.46, 1e-4), 2.0


def _close(a, b, rel=2e-3, extra=4):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.abs(a - b).sum() <= rel * max(np.abs(b).sum(), 1.0) + extra, (a, b)


@pytest.mark.parametrize('law', ['plain_mie', 'hybrid_mie_exact_rate'])
def test_history_by_history_reflecting_models(gpu_available, oracle_lib, law):
    _flush(f'reflecting bores history-by-history {law}')
    kw, mat, conc = _default_cell(law, **ALL_ON)
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 40_000
    Tg, rg, _ = m.run_tallies(N, 77, 0, n_records=N)
    To, ro = oracle_lib.OracleModel(m, mat, conc).run(N, 77, 0, n_records=N)
    same = (rg['status'] == ro['status']) & (rg['native_id'] == ro['native_id']) & \
        (rg['mirror_id'] == ro['mirror_id']) & (rg['draws'] == ro['draws'])
    assert same.mean() >= 0.999, same.mean()
    for k in ('stray', 'face_stray'):
        _close(Tg[k], To[k])
    _close(Tg['det'].sum(0)[:2], To['det'].sum(0)[:2])
    _close(Tg['photodiode'].sum(0), To['photodiode'].sum(0))
    c = abi.C_IDX
    for k in ('src_bore_interactions', 'src_bore_reflections', 'src_reflected_accept'):
        _close([Tg['counters'][c[k]]], [To['counters'][c[k]]])
    assert Tg['counters'][c['src_bore_reflections']] > 0


def test_reflecting_models_within_4_sigma(gpu_available, oracle_lib):
    _flush('reflecting bores 4-sigma config1')
    kw, mat, conc, _, _ = workloads.config1('hybrid_mie_exact_rate', source='beta_radiance_3d')
    kw.update(ALL_ON)
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 300_000
    Tg, _, _ = m.run_tallies(N, 101, 0)
    To, _ = oracle_lib.OracleModel(m, mat, conc).run(N, 202, 0)
    for a, b in ((Tg['det'].sum(0)[:2], To['det'].sum(0)[:2]), (Tg['stray'], To['stray']),
                 (Tg['face_stray'], To['face_stray'])):
        z = np.abs(a - b) / np.sqrt(np.maximum(a + b, 1.0))
        assert np.all(z < 4.0), z.max()
    for k in range(1, 9):
        a, b = Tg['status'][k], To['status'][k]
        assert abs(a - b) <= 4 * np.sqrt(max(a + b, 1)), (k, a, b)


def test_simulate_reports_bore_fields(gpu_available):
    _flush('simulate with reflecting bores')
    kw, mat, conc, n, seed = workloads.config1('plain_mie', source='beta_radiance_3d')
    kw.update(ALL_ON)
    m = ClaritasForwardModel(**kw)
    r = m.simulate(mat, conc, n_rays=256_000, seed=seed)
    np.testing.assert_array_equal(r.hardware_native_direct_hits + r.hardware_native_stray_hits,
                                  r.hardware_native_hits)
    assert np.all(r.hardware_native_ring_face_stray_hits <= r.hardware_native_stray_hits)
    assert r.source_collimator_acceptance_fraction == pytest.approx(
        r.source_direct_acceptance_fraction + r.source_reflected_acceptance_fraction)
    assert r.source_bore_reflections_per_ray > 0 and r.source_reflected_acceptance_fraction > 0
    assert np.all(r.photodiode_symmetry_weighted_scores <= r.hardware_symmetry_scores + 1e-9)
    # the selected score is the photodiode-weighted one for the SFH213 model
    np.testing.assert_allclose(r.symmetry_exact_scores, r.photodiode_symmetry_weighted_scores)
    assert r.sfh213_half_angle_deg == 10.0
    rel = r.photodiode_relative_response_by_channel
    assert np.all((rel >= 0) & (rel <= 1))
