"""GPU parity of the batched (config 3) and full-scale 3-D field (config 5)
workloads against the oracle, and of the reference's analytic sediment
models (CLARITAS.py:1537-1609) history by history.

- Config 3: ten items of the 256-configuration workload (the first six, the
  densest, the densest at 530 nm, the largest size parameter, the sparsest;
  tests/golden/make_config3_fixture.py) run together in one persistent
  batched launch (clr_run_batch) on independent streams; every item's
  channels are compared with the oracle's committed tallies at three
  combined standard errors, family-wise over all items and channels
  (tests/tally_cases.py).
- Config 5: the 128 x 128 x 256 FP32 segregated field at 20 g/L, device
  and oracle on independent streams, same rule.  The device tracks the
  shared field with z-slab majorants and the oracle with one; the tallies'
  law is the same.
- Stokes and Schiller-Naumann drift-diffusion (models 1 and 2) on the
  parabolic and Scully surfaces, including the reference's constructor
  template (hybrid law, rough particles, Scully vortex, Schiller-Naumann
  drift-diffusion, loess PSD): the device and the oracle trace the same
  Philox streams, and >= 99.9 % of histories end identically.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import tally_cases
from tardiis_and_claritas_amd import ClaritasForwardModel, physics, workloads

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / 'golden'


def _flush(msg):
    print(msg, flush=True)


def test_config3_batched_items_vs_oracle(gpu_available):
    from tardiis_and_claritas_amd.batched import BatchedRunner
    fx = json.loads((GOLDEN / 'oracle_config3.json').read_text())['items']
    p = workloads.config3_params()
    idx = sorted(int(k) for k in fx)
    _flush(f'config 3: items {idx} in one batched launch')
    specs = [workloads.config3_item(p, i) for i in idx]
    runner = BatchedRunner()
    runner.prepare(specs)
    budgets = [fx[str(i)]['n'] for i in idx]
    res = runner.run_prepared(budgets, 0, seeds=[fx[str(i)]['key'] + 100_000 for i in idx])
    alpha_item = 1.0 - (1.0 - tally_cases.ALPHA) ** (1.0 / len(idx))
    assert any(fx[str(i)]['wavelength_nm'] == 530.0 for i in idx)
    for k, i in enumerate(idx):
        T, f = res[k], fx[str(i)]
        assert T['status'].sum() == f['n'] and T['counters'][0] == f['n']
        a = tally_cases.tally_vector(T['det'].sum(0), T['aux_window'].sum(0), 1)
        b = tally_cases.tally_vector(f['det'], f['aux_window'], 1)
        ok, rep = tally_cases.compare_counts(a, b, alpha=alpha_item)
        _flush(f'  item {i} (N {f["n"]:,}, {f["wavelength_nm"]:.0f} nm, {f["conc_mg_l"]:.0f} mg/L): {rep}')
        assert ok, (i, rep)
        zk = tally_cases.sidak_z(9 * len(idx))
        for s in range(1, 9):
            z = tally_cases.binomial_z(T['status'][s], f['status'][s], f['n'])
            assert abs(z) <= zk, (i, s, T['status'][s], f['status'][s])


def test_config5_xyz_field_20g_per_L(gpu_available, oracle_lib):
    kw, mat, conc, n, seed = workloads.config5('plain_mie')
    N = 100_000
    _flush(f'config 5: 128x128x256 field, 20 g/L, {N:,} histories each side')
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    Tg, _, _ = m.run_tallies(N, seed + 1, 0)
    To, _ = oracle_lib.OracleModel(m, mat, conc).run(N, seed + 2, 0)
    a = tally_cases.tally_vector(Tg['det'].sum(0), Tg['aux_window'].sum(0), 1)
    b = tally_cases.tally_vector(To['det'].sum(0), To['aux_window'].sum(0), 1)
    ok, rep = tally_cases.compare_counts(a, b)
    _flush(f'  config 5: {rep}')
    assert ok, rep
    zk = tally_cases.sidak_z(8)
    for s in range(1, 9):
        z = tally_cases.binomial_z(Tg['status'][s], To['status'][s], N)
        assert abs(z) <= zk, (s, Tg['status'][s], To['status'][s])
    assert Tg['failures'].sum() == 0 and To['failures'].sum() == 0
    # multiple scattering regime: tens of flights per history on both sides
    ig, io = Tg['counters'][3] / N, To['counters'][3] / N
    assert ig > 5 and abs(ig - io) < 0.02 * io, (ig, io)


def _sediment_case(name):
    if name == 'reference_template':
        tpl = json.loads((GOLDEN / 'model_default_kwargs.json').read_text())['MODEL_DEFAULT_KWARGS']
        return dict(tpl), 'loess', 0.5
    model, surface = name.split(':')
    kw = workloads.base_kwargs('hybrid_mie_exact_rate')
    kw.update(free_surface_model=surface, vortex_wall_center_height_difference_m=0.005, vortex_core_radius_m=0.003,
              sediment_transport_model=model, source_angular_model='beta_radiance_3d')
    return kw, physics.lognormal(20e-6, 1.8, 16, 1.46, 1e-4), 1.0


@pytest.mark.parametrize('name', ['stokes_drift_diffusion:parabolic_vortex',
                                  'stokes_drift_diffusion:localized_scully_vortex',
                                  'schiller_naumann_drift_diffusion:localized_scully_vortex',
                                  'schiller_naumann_drift_diffusion:parabolic_vortex',
                                  'reference_template'])
def test_history_by_history_sediment_models(gpu_available, oracle_lib, name):
    _flush(f'history-by-history sediment {name}')
    kw, mat, conc = _sediment_case(name)
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 20000
    seed = 24620001
    Tg, rg, _ = m.run_tallies(N, seed, 0, n_records=N)
    To, ro = oracle_lib.OracleModel(m, mat, conc).run(N, seed, 0, n_records=N)
    same = np.ones(N, dtype=bool)
    for k in ('status', 'native_id', 'mirror_id', 'interactions', 'draws'):
        same &= rg[k] == ro[k]
    frac = same.mean()
    assert frac >= 0.999, f'{name}: only {frac:.5f} agree; first mismatches {np.nonzero(~same)[0][:5]}'
    assert Tg['counters'][3] > 0.2 * N, 'particle events are exercised'
    assert np.abs(Tg['det'].sum(0) - To['det'].sum(0)).sum() <= 0.002 * N


@pytest.mark.parametrize('conc_mg_l', [5000.0, 316.0])
def test_history_by_history_slab_majorant(gpu_available, oracle_lib, conc_mg_l):
    """Shared FP32 field tracked with z-slab majorants on both sides: the
    device and the oracle (orc_set_slab_majorant) draw the same random
    numbers, so >= 99.9 % of histories end identically."""
    _flush(f'history-by-history slab majorants, config 2 at {conc_mg_l} mg/L')
    kw, mat, conc, n, seed = workloads.config2(conc_mg_l, 'plain_mie')
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 20000
    Tg, rg, _ = m.run_tallies(N, seed, 0, n_records=N)
    To, ro = oracle_lib.OracleModel(m, mat, conc).use_slab_majorant().run(N, seed, 0, n_records=N)
    same = np.ones(N, dtype=bool)
    for k in ('status', 'native_id', 'mirror_id', 'interactions', 'draws'):
        same &= rg[k] == ro[k]
    assert same.mean() >= 0.999, f'only {same.mean():.5f} agree; first mismatches {np.nonzero(~same)[0][:5]}'
    ci = __import__('tardiis_and_claritas_amd.abi', fromlist=['C_IDX']).C_IDX
    assert abs(int(Tg['counters'][ci['n_null']]) - int(To['counters'][ci['n_null']])) <= 0.002 * To['counters'][ci['n_null']] + 5
