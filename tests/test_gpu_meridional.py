"""GPU parity of the meridional advection-diffusion solve (clr_meridional_solve)
against the oracle (oracle/meridional.c), and of transport on the solved
field.

Bars: the fluid mask is identical; coefficients agree to 1e-12 (same
formulas, different FMA contraction); iteration counts are identical; the
converged grid, extrema and point scales agree to 1e-9 relative (SOR on
both sides from the same start, rounding differences in the mass sums damp
out); residuals to 1e-3 relative (they are differences of nearly equal
numbers).  Transport on the solved field: 3-sigma tallies against the
oracle's transport on the same FP32 field.
"""
import sys
import time

import numpy as np
import pytest

from tardiis_and_claritas_amd import meridional, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel
from tardiis_and_claritas_amd.physics import lognormal

pytestmark = pytest.mark.gpu

MER = 'schiller_naumann_meridional_advection_diffusion'


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve_both(m, mat, oracle_lib, raise_on_failure=True):
    mat = m.resolve_material(mat)
    cfg = m.config_struct(mat)
    p, keep = m.particles_struct(mat, 0.0, None, 'mass_fraction')
    sed = m.sediment_struct(mat)
    t0 = time.perf_counter()
    g = meridional.solve(m.ctx, cfg, p, sed, m.meridional, coeff=True, raise_on_failure=raise_on_failure)
    dt = time.perf_counter() - t0
    del keep
    o = oracle_lib.meridional_solve(m, mat, coeff=True)
    return g, o, dt


def _compare(g, o):
    assert np.array_equal(g['active'], o['active'])
    assert g['iterations'] == o['iterations'], (g['iterations'], o['iterations'])
    assert g['status'] == o['status']
    np.testing.assert_allclose(g['coeff'], o['coeff'], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g['grid'], o['grid'], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(g['scale_max'], o['scale_max'], rtol=1e-9)
    np.testing.assert_allclose(g['scale_min'], o['scale_min'], rtol=1e-9)
    np.testing.assert_allclose(g['point_scales'], o['point_scales'], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(g['residual'], o['residual'], rtol=1e-3, atol=1e-13)
    for k in ('grid_zmax', 'vtheta_peak', 'peak_speed'):
        np.testing.assert_allclose(g[k], o[k], rtol=1e-15)


CASES = {
    # the reference's own solver case, validate_claritas.py:1405-1422
    'parabolic_16x32_loess': (dict(free_surface_model='parabolic_vortex', vortex_wall_center_height_difference_m=0.05,
                                   vortex_core_radius_m=0.02, meridional_circulation_ratio=0.05,
                                   meridional_grid_nr=16, meridional_grid_nz=32,
                                   meridional_solver_max_iterations=3000, meridional_solver_check_interval=100,
                                   meridional_solver_tolerance=1e-5, meridional_solver_sor=1.5), 'loess'),
    # reference defaults: Scully vortex, 48 x 96, chi 0.1, SOR 1.75, tol 1e-6
    'scully_48x96_default': ({}, lognormal(20e-6, 1.8, 40, 1.46, 1e-4)),
    # odd plane (colour rows of unequal length)
    'scully_17x33': (dict(meridional_grid_nr=17, meridional_grid_nz=33, meridional_circulation_ratio=0.3),
                     lognormal(8e-6, 2.0, 12, 1.46, 1e-4)),
    # plane above the LDS limit: the global-memory sweep
    'scully_80x96_global': (dict(meridional_grid_nr=80, meridional_grid_nz=96, meridional_solver_check_interval=100),
                            lognormal(30e-6, 1.6, 6, 1.46, 1e-4)),
}


@pytest.mark.parametrize('case', list(CASES))
def test_solve_matches_oracle(gpu_available, oracle_lib, case):
    _flush(f'meridional solve {case}')
    over, mat = CASES[case]
    kw = dict(sediment_transport_model=MER, particle_event_model='geometric')
    kw.update(over)
    m = ClaritasForwardModel(**kw)
    g, o, dt = _solve_both(m, mat, oracle_lib)
    _flush(f'  {case}: {g["iterations"]} iterations, max residual {g["max_residual"]:.3g}, {dt * 1e3:.1f} ms')
    assert g['converged']
    _compare(g, o)


def test_not_converged_reports_partial_result(gpu_available, oracle_lib):
    _flush('meridional solve not converged')
    m = ClaritasForwardModel(sediment_transport_model=MER, particle_event_model='geometric',
                             meridional_solver_max_iterations=250, meridional_solver_check_interval=50,
                             meridional_solver_tolerance=1e-9)
    mat = lognormal(20e-6, 1.8, 8, 1.46, 1e-4)
    with pytest.raises(Exception, match='did not converge'):
        _solve_both(m, mat, oracle_lib)
    g, o, _ = _solve_both(m, mat, oracle_lib, raise_on_failure=False)
    assert g['status'] == meridional.NOT_CONVERGED and g['iterations'] == 250
    _compare(g, o)


def test_repeat_solve_is_deterministic(gpu_available):
    _flush('meridional solve determinism')
    kw, mat, conc, n, seed = workloads.config2_meridional()
    m = ClaritasForwardModel(**kw)
    mat = m.resolve_material(mat)
    cfg = m.config_struct(mat)
    p, keep = m.particles_struct(mat, 0.0, None, 'mass_fraction')
    sed = m.sediment_struct(mat)
    a = meridional.solve(m.ctx, cfg, p, sed, m.meridional)
    b = meridional.solve(m.ctx, cfg, p, sed, m.meridional)
    assert np.array_equal(a['grid'], b['grid']) and a['iterations'] == b['iterations']


def test_transport_on_solved_field_within_3_sigma(gpu_available, oracle_lib):
    _flush('transport on the meridional field, 3 sigma')
    kw, mat, conc, n, seed = workloads.config2_meridional(1000.0)
    kw['aux_detector_angles_deg'] = (135.0,)
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    res = m.meridional_result
    assert res['converged'] and res['iterations'] >= 1000
    f = m.sediment_field
    assert f.dtype == np.float32 and f.shape == (40, 48, 96)
    np.testing.assert_array_equal(f, res['grid'].astype(np.float32))
    o = oracle_lib.meridional_solve(m, mat, coeff=False)
    np.testing.assert_allclose(res['grid'], o['grid'], rtol=1e-9, atol=1e-14)
    N = 200_000
    Tg, _, _ = m.run_tallies(N, seed + 1, 0)
    To, _ = oracle_lib.OracleModel(m, mat, conc).run(N, seed + 2, 0)
    for a, b in ((Tg['det'].sum(0), To['det'].sum(0)),):
        sa, sb = 0.5 * (a[2] + a[3]), 0.5 * (b[2] + b[3])
        z = np.abs(sa - sb) / np.sqrt(np.maximum(sa + sb, 1.0))
        assert np.all(z < 4.0), f'max z {z.max():.2f} at channel {z.argmax()}'
    aw_g, aw_o = Tg['aux_window'].sum(0).sum(0), To['aux_window'].sum(0).sum(0)
    assert np.all(np.abs(aw_g - aw_o) <= 4.0 * np.sqrt(np.maximum(aw_g + aw_o, 1.0))), (aw_g, aw_o)
    for k in range(1, 9):
        a, b = Tg['status'][k], To['status'][k]
        assert abs(a - b) <= 4 * np.sqrt(max(a + b, 1)), f'status {k}: {a} vs {b}'
    others = np.delete(Tg['failures'], 15)
    assert others.sum() == 0, Tg['failures']


def test_simulate_reports_solver(gpu_available):
    _flush('simulate with the meridional field')
    kw, mat, conc, n, seed = workloads.config2_meridional(316.0)
    m = ClaritasForwardModel(**kw)
    r = m.simulate(mat, conc, n_rays=128_000, seed=seed)
    assert r.sediment_transport_model == MER
    assert r.meridional_grid_nr == 48 and r.meridional_grid_nz == 96
    assert r.meridional_solver_iterations >= 1000
    assert 0.0 < r.meridional_solver_max_residual <= 1e-6
    assert r.meridional_peak_speed_m_s > 0.0
    assert r.sediment_max_bin_density_scale == pytest.approx(float(np.max(m.meridional_result['scale_max'])))
    assert r.production_failure_count <= 2


@pytest.mark.parametrize('law,surface', [('stokes_drift_diffusion', 'parabolic_vortex'),
                                         ('schiller_naumann_drift_diffusion', 'localized_scully_vortex'),
                                         ('schiller_naumann_drift_diffusion', 'parabolic_vortex')])
def test_sediment_tables_match_oracle(gpu_available, oracle_lib, law, surface):
    """Per-bin sediment coefficients (the meridional solve's inputs and the
    analytic fields' tables): the device computes the radial integral by a
    chunked workgroup scan, the oracle sequentially."""
    import ctypes as C
    from tardiis_and_claritas_amd import abi
    _flush(f'sediment tables {law} {surface}')
    m = ClaritasForwardModel(sediment_transport_model=law, free_surface_model=surface, particle_event_model='geometric')
    mat = lognormal(30e-6, 2.0, 24, 1.46, 1e-4)
    m.prepare(mat, 0.5)
    n, nr = 24, m.sediment_radial_lookup_points
    t = m.tables(n)
    om = oracle_lib.OracleModel(m, mat, 0.5)
    o = om.tables()
    for k in ('sed_A', 'sed_B', 'sed_norm', 'sed_settling'):
        np.testing.assert_allclose(t[k], o[k], rtol=1e-12, err_msg=k)
    rl = np.zeros((n, nr))
    tt = abi.TablesOut()
    tt.sed_radial_log = rl.ctypes.data_as(C.POINTER(C.c_double))
    m.ctx.check(m.ctx.lib.clr_get_tables(m.ctx.h, 0, C.byref(tt)), 'clr_get_tables')
    np.testing.assert_allclose(rl, om.radial_log(nr), rtol=1e-12, atol=1e-14)
