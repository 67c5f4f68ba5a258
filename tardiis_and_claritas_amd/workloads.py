"""The five benchmark workloads as constructor kwargs + simulate arguments.

Config 1: plain MC, small cell (r 12.5 mm, water 50 mm, flat surface),
monodisperse d=10 um, m=1.46+1e-4i, 0.1 g/L uniform, 860 nm collimated
1 mm beam, 1800-angle Mie table, 1e6 histories, seed 24620001.
Config 2: 40-bin lognormal PSD (median 20 um, sigma_g 1.8), Rouse-profile
FP32 r-z field 64x128, Rankine surface (core 3 mm, depth 5 mm), 860 nm,
concentrations 10..5000 mg/L, 1e8 histories each.
Config 3: 256 sampled configurations in one batched launch.
Config 4: surrogate training/queries on config-3 Monte Carlo.
Config 5: 3-D FP32 field 128x128x256, 20 g/L, detectors 0/90/135 deg.
Each entry states its event law: 'hybrid_mie_exact_rate' with smooth
particles is the reference-parity law; 'plain_mie' samples the Mie phase
function at every extinction event.
"""
from __future__ import annotations

import numpy as np

from .physics import cell_geometry, lognormal, monodisperse, rankine_surface_z, rouse_rz_field, \
    segregated_xyz_field

SEED = 24620001          # CLARITAS.py:61
SMALL_CELL = dict(diameter_m=0.025, water_level_m=0.050, sensor_height_m=0.025, cell_height_m=0.100)
CONFIG2_CONCENTRATIONS_MG_L = (10.0, 31.6, 100.0, 316.0, 1000.0, 3162.0, 5000.0)


def base_kwargs(event_model='hybrid_mie_exact_rate', wavelength_nm=860.0):
    kw = cell_geometry(**SMALL_CELL)
    kw.update(
        n_water=1.33, particle_wavelength_m=wavelength_nm * 1e-9, mie_phase_grid_size=1800,
        particle_event_model=event_model, particle_surface_rms_slope_deg=0.0,
        wave_surface_rms_height_m=0.0, free_surface_model='flat', sediment_transport_model='uniform',
        source_angular_model='collimated_disk', source_beam_radius_m=1.0e-3,
        legacy_detector_acceptance_deg=10.0, density_kg_per_m3=2650.0, max_internal_bounces=4095,
    )
    return kw


def config1(event_model='hybrid_mie_exact_rate', source='collimated_disk'):
    kw = base_kwargs(event_model)
    kw['source_angular_model'] = source
    mat = monodisperse(10e-6, 1.46, 1e-4)
    return kw, mat, 0.1, 1_000_000, SEED


def config1_rough(event_model='hybrid_mie_exact_rate', conc=0.5, aspect=1.5, orientation='isotropic', kappa=0.0):
    """Config-1 cell with the reference's default particle physics
    (CLARITAS.py:5610-5614): Beckmann rough particle surfaces at 25 deg rms
    slope (Smith microsurface walk), 100 nm wave rms height (coherent fraction
    ~0.02 at 860 nm, so wave events are mostly spheroid-morphology events),
    spheroid aspect 1.5 with an isotropic axis."""
    kw, mat, _, n, seed = config1(event_model)
    kw.update(particle_surface_rms_slope_deg=25.0, wave_surface_rms_height_m=100e-9,
              wave_spheroid_aspect_ratio=aspect, wave_spheroid_orientation_model=orientation,
              wave_spheroid_orientation_kappa=kappa)
    return kw, mat, conc, n, seed


def config2(conc_mg_l=1000.0, event_model='plain_mie', rouse_number=0.3, nr=64, nz=128):
    kw = base_kwargs(event_model)
    geo = cell_geometry(**SMALL_CELL)
    rin = geo['tube_inner_radius_m']
    zmin = -geo['sensor_height_above_bottom_m']
    zmean = geo['water_height_m'] - geo['sensor_height_above_bottom_m']
    depth, core = 5e-3, 3e-3
    zs = lambda r: rankine_surface_z(r, zmean, rin, depth, core)
    zmax = float(zs(np.array([rin]))[0]) + 1e-3
    field = rouse_rz_field(rin, zmin, zmax, nr, nz, rouse_number, surface_z=zs)
    kw.update(free_surface_model='rankine_vortex', vortex_wall_center_height_difference_m=depth,
              vortex_core_radius_m=core, sediment_transport_model='rz_field', sediment_field=field,
              sediment_field_zmax_m=zmax)
    mat = lognormal(20e-6, 1.8, 40, 1.46, 1e-4)
    return kw, mat, conc_mg_l * 1e-3, 100_000_000, SEED


def config2_meridional(conc_mg_l=1000.0, event_model='plain_mie', chi=0.10, nr=48, nz=96):
    """Config-2 cell and PSD with the reference's solved sediment field: a
    Scully vortex (wall-centre height 5 mm, core 3 mm) and the meridional
    advection-diffusion model (CLARITAS.py:4040-4226, 6451-6725) on its
    default 48 x 96 grid, SOR 1.75, tolerance 1e-6."""
    kw = base_kwargs(event_model)
    kw.update(free_surface_model='localized_scully_vortex', vortex_wall_center_height_difference_m=5e-3,
              vortex_core_radius_m=3e-3, sediment_transport_model='schiller_naumann_meridional_advection_diffusion',
              meridional_circulation_ratio=chi, meridional_grid_nr=nr, meridional_grid_nz=nz)
    mat = lognormal(20e-6, 1.8, 40, 1.46, 1e-4)
    return kw, mat, conc_mg_l * 1e-3, 100_000_000, SEED


def config3_params(n=256, seed=20261203):
    """256 configurations drawn with PCG64(20261203): PSD median log-U(2,60) um,
    sigma_g U(1.3,2.2), m_r U(1.45,1.60), m_i log-U(1e-4,1e-2), lambda in
    {530,860} nm, concentration log-U(10,10000) mg/L, budget log-U(1e5,1e7);
    vortex depth U(0,6) mm (Rankine, core 3 mm) and Rouse number U(0.05,1.0)
    are the two surrogate inputs config 3 leaves open."""
    g = np.random.Generator(np.random.PCG64(seed))
    lu = lambda a, b: np.exp(g.uniform(np.log(a), np.log(b), n))
    p = dict(
        median_um=lu(2.0, 60.0), sigma_g=g.uniform(1.3, 2.2, n), m_r=g.uniform(1.45, 1.60, n),
        m_i=lu(1e-4, 1e-2), wavelength_nm=np.where(g.uniform(0, 1, n) < 0.5, 530.0, 860.0),
        conc_mg_l=lu(10.0, 10000.0), budget=np.round(lu(1e5, 1e7)).astype(np.int64),
        vortex_depth_mm=g.uniform(0.0, 6.0, n), rouse=g.uniform(0.05, 1.0, n),
    )
    return p


def config3_item(p, i, event_model='plain_mie', nr=64, nz=128, n_bins=40):
    """Constructor kwargs + material + concentration for sample i."""
    kw = base_kwargs(event_model, float(p['wavelength_nm'][i]))
    geo = cell_geometry(**SMALL_CELL)
    rin = geo['tube_inner_radius_m']
    zmin = -geo['sensor_height_above_bottom_m']
    zmean = geo['water_height_m'] - geo['sensor_height_above_bottom_m']
    depth = float(p['vortex_depth_mm'][i]) * 1e-3
    core = 3e-3
    if depth > 0:
        zs = lambda r: rankine_surface_z(r, zmean, rin, depth, core)
        zmax = float(zs(np.array([rin]))[0]) + 1e-3
        kw.update(free_surface_model='rankine_vortex', vortex_wall_center_height_difference_m=depth,
                  vortex_core_radius_m=core)
    else:
        zs, zmax = None, zmean + 1e-3
        zs = lambda r: np.full_like(np.asarray(r, dtype=float), zmean)
    field = rouse_rz_field(rin, zmin, zmax, nr, nz, float(p['rouse'][i]), surface_z=zs)
    kw.update(sediment_transport_model='rz_field', sediment_field=field, sediment_field_zmax_m=zmax,
              aux_detector_angles_deg=(135.0,))
    mat = lognormal(float(p['median_um'][i]) * 1e-6, float(p['sigma_g'][i]), n_bins, float(p['m_r'][i]),
                    float(p['m_i'][i]))
    return kw, mat, float(p['conc_mg_l'][i]) * 1e-3


def config5(event_model='plain_mie', n=(128, 128, 256), seed=SEED):
    kw = base_kwargs(event_model)
    geo = cell_geometry(**SMALL_CELL)
    rin = geo['tube_inner_radius_m']
    zmin = -geo['sensor_height_above_bottom_m']
    zmean = geo['water_height_m'] - geo['sensor_height_above_bottom_m']
    depth, core = 5e-3, 3e-3
    zs = lambda r: rankine_surface_z(r, zmean, rin, depth, core)
    zmax = float(zs(np.array([rin]))[0]) + 1e-3
    field = segregated_xyz_field(rin, zmin, zmax, *n, seed=seed, surface_z=zs)
    kw.update(free_surface_model='rankine_vortex', vortex_wall_center_height_difference_m=depth,
              vortex_core_radius_m=core, sediment_transport_model='xyz_field', sediment_field=field,
              sediment_field_zmax_m=zmax, sediment_field_half_width_m=rin, aux_detector_angles_deg=(135.0,))
    mat = lognormal(20e-6, 1.8, 40, 1.46, 1e-4)
    return kw, mat, 20.0, 10_000_000_000, seed


def measure_causal(case, n):
    """Throughput of the production kernel and of the causal-provenance
    variant (clr_run_causal, records copied and binned) on the config-1
    cell: case 'hybrid' (smooth spheres, hybrid law) or 'rough' (the
    reference's default rough-spheroid physics)."""
    import time
    from . import causal
    from .model import ClaritasForwardModel
    kw, mat, conc, _, seed = config1('hybrid_mie_exact_rate') if case == 'hybrid' else config1_rough()
    m = ClaritasForwardModel(**kw)
    mat, _ = m.prepare(mat, conc)
    m.run_tallies(min(n, 100_000), seed)  # warm-up
    t0 = time.time()
    _, _, ms_prod = m.run_tallies(n, seed)
    wall_prod = time.time() - t0
    acc = causal.CausalAccumulator(m, len(mat.diameter_m))
    t1 = time.time()
    acc.run(n, seed, 0, m._launch())
    wall_causal = time.time() - t1
    ms_causal = float(m.ctx.lib.clr_last_kernel_ms(m.ctx.h))
    return dict(workload=f'causal:{case}', histories=n, records=acc.count,
                production_histories_per_s=n / (ms_prod * 1e-3), production_wall_s=wall_prod,
                causal_last_range_kernel_ms=ms_causal, causal_wall_s=wall_causal,
                causal_histories_per_s_wall=n / wall_causal)


def measure(name, histories=None, event_model='plain_mie'):
    """Run one workload on the GPU and return a summary dict (throughput from
    the library's HIP-event kernel time, detector signals, failure counts)."""
    import time
    from .model import ClaritasForwardModel
    if name == 'meridional':
        from .meridional import measure as measure_solve
        return measure_solve(repeats=int(histories or 5))
    if name.startswith('causal:'):
        return measure_causal(name.split(':')[1], int(histories or 2_000_000))
    if name == 'config1':
        kw, mat, conc, n, seed = config1(event_model)
    elif name == 'config5':
        kw, mat, conc, n, seed = config5(event_model)
    elif name.startswith('config2:'):
        kw, mat, conc, n, seed = config2(float(name.split(':')[1]), event_model)
    else:
        raise ValueError(name)
    n = int(histories or n)
    m = ClaritasForwardModel(**kw)
    t0 = time.time()
    m.prepare(mat, conc)
    prep = time.time() - t0
    t1 = time.time()
    T, _, ms = m.run_tallies(n, seed)
    wall = time.time() - t1
    det = T['det'].sum(0)
    out = dict(workload=name, event_model=event_model, histories=n, kernel_ms=ms, wall_s=wall, prepare_s=prep,
               histories_per_s=n / (ms * 1e-3), steps_per_history=float(T['counters'][30]) / n,
               s0=0.5 * float(det[2, 0] + det[3, 0]) / n, s90=0.5 * float(det[2, 9] + det[3, 9]) / n,
               failures=int(T['failures'].sum()), status=[int(v) for v in T['status']])
    if T['aux_window'].size:
        aw = T['aux_window'].sum(0)
        out['s135'] = 0.5 * float(aw[0, 0] + aw[1, 0]) / n
    return out


if __name__ == '__main__':
    import argparse
    import json
    ap = argparse.ArgumentParser(description='run one benchmark workload on the GPU')
    ap.add_argument('name', help="config1 | config5 | config2:<mg/L> | causal:<hybrid|rough> (production vs "
                                 "causal-provenance throughput) | meridional (device sediment solve, "
                                 "--histories = repeats)")
    ap.add_argument('--histories', type=float, default=None)
    ap.add_argument('--event-model', default='plain_mie')
    a = ap.parse_args()
    print(json.dumps(measure(a.name, a.histories, a.event_model)))
