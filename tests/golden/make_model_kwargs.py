"""Write tests/golden/model_default_kwargs.json: the reference's constructor
template MODEL_DEFAULT_KWARGS (CLARITAS.py:9905-9973) and the keyword set
that its run_one_case path builds from it for the default user
configuration (_resolve_model_kwargs, CLARITAS.py:11467-11516, with
_resolved_geometry CLARITAS.py:10024-10052 and _resolved_sediment_config
CLARITAS.py:10080-10120).

The template is read from the reference's source as text (ast.literal_eval
of the keyword values of the `dict(...)` literal); nothing of the reference
is imported or executed.  The run_one_case overrides are restated here from
the user-configuration constants of CLARITAS.py:40-234 (material 'loess',
622 nm, the default cell); the per-bin optical profile is the loess n(d)
table with the configured scalar k override 0 (CLARITAS.py:162-170,
11428-11464).

    python tests/golden/make_model_kwargs.py [/path/to/CLARITAS.py]
"""
from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / 'model_default_kwargs.json'
MATERIALS = HERE.parents[1] / 'tardiis_and_claritas_amd' / 'data' / 'materials.json'


def template(src: str) -> dict:
    tree = ast.parse(src)
    for node in tree.body:
        if isinstance(node, ast.Assign) and any(getattr(t, 'id', None) == 'MODEL_DEFAULT_KWARGS'
                                                 for t in node.targets):
            call = node.value
            assert isinstance(call, ast.Call) and getattr(call.func, 'id', None) == 'dict'
            return {k.arg: ast.literal_eval(k.value) for k in call.keywords}
    raise RuntimeError('MODEL_DEFAULT_KWARGS not found')


def run_one_case_overrides() -> dict:
    # user configuration, CLARITAS.py:56-234
    diameter, height, level, sensor = 0.098, 0.491, 0.142, 0.093
    wall, clearance, width = 0.0035, 0.0005, 0.0150
    rin = 0.5 * diameter
    rout = rin + wall
    ring_in = rout + clearance
    ring_out = ring_in + width
    n_real = json.loads(MATERIALS.read_text())['loess']['n_real']
    return dict(
        n_water=1.33, n_acrylic=1.4906, n_air=1.0, particle_wavelength_m=622.0 * 1e-9,
        particle_optical_constants_by_material={'loess': {'n_real': n_real, 'k_imag': [0.0] * len(n_real)}},
        particle_surface_rms_slope_deg=25.0, wave_surface_rms_height_m=100.0 * 1e-9,
        tube_inner_radius_m=rin, tube_outer_radius_m=rout, tube_length_m=height, water_height_m=level,
        sensor_height_above_bottom_m=sensor, sensor_ring_inner_radius_m=ring_in,
        sensor_ring_outer_radius_m=ring_out, source_launch_radius_m=ring_out, counterbore_depth_m=0.008,
        through_bore_diameter_m=0.004, counterbore_diameter_m=0.0087, bottom_disc_thickness_m=0.003,
        density_kg_per_m3=2650.0, sediment_transport_model='schiller_naumann_meridional_advection_diffusion',
        sediment_eddy_diffusivity_m2_s=7.0e-4, water_density_kg_per_m3=997.0,
        water_dynamic_viscosity_pa_s=8.9e-4, gravity_m_s2=9.81, free_surface_model='localized_scully_vortex',
        vortex_wall_center_height_difference_m=0.050, vortex_core_radius_m=0.020,
        meridional_circulation_ratio=0.5, sediment_transport_quadrature_points=4096,
        sediment_radial_lookup_points=2049, meridional_grid_nr=48, meridional_grid_nz=96,
        meridional_solver_max_iterations=8000, meridional_solver_check_interval=250,
        meridional_solver_tolerance=1.0e-6, meridional_solver_sor=1.75, chunk_size=250_000,
        statistics_batch_rays=250_000,
    )


def main(path: str) -> None:
    tpl = template(Path(path).read_text())
    resolved = dict(tpl)
    resolved.update(run_one_case_overrides())
    OUT.write_text(json.dumps({'MODEL_DEFAULT_KWARGS': tpl, 'run_one_case_loess_622nm': resolved},
                              indent=1, sort_keys=True) + '\n')
    print('wrote', OUT, len(tpl), 'template keys')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/CLARITAS/CLARITAS.py')
