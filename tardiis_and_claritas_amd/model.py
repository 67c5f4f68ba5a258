"""`ClaritasForwardModel.simulate(...) -> ForwardResult` on the HIP library.

The constructor keeps the reference's keyword names and defaults
(ClaritasForwardModel.__init__, CLARITAS.py:5577-5652) and its validation
rules for the supported physics; `simulate` keeps the reference's signature
(CLARITAS.py:7160-7179) and fills the tally/result layout of ForwardResult
(CLARITAS.py:5343-5533) from device tallies.  All transport runs in
libclaritas_hip.so on the GPU; this module only builds structures and does the
closed-form host arithmetic the reference also does on the host
(normalisation, jackknife, per-ray means).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi, causal, light, meridional
from ._lib import Context
from .meridional import MeridionalSettings
from .physics import (DETECTOR_ANGLES_DEG, Material, reference_material, vortex_surface_geometry)

MERIDIONAL_MODEL = 'schiller_naumann_meridional_advection_diffusion'

CONDITIONAL_MAX_TRIALS = 16384  # CLARITAS.py:1117


def _jackknife_normalized_se(batch_scores):
    """Delete-one-batch jackknife standard error of the normalised response
    r = T / sum(T), T = sum of the per-batch channel scores; the statistic of
    CLARITAS.py:5244-5261.

    With R_b = (T - s_b) / sum(T - s_b) the leave-one-out responses over the
    B' batches whose complement has a positive sum, the estimate is
    sqrt((B' - 1) * Var(R_b)), the population variance taken per channel
    (Quenouille-Tukey jackknife).  NaN per channel when fewer than 3
    batches are given or usable."""
    nch = len(DETECTOR_ANGLES_DEG)
    s = np.asarray(batch_scores, dtype=float)
    if s.ndim != 2 or s.shape[0] < 3:
        return np.full(nch, np.nan)
    comp = s.sum(axis=0)[None, :] - s                  # each batch's complement
    mass = comp.sum(axis=1)
    keep = mass > 0
    if np.count_nonzero(keep) < 3:
        return np.full(nch, np.nan)
    loo = comp[keep] / mass[keep, None]
    return np.sqrt((loo.shape[0] - 1) * loo.var(axis=0))


def _histogram_median(hist):
    """Median of the sample in which value v occurs hist[v] times
    (CLARITAS.py:5231-5242): the mean of the order statistics of ranks
    floor((n+1)/2) and floor(n/2)+1 (1-based), each the smallest v whose
    cumulative count reaches the rank.  NaN for an empty histogram."""
    hist = np.asarray(hist, dtype=np.int64)
    n = int(hist.sum())
    if n <= 0:
        return float('nan')
    cum = np.cumsum(hist)
    lo, hi = np.searchsorted(cum, [(n + 1) // 2, n // 2 + 1])
    return 0.5 * float(lo + hi)


def _norm(v):
    s = float(np.sum(v))
    return v / s if s > 0 else np.zeros_like(v, dtype=float)


@dataclass
class ForwardResult:
    """Tally/result layout of the reference's ForwardResult (CLARITAS.py:5343-5533)
    for the supported physics, plus the extension fields documented below."""

    detector_angles_deg: np.ndarray
    raw_hits: np.ndarray
    normalized_response: np.ndarray
    native_exact_hits: np.ndarray
    mirror_exact_hits: np.ndarray
    symmetry_exact_scores: np.ndarray
    normalized_native_exact_response: np.ndarray
    normalized_mirror_exact_response: np.ndarray
    normalized_symmetry_exact_response: np.ndarray
    normalized_response_jackknife_se: np.ndarray
    hardware_native_hits: np.ndarray
    hardware_mirror_hits: np.ndarray
    hardware_symmetry_scores: np.ndarray
    normalized_hardware_native_response: np.ndarray
    normalized_hardware_mirror_response: np.ndarray
    normalized_hardware_symmetry_response: np.ndarray
    hardware_normalized_jackknife_se: np.ndarray
    legacy_native_channel_counts: np.ndarray
    legacy_mirror_channel_counts: np.ndarray
    legacy_symmetry_channel_scores: np.ndarray
    normalized_legacy_native_response: np.ndarray
    normalized_legacy_mirror_response: np.ndarray
    normalized_legacy_symmetry_response: np.ndarray
    legacy_normalized_jackknife_se: np.ndarray
    legacy_acceptance_deg: float
    valid_exit_count: int
    numerical_failure_count: int
    numerical_failure_fraction: float
    aggregation_model: str
    aggregation_collision_exposure: float
    aggregation_number_reduction_fraction: float
    aggregation_mass_conservation_error: float
    aggregation_projected_area_ratio: float
    aggregation_largest_bin_mass_fraction: float
    aggregation_overflow_event_fraction: float
    aggregation_steps: int
    population_number_reduction_fraction: float
    fragmentation_breakup_exposure: float
    fragmentation_reference_diameter_m: float
    fragmentation_size_exponent: float
    fragmentation_underflow_event_fraction: float
    aggregation_events_per_initial_particle: float
    fragmentation_events_per_initial_particle: float
    wave_coherent_fraction_nominal_cuda: float
    mean_morphology_wave_interactions: float
    morphology_fraction_of_wave_interactions: float
    vortex_asymptotic_height_parameter_mm: float
    vortex_area_mean_uncorrected_rise_mm: float
    vortex_equivalent_core_rpm: float
    vortex_equivalent_rigid_body_rpm: float
    production_failure_count: int
    production_failure_code_counts: dict
    production_failure_subcode_counts: dict
    production_watchdog_count: int
    production_geometry_failure_count: int
    production_nonfinite_count: int
    production_failure_diagnostics: dict
    n_rays: int
    requested_n_rays: int
    stopped_adaptively: bool
    stop_reason: str
    particle_event_model: str
    particle_surface_rms_slope_deg: float
    particle_surface_beckmann_alpha: float
    wave_surface_rms_height_nm: float
    wave_spheroid_aspect_ratio: float
    wave_spheroid_orientation_model: str
    wave_spheroid_orientation_kappa: float
    wave_coherent_fraction_nominal: float
    mu_geom_per_m: float
    mu_wave_per_m: float
    mu_event_per_m: float
    mu_sca_per_m: float
    mu_ext_per_m: float
    geometric_optical_depth_diameter: float
    optical_depth_diameter: float
    mean_interactions: float
    mean_geometric_interactions: float
    mean_wave_interactions: float
    mean_coherent_wave_interactions: float
    median_interactions: float
    ballistic_fraction: float
    mean_fresnel_reflections: float
    entry_reflection_fraction: float
    internal_reflection_fraction: float
    particle_absorption_fraction: float
    particle_absorbed_ray_count: int
    mean_cell_inner_reflections: float
    mean_cell_outer_reflections: float
    cell_tir_fraction: float
    mean_top_surface_reflections: float
    top_surface_tir_fraction: float
    mean_top_surface_interactions: float
    free_surface_model: str
    vortex_wall_center_height_difference_mm: float
    vortex_core_radius_mm: float
    vortex_center_surface_z_mm: float
    vortex_wall_surface_z_mm: float
    sediment_transport_model: str
    sediment_majorant_geom_per_m: float
    sediment_majorant_wave_per_m: float
    sediment_majorant_total_per_m: float
    sediment_eddy_diffusivity_m2_s: float
    sediment_max_bin_density_scale: float
    meridional_circulation_ratio: float
    meridional_peak_speed_m_s: float
    meridional_grid_nr: int
    meridional_grid_nz: int
    meridional_solver_iterations: int
    meridional_solver_max_residual: float
    mean_bottom_inner_reflections: float
    mean_bottom_outer_reflections: float
    bottom_tir_fraction: float
    top_exit_fraction: float
    bottom_exit_fraction: float
    source_bore_surface_model: str
    detector_bore_surface_model: str
    bore_refractive_index: float
    ring_inner_face_surface_model: str
    ring_inner_face_refractive_index: float
    max_ring_face_bounces: int
    source_collimator_acceptance_fraction: float
    source_direct_acceptance_fraction: float
    source_reflected_acceptance_fraction: float
    source_bore_interactions_per_ray: float
    source_bore_reflections_per_ray: float
    hardware_native_direct_hits: np.ndarray
    hardware_mirror_direct_hits: np.ndarray
    hardware_symmetry_direct_scores: np.ndarray
    hardware_native_stray_hits: np.ndarray
    hardware_mirror_stray_hits: np.ndarray
    hardware_symmetry_stray_scores: np.ndarray
    detector_stray_fraction_by_channel: np.ndarray
    detector_stray_detection_fraction: float
    hardware_native_ring_face_stray_hits: np.ndarray
    hardware_mirror_ring_face_stray_hits: np.ndarray
    hardware_symmetry_ring_face_stray_scores: np.ndarray
    detector_ring_face_stray_fraction_by_channel: np.ndarray
    detector_ring_face_stray_detection_fraction: float
    photodiode_response_model: str
    sfh213_half_angle_deg: float
    sfh213_directional_fit_exponent: float
    photodiode_native_weighted_scores: np.ndarray
    photodiode_mirror_weighted_scores: np.ndarray
    photodiode_symmetry_weighted_scores: np.ndarray
    normalized_photodiode_native_response: np.ndarray
    normalized_photodiode_mirror_response: np.ndarray
    normalized_photodiode_symmetry_response: np.ndarray
    photodiode_relative_response_by_channel: np.ndarray
    detector_mean_incidence_angle_deg_by_channel: np.ndarray
    hardware_candidate_detection_fraction: float
    entered_water_fraction: float
    air_exit_fraction: float
    axial_loss_fraction: float
    detector_detection_fraction: float
    native_detector_detection_fraction: float
    mirror_detector_detection_fraction: float
    legacy_channel_score_per_ray: float
    symmetry_variance_reduction: bool
    batch_diagnostics: Optional[list] = None
    particle_diagnostics: Optional[dict] = None
    light_transport_diagnostics: Optional[dict] = None
    detector_particle_path_diagnostics: Optional[dict] = None
    # per-ray outputs and heat maps (collect_rays / heatmap_size) are not produced
    ray_data: Optional[dict] = None
    heatmap_xy: Optional[np.ndarray] = None
    heatmap_zy: Optional[np.ndarray] = None
    # ---- extensions (no reference field) ----
    aux_detector_angles_deg: Optional[np.ndarray] = None
    aux_native_hits: Optional[np.ndarray] = None
    aux_mirror_hits: Optional[np.ndarray] = None
    aux_symmetry_scores: Optional[np.ndarray] = None
    status_counts: Optional[np.ndarray] = None
    work_counters: dict = field(default_factory=dict)
    device_ms: float = float('nan')

    def to_dict(self):
        d = {k: v for k, v in self.__dict__.items() if k != 'particle_diagnostics'}
        for k, v in list(d.items()):
            if isinstance(v, np.ndarray):
                d[k] = v.tolist()
            elif isinstance(v, np.generic):
                d[k] = v.item()
        return d


# failure-code groups, CLARITAS.py:461-463
WATCHDOG_CODES = (2, 4, 6, 8, 10)
GEOMETRY_FAILURE_CODES = (3, 5, 7, 9, 13, 16, 17)
NONFINITE_CODES = (11, 12)
FAILURE_SAMPLE_LIMIT = 512


def _failure_samples(cacc):
    """The first failed histories of a causal run as the reference's sample
    dicts (CLARITAS.py:8505-8640), without the particle drill-down fields."""
    out = []
    for f in cacc.failures:
        for q in f:
            code, sub = int(q['code']), int(q['subcode'])
            out.append({'ray_id': int(q['history']), 'failure_code': code,
                        'failure_name': abi.PRODUCTION_FAILURE_NAMES.get(code, f'CODE_{code}'),
                        'failure_subcode': sub, 'trace_iteration': int(q['trace_iteration']),
                        'x': float(q['x']), 'y': float(q['y']), 'z': float(q['z']),
                        'vx': float(q['vx']), 'vy': float(q['vy']), 'vz': float(q['vz'])})
    return out[:FAILURE_SAMPLE_LIMIT]


def _failure_summary(T, fails, total):
    """Subcode counts, code-group counts and the production-failure
    dictionary of CLARITAS.py:9608-9655.  The failure samples are filled by
    simulate's causal run (without the particle drill-down fields); the
    particle proposal / accessibility counts are not tallied (None)."""
    sub = T['failure_subcodes']
    sub_counts = {str(abi.SUBCODE_VALUES[k]): int(sub[k]) for k in range(len(abi.SUBCODE_VALUES)) if sub[k]}
    if sub[abi.N_SUBCODE - 1]:
        sub_counts['other'] = int(sub[abi.N_SUBCODE - 1])
    sub_counts = dict(sorted(sub_counts.items(), key=lambda kv: (kv[0] == 'other', int(kv[0]) if kv[0] != 'other' else 0)))
    code_counts = {abi.PRODUCTION_FAILURE_NAMES[i]: int(fails[i]) for i in range(1, abi.N_FAILURE) if fails[i]}
    n_fail = int(np.sum(fails[1:]))
    trapped = int(T['status'][8])
    return dict(
        production_failure_subcode_counts=sub_counts,
        production_watchdog_count=int(sum(fails[i] for i in WATCHDOG_CODES)),
        production_geometry_failure_count=int(sum(fails[i] for i in GEOMETRY_FAILURE_CODES)),
        production_nonfinite_count=int(sum(fails[i] for i in NONFINITE_CODES)),
        production_failure_diagnostics={
            'code_counts': code_counts,
            'subcode_counts': dict(sub_counts),
            'samples': [],
            'sample_limit': FAILURE_SAMPLE_LIMIT,
            'count_closure_ok': n_fail == int(T['status'][4]) and int(np.sum(sub)) == n_fail,
            'particle_proposal_count': None,
            'particle_accessible_count': None,
            'particle_null_count': None,
            'acrylic_trapped_tir_count': trapped,
            'acrylic_trapped_tir_fraction': float(trapped / max(int(total), 1)),
        },
    )


class ClaritasForwardModel:
    """GPU light-transport model of the CLARITAS sample cell.

    Keyword arguments and defaults follow CLARITAS.py:5577-5652.  Supported
    physics: every event law except the moment-constrained one, smooth and
    rough (Smith microsurface walk, CLARITAS.py:2314-2430) particle surfaces,
    the coherent Mie and the incoherent spheroid-morphology wave channels
    (CLARITAS.py:2655-2866), absorbing or reflecting (smooth Fresnel nylon)
    collimator bores and ring face, ideal or SFH213-weighted photodiodes
    (CLARITAS.py:1263-1374, 4283-4376).
    Extensions: event law 'plain_mie',
    source 'collimated_disk', surface 'rankine_vortex', sediment models
    'rz_field' / 'xyz_field' / 'grid_per_bin' with a caller-supplied FP32
    field, auxiliary bores at arbitrary angles, 64-bit history ids.
    """

    EVENT_MODEL_CODES = abi.EVENT_MODEL_CODES
    WEIGHT_MODE_CODES = abi.WEIGHT_MODE_CODES
    SOURCE_ANGULAR_MODEL_CODES = abi.SOURCE_ANGULAR_MODEL_CODES
    FREE_SURFACE_MODEL_CODES = abi.FREE_SURFACE_MODEL_CODES
    SEDIMENT_TRANSPORT_MODEL_CODES = abi.SEDIMENT_TRANSPORT_MODEL_CODES
    BORE_SURFACE_MODEL_CODES = {'absorbing': 0, 'smooth_fresnel_nylon': 1}       # CLARITAS.py:5558
    RING_FACE_SURFACE_MODEL_CODES = {'absorbing': 0, 'smooth_fresnel_nylon': 1}  # CLARITAS.py:5559
    PHOTODIODE_RESPONSE_MODEL_CODES = {'ideal_hard_aperture': 0, 'sfh213_datasheet_fit': 1}  # CLARITAS.py:5560
    SPHEROID_ORIENTATION_MODEL_CODES = {'isotropic': 0, 'local_tangential_axial_vmf': 1}   # CLARITAS.py:5564

    def __init__(
        self,
        n_water=1.33,
        n_particle=1.59,
        n_acrylic=1.4906,
        n_air=1.0,
        tube_inner_radius_m=0.0465,
        tube_outer_radius_m=0.0500,
        tube_length_m=0.491,
        water_height_m=0.142,
        sensor_height_above_bottom_m=0.093,
        sensor_ring_inner_radius_m=0.0505,
        sensor_ring_outer_radius_m=0.0655,
        counterbore_depth_m=0.008,
        through_bore_diameter_m=0.004,
        counterbore_diameter_m=0.0087,
        source_launch_radius_m=0.0655,
        legacy_detector_acceptance_deg=6.5,
        source_beam_sigma_m=1.0e-5,
        source_bore_surface_model='absorbing',
        detector_bore_surface_model='absorbing',
        bore_refractive_index=1.53,
        max_bore_bounces=16,
        ring_inner_face_surface_model='absorbing',
        ring_inner_face_refractive_index=1.53,
        max_ring_face_bounces=16,
        photodiode_response_model='ideal_hard_aperture',
        particle_wavelength_m=622e-9,
        particle_optical_constants_by_material=None,
        particle_event_model='geometric',
        hybrid_wave_scale=1.0,
        mie_phase_grid_size=2049,
        moment_constraint_by_material=None,
        particle_surface_rms_slope_deg=25.0,
        wave_surface_rms_height_m=100e-9,
        wave_spheroid_aspect_ratio=1.5,
        wave_spheroid_orientation_model='isotropic',
        wave_spheroid_orientation_kappa=0.0,
        free_surface_model='localized_scully_vortex',
        vortex_wall_center_height_difference_m=0.050,
        vortex_core_radius_m=0.020,
        sediment_transport_model='schiller_naumann_drift_diffusion',
        sediment_eddy_diffusivity_m2_s=1.0e-3,
        water_density_kg_per_m3=997.0,
        water_dynamic_viscosity_pa_s=8.9e-4,
        gravity_m_s2=9.81,
        sediment_transport_quadrature_points=4096,
        sediment_radial_lookup_points=2049,
        meridional_circulation_ratio=0.10,
        meridional_grid_nr=48,
        meridional_grid_nz=96,
        meridional_solver_max_iterations=8000,
        meridional_solver_check_interval=250,
        meridional_solver_tolerance=1.0e-6,
        meridional_solver_sor=1.75,
        alpha1=0.45,
        alpha2=5.0,
        source_angular_model='beta_radiance_3d',
        density_kg_per_m3=2600.0,
        bottom_disc_thickness_m=0.003,
        bottom_external_refractive_index=1.0,
        aggregation_model='none',
        aggregation_collision_exposure=0.0,
        fragmentation_breakup_exposure=0.0,
        fragmentation_reference_diameter_m=100e-6,
        fragmentation_size_exponent=1.0,
        aggregation_step_safety=0.2,
        aggregation_max_steps=200000,
        max_axial_bounces=64,
        max_iterations=int(1e6),
        max_internal_bounces=64,
        max_cell_bounces=64,
        chunk_size=250000,
        statistics_batch_rays=250000,
        symmetry_average_exact=True,
        # ---- extensions ----
        source_beam_radius_m=1.0e-3,
        aux_detector_angles_deg=(),
        jackknife_groups=64,
        sediment_field=None,
        sediment_field_zmax_m=None,
        sediment_field_half_width_m=None,
        device=0,
        launch=None,
    ):
        self.n_water, self.n_particle = float(n_water), float(n_particle)
        self.n_acrylic, self.n_air = float(n_acrylic), float(n_air)
        self.rin, self.rout = float(tube_inner_radius_m), float(tube_outer_radius_m)
        self.tube_length = float(tube_length_m)
        self.water_height = float(water_height_m)
        self.sensor_height = float(sensor_height_above_bottom_m)
        # sensor-plane coordinates, CLARITAS.py:5667-5671
        self.zmin = -self.sensor_height
        self.zmean = self.water_height - self.sensor_height
        self.zwall_top = self.tube_length - self.sensor_height
        self.ring_in = float(sensor_ring_inner_radius_m)
        self.ring_out = float(sensor_ring_outer_radius_m)
        self.counterbore_depth = float(counterbore_depth_m)
        self.throat_out = self.ring_out - self.counterbore_depth
        self.throat_radius = 0.5 * float(through_bore_diameter_m)
        self.counterbore_radius = 0.5 * float(counterbore_diameter_m)
        self.source_launch_radius = float(source_launch_radius_m)
        self.legacy_acceptance_deg = float(legacy_detector_acceptance_deg)
        self.beam_sigma = float(source_beam_sigma_m)
        self.beam_radius = float(source_beam_radius_m)
        self.source_bore_surface_model = str(source_bore_surface_model).strip().lower()
        self.detector_bore_surface_model = str(detector_bore_surface_model).strip().lower()
        self.bore_refractive_index = float(bore_refractive_index)
        self.max_bore_bounces = int(max_bore_bounces)
        self.ring_inner_face_surface_model = str(ring_inner_face_surface_model).strip().lower()
        self.ring_inner_face_refractive_index = float(ring_inner_face_refractive_index)
        self.max_ring_face_bounces = int(max_ring_face_bounces)
        self.photodiode_response_model = str(photodiode_response_model).strip().lower()
        self.particle_wavelength = float(particle_wavelength_m)
        self.particle_optical_constants_by_material = particle_optical_constants_by_material
        self.particle_event_model = str(particle_event_model).strip().lower()
        self.hybrid_wave_scale = float(hybrid_wave_scale)
        self.mie_phase_grid_size = int(mie_phase_grid_size)
        self.particle_surface_rms_slope_deg = float(particle_surface_rms_slope_deg)
        self.particle_surface_beckmann_alpha = math.tan(math.radians(self.particle_surface_rms_slope_deg))
        self.wave_surface_rms_height_m = float(wave_surface_rms_height_m)
        self.wave_spheroid_aspect_ratio = float(wave_spheroid_aspect_ratio)
        self.wave_spheroid_orientation_model = str(wave_spheroid_orientation_model).strip().lower()
        self.wave_spheroid_orientation_kappa = float(wave_spheroid_orientation_kappa)
        self.free_surface_model = str(free_surface_model).strip().lower()
        self.vortex_delta_h = float(vortex_wall_center_height_difference_m)
        self.vortex_core_radius = float(vortex_core_radius_m)
        self.sediment_transport_model = str(sediment_transport_model).strip().lower()
        self.sediment_eddy_diffusivity = float(sediment_eddy_diffusivity_m2_s)
        self.water_density = float(water_density_kg_per_m3)
        self.water_dynamic_viscosity = float(water_dynamic_viscosity_pa_s)
        self.gravity = float(gravity_m_s2)
        self.sediment_transport_quadrature_points = int(sediment_transport_quadrature_points)
        self.sediment_radial_lookup_points = int(sediment_radial_lookup_points)
        self.meridional_circulation_ratio = float(meridional_circulation_ratio)
        self.meridional = MeridionalSettings(
            circulation_ratio=self.meridional_circulation_ratio, nr=int(meridional_grid_nr),
            nz=int(meridional_grid_nz), max_iterations=int(meridional_solver_max_iterations),
            check_interval=int(meridional_solver_check_interval), tolerance=float(meridional_solver_tolerance),
            sor=float(meridional_solver_sor))
        self.meridional_grid_nr, self.meridional_grid_nz = self.meridional.nr, self.meridional.nz
        self.meridional_solver_max_iterations = self.meridional.max_iterations
        self.meridional_solver_check_interval = self.meridional.check_interval
        self.meridional_solver_tolerance, self.meridional_solver_sor = self.meridional.tolerance, self.meridional.sor
        self.meridional_vtheta_peak = 0.0
        self.meridional_peak_speed = 0.0
        self.meridional_result = None
        self.alpha1, self.alpha2 = float(alpha1), float(alpha2)
        self.source_angular_model = str(source_angular_model).strip().lower()
        self.density = float(density_kg_per_m3)
        self.bottom_disc_thickness = float(bottom_disc_thickness_m)
        self.bottom_external_n = float(bottom_external_refractive_index)
        self.max_iterations = int(max_iterations)
        # population balance (CLARITAS.py:3599-3843): only the unaggregated PSD
        # (model 'none', exposure 0: the weights are normalised) is built
        self.aggregation_model = str(aggregation_model).strip().lower()
        self.aggregation_collision_exposure = float(aggregation_collision_exposure)
        self.fragmentation_breakup_exposure = float(fragmentation_breakup_exposure)
        self.fragmentation_reference_diameter_m = float(fragmentation_reference_diameter_m)
        self.fragmentation_size_exponent = float(fragmentation_size_exponent)
        # integrator controls of the population balance; validated as the
        # reference does (CLARITAS.py:5958-5961) although model 'none' never steps
        self.aggregation_step_safety = float(aggregation_step_safety)
        self.aggregation_max_steps = int(aggregation_max_steps)
        self._validate_population()
        self.moment_constraint_by_material = moment_constraint_by_material or {}
        # the acrylic loops are bounded by their own emergency cap (16384,
        # CLARITAS.py:3143-3147); the reference validates this value and never
        # reads it afterwards, and so does this model
        self.max_axial_bounces = int(max_axial_bounces)
        self.max_internal_bounces = int(max_internal_bounces)
        self.max_cell_bounces = int(max_cell_bounces)
        self.chunk_size = int(chunk_size)
        self.statistics_batch_rays = int(statistics_batch_rays)
        self.symmetry_average_exact = bool(symmetry_average_exact)
        self.aux_detector_angles_deg = np.asarray(list(aux_detector_angles_deg), dtype=float)
        self.jackknife_groups = int(jackknife_groups)
        self.sediment_field = None if sediment_field is None else np.ascontiguousarray(sediment_field, dtype=np.float32)
        self.sediment_field_zmax = sediment_field_zmax_m
        self.sediment_field_half_width = sediment_field_half_width_m
        self.launch = launch
        self._validate()
        self.device = int(device)
        self._ctx = None

    @property
    def ctx(self) -> Context:
        """HIP device context, opened on first use (raises without a GPU)."""
        if self._ctx is None:
            self._ctx = Context(self.device)
        return self._ctx

    # ---------------------------------------------------------------- checks
    def _validate_population(self):
        """Population-balance settings: the reference's rules
        (CLARITAS.py:5920-5961), then the one model built here ('none' with
        zero exposures: the PSD weights are only normalised)."""
        if self.aggregation_model not in ('none', 'orthokinetic_compact', 'orthokinetic_compact_fragmentation'):
            raise ValueError("aggregation_model must be 'none', 'orthokinetic_compact', or "
                             "'orthokinetic_compact_fragmentation'")
        if not (math.isfinite(self.aggregation_collision_exposure) and self.aggregation_collision_exposure >= 0.0):
            raise ValueError('aggregation_collision_exposure must be finite and >=0')
        if not (math.isfinite(self.fragmentation_breakup_exposure) and self.fragmentation_breakup_exposure >= 0.0):
            raise ValueError('fragmentation_breakup_exposure must be finite and >=0')
        if not (math.isfinite(self.fragmentation_reference_diameter_m) and self.fragmentation_reference_diameter_m > 0):
            raise ValueError('fragmentation_reference_diameter_m must be finite and >0')
        if not (math.isfinite(self.fragmentation_size_exponent) and 0.0 <= self.fragmentation_size_exponent <= 6.0):
            raise ValueError('fragmentation_size_exponent must be finite and in [0,6]')
        if self.aggregation_model != 'orthokinetic_compact_fragmentation' and self.fragmentation_breakup_exposure > 0:
            raise ValueError('fragmentation_breakup_exposure requires '
                             'aggregation_model=orthokinetic_compact_fragmentation')
        if not (0.0 < self.aggregation_step_safety <= 0.25):
            raise ValueError('aggregation_step_safety must be in (0,0.25]')
        if self.aggregation_max_steps < 1:
            raise ValueError('aggregation_max_steps must be >=1')
        if self.aggregation_model != 'none' or self.aggregation_collision_exposure != 0.0 or \
                self.fragmentation_breakup_exposure != 0.0:
            raise NotImplementedError('orthokinetic aggregation / fragmentation of the PSD is not built: '
                                      "aggregation_model must be 'none' with zero exposures")

    def _validate(self):
        """Fail-closed validation, following CLARITAS.py:5754-5999 for the
        supported physics and refusing what the device path does not model."""
        if self.particle_event_model not in self.EVENT_MODEL_CODES:
            raise ValueError(f'unsupported particle_event_model {self.particle_event_model!r}')
        if self.source_angular_model not in self.SOURCE_ANGULAR_MODEL_CODES:
            raise ValueError(f'unsupported source_angular_model {self.source_angular_model!r}')
        if self.free_surface_model not in self.FREE_SURFACE_MODEL_CODES:
            raise ValueError(f'unsupported free_surface_model {self.free_surface_model!r}')
        if self.sediment_transport_model not in self.SEDIMENT_TRANSPORT_MODEL_CODES:
            raise ValueError(f'unsupported sediment_transport_model {self.sediment_transport_model!r}')
        # CLARITAS.py:5966-5991
        if self.source_bore_surface_model not in self.BORE_SURFACE_MODEL_CODES or \
                self.detector_bore_surface_model not in self.BORE_SURFACE_MODEL_CODES:
            raise ValueError('bore surface model must be absorbing or smooth_fresnel_nylon')
        if self.ring_inner_face_surface_model not in self.RING_FACE_SURFACE_MODEL_CODES:
            raise ValueError('ring inner face surface model must be absorbing or smooth_fresnel_nylon')
        if not (self.ring_inner_face_refractive_index > 0.0 and math.isfinite(self.ring_inner_face_refractive_index)):
            raise ValueError('ring_inner_face_refractive_index must be finite and positive')
        if self.max_ring_face_bounces < 1:
            raise ValueError('max_ring_face_bounces must be >=1')
        if not self.bore_refractive_index > 1.0:
            raise ValueError('bore_refractive_index must be >1')
        if self.max_bore_bounces < 0:
            raise ValueError('max_bore_bounces must be >=0')
        if self.photodiode_response_model not in self.PHOTODIODE_RESPONSE_MODEL_CODES:
            raise ValueError('photodiode_response_model must be ideal_hard_aperture or sfh213_datasheet_fit')
        # CLARITAS.py:5797-5826
        if not (math.isfinite(self.particle_surface_rms_slope_deg) and 0.0 <= self.particle_surface_rms_slope_deg < 80.0):
            raise ValueError('particle_surface_rms_slope_deg must be finite and in [0,80) degrees')
        if not (math.isfinite(self.wave_surface_rms_height_m) and self.wave_surface_rms_height_m >= 0.0):
            raise ValueError('wave_surface_rms_height_m must be finite and >=0')
        if not (math.isfinite(self.wave_spheroid_aspect_ratio) and 0.2 <= self.wave_spheroid_aspect_ratio <= 5.0):
            raise ValueError('wave_spheroid_aspect_ratio must be finite and in [0.2,5.0]')
        if self.wave_spheroid_orientation_model not in self.SPHEROID_ORIENTATION_MODEL_CODES:
            raise ValueError("wave_spheroid_orientation_model must be 'isotropic' or 'local_tangential_axial_vmf'")
        if not (math.isfinite(self.wave_spheroid_orientation_kappa) and 0.0 <= self.wave_spheroid_orientation_kappa <= 100.0):
            raise ValueError('wave_spheroid_orientation_kappa must be finite and in [0,100]')
        if not (0 < self.rin < self.rout < self.ring_in < self.throat_out <= self.source_launch_radius <= self.ring_out):
            raise ValueError('geometry must satisfy 0<Rin<Rout<ring_in<throat_out<=source<=ring_out')
        if not (self.zwall_top > self.zmean > self.zmin):
            raise ValueError('tube top > mean water level > bottom is required')
        if not (0 < self.throat_radius < self.counterbore_radius):
            raise ValueError('bore radii must satisfy 0 < throat < counterbore')
        if self.mie_phase_grid_size < 257:
            raise ValueError('mie_phase_grid_size must be >= 257')
        # CLARITAS.py:5992-5999
        if not (self.zmin < 0 < self.zmean):
            raise ValueError('sensor plane must lie within water column')
        if self.bottom_disc_thickness <= 0.0:
            raise ValueError('bottom_disc_thickness_m must be positive')
        if self.bottom_external_n <= 0.0:
            raise ValueError('bottom_external_refractive_index must be positive')
        if self.max_axial_bounces < 1:
            raise ValueError('max_axial_bounces must be >=1')
        if not (1 <= self.jackknife_groups <= abi.MAX_GROUPS):
            raise ValueError('jackknife_groups must be 1..64')
        if self.aux_detector_angles_deg.size > abi.MAX_AUX:
            raise ValueError('at most 4 auxiliary bores')
        if self.free_surface_model != 'flat' and self.vortex_delta_h > 0.0:
            vg = vortex_surface_geometry(self.free_surface_model, self.zmean, self.rin, self.vortex_delta_h,
                                         self.vortex_core_radius, self.gravity)
            if not (vg['wall_z'] < self.zwall_top and vg['center_z'] > self.zmin):
                raise ValueError('free surface leaves the tube')
        if self.sediment_transport_model in ('rz_field', 'xyz_field', 'grid_per_bin'):
            if self.sediment_field is None:
                raise ValueError(f'{self.sediment_transport_model} needs sediment_field')
        # meridional circulation, CLARITAS.py:5873-5921
        self.meridional.validate()
        if self.free_surface_model in ('localized_scully_vortex', 'parabolic_vortex') and self.vortex_delta_h > 0.0:
            vg = vortex_surface_geometry(self.free_surface_model, self.zmean, self.rin, self.vortex_delta_h,
                                         self.vortex_core_radius, self.gravity)
            self.meridional_vtheta_peak = float(vg['vtheta_peak_m_s'])
            self.meridional_peak_speed = self.meridional_circulation_ratio * self.meridional_vtheta_peak
        if self.sediment_transport_model == MERIDIONAL_MODEL:
            if self.free_surface_model not in ('localized_scully_vortex', 'parabolic_vortex') or \
                    self.vortex_delta_h <= 0.0:
                raise ValueError('meridional transport requires a non-zero supported vortex')

    # --------------------------------------------------------------- structs
    def config_struct(self, material: Material) -> abi.Config:
        c = abi.Config()
        c.n_water, c.n_acrylic, c.n_air = self.n_water, self.n_acrylic, self.n_air
        c.bottom_external_n = self.bottom_external_n
        c.rin, c.rout, c.zmin, c.zmean, c.zwall_top = self.rin, self.rout, self.zmin, self.zmean, self.zwall_top
        c.bottom_disc_thickness = self.bottom_disc_thickness
        c.ring_in, c.throat_out, c.ring_out = self.ring_in, self.throat_out, self.ring_out
        c.source_r, c.throat_r, c.counterbore_r = self.source_launch_radius, self.throat_radius, self.counterbore_radius
        flat = self.free_surface_model == 'flat'
        c.vortex_delta_h = 0.0 if flat else self.vortex_delta_h
        c.vortex_core_radius = self.vortex_core_radius
        c.alpha1, c.alpha2, c.beam_sigma, c.beam_radius = self.alpha1, self.alpha2, self.beam_sigma, self.beam_radius
        c.legacy_accept_deg = self.legacy_acceptance_deg
        c.wavelength = self.particle_wavelength
        c.rough_alpha = self.particle_surface_beckmann_alpha
        c.spheroid_aspect = self.wave_spheroid_aspect_ratio
        c.spheroid_kappa = self.wave_spheroid_orientation_kappa
        c.spheroid_orientation = self.SPHEROID_ORIENTATION_MODEL_CODES[self.wave_spheroid_orientation_model]
        c.bore_n, c.ring_face_n = self.bore_refractive_index, self.ring_inner_face_refractive_index
        c.source_bore_model = self.BORE_SURFACE_MODEL_CODES[self.source_bore_surface_model]
        c.detector_bore_model = self.BORE_SURFACE_MODEL_CODES[self.detector_bore_surface_model]
        c.ring_face_model = self.RING_FACE_SURFACE_MODEL_CODES[self.ring_inner_face_surface_model]
        c.photodiode_model = self.PHOTODIODE_RESPONSE_MODEL_CODES[self.photodiode_response_model]
        c.max_bore_bounces, c.max_ring_face_bounces = self.max_bore_bounces, self.max_ring_face_bounces
        for i, a in enumerate(self.aux_detector_angles_deg):
            c.aux_angles_deg[i] = float(a)
        c.n_aux = int(self.aux_detector_angles_deg.size)
        c.free_surface_model = self.FREE_SURFACE_MODEL_CODES[self.free_surface_model]
        c.source_angular_model = self.SOURCE_ANGULAR_MODEL_CODES[self.source_angular_model]
        c.event_model = self.EVENT_MODEL_CODES[self.particle_event_model]
        c.sediment_model = self.SEDIMENT_TRANSPORT_MODEL_CODES[self.sediment_transport_model]
        c.max_particle_bounces = self.max_internal_bounces
        c.max_cell_bounces = self.max_cell_bounces
        c.max_iterations = self.max_iterations
        c.conditional_max_trials = CONDITIONAL_MAX_TRIALS
        c.n_groups = self.jackknife_groups
        f = self.sediment_field
        if f is not None:
            if c.sediment_model == 5:
                c.field_n0, c.field_n1, c.field_n2 = f.shape[-3], f.shape[-2], f.shape[-1]
            else:
                c.field_n0, c.field_n1 = f.shape[-2], f.shape[-1]
            c.field_nbins = f.shape[0] if (c.sediment_model == 3 and f.ndim == 3) else 1
        c.field_zmax = float(self.sediment_field_zmax if self.sediment_field_zmax is not None else self.zmean)
        c.field_xy_half = float(self.sediment_field_half_width if self.sediment_field_half_width is not None else self.rin)
        return c

    def particles_struct(self, material: Material, concentration_g_per_L: float, weights, weight_mode):
        d = np.ascontiguousarray(material.diameter_m, dtype=np.float64)
        w = np.ascontiguousarray(material.weights if weights is None else weights, dtype=np.float64)
        self.last_psd_weights = w
        if w.ndim != 1 or w.size != d.size:
            raise ValueError(f'{material.name} requires {d.size} PSD weights')
        if np.any(~np.isfinite(w)) or np.any(w < 0.0) or not np.any(w > 0.0):
            raise ValueError('invalid PSD weights')
        if weight_mode not in self.WEIGHT_MODE_CODES:
            raise ValueError("weight_mode must be 'mass_fraction' or 'number_fraction'")
        conc = float(concentration_g_per_L)
        if not np.isfinite(conc) or conc < 0.0:
            raise ValueError('concentration must be finite and >=0')
        nr = np.ascontiguousarray(material.n_real, dtype=np.float64)
        ki = np.ascontiguousarray(material.k_imag, dtype=np.float64)
        keep = (d, w, nr, ki)
        p = abi.Particles()
        PD = C.POINTER(C.c_double)
        p.diameter, p.weight = d.ctypes.data_as(PD), w.ctypes.data_as(PD)
        p.n_real, p.k_imag = nr.ctypes.data_as(PD), ki.ctypes.data_as(PD)
        p.concentration = conc
        p.density = float(material.density_kg_m3 if material.density_kg_m3 else self.density)
        p.n_medium = self.n_water
        p.wavelength = self.particle_wavelength
        p.hybrid_wave_scale = self.hybrid_wave_scale
        p.wave_rms_height_m = self.wave_surface_rms_height_m
        p.n = int(d.size)
        p.n_theta = self.mie_phase_grid_size
        p.weight_mode = self.WEIGHT_MODE_CODES[weight_mode]
        p.event_model = self.EVENT_MODEL_CODES[self.particle_event_model]
        return p, keep

    def sediment_struct(self, material: Material) -> abi.Sediment:
        s = abi.Sediment()
        s.particle_density = float(material.density_kg_m3 if material.density_kg_m3 else self.density)
        s.water_density, s.water_viscosity = self.water_density, self.water_dynamic_viscosity
        s.gravity, s.eddy_diffusivity = self.gravity, self.sediment_eddy_diffusivity
        s.n_quad, s.n_radial = self.sediment_transport_quadrature_points, self.sediment_radial_lookup_points
        return s

    def resolve_material(self, material) -> Material:
        if isinstance(material, Material):
            return material
        return reference_material(material, self.density, self.particle_optical_constants_by_material)

    # -------------------------------------------------------------- prepare
    def prepare(self, material, concentration_g_per_L, weights=None, weight_mode='mass_fraction', slot=0):
        """Device precompute for one configuration (Mie tables, transport
        split, sediment coefficients); replaces _prepare_gpu_particle_physics."""
        mat = self.resolve_material(material)
        cfg = self.config_struct(mat)
        p, keep = self.particles_struct(mat, concentration_g_per_L, weights, weight_mode)
        sed = self.sediment_struct(mat)
        if self.sediment_transport_model == MERIDIONAL_MODEL:
            cfg = self._solve_meridional(cfg, p, sed)
        f = self.sediment_field
        fptr = f.ctypes.data_as(C.POINTER(C.c_float)) if f is not None else None
        flen = int(f.size) if f is not None else 0
        self.ctx.check(self.ctx.lib.clr_prepare(self.ctx.h, int(slot), C.byref(cfg), C.byref(p), C.byref(sed),
                                                fptr, flen), 'clr_prepare')
        del keep
        return mat, cfg

    def _solve_meridional(self, cfg, particles, sed):
        """Device meridional solve for the prepared PSD; its grid becomes the
        FP32 per-bin r-z field the transport reads (sediment model 3, the
        reference's sed_grid, CLARITAS.py:6451-6725, 7719-7722)."""
        res = meridional.solve(self.ctx, cfg, particles, sed, self.meridional)
        self.meridional_result = res
        self.sediment_field = np.ascontiguousarray(res['grid'], dtype=np.float32)
        self.sediment_field_zmax = res['grid_zmax']
        return self.config_struct_from(cfg)

    def config_struct_from(self, cfg):
        """cfg with the current sediment field's shape and extent."""
        f = self.sediment_field
        cfg.field_n0, cfg.field_n1 = f.shape[-2], f.shape[-1]
        cfg.field_nbins = f.shape[0]
        cfg.field_zmax = float(self.sediment_field_zmax)
        return cfg

    def _max_bin_density_scale(self, tables):
        """Largest per-bin density scale (sediment_max_density_scale,
        CLARITAS.py:4021,4202): 1/normalisation for the analytic fields, the
        solved or supplied grid's maximum otherwise."""
        tm = self.sediment_transport_model
        if tm == 'uniform':
            return 1.0
        if tm == MERIDIONAL_MODEL and self.meridional_result is not None:
            return float(np.max(self.meridional_result['scale_max']))
        if self.sediment_field is not None and tm != 'stokes_drift_diffusion' and \
                tm != 'schiller_naumann_drift_diffusion':
            return float(np.max(self.sediment_field))
        return float(np.max(1.0 / np.asarray(tables['sed_norm'])))

    def tables(self, n_bins: int, slot: int = 0):
        """Host copy of the prepared per-bin tables of `slot`."""
        PD = C.POINTER(C.c_double)
        arrs = {k: np.zeros(n_bins) for k in ('radius', 'number_density', 'mu_geom_bin', 'mu_sca_bin',
                                             'mu_ext_bin', 'trace_geom_bin', 'trace_wave_bin', 'geom_cdf',
                                             'wave_cdf', 'sed_A', 'sed_B', 'sed_norm', 'sed_settling')}
        t = abi.TablesOut()
        for k, v in arrs.items():
            setattr(t, k, v.ctypes.data_as(PD))
        self.ctx.check(self.ctx.lib.clr_get_tables(self.ctx.h, int(slot), C.byref(t)), 'clr_get_tables')
        arrs['mu_scalars'] = np.array(list(t.mu_scalars))
        arrs['majorants'] = np.array(list(t.majorants))
        return arrs

    def mie_tables(self, material, n_theta=None):
        """FP64 Mie efficiencies and phase tables on the device
        (phase_table_diagnostics counterpart, CLARITAS.py:7138-7158)."""
        mat = self.resolve_material(material)
        p, keep = self.particles_struct(mat, 0.0, None, 'mass_fraction')
        nt = int(n_theta or self.mie_phase_grid_size)
        p.n_theta = nt
        n = p.n
        out = {k: np.zeros(n) for k in ('size_parameter', 'qext', 'qsca', 'qabs', 'qback', 'asym_g', 'albedo')}
        out['nstop'] = np.zeros(n, dtype=np.int32)
        out['theta'] = np.zeros(nt)
        out['phase_pdf'] = np.zeros((n, nt))
        out['phase_cdf'] = np.zeros((n, nt))
        m = abi.MieOut()
        PD = C.POINTER(C.c_double)
        for k in ('size_parameter', 'qext', 'qsca', 'qabs', 'qback', 'asym_g', 'albedo', 'theta', 'phase_pdf',
                  'phase_cdf'):
            setattr(m, k, out[k].ctypes.data_as(PD))
        m.nstop = out['nstop'].ctypes.data_as(C.POINTER(C.c_int32))
        self.ctx.check(self.ctx.lib.clr_mie_tables(self.ctx.h, C.byref(p), C.byref(m)), 'clr_mie_tables')
        del keep
        return out

    def _launch(self):
        if self.launch is None:
            return None
        l = abi.Launch()
        l.block, l.blocks_per_cu, l.chunk = (int(self.launch.get(k, 0)) for k in ('block', 'blocks_per_cu', 'chunk'))
        return C.byref(l)

    def run_tallies(self, n, seed, history_offset=0, n_records=0, light=None, causal=None):
        """Trace n histories of the prepared slot 0; returns (tallies dict,
        records or None, device ms).  With a light.LightAccumulator the run
        also records the light-transport voxel diagnostics, with a
        causal.CausalAccumulator the detector-path provenance (same tallies)."""
        if causal is not None:
            T = causal.run(n, seed, history_offset, self._launch(), light_acc=light)
            return T, None, float(self.ctx.lib.clr_last_kernel_ms(self.ctx.h))
        if light is not None:
            t = light.run(n, seed, history_offset, self._launch())
            ms = float(self.ctx.lib.clr_last_kernel_ms(self.ctx.h))
            return abi.tallies_to_numpy(t), None, ms
        t = abi.Tallies()
        rec = (abi.Record * int(n_records))() if n_records else None
        self.ctx.check(self.ctx.lib.clr_run(self.ctx.h, int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, int(history_offset),
                                            self._launch(), C.byref(t), rec, int(n_records)), 'clr_run')
        ms = float(self.ctx.lib.clr_last_kernel_ms(self.ctx.h))
        recs = abi.records_to_numpy(rec, int(n_records)) if n_records else None
        return abi.tallies_to_numpy(t), recs, ms

    # ------------------------------------------------------------- simulate
    def simulate(self, material, concentration_g_per_L, weights=None, n_rays=100000, seed=12345,
                 weight_mode='mass_fraction', collect_rays=False, heatmap_size=0, light_transport_config=None,
                 target_detector_score=None, min_rays=None, max_rays=None, min_rear_channel_score=None,
                 min_h170_score=None, stability_l1_tolerance=None, stability_window=3, ray_id_start=0,
                 path_diagnostics=True):
        """Signature of CLARITAS.py:7160-7179.  Per-ray outputs and heat maps
        are not produced (collect_rays and heatmap_size must be off); the
        light-transport voxel diagnostics are (light_transport_config with
        'enabled': True; keys as CLARITAS.py:7232-7247)."""
        if collect_rays:
            raise ValueError('per-ray outputs are not produced by the device path')
        requested = int(n_rays)
        ray_id_start = int(ray_id_start)
        adaptive = target_detector_score is not None
        plan = self._adaptive_plan(requested, ray_id_start, target_detector_score, min_rays, max_rays,
                                   min_rear_channel_score, min_h170_score)
        mat, cfg = self.prepare(material, concentration_g_per_L, weights, weight_mode)
        lacc = None
        vis = int(heatmap_size or 0)
        if vis < 0:
            raise ValueError('heatmap_size must be >= 0')
        if light_transport_config and light_transport_config.get('enabled'):
            lacc = light.LightAccumulator(self, dict(light_transport_config, heatmap_size=vis), int(n_rays),
                                          int(ray_id_start))
        elif vis > 0:  # water-path heat maps alone (CLARITAS.py:7229-7231, 9453-9454)
            lacc = light.LightAccumulator(self, light.heatmap_only_config(vis), int(n_rays), int(ray_id_start))
        # detector-path / causal / H170 diagnostics (always on in the
        # reference; path_diagnostics=False runs the production kernel only)
        cacc = causal.CausalAccumulator(self, len(mat.diameter_m)) if path_diagnostics else None
        if not adaptive:
            T, _, ms = self.run_tallies(requested, seed, ray_id_start, light=lacc, causal=cacc)
            batches = [T['det'][g] for g in range(T['n_groups'])]
            pbatches = [T['photodiode'][g] for g in range(T['n_groups'])]
            group_n = T['group_histories']
            stop_reason, total = 'fixed_n_rays', requested
            det_sum = T['det'].sum(axis=0)
        else:
            run = lambda b, off: self.run_tallies(b, seed, off, light=lacc, causal=cacc)
            T, batches, pbatches, group_n, total, ms, stop_reason = self._run_adaptive(
                run, ray_id_start, plan, stability_l1_tolerance, stability_window)
            det_sum = np.stack(batches).sum(axis=0)
        res = self._result(mat, T, det_sum, batches, pbatches, group_n, requested, total, adaptive, stop_reason, ms)
        if lacc is not None:
            res.heatmap_xy, res.heatmap_zy = lacc.heatmaps()
            if not lacc.cfg.get('heatmap_only'):
                res.light_transport_diagnostics = lacc.result(total)
                light.check_closure(res.light_transport_diagnostics)
        if cacc is not None:
            cond = self.particle_event_model in ('hybrid_mie_exact_rate', 'hybrid_mie_exact_rate_moment_constrained')
            closure = ('Mie extinction and terminal outcomes; conditional finite-particle angular/path law' if cond
                       else self.particle_event_model + '; configured unconditional event law')
            cnt = T['counters']
            counts = causal.conditional_counts(cnt)
            res.detector_particle_path_diagnostics = causal.diagnostics_dict(
                cacc.tally, self._sym(det_sum), closure, counts, CONDITIONAL_MAX_TRIALS)
            res.production_failure_diagnostics['samples'] = _failure_samples(cacc)
        return res

    def _adaptive_plan(self, requested, ray_id_start, target_detector_score, min_rays, max_rays,
                       min_rear_channel_score, min_h170_score):
        """Ray budget and stop targets of simulate, with the reference's
        argument checks (CLARITAS.py:7188-7227): history ids are the
        reference's uint32 ray ids, so a range past 2^32 is refused here (the
        64-bit range is open through run_tallies / clr_run)."""
        if ray_id_start < 0 or ray_id_start + requested > 2 ** 32:
            raise ValueError('ray_id_start range exceeds uint32 global IDs')
        plan = dict(target=None, rear=None, h170=None, maxN=requested, minN=requested)
        if target_detector_score is not None:
            target = float(target_detector_score)
            if target <= 0:
                raise ValueError('target_detector_score must be >0')
            rear = None if min_rear_channel_score is None else float(min_rear_channel_score)
            h170 = None if min_h170_score is None else float(min_h170_score)
            if rear is not None and rear <= 0:
                raise ValueError('min_rear_channel_score must be >0')
            if h170 is not None and h170 <= 0:
                raise ValueError('min_h170_score must be >0')
            maxN = int(max_rays if max_rays is not None else max(requested, self.statistics_batch_rays))
            minN = int(min_rays if min_rays is not None else min(self.statistics_batch_rays, maxN))
            if minN < 0 or maxN < minN:
                raise ValueError('invalid min_rays/max_rays')
            plan.update(target=target, rear=rear, h170=h170, maxN=maxN, minN=minN)
        if plan['maxN'] < 1:
            raise ValueError('number of rays must be positive')
        if ray_id_start + plan['maxN'] > 2 ** 32:
            raise ValueError('global ray range exceeds uint32')
        return plan

    def _run_adaptive(self, run, ray_id_start, plan, stability_l1_tolerance, stability_window):
        """Batches of statistics_batch_rays until the selected detector score
        reaches its target (and the optional rear-channel, H170 and L1
        stability criteria hold) or max_rays is spent; the stop reason names
        the criteria, as CLARITAS.py:8818-8881 does.  `run(b, offset)`
        traces b histories from history id `offset` and returns
        (tallies, records, ms)."""
        total, ms, batches, pbatches, group_n = 0, 0.0, [], [], []
        acc, prev, recent = None, None, []
        stop_reason = 'fixed_n_rays'
        maxN, minN, target = plan['maxN'], plan['minN'], plan['target']
        rear, h170 = plan['rear'], plan['h170']
        sym = np.zeros(len(DETECTOR_ANGLES_DEG))
        while total < maxN:
            b = min(self.statistics_batch_rays, maxN - total)
            Tb, _, msb = run(b, ray_id_start + total)
            ms += msb
            batches.append(Tb['det'].sum(axis=0))
            pbatches.append(Tb['photodiode'].sum(axis=0))
            group_n.append(b)
            acc = Tb if acc is None else self._merge(acc, Tb)
            total += b
            sym = self._selected(np.stack(batches).sum(axis=0), np.stack(pbatches).sum(axis=0))
            cum = _norm(sym)
            if prev is not None:
                recent.append(float(np.abs(cum - prev).sum()))
            prev = cum
            if total < minN or sym.sum() < target:
                continue
            met = (rear is None or float(sym[15:18].min()) >= rear) and (h170 is None or float(sym[17]) >= h170)
            if stability_l1_tolerance is not None:
                w = max(1, int(stability_window))
                met = met and len(recent) >= w and max(recent[-w:]) <= float(stability_l1_tolerance)
            if met:
                names = ['target_detector_score']
                names += ['rear_channels'] if rear is not None else []
                names += ['H170'] if h170 is not None else []
                names += ['stability'] if stability_l1_tolerance is not None else []
                stop_reason = '_and_'.join(names)
                break
        if total >= maxN:
            if sym.sum() < target:
                stop_reason = 'max_rays_before_target'
            elif rear is not None and float(sym[15:18].min()) < rear:
                stop_reason = 'max_rays_before_rear_channel_target'
            elif h170 is not None and float(sym[17]) < h170:
                stop_reason = 'max_rays_before_H170_target'
            elif stop_reason.startswith('fixed'):
                stop_reason = 'max_rays_before_stability' if stability_l1_tolerance is not None else 'max_rays'
        acc['n_groups'] = len(batches)
        return acc, batches, pbatches, np.asarray(group_n), total, ms, stop_reason

    @staticmethod
    def _merge(a, b):
        out = dict(a)
        for k in ('aux', 'incidence_sum', 'status', 'failures', 'failure_subcodes', 'counters', 'ic_hist', 'stray',
                  'face_stray'):
            out[k] = a[k] + b[k]
        return out

    def _sym(self, det):
        """Symmetry score 0.5(native+mirror) per kind, CLARITAS.py:8788-8792."""
        if self.symmetry_average_exact:
            return 0.5 * (det[0].astype(float) + det[1].astype(float))
        return det[0].astype(float)

    def _pair(self, a, b):
        """Symmetry-averaged (or native) pair of per-channel arrays."""
        return 0.5 * (np.asarray(a, float) + np.asarray(b, float)) if self.symmetry_average_exact \
            else np.asarray(a, float)

    def _selected(self, det, pd):
        """The score the reference selects: hardware counts for the ideal
        aperture, photodiode-weighted hits otherwise (CLARITAS.py:8798-8802)."""
        if self.photodiode_response_model == 'ideal_hard_aperture':
            return self._sym(det)
        return self._pair(pd[0], pd[1])

    def _population_fields(self):
        """Population-balance result fields of the unaggregated PSD
        (CLARITAS.py:3627-3631, 6878-6895 with exposure 0): no reduction, no
        events, unchanged area; the largest bin's share of the normalised
        input weights."""
        w = np.asarray(getattr(self, 'last_psd_weights', [1.0]), dtype=np.float64)
        return dict(
            aggregation_model=self.aggregation_model,
            aggregation_collision_exposure=self.aggregation_collision_exposure,
            aggregation_number_reduction_fraction=0.0,
            aggregation_mass_conservation_error=0.0,
            aggregation_projected_area_ratio=1.0,
            aggregation_largest_bin_mass_fraction=float(w[-1] / w.sum()) if w.sum() > 0 else 0.0,
            aggregation_overflow_event_fraction=0.0,
            aggregation_steps=0,
            population_number_reduction_fraction=0.0,
            fragmentation_breakup_exposure=self.fragmentation_breakup_exposure,
            fragmentation_reference_diameter_m=self.fragmentation_reference_diameter_m,
            fragmentation_size_exponent=self.fragmentation_size_exponent,
            fragmentation_underflow_event_fraction=0.0,
            aggregation_events_per_initial_particle=0.0,
            fragmentation_events_per_initial_particle=0.0,
        )

    def _result(self, mat, T, det, batches, pbatches, group_n, requested, total, adaptive, stop_reason, ms):
        cnt = T['counters']
        ci = abi.C_IDX
        native, mirror = det[0].astype(np.int64), det[1].astype(np.int64)
        lnat, lmir = det[2].astype(np.int64), det[3].astype(np.int64)
        sym = self._sym(det)                      # hardware counts
        pd = np.stack(pbatches).sum(axis=0)       # photodiode-weighted [2][18]
        pd_sym = self._pair(pd[0], pd[1])
        sel = self._selected(det, pd)
        lsym = 0.5 * (lnat + lmir) if self.symmetry_average_exact else lnat.astype(float)
        hw_batches = [self._sym(b) for b in batches]
        sel_batches = [self._selected(b, p) for b, p in zip(batches, pbatches)]
        lg_batches = [(0.5 * (b[2] + b[3]) if self.symmetry_average_exact else b[2].astype(float)) for b in batches]
        se = _jackknife_normalized_se(sel_batches)
        hw_se = _jackknife_normalized_se(hw_batches)
        lse = _jackknife_normalized_se(lg_batches)
        stray, face = T['stray'], T['face_stray']
        stray_sym, face_sym = self._pair(stray[0], stray[1]), self._pair(face[0], face[1])
        frac = lambda a: np.divide(a, sym, out=np.zeros(18), where=sym > 0)
        status = T['status']
        fails = T['failures']
        n_fail = int(fails.sum())
        entered = int(cnt[ci['entered']])
        den = max(entered, 1)
        nan = float('nan')

        def per_entered(k):
            return cnt[ci[k]] / den if entered else nan

        tables = self.tables(len(mat.diameter_m))
        mres = self.meridional_result if self.sediment_transport_model == MERIDIONAL_MODEL else None
        mus = tables['mu_scalars']
        vg = vortex_surface_geometry(self.free_surface_model, self.zmean, self.rin,
                                     0.0 if self.free_surface_model == 'flat' else self.vortex_delta_h,
                                     self.vortex_core_radius, self.gravity)
        inc = T['incidence_sum']
        inc_sym = 0.5 * (inc[0] + inc[1]) if self.symmetry_average_exact else inc[0]
        mean_inc = np.divide(inc_sym, sym, out=np.full(18, np.nan), where=sym > 0)
        naux = self.aux_detector_angles_deg.size
        aux = T['aux'].sum(axis=0)[:, :naux]
        batch_rows = [{'batch': g + 1, 'batch_rays': int(group_n[g]),
                       'batch_native_hits': int(batches[g][0].sum()), 'batch_mirror_hits': int(batches[g][1].sum()),
                       'batch_hardware_candidate_score': float(hw_batches[g].sum())}
                      for g in range(len(batches))]
        return ForwardResult(
            detector_angles_deg=DETECTOR_ANGLES_DEG.copy(),
            raw_hits=native.copy(),
            normalized_response=_norm(sel),
            native_exact_hits=native.copy(),
            mirror_exact_hits=mirror.copy(),
            symmetry_exact_scores=sel.copy(),
            normalized_native_exact_response=_norm(native.astype(float)),
            normalized_mirror_exact_response=_norm(mirror.astype(float)),
            normalized_symmetry_exact_response=_norm(sel),
            normalized_response_jackknife_se=se,
            hardware_native_hits=native.copy(),
            hardware_mirror_hits=mirror.copy(),
            hardware_symmetry_scores=sym.copy(),
            normalized_hardware_native_response=_norm(native.astype(float)),
            normalized_hardware_mirror_response=_norm(mirror.astype(float)),
            normalized_hardware_symmetry_response=_norm(sym),
            hardware_normalized_jackknife_se=hw_se,
            legacy_native_channel_counts=lnat,
            legacy_mirror_channel_counts=lmir,
            legacy_symmetry_channel_scores=lsym,
            normalized_legacy_native_response=_norm(lnat.astype(float)),
            normalized_legacy_mirror_response=_norm(lmir.astype(float)),
            normalized_legacy_symmetry_response=_norm(lsym),
            legacy_normalized_jackknife_se=lse,
            legacy_acceptance_deg=self.legacy_acceptance_deg,
            valid_exit_count=int(cnt[ci['valid_air']]),
            numerical_failure_count=n_fail,
            numerical_failure_fraction=n_fail / max(total, 1),
            production_failure_count=n_fail,
            production_failure_code_counts={abi.PRODUCTION_FAILURE_NAMES[i]: int(fails[i])
                                            for i in range(1, abi.N_FAILURE) if fails[i]},
            **_failure_summary(T, fails, total),
            n_rays=int(total),
            requested_n_rays=int(requested),
            stopped_adaptively=bool(adaptive and stop_reason == 'target_detector_score'),
            stop_reason=stop_reason,
            particle_event_model=self.particle_event_model,
            particle_surface_rms_slope_deg=self.particle_surface_rms_slope_deg,
            particle_surface_beckmann_alpha=self.particle_surface_beckmann_alpha,
            wave_surface_rms_height_nm=self.wave_surface_rms_height_m * 1e9,
            wave_spheroid_aspect_ratio=self.wave_spheroid_aspect_ratio,
            wave_spheroid_orientation_model=self.wave_spheroid_orientation_model,
            wave_spheroid_orientation_kappa=self.wave_spheroid_orientation_kappa,
            wave_coherent_fraction_nominal=float(mus[7]),
            mu_geom_per_m=float(mus[2]),
            mu_wave_per_m=float(mus[5]),
            mu_event_per_m=float(mus[6]),
            mu_sca_per_m=float(mus[3]),
            mu_ext_per_m=float(mus[4]),
            geometric_optical_depth_diameter=float(mus[2]) * 2 * self.rin,
            optical_depth_diameter=float(mus[6]) * 2 * self.rin,
            **self._population_fields(),
            mean_interactions=per_entered('sum_ic'),
            mean_geometric_interactions=per_entered('sum_gic'),
            mean_wave_interactions=per_entered('sum_wic'),
            mean_coherent_wave_interactions=per_entered('sum_wcoh'),
            wave_coherent_fraction_nominal_cuda=float(mus[7]),
            mean_morphology_wave_interactions=((cnt[ci['sum_wic']] - cnt[ci['sum_wcoh']]) / den) if entered else nan,
            morphology_fraction_of_wave_interactions=(
                (cnt[ci['sum_wic']] - cnt[ci['sum_wcoh']]) / cnt[ci['sum_wic']] if cnt[ci['sum_wic']] > 0 else 0.0),
            median_interactions=_histogram_median(T['ic_hist']),
            ballistic_fraction=per_entered('ballistic'),
            mean_fresnel_reflections=per_entered('sum_pr'),
            entry_reflection_fraction=per_entered('rays_per'),
            internal_reflection_fraction=per_entered('rays_pir'),
            particle_absorption_fraction=per_entered('rays_pabs'),
            particle_absorbed_ray_count=int(cnt[ci['rays_pabs']]),
            mean_cell_inner_reflections=per_entered('sum_cir'),
            mean_cell_outer_reflections=per_entered('sum_cor'),
            cell_tir_fraction=per_entered('rays_ctir'),
            mean_top_surface_reflections=per_entered('sum_topr'),
            top_surface_tir_fraction=per_entered('rays_toptir'),
            mean_top_surface_interactions=((cnt[ci['sum_topr']] + status[6]) / den) if entered else nan,
            free_surface_model=self.free_surface_model,
            vortex_wall_center_height_difference_mm=self.vortex_delta_h * 1e3,
            vortex_core_radius_mm=self.vortex_core_radius * 1e3,
            vortex_asymptotic_height_parameter_mm=vg['h_inf'] * 1e3,
            vortex_area_mean_uncorrected_rise_mm=vg['h_bar'] * 1e3,
            vortex_equivalent_core_rpm=vg['equivalent_core_rpm'],
            vortex_equivalent_rigid_body_rpm=vg['equivalent_rigid_body_rpm'],
            vortex_center_surface_z_mm=vg['center_z'] * 1e3,
            vortex_wall_surface_z_mm=vg['wall_z'] * 1e3,
            sediment_transport_model=self.sediment_transport_model,
            sediment_majorant_geom_per_m=float(tables['majorants'][0]),
            sediment_majorant_wave_per_m=float(tables['majorants'][1]),
            sediment_majorant_total_per_m=float(tables['majorants'][2]),
            sediment_eddy_diffusivity_m2_s=self.sediment_eddy_diffusivity,
            sediment_max_bin_density_scale=self._max_bin_density_scale(tables),
            meridional_circulation_ratio=self.meridional_circulation_ratio,
            meridional_peak_speed_m_s=self.meridional_peak_speed,
            meridional_grid_nr=self.meridional.nr if mres else 2,
            meridional_grid_nz=self.meridional.nz if mres else 2,
            meridional_solver_iterations=int(mres['iterations']) if mres else 0,
            meridional_solver_max_residual=float(mres['max_residual']) if mres else 0.0,
            mean_bottom_inner_reflections=per_entered('sum_bir'),
            mean_bottom_outer_reflections=per_entered('sum_bor'),
            bottom_tir_fraction=per_entered('rays_btir'),
            top_exit_fraction=status[6] / max(total, 1),
            bottom_exit_fraction=status[7] / max(total, 1),
            source_bore_surface_model=self.source_bore_surface_model,
            detector_bore_surface_model=self.detector_bore_surface_model,
            bore_refractive_index=self.bore_refractive_index,
            ring_inner_face_surface_model=self.ring_inner_face_surface_model,
            ring_inner_face_refractive_index=self.ring_inner_face_refractive_index,
            max_ring_face_bounces=self.max_ring_face_bounces,
            source_collimator_acceptance_fraction=(total - status[3]) / max(total, 1),
            source_direct_acceptance_fraction=(total - status[3] - cnt[ci['src_reflected_accept']]) / max(total, 1),
            source_reflected_acceptance_fraction=cnt[ci['src_reflected_accept']] / max(total, 1),
            source_bore_interactions_per_ray=cnt[ci['src_bore_interactions']] / max(total, 1),
            source_bore_reflections_per_ray=cnt[ci['src_bore_reflections']] / max(total, 1),
            hardware_native_direct_hits=native - stray[0],
            hardware_mirror_direct_hits=mirror - stray[1],
            hardware_symmetry_direct_scores=sym - stray_sym,
            hardware_native_stray_hits=stray[0].copy(),
            hardware_mirror_stray_hits=stray[1].copy(),
            hardware_symmetry_stray_scores=stray_sym,
            detector_stray_fraction_by_channel=frac(stray_sym),
            detector_stray_detection_fraction=float(stray_sym.sum()) / max(total, 1),
            hardware_native_ring_face_stray_hits=face[0].copy(),
            hardware_mirror_ring_face_stray_hits=face[1].copy(),
            hardware_symmetry_ring_face_stray_scores=face_sym,
            detector_ring_face_stray_fraction_by_channel=frac(face_sym),
            detector_ring_face_stray_detection_fraction=float(face_sym.sum()) / max(total, 1),
            photodiode_response_model=self.photodiode_response_model,
            sfh213_half_angle_deg=10.0,
            sfh213_directional_fit_exponent=3.169925001442312,
            photodiode_native_weighted_scores=pd[0].copy(),
            photodiode_mirror_weighted_scores=pd[1].copy(),
            photodiode_symmetry_weighted_scores=pd_sym,
            normalized_photodiode_native_response=_norm(pd[0]),
            normalized_photodiode_mirror_response=_norm(pd[1]),
            normalized_photodiode_symmetry_response=_norm(pd_sym),
            photodiode_relative_response_by_channel=frac(pd_sym),
            detector_mean_incidence_angle_deg_by_channel=mean_inc,
            hardware_candidate_detection_fraction=float(sym.sum()) / max(total, 1),
            entered_water_fraction=entered / max(total, 1),
            air_exit_fraction=int(cnt[ci['valid_air']]) / max(total, 1),
            axial_loss_fraction=(status[6] + status[7]) / max(total, 1),
            detector_detection_fraction=float(sym.sum()) / max(total, 1),
            native_detector_detection_fraction=float(native.sum()) / max(total, 1),
            mirror_detector_detection_fraction=float(mirror.sum()) / max(total, 1),
            legacy_channel_score_per_ray=float(lsym.sum()) / max(total, 1),
            symmetry_variance_reduction=self.symmetry_average_exact,
            batch_diagnostics=batch_rows,
            particle_diagnostics={'diameter_m': mat.diameter_m, 'weights': mat.weights, **tables},
            aux_detector_angles_deg=self.aux_detector_angles_deg.copy(),
            aux_native_hits=aux[0].astype(np.int64),
            aux_mirror_hits=aux[1].astype(np.int64),
            aux_symmetry_scores=(0.5 * (aux[0] + aux[1])) if self.symmetry_average_exact else aux[0].astype(float),
            status_counts=status.copy(),
            work_counters={k: int(cnt[i]) for k, i in ci.items()},
            device_ms=ms,
        )
