"""ctypes mirror of include/claritas.h.

The field order and types follow the C header exactly; `check_sizes` compares
every structure with the library's own `clr_sizeof`, the ctypes counterpart of
the reference's launch-arity check (validate_claritas.py:972-1001).
"""
from __future__ import annotations

import ctypes as C

ABI_VERSION = 11
N_CHANNELS = 18
MAX_GROUPS = 64
MAX_AUX = 4
N_COUNTERS = 64
N_ICHIST = 64
N_STATUS = 9
N_FAILURE = 19
# failure subcode tally slots (include/claritas.h CLR_SUBCODE_VALUES; the
# last slot counts any other subcode), CLARITAS.py:464-521
N_SUBCODE = 64
SUBCODE_VALUES = (0, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37,
                  38, 39, 40, 41, 1001, 1002, 1003, 1004, 1005, 1101, 1102, 1103, 1104, 1105, 1106, 1107, 2001,
                  2002, 2101, 2102, 2103, 2104, 3001, 3002, 3003, 3004, 3005, 3006, 4001, 9001, 9002, 9003)

# error codes (claritas.h CLR_E_*)
E_ARG, E_HIP, E_NOMEM, E_STATE, E_RANGE, E_UNSUPPORTED, E_NOT_CONVERGED = -1, -2, -3, -4, -5, -6, -7

# event laws (CLARITAS.py:5543-5550); 6 = plain Mie (no reference code)
EVENT_MODEL_CODES = {
    'geometric': 0,
    'mie_qsca': 1,
    'mie_qext': 2,
    'hybrid_mie_excess': 3,
    'hybrid_mie_exact_rate': 4,
    'plain_mie': 6,
}
WEIGHT_MODE_CODES = {'mass_fraction': 0, 'number_fraction': 1}
SOURCE_ANGULAR_MODEL_CODES = {
    'legacy_polar_pdf': 0,
    'beta_radiance_3d': 1,
    'thesis_planar_beta': 2,
    'collimated_disk': 3,
}
FREE_SURFACE_MODEL_CODES = {
    'flat': 0,
    'parabolic_vortex': 1,
    'localized_scully_vortex': 2,
    'rankine_vortex': 3,
}
SEDIMENT_TRANSPORT_MODEL_CODES = {
    'uniform': 0,
    'stokes_drift_diffusion': 1,
    'schiller_naumann_drift_diffusion': 2,
    'grid_per_bin': 3,
    'schiller_naumann_meridional_advection_diffusion': 3,
    'rz_field': 4,
    'xyz_field': 5,
}

# clr_probe modes (claritas.h CLR_PROBE_*)
PROBE_MODES = {'surface': 0, 'wall': 1, 'offset': 2, 'sediment': 3, 'detector': 4, 'sampler': 5, 'source': 6,
               'mie_theta': 7, 'conditional': 8}

# counter indices (claritas.h enum)
COUNTER_NAMES = [
    'histories', 'entered', 'valid_air', 'sum_ic', 'sum_gic', 'sum_wic', 'sum_wcoh',
    'ballistic', 'sum_pr', 'rays_per', 'rays_pir', 'rays_pabs', 'sum_cir', 'sum_cor',
    'rays_ctir', 'sum_topr', 'rays_toptir', 'sum_bir', 'sum_bor', 'rays_btir',
    'n_flight', 'n_null', 'n_scat', 'n_iface', 'n_abs', 'n_exit', 'cond_trials',
    'cond_reject', 'cond_inaccessible', 'rng_draws', 'n_steps', 'src_bore_interactions', 'src_bore_reflections',
    'src_reflected_accept', 'cond_accepted_geometric', 'cond_accepted_coherent', 'cond_accepted_morphology',
    'cond_accepted_absorptions', 'cond_accepted_scatterings', 'cond_numerical_failures', 'cond_retry_exhaustions',
]
C_IDX = {name: i for i, name in enumerate(COUNTER_NAMES)}

# production failure names, CLARITAS.py:424-444
PRODUCTION_FAILURE_NAMES = {
    0: 'OK', 1: 'PARTICLE_MICROSURFACE_NUMERICAL', 2: 'PARTICLE_MICROSURFACE_WATCHDOG',
    3: 'PARTICLE_MACRO_GEOMETRY_FAILURE', 4: 'PARTICLE_MACRO_BOUNCE_LIMIT',
    5: 'ACRYLIC_ANNULUS_NO_INTERSECTION', 6: 'ACRYLIC_ANNULUS_BOUNCE_LIMIT',
    7: 'BOTTOM_DISC_GEOMETRY_FAILURE', 8: 'BOTTOM_DISC_BOUNCE_LIMIT',
    9: 'WATER_BOUNDARY_NO_INTERSECTION', 10: 'TRACE_ITERATION_LIMIT',
    11: 'NONFINITE_POSITION', 12: 'NONFINITE_DIRECTION', 13: 'INVALID_MEDIUM_STATE',
    14: 'INVALID_EVENT_STATE', 15: 'OTHER_EXPLICIT_FAILURE', 16: 'SPHEROID_GEOMETRY_FAILURE',
    17: 'HEADSPACE_BOUNDARY_NO_INTERSECTION', 18: 'SPATIAL_MAJORANT_VIOLATION',
}

D = C.c_double
I32 = C.c_int32
U64 = C.c_uint64
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class Config(C.Structure):
    _fields_ = [
        ('n_water', D), ('n_acrylic', D), ('n_air', D), ('bottom_external_n', D),
        ('rin', D), ('rout', D), ('zmin', D), ('zmean', D), ('zwall_top', D),
        ('bottom_disc_thickness', D),
        ('ring_in', D), ('throat_out', D), ('ring_out', D), ('source_r', D), ('throat_r', D),
        ('counterbore_r', D),
        ('vortex_delta_h', D), ('vortex_core_radius', D),
        ('alpha1', D), ('alpha2', D), ('beam_sigma', D), ('beam_radius', D),
        ('legacy_accept_deg', D),
        ('wavelength', D),
        ('rough_alpha', D),
        ('aux_angles_deg', D * MAX_AUX),
        ('field_zmax', D),
        ('field_xy_half', D),
        ('spheroid_aspect', D),
        ('spheroid_kappa', D),
        ('free_surface_model', I32), ('source_angular_model', I32), ('event_model', I32),
        ('sediment_model', I32), ('max_particle_bounces', I32), ('max_cell_bounces', I32),
        ('max_iterations', I32), ('conditional_max_trials', I32), ('n_aux', I32),
        ('n_groups', I32), ('field_n0', I32), ('field_n1', I32), ('field_n2', I32),
        ('field_nbins', I32),
        ('spheroid_orientation', I32), ('pad_', I32),
        ('bore_n', D), ('ring_face_n', D), ('source_bore_model', I32), ('detector_bore_model', I32),
        ('ring_face_model', I32), ('photodiode_model', I32), ('max_bore_bounces', I32),
        ('max_ring_face_bounces', I32),
    ]


class Particles(C.Structure):
    _fields_ = [
        ('diameter', PD), ('weight', PD), ('n_real', PD), ('k_imag', PD),
        ('concentration', D), ('density', D), ('n_medium', D), ('wavelength', D),
        ('hybrid_wave_scale', D), ('wave_rms_height_m', D),
        ('n', I32), ('n_theta', I32), ('weight_mode', I32), ('event_model', I32),
    ]


class Sediment(C.Structure):
    _fields_ = [
        ('particle_density', D), ('water_density', D), ('water_viscosity', D), ('gravity', D),
        ('eddy_diffusivity', D), ('n_quad', I32), ('n_radial', I32),
    ]


class MieOut(C.Structure):
    _fields_ = [
        ('size_parameter', PD), ('qext', PD), ('qsca', PD), ('qabs', PD), ('qback', PD),
        ('asym_g', PD), ('albedo', PD), ('nstop', PI), ('theta', PD), ('phase_pdf', PD),
        ('phase_cdf', PD),
    ]


class TablesOut(C.Structure):
    _fields_ = [
        ('radius', PD), ('number_density', PD), ('mu_geom_bin', PD), ('mu_sca_bin', PD),
        ('mu_ext_bin', PD), ('trace_geom_bin', PD), ('trace_wave_bin', PD), ('geom_cdf', PD),
        ('wave_cdf', PD), ('sed_A', PD), ('sed_B', PD), ('sed_norm', PD), ('sed_settling', PD),
        ('sed_radial_log', PD), ('mu_scalars', D * 8), ('majorants', D * 4),
    ]


class Tallies(C.Structure):
    _fields_ = [
        ('det', U64 * N_CHANNELS * 4 * MAX_GROUPS),
        ('aux', U64 * MAX_AUX * 2 * MAX_GROUPS),
        ('aux_window', U64 * MAX_AUX * 2 * MAX_GROUPS),
        ('group_histories', U64 * MAX_GROUPS),
        ('stray', U64 * N_CHANNELS * 2),
        ('face_stray', U64 * N_CHANNELS * 2),
        ('incidence_sum', D * N_CHANNELS * 2),
        ('photodiode', D * N_CHANNELS * 2 * MAX_GROUPS),
        ('status', U64 * N_STATUS),
        ('failures', U64 * N_FAILURE),
        ('failure_subcodes', U64 * N_SUBCODE),
        ('counters', U64 * N_COUNTERS),
        ('ic_hist', U64 * N_ICHIST),
        ('n_groups', I32),
        ('pad_', I32),
    ]


class Record(C.Structure):
    _fields_ = [
        ('x', C.c_float), ('y', C.c_float), ('z', C.c_float),
        ('vx', C.c_float), ('vy', C.c_float), ('vz', C.c_float),
        ('status', I32), ('native_id', I32), ('mirror_id', I32), ('interactions', I32),
        ('failure', I32), ('draws', I32),
    ]


class BatchItem(C.Structure):
    _fields_ = [('config', I32), ('pad_', I32), ('n_histories', U64), ('seed', U64),
                ('history_offset', U64)]


class Launch(C.Structure):
    _fields_ = [('block', I32), ('blocks_per_cu', I32), ('chunk', I32), ('queue_order', I32)]


class MlpShape(C.Structure):
    _fields_ = [('n_members', I32), ('in_dim', I32), ('hidden', I32), ('out_dim', I32),
                ('n_hidden_layers', I32), ('pad_', I32)]


class Meridional(C.Structure):
    _fields_ = [('circulation_ratio', D), ('tolerance', D), ('sor', D), ('nr', I32), ('nz', I32),
                ('max_iterations', I32), ('check_interval', I32)]


class MeridionalOut(C.Structure):
    _fields_ = [('grid', PD), ('residual', PD), ('scale_max', PD), ('scale_min', PD), ('point_scales', PD),
                ('active', C.POINTER(C.c_uint8)), ('coeff', PD), ('grid_zmax', D), ('vtheta_peak', D),
                ('peak_speed', D), ('iterations', I32), ('pad_', I32)]


def meridional_out(n, nr, nz, coeff=False):
    """Allocate host outputs for one meridional solve; returns (struct, arrays)."""
    import numpy as np

    arrs = {
        'grid': np.zeros((n, nr, nz)), 'residual': np.zeros(n), 'scale_max': np.zeros(n),
        'scale_min': np.zeros(n), 'point_scales': np.zeros((n, 4)), 'active': np.zeros((nr, nz), dtype=np.uint8),
    }
    if coeff:
        arrs['coeff'] = np.zeros((5, n, nr, nz))
    o = MeridionalOut()
    for k, v in arrs.items():
        ct = C.POINTER(C.c_uint8) if v.dtype == np.uint8 else PD
        setattr(o, k, v.ctypes.data_as(ct))
    return o, arrs


PF = C.POINTER(C.c_float)
PU32 = C.POINTER(C.c_uint32)


class LightConfig(C.Structure):
    _fields_ = [('max_rays', U64), ('nx', I32), ('ny', I32), ('nz', I32), ('representative_rays', I32),
                ('max_segments', I32), ('heatmap_size', I32), ('zmax', D)]


class LightOut(C.Structure):
    _fields_ = [('track_length', PF), ('post_scatter_track_length', PF), ('scatter_event_weight', PF),
                ('scatter_event_count', PU32), ('absorption_event_count', PU32),
                ('traj_x0', PF), ('traj_y0', PF), ('traj_z0', PF), ('traj_x1', PF), ('traj_y1', PF), ('traj_z1', PF),
                ('traj_order', PI), ('traj_event', PI), ('traj_count', PI), ('traj_overflow', PI),
                ('segment_total', D), ('deposited_total', D), ('segment_count', U64),
                ('scatter_event_total', U64), ('absorption_event_total', U64), ('error_count', C.c_uint32),
                ('boundary_adjustment_count', C.c_uint32), ('heatmap_xy', PF), ('heatmap_zy', PF),
                ('outside', C.c_uint32 * 8)]


LIGHT_GRIDS = ('track_length', 'post_scatter_track_length', 'scatter_event_weight', 'scatter_event_count',
               'absorption_event_count')
LIGHT_TRAJ = ('traj_x0', 'traj_y0', 'traj_z0', 'traj_x1', 'traj_y1', 'traj_z1', 'traj_order', 'traj_event')
LIGHT_SCALARS = ('segment_total', 'deposited_total', 'segment_count', 'scatter_event_total',
                 'absorption_event_total', 'error_count', 'boundary_adjustment_count')


def light_out(nx, ny, nz, rep, max_segments, vis=0):
    """Allocate host outputs of one light-diagnostic call; returns (struct, arrays)."""
    import numpy as np

    arrs = {}
    if vis > 0:
        arrs['heatmap_xy'] = np.zeros((vis, vis), dtype=np.float32)
        arrs['heatmap_zy'] = np.zeros((vis, vis), dtype=np.float32)
    for k in LIGHT_GRIDS:
        arrs[k] = np.zeros((nz, ny, nx), dtype=np.uint32 if k.endswith('count') else np.float32)
    for k in LIGHT_TRAJ:
        arrs[k] = np.zeros((rep, max_segments), dtype=np.int32 if k in ('traj_order', 'traj_event') else np.float32)
    arrs['traj_count'] = np.zeros(rep, dtype=np.int32)
    arrs['traj_overflow'] = np.zeros(rep, dtype=np.int32)
    o = LightOut()
    for k, v in arrs.items():
        ct = {np.dtype(np.float32): PF, np.dtype(np.uint32): PU32, np.dtype(np.int32): PI}[v.dtype]
        setattr(o, k, v.ctypes.data_as(ct) if v.size else None)
    return o, arrs


def light_scalars(o: LightOut) -> dict:
    d = {k: getattr(o, k) for k in LIGHT_SCALARS}
    d['outside'] = [int(v) for v in o.outside]
    return d


# clr_causal_record field by field (include/claritas.h): the provenance of
# one hardware-detected history, read as a numpy structured array
CAUSAL_INT_FIELDS = (
    'native_id', 'mirror_id', 'ic', 'gic', 'wic', 'wcoh', 'per', 'pir', 'pabs', 'acrylic_reflections',
    'acrylic_returns', 'events_after_return', 'first_type', 'last_type', 'last_order', 'first_backscatter',
    'first_source_hemisphere', 'first_bin', 'last_bin', 'first_geometric_bin', 'last_geometric_bin',
    'first_wave_bin', 'last_wave_bin', 'branch_transitions', 'type_transitions', 'small_deflections',
    'medium_deflections', 'large_deflections', 'max_type', 'max_order', 'max_bin', 'max_macro_bounces',
    'max_micro_order')
CAUSAL_FLOAT_FIELDS = (
    'first_deflection', 'last_deflection', 'internal_path', 'internal_opl', 'water_path_after_last',
    'cumulative_deflection', 'cumulative_deflection_sq', 'max_deflection', 'turn_x', 'turn_y', 'turn_z',
    'first_x', 'first_y', 'first_z', 'last_x', 'last_y', 'last_z', 'mean_r', 'mean_z', 'min_r', 'max_r',
    'min_z', 'max_z', 'water_path_first', 'water_path_last', 'max_internal_path', 'max_path_over_diameter',
    'max_incidence', 'max_exit')


def causal_dtype():
    import numpy as np
    return np.dtype([('history', '<u8')] + [(k, '<i4') for k in CAUSAL_INT_FIELDS] +
                    [(k, '<f4') for k in CAUSAL_FLOAT_FIELDS])


class CausalRecord(C.Structure):
    _fields_ = [('history', C.c_uint64)] + [(k, C.c_int32) for k in CAUSAL_INT_FIELDS] + \
        [(k, C.c_float) for k in CAUSAL_FLOAT_FIELDS]


class FailureSample(C.Structure):
    _fields_ = [('history', C.c_uint64), ('code', C.c_int32), ('subcode', C.c_int32),
                ('trace_iteration', C.c_int32), ('pad_', C.c_int32)] + [(k, C.c_float) for k in
                                                                       ('x', 'y', 'z', 'vx', 'vy', 'vz')]


def failure_dtype():
    import numpy as np
    return np.dtype([('history', '<u8'), ('code', '<i4'), ('subcode', '<i4'), ('trace_iteration', '<i4'),
                     ('pad_', '<i4')] + [(k, '<f4') for k in ('x', 'y', 'z', 'vx', 'vy', 'vz')])


class CausalOut(C.Structure):
    _fields_ = [('records', C.c_void_p), ('capacity', C.c_uint64), ('count', C.c_uint64),
                ('range_start', C.c_uint64), ('range_count', C.c_uint64), ('failures', C.c_void_p),
                ('failure_capacity', C.c_uint64), ('failure_count', C.c_uint64)]


STRUCT_IDS = {
    0: Config, 1: Particles, 2: Sediment, 3: MieOut, 4: TablesOut, 5: Tallies, 6: Record,
    7: BatchItem, 8: Launch, 9: Meridional, 10: MeridionalOut, 11: LightConfig, 12: LightOut,
    13: CausalRecord, 14: FailureSample,
}


def check_sizes(lib) -> None:
    """Raise if any ctypes structure disagrees with the C library's sizeof."""
    lib.clr_sizeof.restype = C.c_size_t
    lib.clr_sizeof.argtypes = [C.c_int]
    for sid, cls in STRUCT_IDS.items():
        n = int(lib.clr_sizeof(sid))
        if n != C.sizeof(cls):
            raise RuntimeError(f'ABI mismatch for {cls.__name__}: C {n} vs ctypes {C.sizeof(cls)}')


def tallies_to_numpy(t: Tallies):
    """Copy a Tallies struct into a dict of numpy arrays."""
    import numpy as np

    g = int(t.n_groups)
    det = np.ctypeslib.as_array(t.det).reshape(MAX_GROUPS, 4, N_CHANNELS).astype(np.int64)
    aux = np.ctypeslib.as_array(t.aux).reshape(MAX_GROUPS, 2, MAX_AUX).astype(np.int64)
    auxw = np.ctypeslib.as_array(t.aux_window).reshape(MAX_GROUPS, 2, MAX_AUX).astype(np.int64)
    return {
        'n_groups': g,
        'det': det[:g].copy(),
        'aux': aux[:g].copy(),
        'aux_window': auxw[:g].copy(),
        'group_histories': np.ctypeslib.as_array(t.group_histories)[:g].astype(np.int64).copy(),
        'incidence_sum': np.ctypeslib.as_array(t.incidence_sum).reshape(2, N_CHANNELS).copy(),
        'stray': np.ctypeslib.as_array(t.stray).reshape(2, N_CHANNELS).astype(np.int64).copy(),
        'face_stray': np.ctypeslib.as_array(t.face_stray).reshape(2, N_CHANNELS).astype(np.int64).copy(),
        'photodiode': np.ctypeslib.as_array(t.photodiode).reshape(MAX_GROUPS, 2, N_CHANNELS)[:g].copy(),
        'status': np.ctypeslib.as_array(t.status).astype(np.int64).copy(),
        'failures': np.ctypeslib.as_array(t.failures).astype(np.int64).copy(),
        'failure_subcodes': np.ctypeslib.as_array(t.failure_subcodes).astype(np.int64).copy(),
        'counters': np.ctypeslib.as_array(t.counters).astype(np.int64).copy(),
        'ic_hist': np.ctypeslib.as_array(t.ic_hist).astype(np.int64).copy(),
    }


def records_to_numpy(recs, n):
    import numpy as np

    dt = np.dtype([('x', 'f4'), ('y', 'f4'), ('z', 'f4'), ('vx', 'f4'), ('vy', 'f4'),
                   ('vz', 'f4'), ('status', 'i4'), ('native_id', 'i4'), ('mirror_id', 'i4'),
                   ('interactions', 'i4'), ('failure', 'i4'), ('draws', 'i4')])
    buf = (C.c_char * (C.sizeof(Record) * n)).from_address(C.addressof(recs))
    return np.frombuffer(bytes(buf), dtype=dt, count=n)
