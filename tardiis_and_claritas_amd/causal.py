"""Detector-path, causal and H170 diagnostics of ClaritasForwardModel.simulate
(CLARITAS.py:7352-7613 accumulators, 8017-8480 per-chunk reductions,
8967-9440 the ``detector_particle_path_diagnostics`` dictionary).

The device keeps each history's causal provenance in the CAUSAL (or LIGHT)
variant of the transport kernel and appends one ``clr_causal_record`` per
history that reaches a hardware detector (clr_run_causal).  The host bins
those records per detector channel, native and mirror, with the symmetry
average the production estimator uses, into the reference's score tables.
Histories that reach no detector contribute to none of the tables, so only
the detected ones travel.  Runs are traced in history ranges of at most
``chunk`` histories; a range's records are binned and dropped before the
next, and the ranges' tallies add up to those of one clr_run call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

N_CH = abi.N_CHANNELS
H170 = 17  # the 170-degree channel

# CLARITAS.py:5325-5334
EVENT_TYPE_NAMES = {
    0: 'none',
    1: 'geometric_external_entry_reflection',
    2: 'geometric_surface_transmission_refraction',
    3: 'geometric_internal_reflection_path',
    4: 'coherent_mie_wave',
    5: 'morphology_wave_external_entry_reflection',
    6: 'morphology_wave_surface_transmission_refraction',
    7: 'morphology_wave_internal_reflection_path',
}
EVENT_CODES = tuple(range(1, 8))

# CLARITAS.py:7354-7380 (overlapping mechanism and exclusive order masks)
PATH_KEYS = (
    'any_particle', 'any_geometric', 'any_wave', 'any_coherent_wave', 'any_morphology_wave',
    'any_entry_reflection', 'any_internal_reflection', 'ballistic', 'single_interaction',
    'single_entry_reflection', 'multiple_interaction', 'interaction_order_2', 'interaction_order_3',
    'interaction_order_4plus', 'exactly_one_geometric', 'exactly_one_wave', 'single_coherent_wave_escape',
    'single_external_entry_reflection_escape', 'single_internal_reflection_escape',
    'single_surface_transmission_escape', 'mixed_geometric_wave', 'geometric_only_multiple',
    'wave_only_multiple', 'multiple_with_entry_reflection', 'multiple_with_internal_reflection',
)
# per-channel sums of history counts, CLARITAS.py:8050-8064
PATH_SUMS = ('interactions', 'geometric_interactions', 'wave_interactions', 'coherent_wave_interactions',
             'morphology_wave_interactions', 'entry_reflections', 'internal_reflections')
# CLARITAS.py:7448-7473
HISTORY_METRICS = (
    'cumulative_particle_deflection_deg', 'cumulative_particle_deflection_sq_deg2', 'max_particle_deflection_deg',
    'first_particle_r_m', 'first_particle_z_m', 'last_particle_r_m', 'last_particle_z_m',
    'particle_event_mean_r_m', 'particle_event_mean_z_m', 'particle_event_min_r_m', 'particle_event_max_r_m',
    'particle_event_min_z_m', 'particle_event_max_z_m', 'water_path_at_first_particle_event_m',
    'water_path_at_last_particle_event_m', 'particle_water_path_span_first_to_last_m', 'particle_radial_span_m',
    'particle_axial_span_m', 'particle_branch_transition_count', 'particle_event_type_transition_count',
    'particle_small_deflection_count', 'particle_medium_deflection_count', 'particle_large_deflection_count',
)
DEFLECTION_CLASSES = ('zero_particle', 'all_lt30', 'has_30_90_no_ge90', 'has_ge90')
BRANCH_CLASSES = ('zero_particle', 'geometric_only', 'wave_only', 'mixed_g_to_g', 'mixed_g_to_w', 'mixed_w_to_g',
                  'mixed_w_to_w')
WALL_CLASSES = ('zero_particle', 'particle_no_acrylic_return', 'particle_with_acrylic_return')
# bin edges, CLARITAS.py:7418-7420, 7500-7517, 7569-7574, 7582-7592
H170_ANGLE_EDGES = np.array([0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0001])
H170_CUMULATIVE_EDGES = np.array([0.0, 30.0, 60.0, 90.0, 120.0, 180.0, 240.0, 360.0, 540.0, 720.0, 1080.0,
                                  1800.0, 3600.0, 1.0e9])
H170_LENGTH_RATIO_EDGES = np.array([0.0, 1.0e-12, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 10.0, 1.0e9])
INTERFACE_ANGLE_EDGES = np.array([0.0, 15.0, 30.0, 45.0, 60.0, 75.0, 90.0001])
INTERFACE_ANGLE_LABELS = ['undefined', '0-15', '15-30', '30-45', '45-60', '60-75', '75-90']
CONDITIONAL_COUNT_NAMES = ('candidate_trials', 'outcome_rejections', 'inaccessible_trials', 'accepted_geometric',
                           'accepted_coherent', 'accepted_morphology', 'accepted_absorptions',
                           'accepted_scatterings', 'numerical_failures', 'retry_exhaustions')


# the device counters behind each name (claritas.h CLR_C_COND_*)
CONDITIONAL_COUNTERS = ('cond_trials', 'cond_reject', 'cond_inaccessible', 'cond_accepted_geometric',
                        'cond_accepted_coherent', 'cond_accepted_morphology', 'cond_accepted_absorptions',
                        'cond_accepted_scatterings', 'cond_numerical_failures', 'cond_retry_exhaustions')


def conditional_counts(counters):
    """The reference's conditional_sampler_counts (CLARITAS.py:9119-9125) from
    a run's counter array: all zero unless the conditional law ran."""
    return {name: int(counters[abi.C_IDX[c]]) for name, c in zip(CONDITIONAL_COUNT_NAMES, CONDITIONAL_COUNTERS)}


def _bin_index(v, edges):
    """Index b with edges[b] <= v < edges[b + 1], else -1 (v compared in FP64)."""
    v = np.asarray(v, dtype=np.float64)
    b = np.searchsorted(edges, v, side='right') - 1
    ok = np.isfinite(v) & (b >= 0) & (b < len(edges) - 1)
    return np.where(ok, b, -1)


def spatial_edges(model):
    """Radial and axial edges of the H170 collision maps: the water domain up
    to the higher of the surface's centre and wall heights (CLARITAS.py:7518-7541)."""
    from .light import box_zmax
    r = np.linspace(0.0, float(model.rin) + 1.0e-9, 6)
    z = np.linspace(float(model.zmin), box_zmax(model) + 1.0e-9, 6)
    return r, z


# tables of the channel-17 (H170) histograms: their shape after the channel is taken
def _h170_shapes(nb):
    return {'h170_cumulative': (13,), 'h170_first_map': (25,), 'h170_last_map': (25,), 'h170_max_map': (25,),
            'h170_type_by_macro': (8, 5), 'h170_ld': (12,), 'h170_incidence': (7,), 'h170_exit': (7,),
            'h170_type_by_incidence': (8, 7), 'h170_type_by_exit': (8, 7), 'h170_psd_by_incidence': (nb, 7),
            'h170_psd_by_exit': (nb, 7), 'h170_type_by_psd': (8, nb), 'h170_first_angle': (6,)}


class PathTally:
    """Score tables over detected histories (the accumulators of
    CLARITAS.py:7352-7600), fed one record array at a time.  engine
    'native' bins in one pass in the library (clr_causal_bin); 'numpy' is
    the same reduction written with numpy masks and bincounts."""

    def __init__(self, n_bins, symmetric, r_edges, z_edges, engine='native'):
        self.engine = engine
        self.nb = int(n_bins)
        self.sym = bool(symmetric)
        self.r_edges, self.z_edges = np.asarray(r_edges, float), np.asarray(z_edges, float)
        self.t = {}
        self.n_records = 0

    # (K, 18) score table: records with key in [0, K) under mask, native and
    # mirror channel each weighted 1/2 (or native alone), CLARITAS.py:8020-8031, 8070-8074
    def _tab(self, rec, key, K, mask, w=None):
        out = np.zeros((K, N_CH))
        key = np.asarray(key, dtype=np.int64)
        kok = mask & (key >= 0) & (key < K)
        ww = None if w is None else np.asarray(w, dtype=np.float64)
        # entries outside the table go to a sink bin (index K * 18), dropped
        for ids, ok_ch, scale in self._sides:
            idx = np.where(kok & ok_ch, key * N_CH + ids, K * N_CH)
            out += scale * np.bincount(idx, weights=ww, minlength=K * N_CH + 1)[:K * N_CH].reshape(K, N_CH)
        return out

    def _add(self, name, val):
        if name in self.t:
            self.t[name] = self.t[name] + val
        else:
            self.t[name] = val

    def add(self, rec):
        """Bin one array of clr_causal_record (numpy, abi.causal_dtype())."""
        if len(rec) == 0:
            return
        if self.engine == 'native':
            self._add_native(rec)
        else:
            self._add_numpy(rec)

    def _layout(self, lib):
        if getattr(self, '_lay', None) is None:
            n = lib.clr_causal_layout(self.nb, -1, None, 0, None)
            lay, off = [], 0
            buf = C.create_string_buffer(128)
            for i in range(n):
                K = C.c_int()
                lib.clr_causal_layout(self.nb, i, buf, 128, C.byref(K))
                lay.append((buf.value.decode(), K.value, off))
                off += K.value * N_CH
            self._lay, self._lay_total = lay, off
        return self._lay, self._lay_total

    def _add_native(self, rec):
        from . import _lib
        lib = _lib.load()
        lay, total = self._layout(lib)
        rec = np.ascontiguousarray(rec)
        tab = np.zeros(total)
        re_, ze_ = np.ascontiguousarray(self.r_edges, float), np.ascontiguousarray(self.z_edges, float)
        pd = C.POINTER(C.c_double)
        rc = lib.clr_causal_bin(C.c_void_p(rec.ctypes.data), len(rec), self.nb, int(self.sym), re_.ctypes.data_as(pd),
                                ze_.ctypes.data_as(pd), tab.ctypes.data_as(pd))
        if rc:
            raise RuntimeError(f'clr_causal_bin failed with code {rc}')
        self.n_records += len(rec)
        h170 = _h170_shapes(self.nb)
        last = None
        for name, K, off in lay:
            a = tab[off:off + K * N_CH].reshape(K, N_CH)
            if name in h170:
                self._add(name, a[:, H170].reshape(h170[name]).copy())
            elif name == 'h170_last_angle_type':
                last = a[:, H170].reshape(8, 6).copy() if last is None else last + a[:, H170].reshape(8, 6)
            elif name == 'h170_last_angle_all':
                row = np.zeros((8, 6))
                row[0] = a[:, H170]
                last = row if last is None else last + row
            else:
                per_bin = name.startswith('bin:') or name == 'max_bin'  # (n_bins, 18) even for one bin
                self._add(name, a[0].copy() if K == 1 and not per_bin else a.copy())
        self._add('h170_last_angle', last)

    def _add_numpy(self, rec):
        # contiguous columns: the record array is 256 bytes per row
        rec = {k: np.ascontiguousarray(rec[k]) for k in rec.dtype.names}
        n = len(rec['history'])
        self.n_records += n
        sides = (('native_id', 0.5), ('mirror_id', 0.5)) if self.sym else (('native_id', 1.0),)
        self._sides = [(rec[k].astype(np.int64), (rec[k] >= 0) & (rec[k] < N_CH), sc) for k, sc in sides]
        zero = np.zeros(n, dtype=np.int64)
        ones = np.ones(n, dtype=bool)
        T = lambda mask, w=None: self._tab(rec, zero, 1, mask, w)[0]
        ic, gic, wic, wcoh = (rec[k].astype(np.int64) for k in ('ic', 'gic', 'wic', 'wcoh'))
        wmorph = wic - wcoh
        per, pir, pabs = rec['per'], rec['pir'], rec['pabs']
        # mechanism and order masks, CLARITAS.py:5265-5322
        single, multiple = ic == 1, ic >= 2
        surface = (gic > 0) | (wmorph > 0)
        masks = {
            'any_particle': ic > 0, 'any_geometric': gic > 0, 'any_wave': wic > 0,
            'any_coherent_wave': wcoh > 0, 'any_morphology_wave': wmorph > 0,
            'any_entry_reflection': per > 0, 'any_internal_reflection': pir > 0, 'ballistic': ic == 0,
            'single_interaction': single, 'single_entry_reflection': single & (per > 0),
            'multiple_interaction': multiple, 'interaction_order_2': ic == 2, 'interaction_order_3': ic == 3,
            'interaction_order_4plus': ic >= 4,
            'exactly_one_geometric': single & (gic == 1) & (wic == 0),
            'exactly_one_wave': single & (gic == 0) & (wic == 1),
            'single_coherent_wave_escape': single & (wcoh == 1),
            'single_external_entry_reflection_escape': single & (per > 0),
            'single_internal_reflection_escape': single & (pir > 0),
            'single_surface_transmission_escape': single & surface & (per == 0) & (pir == 0) & (pabs == 0),
            'mixed_geometric_wave': multiple & (gic > 0) & (wic > 0),
            'geometric_only_multiple': multiple & (gic > 0) & (wic == 0),
            'wave_only_multiple': multiple & (gic == 0) & (wic > 0),
            'multiple_with_entry_reflection': multiple & (per > 0),
            'multiple_with_internal_reflection': multiple & (pir > 0),
        }
        for k in PATH_KEYS:
            self._add('path:' + k, T(masks[k]))
        self._add('hardware', T(ones))
        for k, v in zip(PATH_SUMS, (ic, gic, wic, wcoh, wmorph, per, pir)):
            self._add('sum:' + k, T(ones, v))
        # exact orders and PSD provenance, CLARITAS.py:8076-8101
        self._add('exact_particle_order', self._tab(rec, np.minimum(ic, 5), 6, ones))
        self._add('exact_geometric_order', self._tab(rec, np.minimum(gic, 5), 6, ones))
        nb = self.nb
        for k in ('first_bin', 'last_bin', 'first_geometric_bin', 'last_geometric_bin', 'first_wave_bin',
                  'last_wave_bin'):
            self._add('bin:' + k, self._tab(rec, rec[k], nb, ones))
        # acrylic returns and reflections, CLARITAS.py:8102-8123
        ret = rec['acrylic_returns'] > 0
        wall = np.where(ic == 0, 0, np.where(ret, 2, 1))
        self._add('wall_path', self._tab(rec, wall, 3, ones))
        self._add('particle_after_acrylic_return', T(rec['events_after_return'] > 0))
        self._add('acrylic_reflection_order', self._tab(rec, np.minimum(rec['acrylic_reflections'], 4), 5, ones))
        # causal first / last events, CLARITAS.py:8066-8174
        ft, lt = rec['first_type'], rec['last_type']
        fa = rec['first_deflection'].astype(np.float64)
        la = rec['last_deflection'].astype(np.float64)
        valid = lt > 0
        self._add('last_event', self._tab(rec, lt, 8, ones))
        self._add('first_event', self._tab(rec, ft, 8, ones))
        lav_any = np.isfinite(la) & (la >= 0.0)
        self._add('last_event_angle_sum', self._tab(rec, lt, 8, lav_any, la))
        self._add('last_event_angle_count', self._tab(rec, lt, 8, lav_any))
        self._add('first_backscatter', T(valid & (rec['first_backscatter'] > 0)))
        self._add('first_source_hemisphere', T(valid & (rec['first_source_hemisphere'] > 0)))
        fav = valid & np.isfinite(fa) & (fa >= 0.0)
        lav = valid & lav_any
        self._add('first_angle_sum', T(fav, fa))
        self._add('first_angle_count', T(fav))
        self._add('last_angle_sum', T(lav, la))
        self._add('last_angle_count', T(lav))
        ip, io_, wpa = (rec[k].astype(np.float64) for k in ('internal_path', 'internal_opl', 'water_path_after_last'))
        pv = valid & np.isfinite(ip) & np.isfinite(io_) & np.isfinite(wpa) & (wpa >= 0.0)
        self._add('internal_path_sum', T(pv, ip))
        self._add('internal_opl_sum', T(pv, io_))
        self._add('post_last_water_path_sum', T(pv, wpa))
        self._add('last_order_sum', T(pv, rec['last_order']))
        self._add('path_count', T(pv))
        # per-history means, CLARITAS.py:8158-8238 (float32 radii as the reference)
        first_r = np.sqrt(np.maximum(rec['first_x'] ** 2 + rec['first_y'] ** 2, 0.0))
        last_r = np.sqrt(np.maximum(rec['last_x'] ** 2 + rec['last_y'] ** 2, 0.0))
        f64 = lambda k: rec[k].astype(np.float64)
        values = {
            'cumulative_particle_deflection_deg': rec['cumulative_deflection'],
            'cumulative_particle_deflection_sq_deg2': rec['cumulative_deflection_sq'],
            'max_particle_deflection_deg': rec['max_deflection'],
            'first_particle_r_m': first_r, 'first_particle_z_m': rec['first_z'],
            'last_particle_r_m': last_r, 'last_particle_z_m': rec['last_z'],
            'particle_event_mean_r_m': rec['mean_r'], 'particle_event_mean_z_m': rec['mean_z'],
            'particle_event_min_r_m': rec['min_r'], 'particle_event_max_r_m': rec['max_r'],
            'particle_event_min_z_m': rec['min_z'], 'particle_event_max_z_m': rec['max_z'],
            'water_path_at_first_particle_event_m': rec['water_path_first'],
            'water_path_at_last_particle_event_m': rec['water_path_last'],
            'particle_water_path_span_first_to_last_m': rec['water_path_last'] - rec['water_path_first'],
            'particle_radial_span_m': f64('max_r') - f64('min_r'),
            'particle_axial_span_m': f64('max_z') - f64('min_z'),
            'particle_branch_transition_count': rec['branch_transitions'],
            'particle_event_type_transition_count': rec['type_transitions'],
            'particle_small_deflection_count': rec['small_deflections'],
            'particle_medium_deflection_count': rec['medium_deflections'],
            'particle_large_deflection_count': rec['large_deflections'],
        }
        for k in HISTORY_METRICS:
            v = np.asarray(values[k], dtype=np.float64)
            m = valid & np.isfinite(v)
            self._add('metric_sum:' + k, T(m, v))
            self._add('metric_count:' + k, T(m))
        # deflection-history and branch-sequence classes, CLARITAS.py:8239-8303
        med, lrg = rec['medium_deflections'], rec['large_deflections']
        dcls = np.full(n, -1)
        dcls[ic == 0] = 0
        dcls[valid & (med == 0) & (lrg == 0)] = 1
        dcls[valid & (med > 0) & (lrg == 0)] = 2
        dcls[valid & (lrg > 0)] = 3
        self._add('deflection_history', self._tab(rec, dcls, 4, ones))
        fb = np.where((ft >= 1) & (ft <= 3), 1, np.where(ft >= 4, 2, 0))
        lb = np.where((lt >= 1) & (lt <= 3), 1, np.where(lt >= 4, 2, 0))
        bcls = np.full(n, -1)
        bcls[ic == 0] = 0
        bcls[valid & (gic > 0) & (wic == 0)] = 1
        bcls[valid & (wic > 0) & (gic == 0)] = 2
        mixed = valid & (gic > 0) & (wic > 0)
        for c, (a, b) in enumerate(((1, 1), (1, 2), (2, 1), (2, 2)), start=3):
            bcls[mixed & (fb == a) & (lb == b)] = c
        self._add('branch_sequence', self._tab(rec, bcls, 7, ones))
        self._add('branch_transition', self._tab(rec, np.minimum(rec['branch_transitions'], 5), 6, valid))
        self._add('event_transition', self._tab(rec, np.minimum(rec['type_transitions'], 5), 6, valid))
        # H170 cumulative-deflection histogram and collision maps, CLARITAS.py:8304-8331, 8428-8448
        h = lambda key, K, mask: self._tab(rec, key, K, mask)[:, H170]
        self._add('h170_cumulative', h(_bin_index(rec['cumulative_deflection'], H170_CUMULATIVE_EDGES),
                                       len(H170_CUMULATIVE_EDGES) - 1, valid))

        def cell(r, z):
            ri, zi = _bin_index(r, self.r_edges), _bin_index(z, self.z_edges)
            return np.where((ri >= 0) & (zi >= 0), ri * 5 + zi, -1)

        max_r = np.sqrt(np.maximum(rec['turn_x'] ** 2 + rec['turn_y'] ** 2, 0.0))
        self._add('h170_first_map', h(cell(first_r, rec['first_z']), 25, valid))
        self._add('h170_last_map', h(cell(last_r, rec['last_z']), 25, valid))
        self._add('h170_max_map', h(cell(max_r, rec['turn_z']), 25, valid))
        # provenance of the largest turn, CLARITAS.py:8332-8427
        mt, mo, mbin = rec['max_type'], rec['max_order'], rec['max_bin']
        mac, mic = rec['max_macro_bounces'], rec['max_micro_order']
        self._add('max_type', self._tab(rec, mt, 8, valid))
        self._add('max_order', self._tab(rec, np.where(mo >= 1, np.minimum(mo, 5) - 1, -1), 5, valid))
        self._add('max_bin', self._tab(rec, mbin, nb, valid))
        macro_cls = np.minimum(mac, 4)
        self._add('max_macro', self._tab(rec, macro_cls, 5, valid))
        self._add('h170_type_by_macro', h(np.where(mt >= 0, mt * 5 + macro_cls, -1), 40, valid).reshape(8, 5))
        self._add('max_micro', self._tab(rec, np.minimum(mic, 5), 6, valid))
        ld = rec['max_path_over_diameter'].astype(np.float64)
        self._add('max_ld_sum', T(valid, ld))
        self._add('max_macro_sum', T(valid, mac))
        self._add('max_micro_sum', T(valid, mic))
        self._add('h170_ld', h(_bin_index(ld, H170_LENGTH_RATIO_EDGES), len(H170_LENGTH_RATIO_EDGES) - 1, valid))
        inc, ext = rec['max_incidence'].astype(np.float64), rec['max_exit'].astype(np.float64)
        inc_ok = valid & np.isfinite(inc) & (inc >= 0.0) & (inc <= 90.0001)
        ext_ok = valid & np.isfinite(ext) & (ext >= 0.0) & (ext <= 90.0001)
        self._add('incidence_sum', T(inc_ok, inc))
        self._add('incidence_defined', T(inc_ok))
        self._add('exit_sum', T(ext_ok, ext))
        self._add('exit_defined', T(ext_ok))
        inc_cls = np.where(inc_ok, _bin_index(inc, INTERFACE_ANGLE_EDGES) + 1, 0)
        ext_cls = np.where(ext_ok, _bin_index(ext, INTERFACE_ANGLE_EDGES) + 1, 0)
        inc_cls[inc_cls < 0] = 0
        ext_cls[ext_cls < 0] = 0
        self._add('h170_incidence', h(inc_cls, 7, valid))
        self._add('h170_exit', h(ext_cls, 7, valid))
        self._add('h170_type_by_incidence', h(mt * 7 + inc_cls, 56, valid & (mt >= 1)).reshape(8, 7))
        self._add('h170_type_by_exit', h(mt * 7 + ext_cls, 56, valid & (mt >= 1)).reshape(8, 7))
        self._add('h170_psd_by_incidence', h(mbin * 7 + inc_cls, nb * 7, valid & (mbin >= 0)).reshape(nb, 7))
        self._add('h170_psd_by_exit', h(mbin * 7 + ext_cls, nb * 7, valid & (mbin >= 0)).reshape(nb, 7))
        self._add('h170_type_by_psd', h(mt * nb + mbin, 8 * nb, valid & (mt >= 1) & (mbin >= 0)).reshape(8, nb))
        # H170 first / last angle histograms, CLARITAS.py:8449-8481
        lbin = _bin_index(la, H170_ANGLE_EDGES)
        last_hist = np.zeros((8, 6))
        last_hist[0] = h(lbin, 6, lav)
        last_hist += h(np.where(lt >= 1, lt * 6 + lbin, -1), 48, lav & (lbin >= 0)).reshape(8, 6)
        self._add('h170_last_angle', last_hist)
        self._add('h170_first_angle', h(_bin_index(fa, H170_ANGLE_EDGES), 6, fav))

    def get(self, name, shape):
        return self.t.get(name, np.zeros(shape))


def _frac(num, den):
    num = np.asarray(num, dtype=np.float64)
    return np.divide(num, den, out=np.zeros_like(num), where=den > 0)


def _mean(num, den):
    return np.divide(num, den, out=np.full(N_CH, np.nan), where=den > 0)


def diagnostics_dict(P: PathTally, hardware_sym, closure, conditional_counts, max_trials):
    """The reference's ``detector_particle_path_diagnostics`` (CLARITAS.py:8967-9440):
    scores, fractions of the hardware symmetry score, conditional means,
    labels, edges and the partition-closure errors (raising on a failed
    closure as the reference does)."""
    nb = P.nb
    g = lambda k, shape=(N_CH,): P.get(k, shape)
    den = np.asarray(hardware_sym, dtype=np.float64)
    d = {'hardware_symmetry_scores': den.copy()}
    ps = {k: g('path:' + k) for k in PATH_KEYS}
    for k in PATH_KEYS:
        d[k + '_scores'] = ps[k].copy()
        d[k + '_fraction'] = _frac(ps[k], den)
    names = {'interactions': 'particle_interactions', 'entry_reflections': 'entry_reflections',
             'internal_reflections': 'internal_reflections'}
    for k in PATH_SUMS:
        d['mean_' + names.get(k, k) + '_by_channel'] = _mean(g('sum:' + k), den)
    last, first = g('last_event', (8, N_CH)), g('first_event', (8, N_CH))
    las, lac = g('last_event_angle_sum', (8, N_CH)), g('last_event_angle_count', (8, N_CH))
    for c in EVENT_CODES:
        nm = EVENT_TYPE_NAMES[c]
        d[f'last_event_{nm}_scores'] = last[c].copy()
        d[f'last_event_{nm}_fraction'] = _frac(last[c], den)
        d[f'first_event_{nm}_scores'] = first[c].copy()
        d[f'first_event_{nm}_fraction'] = _frac(first[c], den)
        d[f'last_event_{nm}_mean_deflection_deg_by_channel'] = _mean(las[c], lac[c])
    for k, src in (('first_event_backscatter', 'first_backscatter'),
                   ('first_event_source_hemisphere', 'first_source_hemisphere')):
        d[k + '_scores'] = g(src).copy()
        d[k + '_fraction'] = _frac(g(src), den)
    d['mean_first_particle_deflection_deg_by_channel'] = _mean(g('first_angle_sum'), g('first_angle_count'))
    d['mean_last_particle_deflection_deg_by_channel'] = _mean(g('last_angle_sum'), g('last_angle_count'))
    pc = g('path_count')
    d['mean_total_internal_particle_path_m_by_channel'] = _mean(g('internal_path_sum'), pc)
    d['mean_total_internal_particle_opl_m_by_channel'] = _mean(g('internal_opl_sum'), pc)
    d['mean_water_path_after_last_particle_event_m_by_channel'] = _mean(g('post_last_water_path_sum'), pc)
    d['mean_last_particle_event_order_by_channel'] = _mean(g('last_order_sum'), pc)
    anyp = ps['any_particle']
    d['causal_last_event_partition_closure_error_scores'] = anyp - last[1:].sum(axis=0)
    d['causal_first_event_partition_closure_error_scores'] = anyp - first[1:].sum(axis=0)
    d['H170_last_event_angle_bin_edges_deg'] = H170_ANGLE_EDGES.copy()
    d['H170_last_event_angle_histogram_scores_by_event_code'] = g('h170_last_angle', (8, 6)).copy()
    d['H170_first_event_angle_histogram_scores'] = g('h170_first_angle', (6,)).copy()
    d['interaction_order_partition_closure_error_scores'] = anyp - (
        ps['single_interaction'] + ps['interaction_order_2'] + ps['interaction_order_3'] +
        ps['interaction_order_4plus'])
    d['multiple_branch_partition_closure_error_scores'] = ps['multiple_interaction'] - (
        ps['mixed_geometric_wave'] + ps['geometric_only_multiple'] + ps['wave_only_multiple'])
    d['single_branch_partition_closure_error_scores'] = ps['single_interaction'] - (
        ps['single_external_entry_reflection_escape'] + ps['single_internal_reflection_escape'] +
        ps['single_surface_transmission_escape'] + ps['single_coherent_wave_escape'])
    order6 = ['0', '1', '2', '3', '4', '5+']
    d['exact_particle_order_labels'] = list(order6)
    d['conditional_sampler_counts'] = dict(conditional_counts)
    d['closure'] = closure
    d['conditional_max_trials'] = int(max_trials)
    epo, ego = g('exact_particle_order', (6, N_CH)), g('exact_geometric_order', (6, N_CH))
    d['exact_particle_order_scores'] = epo.copy()
    d['exact_geometric_order_labels'] = list(order6)
    d['exact_geometric_order_scores'] = ego.copy()
    d['exact_particle_order_partition_closure_error_scores'] = den - epo.sum(axis=0)
    d['exact_geometric_order_partition_closure_error_scores'] = den - ego.sum(axis=0)
    for k in ('first_particle_bin', 'last_particle_bin', 'first_geometric_bin', 'last_geometric_bin',
              'first_wave_bin', 'last_wave_bin'):
        d[k + '_scores'] = g('bin:' + k.replace('_particle', ''), (nb, N_CH)).copy()
    d['first_particle_bin_partition_closure_error_scores'] = anyp - d['first_particle_bin_scores'].sum(axis=0)
    d['last_particle_bin_partition_closure_error_scores'] = anyp - d['last_particle_bin_scores'].sum(axis=0)
    wall = g('wall_path', (3, N_CH))
    d['wall_path_scores'] = {k: wall[i].copy() for i, k in enumerate(WALL_CLASSES)}
    d['wall_path_partition_closure_error_scores'] = den - wall.sum(axis=0)
    d['particle_after_acrylic_return_scores'] = g('particle_after_acrylic_return').copy()
    aro = g('acrylic_reflection_order', (5, N_CH))
    d['acrylic_reflection_order_labels'] = ['0', '1', '2', '3', '4+']
    d['acrylic_reflection_order_scores'] = aro.copy()
    d['acrylic_reflection_order_partition_closure_error_scores'] = den - aro.sum(axis=0)
    for k in HISTORY_METRICS:
        d['mean_' + k + '_by_channel'] = _mean(g('metric_sum:' + k), g('metric_count:' + k))
    dh = g('deflection_history', (4, N_CH))
    d['deflection_history_class_scores'] = {k: dh[i].copy() for i, k in enumerate(DEFLECTION_CLASSES)}
    d['deflection_history_class_partition_closure_error_scores'] = den - dh.sum(axis=0)
    bs = g('branch_sequence', (7, N_CH))
    d['branch_sequence_class_scores'] = {k: bs[i].copy() for i, k in enumerate(BRANCH_CLASSES)}
    d['branch_sequence_class_partition_closure_error_scores'] = den - bs.sum(axis=0)
    bt, et = g('branch_transition', (6, N_CH)), g('event_transition', (6, N_CH))
    d['branch_transition_order_labels'] = list(order6)
    d['branch_transition_order_scores'] = bt.copy()
    d['branch_transition_order_partition_closure_error_scores'] = anyp - bt.sum(axis=0)
    d['event_type_transition_order_labels'] = list(order6)
    d['event_type_transition_order_scores'] = et.copy()
    d['event_type_transition_order_partition_closure_error_scores'] = anyp - et.sum(axis=0)
    h170 = float(anyp[H170])
    d['H170_cumulative_particle_deflection_bin_edges_deg'] = H170_CUMULATIVE_EDGES.copy()
    d['H170_cumulative_particle_deflection_histogram_scores'] = g('h170_cumulative', (13,)).copy()
    d['H170_spatial_radial_bin_edges_m'] = P.r_edges.copy()
    d['H170_spatial_axial_bin_edges_m'] = P.z_edges.copy()
    d['H170_first_particle_spatial_score_map'] = g('h170_first_map', (25,)).reshape(5, 5).copy()
    d['H170_last_particle_spatial_score_map'] = g('h170_last_map', (25,)).reshape(5, 5).copy()
    d['H170_cumulative_deflection_histogram_closure_error_score'] = \
        h170 - float(d['H170_cumulative_particle_deflection_histogram_scores'].sum())
    d['H170_first_particle_spatial_map_closure_error_score'] = \
        h170 - float(d['H170_first_particle_spatial_score_map'].sum())
    d['H170_last_particle_spatial_map_closure_error_score'] = \
        h170 - float(d['H170_last_particle_spatial_score_map'].sum())
    mt = g('max_type', (8, N_CH))
    d['max_deflection_event_type_scores'] = {EVENT_TYPE_NAMES[c]: mt[c].copy() for c in EVENT_CODES}
    d['max_deflection_event_type_partition_closure_error_scores'] = anyp - mt[1:].sum(axis=0)
    mo = g('max_order', (5, N_CH))
    d['max_deflection_event_order_labels'] = ['1', '2', '3', '4', '5+']
    d['max_deflection_event_order_scores'] = mo.copy()
    d['max_deflection_event_order_partition_closure_error_scores'] = anyp - mo.sum(axis=0)
    mb = g('max_bin', (nb, N_CH))
    d['max_deflection_event_bin_scores'] = mb.copy()
    d['max_deflection_event_bin_partition_closure_error_scores'] = anyp - mb.sum(axis=0)
    d['H170_max_deflection_event_spatial_score_map'] = g('h170_max_map', (25,)).reshape(5, 5).copy()
    d['H170_max_deflection_event_spatial_map_closure_error_score'] = \
        h170 - float(d['H170_max_deflection_event_spatial_score_map'].sum())
    mac, mic = g('max_macro', (5, N_CH)), g('max_micro', (6, N_CH))
    d['max_deflection_event_macro_bounce_labels'] = ['0', '1', '2', '3', '4+']
    d['max_deflection_event_macro_bounce_scores'] = mac.copy()
    d['max_deflection_event_macro_bounce_partition_closure_error_scores'] = anyp - mac.sum(axis=0)
    d['max_deflection_event_terminal_micro_order_labels'] = list(order6)
    d['max_deflection_event_terminal_micro_order_scores'] = mic.copy()
    d['max_deflection_event_terminal_micro_order_partition_closure_error_scores'] = anyp - mic.sum(axis=0)
    d['mean_max_deflection_event_internal_path_over_diameter_by_channel'] = _mean(g('max_ld_sum'), anyp)
    d['mean_max_deflection_event_macro_bounces_by_channel'] = _mean(g('max_macro_sum'), anyp)
    d['mean_max_deflection_event_terminal_micro_order_by_channel'] = _mean(g('max_micro_sum'), anyp)
    d['H170_max_deflection_internal_path_over_diameter_edges'] = H170_LENGTH_RATIO_EDGES.copy()
    d['H170_max_deflection_internal_path_over_diameter_scores'] = g('h170_ld', (12,)).copy()
    d['H170_max_deflection_internal_path_over_diameter_histogram_closure_error_score'] = \
        h170 - float(d['H170_max_deflection_internal_path_over_diameter_scores'].sum())
    d['H170_max_deflection_event_type_by_macro_bounce_scores'] = g('h170_type_by_macro', (8, 5)).copy()
    d['H170_max_deflection_event_type_by_macro_bounce_closure_error_score'] = \
        h170 - float(d['H170_max_deflection_event_type_by_macro_bounce_scores'].sum())
    d['max_deflection_interface_angle_class_labels'] = list(INTERFACE_ANGLE_LABELS)
    d['max_deflection_interface_angle_edges_deg'] = INTERFACE_ANGLE_EDGES.copy()
    for k, src, shape in (('H170_max_deflection_incidence_angle_scores', 'h170_incidence', (7,)),
                          ('H170_max_deflection_exit_angle_scores', 'h170_exit', (7,)),
                          ('H170_max_deflection_event_type_by_psd_bin_scores', 'h170_type_by_psd', (8, nb)),
                          ('H170_max_deflection_event_type_by_incidence_angle_scores', 'h170_type_by_incidence',
                           (8, 7)),
                          ('H170_max_deflection_event_type_by_exit_angle_scores', 'h170_type_by_exit', (8, 7)),
                          ('H170_max_deflection_psd_bin_by_incidence_angle_scores', 'h170_psd_by_incidence',
                           (nb, 7)),
                          ('H170_max_deflection_psd_bin_by_exit_angle_scores', 'h170_psd_by_exit', (nb, 7))):
        d[k] = g(src, shape).copy()
    d['mean_max_deflection_event_incidence_deg_by_channel'] = _mean(g('incidence_sum'), g('incidence_defined'))
    d['mean_max_deflection_event_exit_deg_by_channel'] = _mean(g('exit_sum'), g('exit_defined'))
    d['max_deflection_event_incidence_defined_scores'] = g('incidence_defined').copy()
    d['max_deflection_event_exit_defined_scores'] = g('exit_defined').copy()
    for k, src in (('incidence_angle_histogram', 'H170_max_deflection_incidence_angle_scores'),
                   ('exit_angle_histogram', 'H170_max_deflection_exit_angle_scores'),
                   ('event_type_by_psd_bin', 'H170_max_deflection_event_type_by_psd_bin_scores'),
                   ('event_type_by_incidence_angle', 'H170_max_deflection_event_type_by_incidence_angle_scores'),
                   ('event_type_by_exit_angle', 'H170_max_deflection_event_type_by_exit_angle_scores'),
                   ('psd_bin_by_incidence_angle', 'H170_max_deflection_psd_bin_by_incidence_angle_scores'),
                   ('psd_bin_by_exit_angle', 'H170_max_deflection_psd_bin_by_exit_angle_scores')):
        d[f'H170_max_deflection_{k}_closure_error_score'] = h170 - float(d[src].sum())
    check_closure(d)
    return d


CLOSURE_VECTORS = (
    'exact_particle_order_partition_closure_error_scores', 'exact_geometric_order_partition_closure_error_scores',
    'first_particle_bin_partition_closure_error_scores', 'last_particle_bin_partition_closure_error_scores',
    'wall_path_partition_closure_error_scores', 'acrylic_reflection_order_partition_closure_error_scores',
    'deflection_history_class_partition_closure_error_scores',
    'branch_sequence_class_partition_closure_error_scores',
    'branch_transition_order_partition_closure_error_scores',
    'event_type_transition_order_partition_closure_error_scores',
    'max_deflection_event_type_partition_closure_error_scores',
    'max_deflection_event_order_partition_closure_error_scores',
    'max_deflection_event_bin_partition_closure_error_scores',
    'max_deflection_event_macro_bounce_partition_closure_error_scores',
    'max_deflection_event_terminal_micro_order_partition_closure_error_scores',
)
CLOSURE_H170 = (
    'H170_cumulative_deflection_histogram_closure_error_score',
    'H170_first_particle_spatial_map_closure_error_score', 'H170_last_particle_spatial_map_closure_error_score',
    'H170_max_deflection_event_spatial_map_closure_error_score',
    'H170_max_deflection_internal_path_over_diameter_histogram_closure_error_score',
    'H170_max_deflection_event_type_by_macro_bounce_closure_error_score',
    'H170_max_deflection_incidence_angle_histogram_closure_error_score',
    'H170_max_deflection_exit_angle_histogram_closure_error_score',
    'H170_max_deflection_event_type_by_psd_bin_closure_error_score',
    'H170_max_deflection_event_type_by_incidence_angle_closure_error_score',
    'H170_max_deflection_event_type_by_exit_angle_closure_error_score',
    'H170_max_deflection_psd_bin_by_incidence_angle_closure_error_score',
    'H170_max_deflection_psd_bin_by_exit_angle_closure_error_score',
)


def check_closure(d, tol=1.0e-9):
    """The reference's decomposition checks (CLARITAS.py:9398-9440): every
    exclusive partition recovers its parent score."""
    for k in CLOSURE_VECTORS:
        e = np.asarray(d[k], dtype=np.float64)
        if np.any(np.abs(e) > tol):
            raise RuntimeError(f'detector-score decomposition closure failure {k}: max_abs={float(np.abs(e).max())}')
    for k in CLOSURE_H170:
        if abs(float(d[k])) > tol:
            raise RuntimeError(f'H170 spatial/angular diagnostic closure failure {k}: error={float(d[k])}')


def sum_tallies(a, b):
    """Add the tally dicts of two history ranges of one run."""
    if a is None:
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in b.items()}
    return {k: (a[k] + b[k] if k != 'n_groups' else a[k]) for k in a}


class CausalAccumulator:
    """Runs clr_run_causal over history ranges of at most `chunk` histories
    and bins each range's records into a PathTally."""

    def __init__(self, model, n_bins, chunk=1 << 20):
        self.model = model
        self.chunk = int(chunk)
        r, z = spatial_edges(model)
        self.tally = PathTally(n_bins, model.symmetry_average_exact, r, z)
        self.count = 0
        self.sample_limit = 512
        self.failures = []  # the failed histories' samples with the lowest ids (numpy rows)
        self.failure_count = 0
        self.failure_device_capacity = None  # samples one launch may write (None: 4x the room, >= 1024)

    def run(self, n, seed, history_offset, launch, light_acc=None):
        """clr_run_causal over histories [history_offset, history_offset + n):
        returns the tally dict of the whole run."""
        ctx = self.model.ctx
        total = None
        dt = abi.causal_dtype()
        for r0 in range(0, int(n), self.chunk):
            rn = min(self.chunk, int(n) - r0)
            recs = np.zeros(rn, dtype=dt)
            co = abi.CausalOut()
            co.records = recs.ctypes.data
            co.capacity = rn
            co.range_start, co.range_count = r0, rn
            room = max(self.sample_limit - sum(len(f) for f in self.failures), 0)
            cap = self._failure_cap(room, rn)
            fails = np.zeros(max(cap, 1), dtype=abi.failure_dtype())
            co.failures, co.failure_capacity = fails.ctypes.data, cap
            t = abi.Tallies()
            if light_acc is not None:
                out, arrs = abi.light_out(*light_acc.shape())
                ctx.check(ctx.lib.clr_run_causal(ctx.h, int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, int(history_offset),
                                                 launch, C.byref(light_acc.struct), C.byref(t), C.byref(out),
                                                 C.byref(co)), 'clr_run_causal')
                light_acc.add(arrs, abi.light_scalars(out))
            else:
                ctx.check(ctx.lib.clr_run_causal(ctx.h, int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, int(history_offset),
                                                 launch, None, C.byref(t), None, C.byref(co)), 'clr_run_causal')
            if co.count > rn:
                raise RuntimeError('causal record buffer overflow')
            self.tally.add(recs[:int(co.count)])
            self.count += int(co.count)
            self.failure_count += int(co.failure_count)
            if room and co.failure_count:
                if co.failure_count <= cap:
                    f = np.sort(fails[:int(co.failure_count)], order='history')[:room]
                else:
                    f = self._lowest_failures(n, seed, history_offset, launch, r0, rn, room, cap)
                self.failures.append(f)
            total = sum_tallies(total, abi.tallies_to_numpy(t))
        return total

    def _failure_cap(self, room, rn):
        if not room:
            return 0
        c = self.failure_device_capacity or max(4 * room, 1024)
        return int(min(max(c, room), rn))

    def _lowest_failures(self, n, seed, history_offset, launch, r0, rn, room, cap):
        """The `room` failed histories with the lowest ids in the range
        [r0, r0 + rn) of a run whose launch wrote more failures than `cap`
        (in completion order).  The range's histories are traced again in
        prefixes short enough for all their failures to fit (the traces are
        the same: one Philox stream per history); tallies and records of
        these passes are discarded."""
        ctx = self.model.ctx
        out, need, start, end = [], room, r0, r0 + rn
        while need > 0 and start < end:
            length = end - start
            while True:
                fails = np.zeros(cap, dtype=abi.failure_dtype())
                co = abi.CausalOut()
                co.records, co.capacity = None, 0
                co.range_start, co.range_count = start, length
                co.failures, co.failure_capacity = fails.ctypes.data, cap
                t = abi.Tallies()
                ctx.check(ctx.lib.clr_run_causal(ctx.h, int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, int(history_offset),
                                                 launch, None, C.byref(t), None, C.byref(co)), 'clr_run_causal')
                if co.failure_count <= cap or length == 1:
                    break
                length = max(length // 2, 1)
            f = np.sort(fails[:min(int(co.failure_count), cap)], order='history')[:need]
            out.append(f)
            need -= len(f)
            start += length
        return np.concatenate(out) if out else np.zeros(0, dtype=abi.failure_dtype())
