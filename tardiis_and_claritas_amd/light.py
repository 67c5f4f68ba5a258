"""Light-transport voxel diagnostics: configuration, device runs, the
reference's result dictionary, cross sections and on-disk format.

The device records, for the histories with id < max_rays, the exact water
track length per voxel (all path and the part after the first particle
event), the sites of surviving particle events and of absorbing extinction
encounters, closure totals, and the segments of the first representative
histories (clr_run_light; CLARITAS.py:3296-3485 device side, 7232-7300
configuration, 9455-9560 readback).  `LightAccumulator` merges the calls of
one simulate(), `cross_sections` and `write_outputs` restate the data part
of the reference's writer (CLARITAS.py:11645-11694, 12108-12330; plots are
not produced).
"""
from __future__ import annotations

import csv
import ctypes as C
import json
import math
import time
from pathlib import Path

import numpy as np

from . import abi
from .physics import vortex_surface_geometry

# the reference's production defaults, CLARITAS.py:198-207
DEFAULTS = dict(enabled=False, save_data=True, max_rays=None, nx=256, ny=256, nz=384, representative_rays=256,
                max_segments_per_ray=512, vertical_slice_thickness_m=0.001, sensor_slice_thickness_m=0.001)

# terminal event codes of representative segments, CLARITAS.py:12117-12129
EVENT_NAMES = {
    0: 'propagation_or_boundary', 1: 'geometric_external_reflection', 2: 'geometric_transmission',
    3: 'geometric_internal_reflection', 4: 'coherent_mie', 5: 'morphology_external_reflection',
    6: 'morphology_transmission', 7: 'morphology_internal_reflection', 8: 'null_collision',
    9: 'particle_absorption', 10: 'particle_transport_failure',
}


def resolve(config, n_rays):
    """Normalised configuration dict; raises on the reference's invalid
    settings (CLARITAS.py:7233-7253).  max_rays None follows n_rays."""
    c = dict(DEFAULTS)
    c.update(dict(config or {}))
    c['enabled'] = bool(c['enabled'])
    if not c['enabled']:
        return c
    c['max_rays'] = int(n_rays if c.get('max_rays') is None else c['max_rays'])
    c['nx'], c['ny'], c['nz'] = int(c['nx']), int(c['ny']), int(c['nz'])
    c['max_segments_per_ray'] = int(c['max_segments_per_ray'])
    c['representative_rays'] = min(int(c['representative_rays']), c['max_rays'])
    c['heatmap_size'] = int(c.get('heatmap_size', 0))
    if c['max_rays'] < 1 and not c.get('heatmap_only'):
        raise ValueError('light diagnostic max_rays must be >=1')
    if min(c['nx'], c['ny'], c['nz']) < 4:
        raise ValueError('light diagnostic grid dimensions must each be >=4')
    if c['representative_rays'] < 0 or c['max_segments_per_ray'] < 1:
        raise ValueError('invalid representative-ray diagnostic configuration')
    return c


def memory_bytes(c):
    """Diagnostic allocation of the reference's accounting, CLARITAS.py:7262-7267."""
    vox = c['nx'] * c['ny'] * c['nz']
    rep, seg = c['representative_rays'], c['max_segments_per_ray']
    return vox * 20 + max(rep * seg, 1) * 32 + max(rep, 1) * 8 + 8


def box_zmax(model):
    """Top of the diagnostic box: the higher of the free surface's centre and
    wall heights (CLARITAS.py:7254-7261)."""
    fs = model.free_surface_model
    dh = 0.0 if fs == 'flat' else model.vortex_delta_h
    vg = vortex_surface_geometry(fs, model.zmean, model.rin, dh, model.vortex_core_radius, model.gravity)
    return max(float(vg['center_z']), float(vg['wall_z']))


def config_struct(c, zmax) -> abi.LightConfig:
    s = abi.LightConfig()
    s.max_rays = int(c['max_rays'])
    s.nx, s.ny, s.nz = c['nx'], c['ny'], c['nz']
    s.representative_rays = int(c['representative_rays'])
    s.max_segments = int(c['max_segments_per_ray'])
    s.heatmap_size = int(c.get('heatmap_size', 0))
    s.zmax = float(zmax)
    return s


def heatmap_only_config(vis):
    """Configuration of a run that keeps only the water-path heat maps
    (simulate's heatmap_size without light diagnostics): no voxel histories."""
    return dict(DEFAULTS, enabled=True, max_rays=0, nx=4, ny=4, nz=4, representative_rays=0,
                max_segments_per_ray=1, heatmap_size=int(vis), heatmap_only=True)


def probe(ctx, mode, data, light_config, rin, zmin):
    """clr_light_probe: the deposition on explicit inputs (mode 0 segments
    [n][7], mode 1 event sites [n][3], mode 2 the heat-map raster of segments
    [n][7]); returns (arrays, scalars)."""
    lc = light_config
    data = np.ascontiguousarray(data, dtype=np.float32)
    out, arrs = abi.light_out(lc.nx, lc.ny, lc.nz, lc.representative_rays, lc.max_segments, lc.heatmap_size)
    ctx.check(ctx.lib.clr_light_probe(ctx.h, int(mode), int(data.shape[0]), data.ctypes.data_as(abi.PF), C.byref(lc),
                                      float(rin), float(zmin), C.byref(out)), 'clr_light_probe')
    return arrs, abi.light_scalars(out)


class LightAccumulator:
    """Device diagnostics of one simulate(): each run over a history range
    adds its grids and totals; representative segments belong to fixed
    history ids, so each range fills its own rows."""

    def __init__(self, model, config, n_rays, ray_id_start=0):
        self.model = model
        self.cfg = resolve(config, n_rays)
        self.zmax = box_zmax(model)
        self.struct = config_struct(self.cfg, self.zmax)
        c = self.cfg
        self.sum = None
        self.scalars = None
        self.ray_id_start = int(ray_id_start)
        self.t0 = time.perf_counter()
        if memory_bytes(c) > 1_000_000_000:
            raise ValueError(f'light diagnostic allocation is too large ({memory_bytes(c) / 1e9:.3g} GB); reduce '
                             'grid dimensions, representative rays, or segments per ray')

    def shape(self):
        """(nx, ny, nz, representative rays, segments per ray, heat-map side) of the output arrays."""
        c = self.cfg
        return c['nx'], c['ny'], c['nz'], c['representative_rays'], c['max_segments_per_ray'], c['heatmap_size']

    def heatmaps(self):
        """(x-y, z-y) water-path heat maps summed over the runs, or (None, None)."""
        if not self.cfg['heatmap_size'] or self.sum is None:
            return None, None
        return self.sum['heatmap_xy'], self.sum['heatmap_zy']

    def run(self, n, seed, history_offset, launch):
        """clr_run_light over histories [history_offset, history_offset + n)."""
        c = self.cfg
        out, arrs = abi.light_out(*self.shape())
        t = abi.Tallies()
        ctx = self.model.ctx
        ctx.check(ctx.lib.clr_run_light(ctx.h, int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, int(history_offset), launch,
                                        C.byref(self.struct), C.byref(t), C.byref(out)), 'clr_run_light')
        self.add(arrs, abi.light_scalars(out))
        return t

    def add(self, arrs, sc):
        if self.sum is None:
            self.sum, self.scalars = {k: v.copy() for k, v in arrs.items()}, dict(sc)
            return
        for k in abi.LIGHT_GRIDS + (('heatmap_xy', 'heatmap_zy') if 'heatmap_xy' in arrs else ()):
            self.sum[k] = self.sum[k] + arrs[k]
        rows = arrs['traj_count'] > 0
        for k in abi.LIGHT_TRAJ:
            self.sum[k][rows] = arrs[k][rows]
        self.sum['traj_count'][rows] = arrs['traj_count'][rows]
        self.sum['traj_overflow'] |= arrs['traj_overflow']
        for k in abi.LIGHT_SCALARS:
            self.scalars[k] += sc[k]
        first = self.scalars['outside']
        if sc['outside'][0]:
            if not first[6] and sc['outside'][6]:
                first[1:8] = sc['outside'][1:8]
            first[0] += sc['outside'][0]

    def result(self, total_rays):
        """The reference's light_transport_diagnostics dict (CLARITAS.py:9455-9560)."""
        return diagnostics_dict(self.model, self.cfg, self.zmax, self.sum, self.scalars, total_rays,
                                self.ray_id_start, time.perf_counter() - self.t0)


def diagnostics_dict(model, c, zmax, a, sc, total_rays, ray_id_start, wall_s):
    grid_total = np.asarray(a['track_length'], dtype=np.float32)
    grid_post = np.asarray(a['post_scatter_track_length'], dtype=np.float32)
    grid_ew = np.asarray(a['scatter_event_weight'], dtype=np.float32)
    grid_ec = np.asarray(a['scatter_event_count'], dtype=np.uint32)
    grid_ac = np.asarray(a['absorption_event_count'], dtype=np.uint32)
    seg_total, dep_total = float(sc['segment_total']), float(sc['deposited_total'])
    abs_total = int(sc['absorption_event_total'])
    abs_grid = int(np.sum(grid_ac, dtype=np.uint64))
    ev_total = int(sc['scatter_event_total'])
    ev_grid = int(np.sum(grid_ec, dtype=np.uint64))
    nonfinite = sum(int(np.count_nonzero(~np.isfinite(q))) for q in (grid_total, grid_post, grid_ew))
    negative = sum(int(np.count_nonzero(q < 0)) for q in (grid_total, grid_post, grid_ew))
    out = sc['outside']
    first_outside = None
    if out[6]:
        xyz = np.array(out[1:4], dtype=np.uint32).view(np.float32).astype(float)
        first_outside = {
            'global_ray_id': int(out[4]), 'x_m': float(xyz[0]), 'y_m': float(xyz[1]), 'z_m': float(xyz[2]),
            'coordinate_tolerance_m': float(np.array([out[5]], dtype=np.uint32).view(np.float32)[0]),
            'x_bounds_m': [-float(model.rin), float(model.rin)], 'y_bounds_m': [-float(model.rin), float(model.rin)],
            'z_bounds_m': [float(model.zmin), float(zmax)],
        }
    rep, seg = int(c['representative_rays']), int(c['max_segments_per_ray'])
    return {
        'track_length_m': grid_total,
        'post_scatter_track_length_m': grid_post,
        'absorption_event_count': grid_ac,
        'absorption_event_total': abs_total,
        'deposited_absorption_event_total': abs_grid,
        'absorption_event_closure_error': abs_grid - abs_total,
        'absorption_grid_basis': 'accepted extinction encounter site; not volumetric absorbed energy',
        'scatter_event_weight': grid_ew,
        'scatter_event_count': grid_ec,
        'nx': int(c['nx']), 'ny': int(c['ny']), 'nz': int(c['nz']),
        'x_min_m': -model.rin, 'x_max_m': model.rin, 'y_min_m': -model.rin, 'y_max_m': model.rin,
        'z_min_m': model.zmin, 'z_max_m': zmax,
        'diagnostic_ray_limit': int(c['max_rays']),
        'diagnostic_rays_processed': max(0, min(int(total_rays), int(c['max_rays']) - int(ray_id_start))),
        'representative_ray_limit': rep,
        'max_segments_per_ray': seg,
        'segment_count': int(sc['segment_count']),
        'scatter_event_total': ev_total,
        'deposited_scatter_event_total': ev_grid,
        'scatter_event_closure_error': ev_grid - ev_total,
        'error_count': int(sc['error_count']),
        'hard_error_count': int(sc['error_count']),
        'boundary_adjustment_count': int(sc['boundary_adjustment_count']),
        'event_outside_count': int(out[0]),
        'first_outside_event': first_outside,
        'nonfinite_count': nonfinite,
        'negative_value_count': negative,
        'diagnostic_memory_bytes': int(memory_bytes(c)),
        'diagnostic_active_wall_time_s': float(wall_s),
        'segment_length_total_m': seg_total,
        'deposited_length_total_m': dep_total,
        'deposition_closure_relative_error': abs(dep_total - seg_total) / max(abs(seg_total), 1e-30),
        'grid_closure_relative_error': abs(float(np.sum(grid_total, dtype=np.float64)) - seg_total)
        / max(abs(seg_total), 1e-30),
        'trajectory_count': a['traj_count'][:rep].copy(),
        'trajectory_overflow': a['traj_overflow'][:rep].copy(),
        'trajectory_x0_m': a['traj_x0'][:rep].copy(), 'trajectory_y0_m': a['traj_y0'][:rep].copy(),
        'trajectory_z0_m': a['traj_z0'][:rep].copy(), 'trajectory_x1_m': a['traj_x1'][:rep].copy(),
        'trajectory_y1_m': a['traj_y1'][:rep].copy(), 'trajectory_z1_m': a['traj_z1'][:rep].copy(),
        'trajectory_particle_order': a['traj_order'][:rep].copy(),
        'trajectory_event_code': a['traj_event'][:rep].copy(),
        'selection_method': 'first global history ids; deterministic and independent of the transport random numbers',
        'normalization': 'raw metres of water-path track length or raw accepted scatter-event counts; normalized '
                         'views divide by diagnostic_rays_processed',
        'ray_statistical_weight': 1.0,
    }


def check_closure(diag):
    """The reference's integrity stops (CLARITAS.py:9466-9471): every
    absorbing encounter is deposited."""
    if diag['absorption_event_total'] != diag['deposited_absorption_event_total']:
        raise RuntimeError('accepted absorption-event deposition does not close')


def cross_sections(diag, cfg):
    """Source-axis (y-z about x = 0) and sensor-plane (x-y about z = 0) slabs
    of the 3-D grids, raw and per diagnostic ray and volume
    (CLARITAS.py:11645-11694)."""
    nx, ny, nz = (int(diag[k]) for k in ('nx', 'ny', 'nz'))
    span = lambda lo, hi, n: np.linspace(lo, hi, n, endpoint=False) + (hi - lo) / (2 * n)
    x = span(diag['x_min_m'], diag['x_max_m'], nx)
    y = span(diag['y_min_m'], diag['y_max_m'], ny)
    z = span(diag['z_min_m'], diag['z_max_m'], nz)
    dx = (diag['x_max_m'] - diag['x_min_m']) / nx
    dy = (diag['y_max_m'] - diag['y_min_m']) / ny
    dz = (diag['z_max_m'] - diag['z_min_m']) / nz
    rays = max(int(diag['diagnostic_rays_processed']), 1)
    vol = dx * dy * dz
    xi = np.flatnonzero(np.abs(x) <= 0.5 * cfg['vertical_slice_thickness_m'])
    zi = np.flatnonzero(np.abs(z) <= 0.5 * cfg['sensor_slice_thickness_m'])
    if not len(xi):
        xi = np.array([int(np.argmin(np.abs(x)))])
    if not len(zi):
        zi = np.array([int(np.argmin(np.abs(z)))])
    total = np.asarray(diag['track_length_m'], dtype=float)
    post = np.asarray(diag['post_scatter_track_length_m'], dtype=float)
    events = np.asarray(diag['scatter_event_count'], dtype=float)
    vr, vp, ve = total[:, :, xi].sum(2), post[:, :, xi].sum(2), events[:, :, xi].sum(2)
    hr, hp, he = total[zi].sum(0), post[zi].sum(0), events[zi].sum(0)
    nv, nh = rays * vol * len(xi), rays * vol * len(zi)
    return dict(x=x, y=y, z=z, xi=xi, zi=zi, dx=dx, dy=dy, dz=dz, voxel_volume=vol, vr=vr, vp=vp, ve=ve, hr=hr,
                hp=hp, he=he, vd=vr / nv, vpd=vp / nv, ved=ve / nv, hd=hr / nh, hpd=hp / nh, hed=he / nh)


def _safe(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, dict):
        return {str(k): _safe(q) for k, q in v.items()}
    if isinstance(v, (list, tuple)):
        return [_safe(q) for q in v]
    if isinstance(v, float) and not math.isfinite(v):
        return None
    return v


def _csv(path, fields, rows):
    with Path(path).open('w', newline='', encoding='utf-8') as f:
        w = csv.DictWriter(f, fieldnames=list(fields))
        w.writeheader()
        for r in rows:
            w.writerow({k: _safe(r.get(k)) for k in fields})


TRAJ_FIELDS = ('ray_id', 'segment_index', 'x0_m', 'y0_m', 'z0_m_sensor_frame', 'x1_m', 'y1_m', 'z1_m_sensor_frame',
               'start_height_above_bottom_m', 'end_height_above_bottom_m', 'segment_length_m', 'ray_weight',
               'particle_order_before_segment', 'segment_after_particle_scatter', 'terminal_event_code',
               'terminal_event_name', 'ray_truncated')


def write_outputs(out_dir, diag, sensor_height_m, cfg=None):
    """The reference's light-transport data files (CLARITAS.py:12108-12330,
    data part): light_transport_vertical_cross_section.csv,
    light_transport_sensor_plane_cross_section.csv,
    light_transport_representative_trajectories.csv,
    light_transport_volume.npz and light_transport_summary.json.  Heights
    'above bottom' add the sensor height to the sensor-frame z."""
    cfg = dict(DEFAULTS, **(cfg or {}))
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    q = cross_sections(diag, cfg)
    rays = max(int(diag['diagnostic_rays_processed']), 1)
    counts = np.asarray(diag['trajectory_count'], dtype=int)
    overflow = np.asarray(diag['trajectory_overflow'], dtype=int)
    tr = []
    for ray, n in enumerate(counts):
        for j in range(int(n)):
            p0 = np.array([diag['trajectory_x0_m'][ray, j], diag['trajectory_y0_m'][ray, j],
                           diag['trajectory_z0_m'][ray, j]], dtype=float)
            p1 = np.array([diag['trajectory_x1_m'][ray, j], diag['trajectory_y1_m'][ray, j],
                           diag['trajectory_z1_m'][ray, j]], dtype=float)
            code = int(diag['trajectory_event_code'][ray, j])
            order = int(diag['trajectory_particle_order'][ray, j])
            tr.append(dict(ray_id=ray, segment_index=j, x0_m=p0[0], y0_m=p0[1], z0_m_sensor_frame=p0[2], x1_m=p1[0],
                           y1_m=p1[1], z1_m_sensor_frame=p1[2], start_height_above_bottom_m=p0[2] + sensor_height_m,
                           end_height_above_bottom_m=p1[2] + sensor_height_m,
                           segment_length_m=float(np.sqrt(np.sum((p1 - p0) ** 2))),
                           ray_weight=float(diag['ray_statistical_weight']), particle_order_before_segment=order,
                           segment_after_particle_scatter=order > 0, terminal_event_code=code,
                           terminal_event_name=EVENT_NAMES.get(code, 'unknown'), ray_truncated=bool(overflow[ray])))
    vr = [dict(y_m=y, height_above_bottom_m=z + sensor_height_m, z_m_sensor_frame=z, track_length_m=q['vr'][iz, iy],
               post_scatter_track_length_m=q['vp'][iz, iy], scatter_event_count=int(q['ve'][iz, iy]),
               track_length_density_per_ray_m2=q['vd'][iz, iy], post_scatter_density_per_ray_m2=q['vpd'][iz, iy],
               scatter_event_density_per_ray_m3=q['ved'][iz, iy])
          for iz, z in enumerate(q['z']) for iy, y in enumerate(q['y'])]
    hr = [dict(x_m=x, y_m=y, track_length_m=q['hr'][iy, ix], post_scatter_track_length_m=q['hp'][iy, ix],
               scatter_event_count=int(q['he'][iy, ix]), track_length_density_per_ray_m2=q['hd'][iy, ix],
               post_scatter_density_per_ray_m2=q['hpd'][iy, ix], scatter_event_density_per_ray_m3=q['hed'][iy, ix])
          for iy, y in enumerate(q['y']) for ix, x in enumerate(q['x'])]
    _csv(out / 'light_transport_vertical_cross_section.csv', vr[0].keys(), vr)
    _csv(out / 'light_transport_sensor_plane_cross_section.csv', hr[0].keys(), hr)
    _csv(out / 'light_transport_representative_trajectories.csv', TRAJ_FIELDS, tr)
    np.savez_compressed(
        out / 'light_transport_volume.npz',
        total_track_length_weighted_m=diag['track_length_m'],
        post_scatter_track_length_weighted_m=diag['post_scatter_track_length_m'],
        particle_scatter_event_weight=diag['scatter_event_weight'],
        particle_scatter_event_count=diag['scatter_event_count'],
        particle_absorption_event_count=diag['absorption_event_count'],
        x_edges_m=np.linspace(diag['x_min_m'], diag['x_max_m'], int(diag['nx']) + 1),
        y_edges_m=np.linspace(diag['y_min_m'], diag['y_max_m'], int(diag['ny']) + 1),
        z_edges_above_bottom_m=np.linspace(diag['z_min_m'] + sensor_height_m, diag['z_max_m'] + sensor_height_m,
                                           int(diag['nz']) + 1),
        x_centres_m=q['x'], y_centres_m=q['y'], z_centres_above_bottom_m=q['z'] + sensor_height_m,
        diagnostic_rays_processed=np.asarray(rays), ray_statistical_weight=np.asarray(diag['ray_statistical_weight']),
    )
    summary = {
        'diagnostics_enabled': True,
        'provenance': 'device production histories in water; no synthetic rays',
        'diagnostic_rays_processed': int(diag['diagnostic_rays_processed']),
        'diagnostic_ray_limit': int(diag['diagnostic_ray_limit']),
        'grid_shape_zyx': [int(diag['nz']), int(diag['ny']), int(diag['nx'])],
        'grid_bounds_m': {'x': [diag['x_min_m'], diag['x_max_m']], 'y': [diag['y_min_m'], diag['y_max_m']],
                          'z_sensor_frame': [diag['z_min_m'], diag['z_max_m']]},
        'voxel_size_m': [q['dx'], q['dy'], q['dz']],
        'raw_quantity': 'water track length (m) and accepted scatter-event count',
        'normalized_quantity': 'track-length density per diagnostic ray (m^-2 ray^-1); event density per diagnostic '
                               'ray (m^-3 ray^-1)',
        'diagnostic_memory_bytes': int(diag['diagnostic_memory_bytes']),
        'diagnostic_active_wall_time_s': float(diag['diagnostic_active_wall_time_s']),
    }
    for k in ('segment_count', 'absorption_event_total', 'deposited_absorption_event_total',
              'absorption_event_closure_error', 'scatter_event_total', 'deposited_scatter_event_total',
              'scatter_event_closure_error', 'boundary_adjustment_count', 'event_outside_count', 'nonfinite_count',
              'negative_value_count'):
        summary[k] = int(diag[k])
    for k in ('absorption_grid_basis', 'segment_length_total_m', 'deposited_length_total_m',
              'deposition_closure_relative_error', 'grid_closure_relative_error'):
        summary[k] = diag[k]
    summary.update(
        diagnostic_error_count=int(diag['error_count']), diagnostic_hard_error_count=int(diag['hard_error_count']),
        representative_rays_requested=int(diag['representative_ray_limit']),
        representative_rays_with_segments=int(np.count_nonzero(counts)),
        representative_trajectory_segments=int(np.sum(counts)),
        representative_trajectory_truncated_rays=int(np.count_nonzero(overflow)),
        max_segments_per_ray=int(diag['max_segments_per_ray']),
        representative_ray_selection=diag['selection_method'],
        vertical_slice_thickness_requested_m=cfg['vertical_slice_thickness_m'],
        vertical_slice_thickness_actual_m=len(q['xi']) * q['dx'],
        sensor_slice_thickness_requested_m=cfg['sensor_slice_thickness_m'],
        sensor_slice_thickness_actual_m=len(q['zi']) * q['dz'],
    )
    (out / 'light_transport_summary.json').write_text(json.dumps(_safe(summary), indent=2))
    return summary
