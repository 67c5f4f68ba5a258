"""Particle populations, cell geometry and sediment fields (host side).

Pure numpy helpers that turn the reference's user-facing parameter set into
the POD structures of the C ABI.  No transport physics runs here.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from functools import lru_cache
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / 'data' / 'materials.json'
DETECTOR_ANGLES_DEG = np.arange(0, 180, 10, dtype=float)  # CLARITAS.py:420


@dataclass
class Material:
    """One particle population: size bins, PSD weights and complex index."""

    name: str
    diameter_m: np.ndarray
    weights: np.ndarray
    n_real: np.ndarray
    k_imag: np.ndarray
    density_kg_m3: float = 2650.0
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.diameter_m = np.asarray(self.diameter_m, dtype=np.float64).ravel()
        n = self.diameter_m.size
        self.weights = np.broadcast_to(np.asarray(self.weights, dtype=np.float64), (n,)).copy()
        self.n_real = np.broadcast_to(np.asarray(self.n_real, dtype=np.float64), (n,)).copy()
        self.k_imag = np.broadcast_to(np.asarray(self.k_imag, dtype=np.float64), (n,)).copy()
        if n < 1 or n > 4096:
            raise ValueError('a material needs 1..4096 size bins')
        if np.any(~np.isfinite(self.diameter_m)) or np.any(self.diameter_m <= 0):
            raise ValueError('diameters must be finite and positive')
        if np.any(~np.isfinite(self.weights)) or np.any(self.weights < 0) or not np.any(self.weights > 0):
            raise ValueError('invalid PSD weights')
        if np.any(self.n_real <= 0) or np.any(self.k_imag < 0):
            raise ValueError('invalid complex refractive index')


@lru_cache(maxsize=1)
def _tables():
    return json.loads(DATA.read_text())


def reference_material(name: str, density_kg_m3: float = 2650.0, optical_override=None) -> Material:
    """The reference's measured PSDs with their exploratory n(d), k(d)
    profiles (CLARITAS.py:248-705, get_material_psd_raw CLARITAS.py:5215,
    get_material_optical_constants CLARITAS.py:1082-1112)."""
    m = str(name).strip().lower()
    t = _tables()
    if m not in t:
        raise ValueError("material must be 'loess' or 'kaolin'")
    d = t[m]
    n_real, k_imag = np.asarray(d['n_real']), np.asarray(d['k_imag'])
    if optical_override:
        q = optical_override.get(m, optical_override.get(name))
        if q is not None:
            n_real = np.asarray(q.get('n_real', n_real), dtype=np.float64)
            k_imag = np.asarray(q.get('k_imag', k_imag), dtype=np.float64)
    return Material(m, d['diameter_m'], d['weights'], n_real, k_imag, density_kg_m3)


def monodisperse(diameter_m: float, n_real: float, k_imag: float, density_kg_m3: float = 2650.0) -> Material:
    return Material('monodisperse', [diameter_m], [1.0], [n_real], [k_imag], density_kg_m3,
                    {'diameter_m': diameter_m})


def lognormal(median_m: float, sigma_g: float, n_bins: int = 40, n_real: float = 1.46,
              k_imag: float = 1e-4, span: float = 3.0, density_kg_m3: float = 2650.0) -> Material:
    """Mass-weighted lognormal PSD: n_bins log-spaced bin centres over
    median * sigma_g**[-span, span]; each bin's mass fraction is the lognormal
    probability of its log-width cell (the median is the mass median)."""
    if not (median_m > 0 and sigma_g > 1.0 and n_bins >= 1):
        raise ValueError('lognormal PSD needs median > 0, sigma_g > 1, n_bins >= 1')
    ls = math.log(sigma_g)
    edges = np.linspace(-span * ls, span * ls, n_bins + 1)
    centres = 0.5 * (edges[1:] + edges[:-1])
    cdf = 0.5 * (1.0 + np.vectorize(math.erf)(edges / (ls * math.sqrt(2.0))))
    w = np.diff(cdf)
    return Material('lognormal', median_m * np.exp(centres), w / w.sum(), n_real, k_imag, density_kg_m3,
                    {'median_m': median_m, 'sigma_g': sigma_g, 'n_bins': n_bins})


def cell_geometry(diameter_m=0.098, water_level_m=0.142, sensor_height_m=0.093, cell_height_m=0.491,
                  wall_m=0.0035, ring_clearance_m=0.0005, ring_width_m=0.015, counterbore_depth_m=0.008,
                  through_bore_diameter_m=0.004, counterbore_diameter_m=0.0087,
                  bottom_disc_thickness_m=0.003, source_launch_radius_m=None) -> dict:
    """Constructor kwargs of a cell from its bore diameter and heights,
    following _resolved_geometry (CLARITAS.py:10024-10052): tube outer
    radius = bore radius + wall, ring radii follow the wall, the source
    launches at the ring's outer edge."""
    rin = 0.5 * diameter_m
    rout = rin + wall_m
    ring_in = rout + ring_clearance_m
    ring_out = ring_in + ring_width_m
    return dict(
        tube_inner_radius_m=rin, tube_outer_radius_m=rout, tube_length_m=cell_height_m,
        water_height_m=water_level_m, sensor_height_above_bottom_m=sensor_height_m,
        sensor_ring_inner_radius_m=ring_in, sensor_ring_outer_radius_m=ring_out,
        counterbore_depth_m=counterbore_depth_m, through_bore_diameter_m=through_bore_diameter_m,
        counterbore_diameter_m=counterbore_diameter_m,
        source_launch_radius_m=ring_out if source_launch_radius_m is None else source_launch_radius_m,
        bottom_disc_thickness_m=bottom_disc_thickness_m,
    )


def vortex_surface_geometry(model, zmean, radius, delta_h, core_radius, gravity_m_s2=9.81):
    """Centre/wall heights and equivalent rotation of the free surface.
    Flat, parabolic and Scully forms follow CLARITAS.py:708-792; the Rankine
    form (solid core of radius a, potential 1/r outside) is an extension with
    the same volume-preserving convention (mean height = zmean)."""
    model = str(model).strip().lower()
    R, D, a, g = float(radius), float(delta_h), float(core_radius), float(gravity_m_s2)
    if not (R > 0 and D >= 0 and g > 0):
        raise ValueError('invalid vortex parameters')
    nan = float('nan')
    if model == 'flat' or D == 0.0:
        return dict(center_z=zmean, wall_z=zmean, h_inf=0.0, h_bar=0.0, angular_velocity_rad_s=0.0,
                    vtheta_peak_m_s=0.0, equivalent_rigid_body_rpm=0.0, equivalent_core_rpm=0.0)
    if model == 'parabolic_vortex':
        om = math.sqrt(2.0 * g * D / (R * R))
        return dict(center_z=zmean - 0.5 * D, wall_z=zmean + 0.5 * D, h_inf=D, h_bar=0.5 * D,
                    angular_velocity_rad_s=om, vtheta_peak_m_s=om * R,
                    equivalent_rigid_body_rpm=om * 60.0 / (2.0 * math.pi), equivalent_core_rpm=nan)
    if model == 'localized_scully_vortex':
        if not a > 0:
            raise ValueError('Scully vortex needs a positive core radius')
        a2, R2 = a * a, R * R
        hinf = D * (a2 + R2) / R2
        hbar = hinf * (1.0 - (a2 / R2) * math.log1p(R2 / a2))
        om = math.sqrt(2.0 * g * hinf / a2)
        return dict(center_z=zmean - hbar, wall_z=zmean - hbar + D, h_inf=hinf, h_bar=hbar,
                    angular_velocity_rad_s=om, vtheta_peak_m_s=0.5 * om * a,
                    equivalent_rigid_body_rpm=nan, equivalent_core_rpm=om * 60.0 / (2.0 * math.pi))
    if model == 'rankine_vortex':
        if not (0 < a < R):
            raise ValueError('Rankine vortex needs 0 < core radius < vessel radius')
        a2, R2 = a * a, R * R
        K = D / (2.0 - a2 / R2)                     # centre-to-core-edge rise
        hbar = (2.0 * K / R2) * (R2 - 0.75 * a2 - a2 * math.log(R / a))
        om = math.sqrt(2.0 * g * K) / a             # core angular velocity
        return dict(center_z=zmean - hbar, wall_z=zmean - hbar + D, h_inf=2.0 * K, h_bar=hbar,
                    angular_velocity_rad_s=om, vtheta_peak_m_s=om * a,
                    equivalent_rigid_body_rpm=nan, equivalent_core_rpm=om * 60.0 / (2.0 * math.pi))
    raise ValueError(f'unknown free-surface model {model!r}')


def rankine_surface_z(r, zmean, radius, delta_h, core_radius):
    """Height of the Rankine free surface at radius r (numpy)."""
    v = vortex_surface_geometry('rankine_vortex', zmean, radius, delta_h, core_radius)
    a = float(core_radius)
    K = 0.5 * v['h_inf']
    r = np.asarray(r, dtype=np.float64)
    inner = K * r * r / (a * a)
    outer = K * (2.0 - a * a / np.maximum(r * r, 1e-300))
    return v['center_z'] + np.where(r < a, inner, outer)


def rouse_rz_field(rin, zmin, zmax_field, nr, nz, rouse_number, surface_z=None, ref_fraction=0.05):
    """Steady Rouse profile of suspended sediment on a cell-centred r-z grid.

    c(h)/c(a) = [((H - h)/h) (a/(H - a))]^P with h the height above the bed,
    H the local water depth, a = ref_fraction*H the reference level and P the
    Rouse number (Rouse 1937).  Below a the concentration is held at c(a);
    above the local surface it is extended from the last wet cell.  The field
    is normalised to a volume mean of 1 over the wet region, so the transport
    kernel can multiply the bulk rates by it.  Returns float32 [nr][nz].
    """
    dr = rin / nr
    dz = (zmax_field - zmin) / nz
    r = (np.arange(nr) + 0.5) * dr
    z = zmin + (np.arange(nz) + 0.5) * dz
    zs = np.full(nr, zmin + 0.9 * (zmax_field - zmin)) if surface_z is None else np.asarray(surface_z(r))
    H = np.maximum(zs - zmin, 1e-9)[:, None]
    h = np.maximum(z[None, :] - zmin, 1e-12)
    a = ref_fraction * H
    hh = np.clip(h, a, 0.999 * H)
    prof = (((H - hh) / hh) * (a / (H - a))) ** float(rouse_number)
    wet = h < H
    # extend above the surface from the last wet cell in each column
    for i in range(nr):
        idx = np.nonzero(wet[i])[0]
        if idx.size:
            prof[i, idx[-1] + 1:] = prof[i, idx[-1]]
    vol = (2.0 * np.pi * r * dr)[:, None] * dz * wet
    mean = float(np.sum(prof * vol) / np.sum(vol))
    return (prof / mean).astype(np.float32)


def segregated_xyz_field(rin, zmin, zmax_field, nx, ny, nz, seed=0, rouse_number=0.3, vortex_strength=0.5,
                         surface_z=None):
    """3-D FP32 scale field with radial vortex segregation: a Rouse column
    modulated by exp(-k r^2) radial depletion toward the axis plus a seeded
    low-amplitude azimuthal perturbation; volume mean 1 over the wet cylinder.
    Layout [nx][ny][nz] over [-rin, rin]^2 x [zmin, zmax_field]."""
    rng = np.random.default_rng(seed)
    x = -rin + (np.arange(nx) + 0.5) * (2 * rin / nx)
    y = -rin + (np.arange(ny) + 0.5) * (2 * rin / ny)
    z = zmin + (np.arange(nz) + 0.5) * ((zmax_field - zmin) / nz)
    X, Y = np.meshgrid(x, y, indexing='ij')
    R = np.hypot(X, Y)
    th = np.arctan2(Y, X)
    zs = np.full_like(R, zmin + 0.9 * (zmax_field - zmin)) if surface_z is None else surface_z(R)
    H = np.maximum(zs - zmin, 1e-9)[:, :, None]
    h = np.maximum(z[None, None, :] - zmin, 1e-12)
    a = 0.05 * H
    hh = np.clip(h, a, 0.999 * H)
    col = (((H - hh) / hh) * (a / (H - a))) ** float(rouse_number)
    radial = 1.0 - vortex_strength * np.exp(-(R / (0.35 * rin)) ** 2)
    phases = rng.uniform(0, 2 * np.pi, 3)
    az = 1.0 + 0.05 * sum(np.cos((k + 2) * th + phases[k]) for k in range(3))
    f = col * (radial * az)[:, :, None]
    inside = (R < rin)[:, :, None] & (h < H)
    f = np.where(R[:, :, None] < rin, f, f)  # outside the bore the kernel never samples
    mean = float(np.sum(f * inside) / np.sum(inside))
    return (f / mean).astype(np.float32)
