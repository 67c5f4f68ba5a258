"""Meridional advection-diffusion sediment field on the device.

The reference's sediment model 'schiller_naumann_meridional_advection_diffusion'
solves, per particle bin, the stationary axisymmetric transport of the
number density in the vortex's secondary circulation plus the particles'
settling and finite-Re radial slip (CLARITAS.py:4040-4226 kernels,
6451-6725 driver).  `solve` runs that solve through `clr_meridional_solve`
and returns the host copy; the transport then reads the grid as the
per-bin r-z field of sediment model 3.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from ._lib import ClaritasError

NOT_CONVERGED = -7


@dataclass
class MeridionalSettings:
    """Solver controls, defaults and limits of CLARITAS.py:5625-5631,5873-5887."""
    circulation_ratio: float = 0.10
    nr: int = 48
    nz: int = 96
    max_iterations: int = 8000
    check_interval: int = 250
    tolerance: float = 1.0e-6
    sor: float = 1.75

    def validate(self):
        if not (np.isfinite(self.circulation_ratio) and 0.0 <= self.circulation_ratio <= 0.5):
            raise ValueError('meridional_circulation_ratio must be finite and in [0,0.5]')
        if self.nr < 16 or self.nz < 32:
            raise ValueError('meridional solver grid must be at least 16 x 32')
        if self.max_iterations < 250 or self.check_interval < 10:
            raise ValueError('invalid meridional solver iteration controls')
        if not (np.isfinite(self.tolerance) and 0.0 < self.tolerance < 1e-2):
            raise ValueError('meridional_solver_tolerance must be in (0,1e-2)')
        if not (np.isfinite(self.sor) and 1.0 <= self.sor < 1.95):
            raise ValueError('meridional_solver_sor must be in [1,1.95)')

    def struct(self) -> abi.Meridional:
        m = abi.Meridional()
        m.circulation_ratio, m.tolerance, m.sor = float(self.circulation_ratio), float(self.tolerance), float(self.sor)
        m.nr, m.nz = int(self.nr), int(self.nz)
        m.max_iterations, m.check_interval = int(self.max_iterations), int(self.check_interval)
        return m


def result_dict(out: abi.MeridionalOut, arrs: dict, rc: int, message: str = '') -> dict:
    """Host view of one solve (the reference's meridional host diagnostics,
    CLARITAS.py:6915-6929, plus the grid and extrema)."""
    res = dict(arrs)
    res.update(grid_zmax=float(out.grid_zmax), vtheta_peak=float(out.vtheta_peak), peak_speed=float(out.peak_speed),
               iterations=int(out.iterations), converged=(rc == 0), status=int(rc), message=message)
    res['max_residual'] = float(np.max(arrs['residual'])) if arrs['residual'].size else 0.0
    return res


def solve(ctx, cfg: abi.Config, particles: abi.Particles, sediment: abi.Sediment, settings: MeridionalSettings,
          coeff: bool = False, raise_on_failure: bool = True) -> dict:
    """Run the device solve.  Raises ClaritasError when it fails or ends above
    its tolerance (as the reference raises), unless raise_on_failure is False,
    in which case the partial result carries status and message."""
    settings.validate()
    out, arrs = abi.meridional_out(int(particles.n), int(settings.nr), int(settings.nz), coeff=coeff)
    mp = settings.struct()
    rc = int(ctx.lib.clr_meridional_solve(ctx.h, C.byref(cfg), C.byref(particles), C.byref(sediment), C.byref(mp),
                                          C.byref(out)))
    msg = ''
    if rc:
        raw = ctx.lib.clr_last_error(ctx.h)
        msg = raw.decode() if raw else ''
        if raise_on_failure:
            raise ClaritasError(rc, f'clr_meridional_solve: {msg}')
    return result_dict(out, arrs, rc, msg)


def measure(repeats=5, n_bins=40, nr=48, nz=96, chi=0.10, device=0):
    """Time the device solve on the reference's default cell (Scully vortex,
    wall-centre height 50 mm, core 20 mm, CLARITAS.py:5577-5631) for a
    lognormal PSD (median 20 um, sigma_g 1.8).  Returns a summary dict."""
    import time

    from .model import ClaritasForwardModel
    from .physics import lognormal

    m = ClaritasForwardModel(sediment_transport_model='schiller_naumann_meridional_advection_diffusion',
                             particle_event_model='geometric', meridional_grid_nr=nr, meridional_grid_nz=nz,
                             meridional_circulation_ratio=chi, device=device)
    mat = lognormal(20e-6, 1.8, n_bins, 1.46, 1e-4)
    cfg = m.config_struct(mat)
    p, keep = m.particles_struct(mat, 0.0, None, 'mass_fraction')
    sed = m.sediment_struct(mat)
    times, res = [], None
    for _ in range(int(repeats)):
        t0 = time.perf_counter()
        res = solve(m.ctx, cfg, p, sed, m.meridional)
        times.append(time.perf_counter() - t0)
    del keep
    it = res['iterations']
    best = min(times)
    return dict(n_bins=n_bins, nr=nr, nz=nz, iterations=it, max_residual=res['max_residual'],
                solve_ms_best=best * 1e3, solve_ms_median=float(np.median(times)) * 1e3,
                cell_updates_per_s=n_bins * nr * nz * it / best)

