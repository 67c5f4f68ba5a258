"""Inputs with answers fixed by construction or by independent mathematics,
shared by the oracle pins (CPU) and the device probe tests (GPU).

Nothing here restates the transport's formulas: detector rays are built so
that geometry alone decides acceptance (on a bore axis, or just inside /
outside one aperture while the others are clear), free-surface hits come
from mpmath roots of z(t) - z_s(r(t)), wall hits from the exact quadratic in
mpmath, and sampler checks from the closed-form distributions (scipy).
"""
from __future__ import annotations

import math

import mpmath as mp
import numpy as np
from scipy import integrate, stats

from tardiis_and_claritas_amd import physics

# ------------------------------------------------------------- detectors


def _bore_frame(j):
    th = math.radians(10.0 * j)
    u = np.array([math.sin(th), math.cos(th), 0.0])   # bore axis (CLARITAS.py:4242)
    lat = np.array([math.cos(th), -math.sin(th), 0.0])  # horizontal, perpendicular to the axis
    return u, lat


def _ray_through(point, direction, back):
    d = np.asarray(direction, float)
    d /= np.linalg.norm(d)
    p = np.asarray(point, float) - back * d
    return np.r_[p, d]


def detector_rays(m, rel=1e-3):
    """Exit rays whose hardware bore follows from the aperture stack's
    geometry: returns (rays [n][6] float32, native ids, mirror ids, labels).

    Per bore j (axis u_j, apertures: radius throat_r at u.p = ring_in and at
    throat_out, counterbore_r at ring_out):
      - the axis itself, started inside r = ring_in: bore j;
      - parallel to the axis, offset vertically by throat_r (1 -+ rel): j / none;
      - pivoting about the throat_out axis point, tilted vertically so the
        offset at the ring_in plane is throat_r (1 -+ rel) (ring_in binds,
        the ring_out offset stays inside counterbore_r): j / none;
      - starting on the axis at the ring_in plane, tilted so the throat_out
        offset is throat_r (1 -+ rel) (throat_out binds when
        ring_out - ring_in times the tilt stays inside counterbore_r): j / none;
      - entering the ring_in plane at z0 = -0.75 throat_r and tilted so the
        ring_out offset is counterbore_r (1 -+ rel) while the throat offset
        stays inside throat_r (ring_out binds): j / none;
      - every native ray reflected x -> -x is the mirror image: it is
        accepted by the mirror scan at j (and natively only by bore 0).
    Vertical offsets and tilts move the ray away from every other bore, whose
    axes are at >= 10 degrees horizontally (a gap of at least
    ring_in tan 10 deg at their ring_in plane)."""
    ri, to, ro = m.ring_in, m.throat_out, m.ring_out
    rt, rc = m.throat_radius, m.counterbore_radius
    rays, nat, mir, lab = [], [], [], []

    def add(ray, j_ok, label):
        rays.append(ray)
        nat.append(j_ok)
        mir.append(0 if j_ok == 0 else -1)
        lab.append(label)
        r2 = ray.copy()
        r2[0] *= -1.0
        r2[3] *= -1.0
        rays.append(r2)
        nat.append(0 if j_ok == 0 else -1)
        mir.append(j_ok)
        lab.append(label + ' (mirror image)')

    back = 0.5 * (ri - m.rout)
    for j in range(18):
        u, lat = _bore_frame(j)
        ez = np.array([0.0, 0.0, 1.0])
        add(_ray_through(ri * u, u, back), j, f'axis {j}')
        for s, ok in ((1 - rel, j), (1 + rel, -1)):
            add(_ray_through(ri * u + s * rt * ez, u, back), ok, f'offset {j} {s}')
            # ring_in binds: pivot at the throat_out axis point
            tan_a = s * rt / (to - ri)
            if (ro - to) * tan_a < rc:
                d = u + tan_a * ez
                add(_ray_through(to * u, d, back + (to - ri) * np.linalg.norm(d)), ok, f'ring_in {j} {s}')
            # throat_out binds: start on the axis at the ring_in plane
            tan_b = s * rt / (to - ri)
            if (ro - ri) * tan_b < rc * (1 - 2 * rel):
                d = u + tan_b * ez
                add(_ray_through(ri * u, d, back), ok, f'throat {j} {s}')
            # ring_out binds
            z0 = -0.75 * rt
            tan_c = (s * rc - z0) / (ro - ri)
            if abs(z0 + (to - ri) * tan_c) < rt * (1 - 2 * rel):
                d = u + tan_c * ez
                add(_ray_through(ri * u + z0 * ez, d, back), ok, f'ring_out {j} {s}')
    del lat
    return np.asarray(rays, np.float32), np.asarray(nat), np.asarray(mir), lab


def random_exit_rays(m, n, seed):
    """Rays leaving the outer acrylic wall outwards in random directions,
    aimed near the bores (for device-vs-oracle agreement of the full scan)."""
    g = np.random.default_rng(seed)
    ph = np.radians(g.uniform(-185.0, 185.0, n))
    z = g.uniform(-1.5, 1.5, n) * m.counterbore_radius
    p = np.stack([m.rout * np.sin(ph), m.rout * np.cos(ph), z], 1)
    aim = np.radians(np.round(np.degrees(ph) / 10.0) * 10.0 + g.normal(0.0, 3.0, n))
    tgt = np.stack([m.ring_out * np.sin(aim), m.ring_out * np.cos(aim), g.normal(0.0, 1.0, n) * m.throat_radius], 1)
    d = tgt - p
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([p, d], 1).astype(np.float32)


# ----------------------------------------------------------- free surface


def surface_rays(m, n, seed):
    """Upward rays from below the free surface of model m's cell."""
    R = m.rin
    g = physics.vortex_surface_geometry(m.free_surface_model, m.zmean, R,
                                        0.0 if m.free_surface_model == 'flat' else m.vortex_delta_h,
                                        m.vortex_core_radius, m.gravity)
    rng = np.random.default_rng(seed)
    rr = R * np.sqrt(rng.uniform(0, 0.8, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    z0 = g['center_z'] - rng.uniform(0.002, 0.01, n)
    th = rng.uniform(0.0, 1.2, n)
    az = rng.uniform(0, 2 * np.pi, n)
    return np.stack([rr * np.cos(ph), rr * np.sin(ph), z0, np.sin(th) * np.cos(az), np.sin(th) * np.sin(az),
                     np.cos(th)], 1).astype(np.float32), g


def surface_height_mp(m, g, q):
    """Free-surface height at r^2 = q (mpmath), from the held geometry."""
    surface, zm, R, D, a = m.free_surface_model, m.zmean, m.rin, m.vortex_delta_h, m.vortex_core_radius
    if surface == 'flat' or D <= 0:
        return mp.mpf(zm)
    if surface == 'parabolic_vortex':
        return mp.mpf(zm) + D * (q / R ** 2 - mp.mpf(0.5))
    if surface == 'localized_scully_vortex':
        return mp.mpf(g['center_z']) + g['h_inf'] * q / (a * a + q)
    K = g['h_inf'] / 2
    return mp.mpf(g['center_z']) + (K * q / a ** 2 if q < a * a else K * (2 - a * a / q))


def surface_roots_mp(m, g, rays, step=2e-4, n_steps=400):
    """First root of z(t) - z_s(r(t)) inside the cell for each ray (None if
    the ray reaches the wall first); mpmath at 30 digits."""
    mp.mp.dps = 30
    out = []
    for ray in rays:
        x, y, z, vx, vy, vz = (mp.mpf(float(v)) for v in ray)
        f = lambda t: z + t * vz - surface_height_mp(m, g, (x + t * vx) ** 2 + (y + t * vy) ** 2)
        prev, root = f(mp.mpf(0)), None
        for k in range(1, n_steps):
            t = mp.mpf(k) * mp.mpf(step)
            if (x + t * vx) ** 2 + (y + t * vy) ** 2 > mp.mpf(m.rin) ** 2:
                break
            ft = f(t)
            if ft * prev <= 0:
                root = float(mp.findroot(f, (t - mp.mpf(step), t), solver='bisect'))
                break
            prev = ft
        out.append(root)
    return out


def wall_exit_mp(rays, R):
    """Forward distance to the cylinder r = R from inside (mpmath)."""
    mp.mp.dps = 30
    t = []
    for ray in rays:
        X, Y, _, VX, VY, _ = (mp.mpf(float(v)) for v in ray)
        a, b, c = VX * VX + VY * VY, X * VX + Y * VY, X * X + Y * Y - mp.mpf(R) ** 2
        t.append(float((-b + mp.sqrt(b * b - a * c)) / a))
    return np.asarray(t)


def water_rays(m, n, seed):
    """Rays from random points inside the water in random directions."""
    rng = np.random.default_rng(seed)
    r0 = m.rin * np.sqrt(rng.uniform(0, 0.95, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(m.zmin + 1e-3, m.zmean - 1e-3, n)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return np.stack([r0 * np.cos(ph), r0 * np.sin(ph), z, v[:, 0], v[:, 1], v[:, 2]], 1).astype(np.float32)


# --------------------------------------------------------------- samplers


def beta_radiance_cdf(a, b):
    """CDF of the polar angle x in [0, pi/2] with density
    proportional to BetaPDF(x / (pi/2); a, b) sin x (the 3-D Beta radiance
    law, CLARITAS.py:1193-1198), by adaptive quadrature of the closed-form
    density."""
    c = 0.5 * math.pi
    pdf = lambda x: stats.beta.pdf(x / c, a, b) * math.sin(x)
    grid = np.concatenate([np.linspace(0.0, 0.05, 400, endpoint=False), np.linspace(0.05, c, 2000)])
    cum = np.zeros(grid.size)
    for k in range(1, grid.size):
        cum[k] = cum[k - 1] + integrate.quad(pdf, grid[k - 1], grid[k], limit=200)[0]
    cum /= cum[-1]
    return lambda x: np.interp(x, grid, cum)


def chi2_pvalue(counts, probs):
    counts = np.asarray(counts, float)
    exp = np.asarray(probs, float) * counts.sum()
    chi2 = float(((counts - exp) ** 2 / exp).sum())
    return stats.chi2.sf(chi2, counts.size - 1), chi2


def theta_bins(theta, cdf, n_bins=100):
    """Edges (indices of the theta grid) of about n_bins bins of equal
    probability under the FP64 CDF, and their probabilities."""
    targets = np.linspace(0.0, 1.0, n_bins + 1)[1:-1]
    idx = np.unique(np.r_[0, np.searchsorted(cdf, targets), theta.size - 1])
    probs = np.diff(cdf[idx])
    keep = probs > 0
    return idx, probs, keep
