"""Arbitrary-precision Lorenz-Mie fixtures (mpmath), independent of the code
under test.

a_n, b_n come straight from the Riccati-Bessel definitions
    psi_n(z) = z j_n(z),  xi_n(z) = z h_n^(1)(z),
    a_n = [m psi_n(mx) psi_n'(x) - psi_n(x) psi_n'(mx)] /
          [m psi_n(mx) xi_n'(x)  - xi_n(x)  psi_n'(mx)],
    b_n = [psi_n(mx) psi_n'(x) - m psi_n(x) psi_n'(mx)] /
          [psi_n(mx) xi_n'(x)  - m xi_n(x)  psi_n'(mx)]
(Bohren & Huffman 1983, eqs. 4.53, 4.56-4.57), with spherical Bessel
functions from mpmath's cylinder functions of half-integer order at 40
digits, and the efficiencies / amplitude functions from their series
(B&H eqs. 4.61-4.62, 4.74, 4.79).  This shares no recurrence with the
logarithmic-derivative algorithm it checks.

Run: python tests/golden/make_mie_golden.py  -> tests/golden/mie_mpmath.json
"""
import json
from pathlib import Path

import mpmath as mp

mp.mp.dps = 40


def sph_jn(n, z):
    return mp.sqrt(mp.pi / (2 * z)) * mp.besselj(n + mp.mpf(1) / 2, z)


def sph_yn(n, z):
    return mp.sqrt(mp.pi / (2 * z)) * mp.bessely(n + mp.mpf(1) / 2, z)


def psi(n, z):
    return z * sph_jn(n, z)


def xi(n, z):
    return z * (sph_jn(n, z) + 1j * sph_yn(n, z))


def dpsi(n, z):
    # psi_n'(z) = psi_{n-1}(z) - n psi_n(z)/z
    return psi(n - 1, z) - n * psi(n, z) / z


def dxi(n, z):
    return xi(n - 1, z) - n * xi(n, z) / z


def mie(d, nr, k, n_med, wl, angles_deg=()):
    x = mp.pi * mp.mpf(n_med) * mp.mpf(d) / mp.mpf(wl)
    m = mp.mpc(nr, k) / mp.mpf(n_med)
    mx = m * x
    nstop = int(float(x) + 4 * float(x) ** (1 / 3) + 2)
    a, b = [], []
    for n in range(1, nstop + 1):
        pmx, dpmx = psi(n, mx), dpsi(n, mx)
        px, dpx, xx, dxx = psi(n, x), dpsi(n, x), xi(n, x), dxi(n, x)
        a.append((m * pmx * dpx - px * dpmx) / (m * pmx * dxx - xx * dpmx))
        b.append((pmx * dpx - m * px * dpmx) / (pmx * dxx - m * xx * dpmx))
    x2 = x * x
    qe = 2 / x2 * sum((2 * n + 1) * mp.re(a[n - 1] + b[n - 1]) for n in range(1, nstop + 1))
    qs = 2 / x2 * sum((2 * n + 1) * (abs(a[n - 1]) ** 2 + abs(b[n - 1]) ** 2) for n in range(1, nstop + 1))
    sb = sum((2 * n + 1) * (-1) ** n * (a[n - 1] - b[n - 1]) for n in range(1, nstop + 1))
    qb = abs(sb) ** 2 / x2
    # asymmetry parameter, B&H eq. 4.79
    g = mp.mpf(0)
    for n in range(1, nstop + 1):
        if n < nstop:
            g += mp.mpf(n) * (n + 2) / (n + 1) * mp.re(a[n - 1] * mp.conj(a[n]) + b[n - 1] * mp.conj(b[n]))
        g += mp.mpf(2 * n + 1) / (n * (n + 1)) * mp.re(a[n - 1] * mp.conj(b[n - 1]))
    g = 4 / (x2 * qs) * g
    s11 = []
    for ang in angles_deg:
        mu = mp.cos(mp.radians(ang))
        pm1, pn = mp.mpf(0), mp.mpf(1)
        s1 = s2 = mp.mpc(0)
        for n in range(1, nstop + 1):
            tau = n * mu * pn - (n + 1) * pm1
            c = mp.mpf(2 * n + 1) / (n * (n + 1))
            s1 += c * (a[n - 1] * pn + b[n - 1] * tau)
            s2 += c * (a[n - 1] * tau + b[n - 1] * pn)
            pm1, pn = pn, ((2 * n + 1) * mu * pn - (n + 1) * pm1) / n
        s11.append(float((abs(s1) ** 2 + abs(s2) ** 2) / 2))
    return dict(x=float(x), nstop=nstop, qext=float(qe), qsca=float(qs), qabs=float(qe - qs), qback=float(qb),
                g=float(g), angles_deg=list(angles_deg), s11=s11,
                a1=[float(mp.re(a[0])), float(mp.im(a[0]))], b1=[float(mp.re(b[0])), float(mp.im(b[0]))])


CASES = [
    # name, diameter m, n_real, k_imag, n_medium, wavelength m
    # Bohren & Huffman (1983) Appendix A test case: radius 0.525 um, 0.6328 um, m=1.55
    ('bohren_huffman_appendix_a', 1.05e-6, 1.55, 0.0, 1.0, 0.6328e-6),
    ('bohren_huffman_x3_m155', 3.0 * 0.6328e-6 / 3.141592653589793, 1.55, 0.0, 1.0, 0.6328e-6),
    ('config1_sphere', 10e-6, 1.46, 1e-4, 1.33, 860e-9),
    ('rayleigh_x005', 0.05 * 0.5e-6 / 3.141592653589793, 1.5, 0.01, 1.0, 0.5e-6),
    ('strong_absorber_x10', 10.0 * 0.5e-6 / 3.141592653589793, 1.6, 0.5, 1.0, 0.5e-6),
    ('loess_coarse_530nm', 60e-6, 1.55, 1e-3, 1.33, 530e-9),
]
ANGLES = (0.0, 10.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0)

if __name__ == '__main__':
    out = []
    for name, d, nr, k, nm, wl in CASES:
        r = mie(d, nr, k, nm, wl, ANGLES)
        r.update(name=name, diameter_m=d, n_real=nr, k_imag=k, n_medium=nm, wavelength_m=wl)
        out.append(r)
        print(name, 'x=%.4f qext=%.10f qsca=%.10f g=%.8f' % (r['x'], r['qext'], r['qsca'], r['g']))
    Path(__file__).with_name('mie_mpmath.json').write_text(json.dumps(out, indent=1))
