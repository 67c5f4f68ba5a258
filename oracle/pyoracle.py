"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  It builds the same ABI structs
as the product through ClaritasForwardModel's struct builders, so the oracle
and the HIP library see identical inputs.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from tardiis_and_claritas_amd import abi

HERE = Path(__file__).resolve().parent
LIB = HERE / 'liboracle.so'
_lib = None


def build() -> Path:
    subprocess.run(['make', '-s', '-C', str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    P = C.POINTER
    lib.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.orc_uniform_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_int, P(C.c_float)]
    lib.orc_mie.argtypes = [P(abi.Particles), P(abi.MieOut)]
    lib.orc_mie.restype = C.c_int
    lib.orc_prepare.argtypes = [P(abi.Config), P(abi.Particles), P(abi.Sediment), P(C.c_float), C.c_size_t,
                                P(C.c_void_p)]
    lib.orc_prepare.restype = C.c_int
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_get_tables.argtypes = [C.c_void_p, P(abi.TablesOut)]
    lib.orc_set_phase_table.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_int]
    lib.orc_run.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, P(abi.Tallies),
                            P(abi.Record), C.c_uint64]
    lib.orc_run.restype = C.c_int
    lib.orc_free_surface_probe.argtypes = [P(abi.Config), P(C.c_float), C.c_int] + [P(C.c_float)] * 6
    lib.orc_sediment_probe.argtypes = [C.c_void_p, C.c_int, P(C.c_float), C.c_int, P(C.c_double)]
    lib.orc_fresnel_probe.argtypes = [C.c_int, P(C.c_float), C.c_float, C.c_float, P(C.c_float), P(C.c_int)]
    lib.orc_complex_probe.argtypes = [C.c_int, P(C.c_float), C.c_float, C.c_float, C.c_float, C.c_float,
                                      P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_int)]
    lib.orc_refract_probe.argtypes = [C.c_int, P(C.c_float), P(C.c_float), C.c_float, C.c_float, P(C.c_float),
                                      P(C.c_int)]
    lib.orc_wall_probe.argtypes = [C.c_int, P(C.c_float), C.c_float, C.c_float, C.c_float, P(C.c_double)]
    lib.orc_erfinv.argtypes = [C.c_double]
    lib.orc_erfinv.restype = C.c_double
    lib.orc_beckmann_B_probe.argtypes = [C.c_int, P(C.c_double), P(C.c_double)]
    lib.orc_smith_lambda.argtypes = [C.c_float] * 7
    lib.orc_smith_lambda.restype = C.c_double
    lib.orc_smith_height.argtypes = [C.c_double] + [C.c_float] * 8 + [P(C.c_float), P(C.c_int)]
    lib.orc_smith_height.restype = C.c_int
    lib.orc_beckmann_slopes_probe.argtypes = [C.c_int] + [P(C.c_double)] * 5 + [P(C.c_int)]
    lib.orc_walk_probe.argtypes = [C.c_int, C.c_uint64, C.c_float, P(C.c_float), C.c_float, C.c_float, C.c_float,
                                   C.c_float, P(C.c_float), P(C.c_int), P(C.c_int), P(C.c_int)]
    lib.orc_sediment_bins.argtypes = [P(abi.Config), P(abi.Sediment), C.c_int, P(C.c_double)] + [P(C.c_double)] * 6
    lib.orc_sediment_bins.restype = C.c_int
    lib.orc_vortex_geometry.argtypes = [C.c_int] + [C.c_double] * 5 + [P(C.c_double)]
    lib.orc_vortex_geometry.restype = C.c_int
    lib.orc_meridional_velocity.argtypes = [C.c_int, P(C.c_double), P(C.c_double)] + [C.c_double] * 3 + \
        [C.c_int] + [C.c_double] * 3 + [P(C.c_double), P(C.c_double)]
    lib.orc_meridional_velocity.restype = None
    lib.orc_meridional_solve.argtypes = [P(abi.Config), P(abi.Particles), P(abi.Sediment), P(abi.Meridional),
                                         P(abi.MeridionalOut)]
    lib.orc_meridional_solve.restype = C.c_int
    lib.orc_run_light.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, P(abi.LightConfig),
                                  P(abi.Tallies), P(abi.LightOut)]
    lib.orc_run_light.restype = C.c_int
    lib.orc_run_causal.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, P(abi.Tallies),
                                   C.c_void_p, P(C.c_uint8)]
    lib.orc_run_causal.restype = C.c_int
    for name in ('orc_light_segment_probe', 'orc_light_event_probe', 'orc_light_heat_probe'):
        fn = getattr(lib, name)
        fn.argtypes = [P(abi.LightConfig), C.c_float, C.c_float, C.c_int, P(C.c_float), P(abi.LightOut)]
        fn.restype = C.c_int
    lib.orc_bore_probe.argtypes = [P(abi.Config), C.c_int, C.c_uint64, P(C.c_float), C.c_float, C.c_float,
                                   P(C.c_int), P(C.c_int), P(C.c_int)]
    lib.orc_bore_probe.restype = C.c_int
    lib.orc_set_slab_majorant.argtypes = [C.c_void_p, C.c_int]
    lib.orc_set_slab_majorant.restype = C.c_int
    lib.orc_probe_shape.argtypes = [C.c_int, P(C.c_int), P(C.c_int)]
    lib.orc_probe_shape.restype = C.c_int
    lib.orc_probe.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint64, P(C.c_float), P(C.c_double)]
    lib.orc_probe.restype = C.c_int
    lib.orc_sfh213.argtypes = [C.c_float]
    lib.orc_sfh213.restype = C.c_float
    _lib = lib
    return lib


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def fresnel(ci, n1, n2):
    ci = np.ascontiguousarray(ci, dtype=np.float32)
    R = np.zeros_like(ci)
    tir = np.zeros(ci.size, dtype=np.int32)
    load().orc_fresnel_probe(ci.size, _pf(ci), n1, n2, _pf(R), tir.ctypes.data_as(C.POINTER(C.c_int)))
    return R, tir


def complex_interface(ci, n1r, n1k, n2r, n2k):
    ci = np.ascontiguousarray(ci, dtype=np.float32)
    R, tt, tn = (np.zeros_like(ci) for _ in range(3))
    ok = np.zeros(ci.size, dtype=np.int32)
    load().orc_complex_probe(ci.size, _pf(ci), n1r, n1k, n2r, n2k, _pf(R), _pf(tt), _pf(tn),
                             ok.ctypes.data_as(C.POINTER(C.c_int)))
    return R, tt, tn, ok


def refract(v, nrm, n1, n2):
    v = np.ascontiguousarray(v, dtype=np.float32)
    nrm = np.ascontiguousarray(nrm, dtype=np.float32)
    out = np.zeros_like(v)
    ok = np.zeros(len(v), dtype=np.int32)
    load().orc_refract_probe(len(v), _pf(v), _pf(nrm), n1, n2, _pf(out), ok.ctypes.data_as(C.POINTER(C.c_int)))
    return out, ok


def wall_exit(rays6, R, zlo, zhi):
    rays6 = np.ascontiguousarray(rays6, dtype=np.float32)
    t = np.zeros(len(rays6))
    load().orc_wall_probe(len(rays6), _pf(rays6), R, zlo, zhi, t.ctypes.data_as(C.POINTER(C.c_double)))
    return t


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def erfinv(y):
    lib = load()
    return np.array([lib.orc_erfinv(float(v)) for v in np.ravel(y)])


def beckmann_B(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    load().orc_beckmann_B_probe(x.size, _pd(x), _pd(out))
    return out


def smith_lambda(w, b, alpha):
    return load().orc_smith_lambda(*(float(v) for v in w), *(float(v) for v in b), float(alpha))


def smith_height(U, w, b, alpha, h):
    """(status, next height, failure code) of one Smith height step."""
    nh = C.c_float(0.0)
    fc = C.c_int(0)
    st = load().orc_smith_height(float(U), *(float(v) for v in w), *(float(v) for v in b), float(alpha), float(h),
                                 C.byref(nh), C.byref(fc))
    return st, nh.value, fc.value


def beckmann_slopes(si, u1, u2):
    si, u1, u2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (si, u1, u2))
    sx, sy = np.zeros_like(si), np.zeros_like(si)
    ok = np.zeros(si.size, dtype=np.int32)
    load().orc_beckmann_slopes_probe(si.size, _pd(si), _pd(u1), _pd(u2), _pd(sx), _pd(sy), _pi(ok))
    return sx, sy, ok


def micro_walk(seed, alpha, b, d, n_top, k_top, n_bot, k_bot):
    """Microsurface walks of directions d[n][3] on interfaces with top
    normals b[n][3]: (outgoing directions, status, top, failure code)."""
    b = np.ascontiguousarray(b, dtype=np.float32)
    d = np.array(d, dtype=np.float32, copy=True, order='C')
    n = len(d)
    st, top, fc = (np.zeros(n, dtype=np.int32) for _ in range(3))
    load().orc_walk_probe(n, seed, alpha, _pf(b), n_top, k_top, n_bot, k_bot, _pf(d), _pi(st), _pi(top), _pi(fc))
    return d, st, top, fc


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def uniform_stream(seed, history, n):
    out = np.zeros(n, dtype=np.float32)
    load().orc_uniform_stream(seed, history, n, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def mie(model, material, n_theta=None):
    """FP64 Mie tables of the oracle for the model's medium/wavelength."""
    p, keep = model.particles_struct(material, 0.0, None, 'mass_fraction')
    nt = int(n_theta or model.mie_phase_grid_size)
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
    rc = load().orc_mie(C.byref(p), C.byref(m))
    if rc:
        raise RuntimeError(f'oracle Mie failed ({rc})')
    del keep
    return out


class OracleModel:
    """Prepared oracle model for one configuration of a ClaritasForwardModel."""

    def __init__(self, model, material, concentration_g_per_L, weights=None, weight_mode='mass_fraction'):
        self.lib = load()
        self.material = model.resolve_material(material)
        cfg = model.config_struct(self.material)
        p, keep = model.particles_struct(self.material, concentration_g_per_L, weights, weight_mode)
        sed = model.sediment_struct(self.material)
        f = model.sediment_field
        fptr = f.ctypes.data_as(C.POINTER(C.c_float)) if f is not None else None
        h = C.c_void_p()
        rc = self.lib.orc_prepare(C.byref(cfg), C.byref(p), C.byref(sed), fptr, int(f.size) if f is not None else 0,
                                  C.byref(h))
        del keep
        if rc:
            raise RuntimeError(f'oracle prepare failed with code {rc}')
        self.h = h
        self.n_bins = p.n

    def tables(self):
        PD = C.POINTER(C.c_double)
        arrs = {k: np.zeros(self.n_bins) for k in ('radius', 'number_density', 'mu_geom_bin', 'mu_sca_bin',
                                                  'mu_ext_bin', 'trace_geom_bin', 'trace_wave_bin', 'geom_cdf',
                                                  'wave_cdf', 'sed_A', 'sed_B', 'sed_norm', 'sed_settling')}
        t = abi.TablesOut()
        for k, v in arrs.items():
            setattr(t, k, v.ctypes.data_as(PD))
        self.lib.orc_get_tables(self.h, C.byref(t))
        arrs['mu_scalars'] = np.array(list(t.mu_scalars))
        arrs['majorants'] = np.array(list(t.majorants))
        return arrs

    def radial_log(self, n_radial):
        PD = C.POINTER(C.c_double)
        out = np.zeros((self.n_bins, n_radial))
        t = abi.TablesOut()
        t.sed_radial_log = out.ctypes.data_as(PD)
        self.lib.orc_get_tables(self.h, C.byref(t))
        return out

    def use_slab_majorant(self, on=True):
        """Track the shared field with the device's z-slab majorants."""
        if self.lib.orc_set_slab_majorant(self.h, int(bool(on))):
            raise ValueError('slab majorants need a shared r-z or x-y-z field')
        return self

    def probe(self, mode, items, seed=0):
        """orc_probe `mode` (abi.PROBE_MODES name or code) on items [n][IN];
        returns [n][OUT] doubles (clr_probe's layouts)."""
        code = abi.PROBE_MODES[mode] if isinstance(mode, str) else int(mode)
        si, so = C.c_int(), C.c_int()
        if self.lib.orc_probe_shape(code, C.byref(si), C.byref(so)):
            raise ValueError(f'unknown probe mode {mode!r}')
        a = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, si.value)
        out = np.zeros((a.shape[0], so.value), dtype=np.float64)
        rc = self.lib.orc_probe(self.h, code, a.shape[0], int(seed) & 0xFFFFFFFFFFFFFFFF,
                                a.ctypes.data_as(C.POINTER(C.c_float)), out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError(f'oracle probe failed with code {rc}')
        return out

    def set_phase_table(self, theta, cdf):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        cdf = np.ascontiguousarray(cdf, dtype=np.float64)
        PD = C.POINTER(C.c_double)
        self.lib.orc_set_phase_table(self.h, theta.ctypes.data_as(PD), cdf.ctypes.data_as(PD), theta.size)

    def run(self, n, seed, history_offset=0, threads=0, n_records=0):
        t = abi.Tallies()
        rec = (abi.Record * int(n_records))() if n_records else None
        rc = self.lib.orc_run(self.h, int(n), int(seed), int(history_offset), int(threads), C.byref(t), rec,
                              int(n_records))
        if rc:
            raise RuntimeError(f'oracle run failed with code {rc}')
        recs = abi.records_to_numpy(rec, int(n_records)) if n_records else None
        return abi.tallies_to_numpy(t), recs

    def run_light(self, n, seed, light_config, history_offset=0, threads=0):
        """orc_run_light: (tallies dict, arrays dict, scalars dict) with the
        clr_run_light contract."""
        lc = light_config
        out, arrs = abi.light_out(lc.nx, lc.ny, lc.nz, lc.representative_rays, lc.max_segments, lc.heatmap_size)
        t = abi.Tallies()
        rc = self.lib.orc_run_light(self.h, int(n), int(seed), int(history_offset), int(threads), C.byref(lc),
                                    C.byref(t), C.byref(out))
        if rc:
            raise RuntimeError(f'oracle light run failed with code {rc}')
        return abi.tallies_to_numpy(t), arrs, abi.light_scalars(out)

    def run_causal(self, n, seed, history_offset=0, threads=0):
        """orc_run_causal: (tallies dict, records of the detected histories
        in history order)."""
        recs = np.zeros(int(n), dtype=abi.causal_dtype())
        det = np.zeros(int(n), dtype=np.uint8)
        t = abi.Tallies()
        rc = self.lib.orc_run_causal(self.h, int(n), int(seed), int(history_offset), int(threads), C.byref(t),
                                     C.c_void_p(recs.ctypes.data), det.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            raise RuntimeError(f'oracle causal run failed with code {rc}')
        return abi.tallies_to_numpy(t), recs[det.astype(bool)]

    def sediment_scale(self, bin_index, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        out = np.zeros(len(xyz))
        self.lib.orc_sediment_probe(self.h, int(bin_index), xyz.ctypes.data_as(C.POINTER(C.c_float)), len(xyz),
                                    out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def __del__(self):
        if getattr(self, 'h', None):
            self.lib.orc_free(self.h)
            self.h = None


def free_surface_probe(model, rays6):
    rays6 = np.ascontiguousarray(rays6, dtype=np.float32)
    n = len(rays6)
    outs = [np.zeros(n, dtype=np.float32) for _ in range(6)]
    cfg = model.config_struct(None)
    PF = C.POINTER(C.c_float)
    load().orc_free_surface_probe(C.byref(cfg), rays6.ctypes.data_as(PF), n, *[o.ctypes.data_as(PF) for o in outs])
    return dict(zip(['zs_start', 't_hit', 'zs_hit', 'nx', 'ny', 'nz'], outs))


# ---------------------------------------------------------------- meridional
SURFACE_CODES = {'flat': 0, 'parabolic_vortex': 1, 'localized_scully_vortex': 2}


def vortex_geometry(model, zmean, R, D, a, g=9.81):
    """(wall_z, center_z, h_inf, h_bar, vtheta_peak) of the oracle."""
    out = np.zeros(5)
    rc = load().orc_vortex_geometry(SURFACE_CODES[model], zmean, R, D, a, g, _pd(out))
    if rc:
        raise ValueError('invalid vortex')
    return out


def meridional_velocity(r, z, zmean, zmin, R, D, a, chi, g=9.81, model='localized_scully_vortex'):
    """(ur, uz) of the secondary circulation, with the host signature of the
    reference's meridional_velocity_host (CLARITAS.py:795-867)."""
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    z = np.ascontiguousarray(z, dtype=np.float64).ravel()
    vt = vortex_geometry(model, zmean, R, D, a, g)[4]
    ur, uz = np.zeros_like(r), np.zeros_like(r)
    load().orc_meridional_velocity(r.size, _pd(r), _pd(z), zmin, zmean, R, SURFACE_CODES[model], D, a,
                                   chi * vt, _pd(ur), _pd(uz))
    return ur, uz


def sediment_bins(model, material):
    """Per-bin sediment tables (A, B, norm, ws, radlog, radwr) for the
    model's sediment law (the meridional law uses model 3's tables)."""
    mat = model.resolve_material(material)
    cfg = model.config_struct(mat)
    sed = model.sediment_struct(mat)
    d = np.ascontiguousarray(mat.diameter_m, dtype=np.float64)
    n, nr = d.size, sed.n_radial
    out = [np.zeros(n) for _ in range(4)] + [np.zeros((n, nr)), np.zeros((n, nr))]
    rc = load().orc_sediment_bins(C.byref(cfg), C.byref(sed), n, _pd(d), *[_pd(a) for a in out])
    if rc:
        raise RuntimeError(f'oracle sediment bins failed with code {rc}')
    return dict(zip(('A', 'B', 'norm', 'ws', 'radlog', 'radwr'), out))


def meridional_solve(model, material, coeff=True, settings=None):
    """Oracle meridional solve for `model`'s settings and the material's PSD;
    returns the same dict as tardiis_and_claritas_amd.meridional.solve."""
    from tardiis_and_claritas_amd.meridional import result_dict

    mat = model.resolve_material(material)
    cfg = model.config_struct(mat)
    p, keep = model.particles_struct(mat, 0.0, None, 'mass_fraction')
    sed = model.sediment_struct(mat)
    st = settings or model.meridional
    out, arrs = abi.meridional_out(int(p.n), int(st.nr), int(st.nz), coeff=coeff)
    mp = st.struct()
    rc = load().orc_meridional_solve(C.byref(cfg), C.byref(p), C.byref(sed), C.byref(mp), C.byref(out))
    del keep
    if rc and rc != -7:
        raise RuntimeError(f'oracle meridional solve failed with code {rc}')
    return result_dict(out, arrs, rc)


# ----------------------------------------------------------- light probes
def light_probe(mode, data, light_config, rin, zmin):
    """The reference's diagnostic test kernels on the oracle: mode 0 deposits
    segments [n][7], mode 1 event sites [n][3], mode 2 rasters segments [n][7]
    into the heat maps; returns (arrays, scalars)."""
    lc = light_config
    data = np.ascontiguousarray(data, dtype=np.float32)
    out, arrs = abi.light_out(lc.nx, lc.ny, lc.nz, lc.representative_rays, lc.max_segments, lc.heatmap_size)
    for k in ('traj_x0', 'traj_y0', 'traj_z0', 'traj_x1', 'traj_y1', 'traj_z1'):
        arrs[k][:] = np.nan
    for k in ('traj_order', 'traj_event'):
        arrs[k][:] = -1
    fn = {0: load().orc_light_segment_probe, 1: load().orc_light_event_probe, 2: load().orc_light_heat_probe}[mode]
    fn(C.byref(lc), float(rin), float(zmin), int(data.shape[0]), _pf(data), C.byref(out))
    return arrs, abi.light_scalars(out)


def bore_probe(cfg, rays6, s_end, radius, seed=1):
    """One reflecting bore section per ray: (pass, interactions, reflections)."""
    rays6 = np.ascontiguousarray(rays6, dtype=np.float32)
    n = rays6.shape[0]
    out = [np.zeros(n, dtype=np.int32) for _ in range(3)]
    load().orc_bore_probe(C.byref(cfg), n, int(seed), _pf(rays6), float(s_end), float(radius), *[_pi(a) for a in out])
    return out
