"""Pins of the oracle's light-transport voxel diagnostics (no GPU).

The reference's deposition functions (CLARITAS.py:3312-3467) are restated in
oracle/oracle.c (light_segment, light_site, light_mark) and hooked into its
transport loop at the reference's call sites.  Checks:

- the reference's own test vectors (validate_claritas.py:1290-1402): nine
  segments through a 12 x 10 x 14 box including wall-to-wall, end-on and
  overshooting ones deposit exactly their clipped lengths with no errors;
  five event sites give 4 counts, 1 outside error and >= 1 edge snap;
- the per-voxel lengths of random segments equal an independent
  computation (all plane crossings sorted, each sub-interval assigned to
  the voxel holding its midpoint, with the reference's snap of a crossing
  within 2e-6 of the segment end onto the end);
- transport invariants: deposition and grid closure, post-scatter <= total,
  event and absorption closure, representative segments chained across
  scatter events;
- the collision-density identity of a uniform medium: the expected number
  of real events equals mu_t times the expected water track length
  (Poisson 4-sigma);
- the on-disk format (npz, CSV cross sections, trajectories, summary)
  round-trips the grids.
"""
import csv
import json
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tardiis_and_claritas_amd import abi, light, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel

R, ZMIN, ZMAX = 0.0465, -0.093, 0.074


def _lc(nx, ny, nz, rep=0, seg=1, max_rays=10 ** 12, zmax=ZMAX):
    lc = abi.LightConfig()
    lc.max_rays = max_rays
    lc.nx, lc.ny, lc.nz = nx, ny, nz
    lc.representative_rays, lc.max_segments = rep, seg
    lc.zmax = zmax
    return lc


def reference_segments():
    inv = 1 / math.sqrt(3)
    return np.asarray([
        [0, 0, -0.08, 0, 0, 1, 0.12], [-0.04, 0, 0, 1, 0, 0, 0.08], [-0.03, -0.03, -0.08, inv, inv, inv, 0.08],
        [0, 0, -0.02, 1, 0, 0, 0.03], [-0.06, 0, 0, 1, 0, 0, 0.03], [0, 0, 0, 1, 0, 0, R], [R, 0, 0, -1, 0, 0, 2 * R],
        [0, 0, ZMIN, 0, 0, 1, ZMAX - ZMIN], [0, 0, 0, 1, 0, 0, R + 1e-07],
    ], dtype=np.float32)


def reference_sites():
    return np.asarray([[0, 0, 0], [0.02, -0.01, -0.04], [-0.03, 0.02, 0.05], [R + 5e-08, 0, 0], [0.2, 0, 0]],
                      dtype=np.float32)


def test_reference_segment_vectors():
    seg = reference_segments()
    a, sc = po.light_probe(0, seg, _lc(12, 10, 14), R, ZMIN)
    t, p = a['track_length'], a['post_scatter_track_length']
    assert sc['error_count'] == 0 and sc['segment_count'] == len(seg)
    assert np.all(np.isfinite(t)) and np.all(t >= 0) and np.all(p >= 0) and np.all(p <= t + 1e-8)
    expected = float(seg[:4, 6].sum(dtype=np.float64)) + 0.0165 + R + 2 * R + (ZMAX - ZMIN) + R
    np.testing.assert_allclose(sc['segment_total'], expected, rtol=2e-7, atol=2e-8)
    np.testing.assert_allclose(sc['deposited_total'], sc['segment_total'], rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(t.sum(dtype=np.float64), sc['segment_total'], rtol=2e-6, atol=2e-7)
    # odd-indexed segments are post-scatter (particle order i & 1)
    odd = sum(min(float(s[6]), 1e9) for s in seg[1::2])
    assert p.sum(dtype=np.float64) <= odd + 1e-6


def test_reference_event_vectors():
    a, sc = po.light_probe(1, reference_sites(), _lc(12, 10, 14), R, ZMIN)
    assert sc['scatter_event_total'] == 5 and int(a['scatter_event_count'].sum()) == 4
    assert sc['error_count'] == 1 and sc['outside'][0] == 1 and sc['outside'][6] == 1
    assert sc['boundary_adjustment_count'] >= 1
    np.testing.assert_array_equal(a['scatter_event_weight'], a['scatter_event_count'].astype(np.float32))
    first = np.array(sc['outside'][1:4], dtype=np.uint32).view(np.float32)
    np.testing.assert_array_equal(first, reference_sites()[4])
    assert sc['outside'][4] == 4


def brute_force_lengths(seg, nx, ny, nz, zmax=ZMAX):
    """Per-voxel lengths from all plane crossings of the clipped segment.
    The box bounds are FP32 values, as the reference passes them."""
    out = np.zeros((nz, ny, nx))
    r32, z0, z1 = float(np.float32(R)), float(np.float32(ZMIN)), float(np.float32(zmax))
    lo = np.array([-r32, -r32, z0])
    hi = np.array([r32, r32, z1])
    n = np.array([nx, ny, nz])
    h = (hi - lo) / n
    for s in seg.astype(np.float64):
        p0, v, d = s[:3], s[3:6], s[6]
        t0, t1 = 0.0, 1.0
        for k in range(3):
            dk = v[k] * d
            if abs(dk) < 1e-30:
                if not (lo[k] <= p0[k] <= hi[k]):
                    t0, t1 = 1.0, 0.0
                continue
            a, b = sorted(((lo[k] - p0[k]) / dk, (hi[k] - p0[k]) / dk))
            t0, t1 = max(t0, a), min(t1, b)
        if not t1 > t0:
            continue
        ts = [t0, t1]
        # the reference snaps a face crossing within 2e-6 (as a fraction of
        # the clipped segment) of its end onto the end (CLARITAS.py:3383)
        snap = t0 + (t1 - t0) * float(np.float32(1.0) - np.float32(2e-6))
        for k in range(3):
            dk = v[k] * d
            if abs(dk) < 1e-30:
                continue
            planes = lo[k] + h[k] * np.arange(n[k] + 1)
            tk = (planes - p0[k]) / dk
            ts += list(tk[(tk > t0) & (tk < snap)])
        ts = np.sort(np.array(ts))
        for ta, tb in zip(ts[:-1], ts[1:]):
            if tb <= ta:
                continue
            mid = p0 + v * d * (0.5 * (ta + tb))
            idx = np.clip(np.floor((mid - lo) / h).astype(int), 0, n - 1)
            out[idx[2], idx[1], idx[0]] += (tb - ta) * d
    return out


def test_dda_matches_brute_force_per_voxel():
    g = np.random.Generator(np.random.PCG64(11))
    n = 300
    p0 = np.stack([g.uniform(-1.2 * R, 1.2 * R, n), g.uniform(-1.2 * R, 1.2 * R, n),
                   g.uniform(ZMIN - 0.02, ZMAX + 0.02, n)], 1)
    v = g.normal(size=(n, 3))
    v[: n // 6, 0] = 0.0          # axis-parallel cases
    v[n // 6: n // 3, 1:] = 0.0
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    d = g.uniform(1e-4, 0.2, n)
    seg = np.concatenate([p0, v, d[:, None]], 1).astype(np.float32)
    nx, ny, nz = 9, 13, 17
    a, sc = po.light_probe(0, seg, _lc(nx, ny, nz), R, ZMIN)
    ref = brute_force_lengths(seg, nx, ny, nz)
    assert sc['error_count'] == 0
    np.testing.assert_allclose(a['track_length'], ref, rtol=1e-6, atol=2e-10)
    np.testing.assert_allclose(sc['segment_total'], ref.sum(), rtol=1e-9)


def _oracle_light_run(builder, n, grid=(24, 24, 32), rep=16, seg=64, seed=7):
    kw, mat, conc, _, _ = builder
    m = ClaritasForwardModel(**kw)
    om = po.OracleModel(m, mat, conc)
    lc = _lc(*grid, rep=rep, seg=seg, max_rays=n, zmax=light.box_zmax(m))
    T, arrs, sc = om.run_light(n, seed, lc)
    return m, om, T, arrs, sc, lc


def test_transport_invariants_and_collision_identity():
    m, om, T, a, sc, lc = _oracle_light_run(workloads.config1('plain_mie'), 40_000)
    assert sc['error_count'] == 0
    np.testing.assert_allclose(sc['deposited_total'], sc['segment_total'], rtol=1e-12)
    # FP32 voxels accumulate thousands of deposits each on the beam axis: the
    # grid sum carries their rounding (the reference reports this closure,
    # grid_closure_relative_error, without a bar)
    np.testing.assert_allclose(a['track_length'].sum(dtype=np.float64), sc['segment_total'], rtol=5e-5)
    assert np.all(a['post_scatter_track_length'] <= a['track_length'] * (1 + 1e-4) + 1e-9)
    assert int(a['scatter_event_count'].sum()) == sc['scatter_event_total']
    assert int(a['absorption_event_count'].sum()) == sc['absorption_event_total']
    # every real event of the plain law is a surviving scatter or an absorption
    assert sc['scatter_event_total'] + sc['absorption_event_total'] == int(T['counters'][abi.C_IDX['sum_ic']])
    mu = float(om.tables()['mu_scalars'][0] + om.tables()['mu_scalars'][1])
    events = sc['scatter_event_total'] + sc['absorption_event_total']
    expected = mu * sc['segment_total']
    assert abs(events - expected) < 4.0 * math.sqrt(expected), (events, expected)
    assert events > 1000
    # representative segments: ordered, chained across scatters (the plain
    # law does not move the photon at an event), event codes in range
    cnt = a['traj_count']
    assert np.all(cnt <= lc.max_segments) and cnt.sum() > 0
    for r in range(lc.representative_rays):
        k = cnt[r]
        if k == 0:
            continue
        assert np.all(np.isfinite(a['traj_x0'][r, :k])) and np.all(np.isnan(a['traj_x0'][r, k:]))
        ends = np.stack([a['traj_x1'][r, :k - 1], a['traj_y1'][r, :k - 1], a['traj_z1'][r, :k - 1]], 1)
        starts = np.stack([a['traj_x0'][r, 1:k], a['traj_y0'][r, 1:k], a['traj_z0'][r, 1:k]], 1)
        scatter = a['traj_event'][r, :k - 1] == 4   # boundary ends (code 0) are followed by the wall step
        assert np.all(np.abs(ends - starts)[scatter] < 2e-6)
        assert np.all(np.diff(a['traj_order'][r, :k]) >= 0)
        assert set(a['traj_event'][r, :k].tolist()) <= {0, 4, 9}
        assert np.all(a['traj_event'][r, k:] == -1)


def test_hybrid_law_event_codes():
    m, om, T, a, sc, lc = _oracle_light_run(workloads.config1('hybrid_mie_exact_rate'), 20_000)
    assert sc['error_count'] == 0
    codes = a['traj_event'][a['traj_event'] >= 0]
    assert set(codes.tolist()) <= set(range(11))
    assert (codes == 2).any() and (codes == 4).any()   # geometric transmissions and coherent Mie events
    assert int(a['absorption_event_count'].sum()) == sc['absorption_event_total']
    assert int(a['scatter_event_count'].sum()) == sc['scatter_event_total']


def test_writer_round_trip(tmp_path):
    m, om, T, a, sc, lc = _oracle_light_run(workloads.config1('plain_mie'), 5_000, grid=(16, 16, 20))
    cfg = light.resolve({'enabled': True, 'nx': 16, 'ny': 16, 'nz': 20, 'representative_rays': 16,
                         'max_segments_per_ray': 64, 'vertical_slice_thickness_m': 0.002}, 5_000)
    diag = light.diagnostics_dict(m, cfg, light.box_zmax(m), a, sc, 5_000, 0, 0.0)
    light.check_closure(diag)
    summary = light.write_outputs(tmp_path, diag, m.sensor_height, cfg)
    z = np.load(tmp_path / 'light_transport_volume.npz')
    np.testing.assert_array_equal(z['total_track_length_weighted_m'], a['track_length'])
    assert z['x_edges_m'].shape == (17,) and z['z_centres_above_bottom_m'].shape == (20,)
    np.testing.assert_allclose(z['z_edges_above_bottom_m'][0], 0.0, atol=1e-15)
    with open(tmp_path / 'light_transport_vertical_cross_section.csv') as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 20 * 16
    q = light.cross_sections(diag, cfg)
    vr = np.array([float(r['track_length_m']) for r in rows]).reshape(20, 16)
    np.testing.assert_allclose(vr, q['vr'], rtol=1e-12)
    np.testing.assert_allclose(q['vr'], a['track_length'][:, :, q['xi']].sum(2), rtol=1e-6)
    with open(tmp_path / 'light_transport_representative_trajectories.csv') as f:
        tr = list(csv.DictReader(f))
    assert len(tr) == int(a['traj_count'].sum())
    s = json.loads((tmp_path / 'light_transport_summary.json').read_text())
    assert s['segment_count'] == sc['segment_count'] and summary['grid_shape_zyx'] == [20, 16, 16]


def test_config_validation():
    with pytest.raises(ValueError):
        light.resolve({'enabled': True, 'nx': 3}, 10)
    with pytest.raises(ValueError):
        light.resolve({'enabled': True, 'max_rays': 0}, 10)
    with pytest.raises(ValueError):
        light.resolve({'enabled': True, 'max_segments_per_ray': 0}, 10)
    c = light.resolve({'enabled': True, 'representative_rays': 500}, 100)
    assert c['max_rays'] == 100 and c['representative_rays'] == 100
    assert light.memory_bytes(light.resolve({'enabled': True}, 10)) < 1e9


# ---- water-path heat maps, CLARITAS.py:3296-3309 ----
def _heat_lc(vis, zmax=0.074):
    from tardiis_and_claritas_amd import abi
    lc = abi.LightConfig()
    lc.max_rays, lc.nx, lc.ny, lc.nz, lc.representative_rays, lc.max_segments = 0, 4, 4, 4, 0, 1
    lc.heatmap_size, lc.zmax = vis, zmax
    return lc


def test_heat_raster_samples_and_bins():
    """A segment gets ceil(dist / 10 um) samples (at most 20000) at
    fractions k/n; each lands in bin floor((x + R) / 2R * vis) of the x-y map
    (rows from +y down) and floor((z - zmin) / (zmax - zmin) * vis) of the
    z-y map.  Brute force on tiny inputs, FP32 arithmetic, +-1 bin for
    rounding at a face."""
    Rr, zmin, zmax, vis = 0.0465, -0.093, 0.074, 16
    g = np.random.Generator(np.random.PCG64(8))
    n = 40
    p0 = np.stack([g.uniform(-0.9 * Rr, 0.9 * Rr, n), g.uniform(-0.9 * Rr, 0.9 * Rr, n),
                   g.uniform(zmin, zmax, n)], 1)
    v = g.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    seg = np.concatenate([p0, v, g.uniform(1e-4, 3e-3, n)[:, None]], 1).astype(np.float32)
    seg[0, 6] = 0.5  # longer than 20000 samples: capped
    arrs, _ = po.light_probe(2, seg, _heat_lc(vis, zmax), Rr, zmin)
    xy = np.zeros((vis, vis))
    zy = np.zeros((vis, vis))
    for q in seg.astype(np.float32):
        dist = np.float32(q[6])
        m = min(max(int(np.ceil(dist / np.float32(1e-5))), 1), 20000)
        f = (np.arange(1, m + 1, dtype=np.float32) / np.float32(m)).astype(np.float32)
        x, y, z = (q[i] + q[3 + i] * dist * f for i in range(3))
        ix = ((x + np.float32(Rr)) / np.float32(2 * Rr) * vis).astype(np.float32)
        iy = ((y + np.float32(Rr)) / np.float32(2 * Rr) * vis).astype(np.float32)
        iz = ((z - np.float32(zmin)) / np.float32(zmax - zmin) * vis).astype(np.float32)
        ix, iy, iz = np.trunc(ix).astype(int), vis - 1 - np.trunc(iy).astype(int), np.trunc(iz).astype(int)
        ok = (ix >= 0) & (ix < vis) & (iy >= 0) & (iy < vis)
        np.add.at(xy, (iy[ok], ix[ok]), 1.0)
        ok = (iz >= 0) & (iz < vis) & (iy >= 0) & (iy < vis)
        np.add.at(zy, (iy[ok], iz[ok]), 1.0)
    for got, want in ((arrs['heatmap_xy'], xy), (arrs['heatmap_zy'], zy)):
        assert abs(got.sum() - want.sum()) <= 4
        assert np.abs(got - want).sum() <= 2e-3 * want.sum() + 8


def test_heat_map_total_matches_track_length():
    """Over a transport run, samples x 10 um equal the water track length in
    the box up to the per-segment rounding (ceil) and the 20000-sample cap."""
    from tardiis_and_claritas_amd import workloads
    from tardiis_and_claritas_amd.light import box_zmax
    from tardiis_and_claritas_amd.model import ClaritasForwardModel
    kw, mat, conc, _, _ = workloads.config1('plain_mie')
    m = ClaritasForwardModel(**kw)
    lc = _heat_lc(32, box_zmax(m))
    lc.max_rays, lc.nx, lc.ny, lc.nz = 10 ** 9, 8, 8, 8
    T, a, sc = po.OracleModel(m, mat, conc).run_light(3000, 5, lc)
    samples = float(a['heatmap_xy'].sum())
    assert samples > 0 and float(a['heatmap_zy'].sum()) > 0.9 * samples
    track = sc['segment_total']
    assert track <= samples * 1e-5 + 1e-9
    assert samples * 1e-5 - track <= sc['segment_count'] * 1e-5 + 1e-6
