"""Production failure subcodes (CLARITAS.py:464-521, 8483-8505, 9608-9655):
the tallies count failed histories by subcode slot; the slots add up to the
failure codes and to status 4; a forced trace-iteration limit lands in
subcode 3002 (TRACE_ITERATION_LIMIT) and a forced sphere bounce limit in
1005 (SPHERE_BOUNCE_LIMIT); device and oracle agree on the same histories.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tardiis_and_claritas_amd import abi, workloads
from tardiis_and_claritas_amd.model import ClaritasForwardModel

SLOT = {v: k for k, v in enumerate(abi.SUBCODE_VALUES)}


def _forced():
    """config-1 cell, geometric law on 40 um smooth spheres at 5 g/L, with a
    trace-iteration cap of 6 and no internal reflection allowed"""
    kw, _, _, _, _ = workloads.config1('geometric')
    kw.update(particle_surface_rms_slope_deg=0.0, wave_surface_rms_height_m=0.0, max_iterations=6,
              max_internal_bounces=0)
    return kw, workloads.monodisperse(40e-6, 1.46, 0.0), 5.0


def _check(T):
    sub, fails = T['failure_subcodes'], T['failures']
    assert int(sub.sum()) == int(fails.sum()) == int(T['status'][4]) > 0
    return sub


def test_oracle_subcode_slots_close():
    kw, mat, conc = _forced()
    m = ClaritasForwardModel(**kw)
    T, _ = po.OracleModel(m, mat, conc).run(20_000, 3, 0)
    sub = _check(T)
    assert sub[SLOT[3002]] > 0 and sub[SLOT[1005]] > 0
    assert sub[abi.N_SUBCODE - 1] == 0


@pytest.mark.gpu
def test_device_subcodes_match_oracle(gpu_available, oracle_lib):
    kw, mat, conc = _forced()
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 40_000
    Tg, _, _ = m.run_tallies(N, 3, 0)
    To, _ = oracle_lib.OracleModel(m, mat, conc).run(N, 3, 0)
    sg, so = _check(Tg), _check(To)
    assert np.abs(sg - so).sum() <= 1e-3 * N, (np.flatnonzero(sg), sg[sg > 0], so[so > 0])
    r = m.simulate(mat, conc, n_rays=N, seed=3, path_diagnostics=False)
    d = r.production_failure_diagnostics
    assert d['count_closure_ok'] and r.production_failure_subcode_counts['3002'] > 0
    assert r.production_watchdog_count + r.production_geometry_failure_count + r.production_nonfinite_count <= \
        r.production_failure_count


def test_population_balance_is_the_unaggregated_psd():
    """Only aggregation_model 'none' with zero exposures is built
    (CLARITAS.py:5632-5636 defaults); anything else raises."""
    m = ClaritasForwardModel()
    assert m.aggregation_model == 'none' and m.fragmentation_reference_diameter_m == 100e-6
    with pytest.raises(NotImplementedError):
        ClaritasForwardModel(aggregation_model='orthokinetic_compact')
    with pytest.raises(NotImplementedError):
        ClaritasForwardModel(aggregation_collision_exposure=0.5)


@pytest.mark.gpu
def test_failure_samples(gpu_available, oracle_lib):
    """Failed-history samples of the causal run: with room for all of them,
    their codes and subcodes histogram to the tallies' failure and subcode
    counts, and each sampled history failed with the same code in the
    oracle's trace of the same history (up to < 0.1 % of histories)."""
    import ctypes as C
    kw, mat, conc = _forced()
    m = ClaritasForwardModel(**kw)
    m.prepare(mat, conc)
    N = 20_000
    recs = np.zeros(N, dtype=abi.causal_dtype())
    fails = np.zeros(N, dtype=abi.failure_dtype())
    co = abi.CausalOut()
    co.records, co.capacity, co.failures, co.failure_capacity = recs.ctypes.data, N, fails.ctypes.data, N
    t = abi.Tallies()
    m.ctx.check(m.ctx.lib.clr_run_causal(m.ctx.h, N, 3, 0, m._launch(), None, C.byref(t), None, C.byref(co)),
                'clr_run_causal')
    T = abi.tallies_to_numpy(t)
    f = fails[:co.failure_count]
    assert co.failure_count == T['status'][4] > 0
    np.testing.assert_array_equal(np.bincount(f['code'], minlength=abi.N_FAILURE), T['failures'])
    slots = np.array([SLOT.get(int(s), abi.N_SUBCODE - 1) for s in f['subcode']])
    np.testing.assert_array_equal(np.bincount(slots, minlength=abi.N_SUBCODE), T['failure_subcodes'])
    To, ro = oracle_lib.OracleModel(m, mat, conc).run(N, 3, 0, n_records=N)
    o_code = ro['failure'][f['history'].astype(np.int64)]
    o_st = ro['status'][f['history'].astype(np.int64)]
    assert np.mean((o_st == 4) & (o_code == f['code'])) >= 0.999
    r = m.simulate(mat, conc, n_rays=N, seed=3)
    s = r.production_failure_diagnostics['samples']
    assert 0 < len(s) <= 512 and all(q['failure_code'] > 0 for q in s)
    # more failures than one launch may write: the samples kept are those of
    # the lowest failed history ids (CLARITAS.py:8504-8514), also across chunks
    from tardiis_and_claritas_amd import causal
    assert co.failure_count > 40
    want = np.sort(f['history'])[:24]
    nb = len(m.resolve_material(mat).diameter_m)
    for chunk in (1 << 20, 3000):
        acc = causal.CausalAccumulator(m, nb, chunk=chunk)
        acc.sample_limit, acc.failure_device_capacity = 24, 32
        acc.run(N, 3, 0, m._launch())
        got = np.concatenate(acc.failures)['history']
        np.testing.assert_array_equal(got, want)
