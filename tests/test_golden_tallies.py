"""The committed oracle tally fixtures (golden/oracle_tallies.json,
golden/oracle_config3.json) belong to the oracle in this tree: the tallies
of their first histories are recomputed here and must be equal, and the
fixtures hold the photon counts the GPU tally tests rely on."""
import json
from pathlib import Path

import numpy as np
import pytest

import tally_cases
from tardiis_and_claritas_amd import ClaritasForwardModel, workloads

GOLDEN = Path(__file__).parent / 'golden'
TALLIES = json.loads((GOLDEN / 'oracle_tallies.json').read_text())


@pytest.mark.parametrize('case', tally_cases.CASES)
def test_tally_fixture_reproduces(oracle_lib, case):
    fx = TALLIES[case]
    kw, mat, conc, seed, N = tally_cases.case_spec(case)
    assert fx['n'] == N and fx['key'] == seed + 2
    om = oracle_lib.OracleModel(ClaritasForwardModel(**kw), mat, conc)
    T, _ = om.run(fx['check_n'], fx['key'], 0)
    assert np.array_equal(T['det'].sum(0), np.asarray(fx['check']['det']))
    assert np.array_equal(T['status'], np.asarray(fx['check']['status']))
    det = np.asarray(fx['det'])
    assert tally_cases.power_ok(det), det[:2, 9]
    assert sum(fx['status']) == N


def test_config3_fixture_reproduces(oracle_lib):
    fx = json.loads((GOLDEN / 'oracle_config3.json').read_text())['items']
    p = workloads.config3_params()
    assert len(fx) >= 8 and any(v['wavelength_nm'] == 530.0 for v in fx.values())
    for k in sorted(fx, key=int)[:4]:
        f = fx[k]
        kw, mat, conc = workloads.config3_item(p, int(k))
        om = oracle_lib.OracleModel(ClaritasForwardModel(**kw), mat, conc)
        T, _ = om.run(f['check_n'], f['key'], 0)
        assert np.array_equal(T['det'].sum(0), np.asarray(f['check_det'])), k
        assert np.array_equal(T['status'], np.asarray(f['check_status'])), k
        assert sum(f['status']) == f['n']
