"""Write tests/golden/oracle_config3.json: the oracle's tallies for ten
items of the 256-configuration batched workload (workloads.config3_params,
PCG64(20261203)), for the GPU test that runs the same items in one
persistent batched launch.

Items: the first six, the densest one (largest estimated optical depth
across the cell: mu_ext from the inputs with Q_ext = 2, times the field
maximum, times 2 R_in), the densest 530 nm one, the item with
the largest size parameter, and the sparsest one.  Each runs N histories
with Philox key 7000 + item: its budget, capped so that the oracle spends
about 8 s of 8-thread CPU time on it (N is measured from a pilot run and
recorded).  The tallies of the first CHECK histories are kept as well, for
test_golden_tallies.py to recompute.

    python tests/golden/make_config3_fixture.py
"""
from __future__ import annotations

import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from tardiis_and_claritas_amd import ClaritasForwardModel, workloads  # noqa: E402

OUT = HERE / 'oracle_config3.json'
TARGET_S = 8.0
KEY0 = 7000
CHECK = 5000


def pick_items(p):
    n = len(p['conc_mg_l'])
    tau = np.zeros(n)
    xmax = np.zeros(n)
    for i in range(n):
        kw, mat, conc = workloads.config3_item(p, i)
        m = ClaritasForwardModel(**kw)
        # optical depth of the field maximum across the cell from the inputs:
        # mu_ext = sum 1.5 C w Q_ext / (rho d); Q_ext <= ~2.5 for these sizes,
        # so the large-particle limit Q = 2 ranks the items
        d, w = np.asarray(mat.diameter_m), np.asarray(mat.weights) / np.sum(mat.weights)
        tau[i] = np.sum(1.5 * conc * w * 2.0 / (m.density * d)) * float(np.max(kw['sediment_field'])) * 2 * m.rin
        xmax[i] = math.pi * m.n_water * d.max() / m.particle_wavelength
    sel = list(range(6))
    dense = int(np.argmax(tau))
    green = np.nonzero(p['wavelength_nm'] == 530.0)[0]
    dense_green = int(green[np.argmax(tau[green])])
    big = int(np.argmax(xmax))
    sparse = int(np.argmin(tau))
    for k in (dense, dense_green, big, sparse):
        if k not in sel:
            sel.append(k)
    return sel, tau


def main():
    from oracle import pyoracle
    p = workloads.config3_params()
    sel, tau = pick_items(p)
    out = {'items': {}}
    for i in sel:
        kw, mat, conc = workloads.config3_item(p, i)
        om = pyoracle.OracleModel(ClaritasForwardModel(**kw), mat, conc)
        key = KEY0 + i
        t0 = time.time()
        pilot = 2000
        om.run(pilot, key + 10 ** 6, 0)
        rate = pilot / max(time.time() - t0, 1e-3)
        N = int(min(p['budget'][i], max(20_000, rate * TARGET_S)))
        t1 = time.time()
        T, _ = om.run(N, key, 0)
        C, _ = om.run(CHECK, key, 0)
        d = T['det'].sum(0)
        out['items'][str(i)] = {
            'n': N, 'key': key, 'tau_estimate': float(tau[i]), 'wavelength_nm': float(p['wavelength_nm'][i]),
            'conc_mg_l': float(p['conc_mg_l'][i]), 'median_um': float(p['median_um'][i]),
            'det': d.tolist(), 'aux_window': T['aux_window'].sum(0).tolist(), 'status': T['status'].tolist(),
            'failures': T['failures'].tolist(), 'check_n': CHECK, 'check_det': C['det'].sum(0).tolist(),
            'check_status': C['status'].tolist(),
        }
        print(f'item {i}: N {N:,} in {time.time() - t1:.1f} s (tau ~{tau[i]:.1f}, {p["wavelength_nm"][i]:.0f} nm)',
              flush=True)
    OUT.write_text(json.dumps(out, indent=1, sort_keys=True) + '\n')


if __name__ == '__main__':
    main()
