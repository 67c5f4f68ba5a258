"""Write tests/golden/oracle_tallies.json: the oracle's detector tallies at
the photon counts the GPU tally tests need (>= 2,000 hardware 90-degree
hits per side), so that those tests compare against the oracle without
running it for minutes on the GPU host.

Each case is a workload of tardiis_and_claritas_amd.workloads traced by the
CPU oracle (oracle/liboracle.so) over histories [0, N) with Philox key
seed + 2.  Stored per case: N, the key, the summed detector tallies
(hardware native/mirror, legacy native/mirror), the exit-position windows
at the auxiliary angles, status and failure counts, and the tallies of
the first CHECK histories, which test_golden_tallies.py recomputes to show
the fixture still belongs to the oracle in the tree.

    python tests/golden/make_oracle_tallies.py [case ...]
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE.parent))

from tally_cases import CASES, case_spec  # noqa: E402
from tardiis_and_claritas_amd import ClaritasForwardModel  # noqa: E402

OUT = HERE / 'oracle_tallies.json'
CHECK = 200_000
CHUNK = 10_000_000


def summary(T):
    d = T['det'].sum(0)
    return {'det': d.tolist(), 'aux_window': T['aux_window'].sum(0).tolist(), 'status': T['status'].tolist(),
            'failures': T['failures'].tolist()}


def run_case(name):
    from oracle import pyoracle
    kw, mat, conc, seed, N = case_spec(name)
    om = pyoracle.OracleModel(ClaritasForwardModel(**kw), mat, conc)
    key = seed + 2
    first, _ = om.run(CHECK, key, 0)
    acc = None
    t0 = time.time()
    for off in range(0, N, CHUNK):
        T, _ = om.run(min(CHUNK, N - off), key, off)
        if acc is None:
            acc = T
        else:
            for k in ('det', 'aux_window'):
                acc[k] = acc[k] + T[k]
            for k in ('status', 'failures'):
                acc[k] = acc[k] + T[k]
        print(f'{name}: {off + min(CHUNK, N - off):,} / {N:,} histories, {time.time() - t0:.0f} s', flush=True)
    out = {'n': N, 'key': key, 'check_n': CHECK, 'check': summary(first), **summary(acc)}
    return out


def main(names):
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in names:
        data[name] = run_case(name)
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + '\n')


if __name__ == '__main__':
    main(sys.argv[1:] or CASES)
