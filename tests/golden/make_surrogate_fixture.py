"""Write tests/golden/surrogate_config3.npz: a small training set and a
spot-check set of config-3 configurations with their batched Monte Carlo
signals, for the GPU test of what the trained surrogate gains over Monte
Carlo.

The inputs are dataset.npz files written by `python -m
tardiis_and_claritas_amd.surrogate --generate-only` on one MI355X:

    train: --n-optical 1024 --n-variants 48 --counts 150 --seed 20261401
    spot:  --n-test-optical 64 --n-test-variants 32 --test-counts 4000
           --max-photons 8e8 --seed 20261520

The first N_TRAIN training configurations and the first N_SPOT spot-check
configurations are kept (parameters, signals S90/S0/S135 and the histories
behind them).

    python tests/golden/make_surrogate_fixture.py TRAIN.npz SPOT.npz
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / 'surrogate_config3.npz'
N_TRAIN = 8192
N_SPOT = 256
KEYS = ('median_um', 'sigma_g', 'm_r', 'm_i', 'wavelength_nm', 'conc_mg_l', 'vortex_depth_mm', 'rouse')


def main(train_path, spot_path):
    a, b = np.load(train_path), np.load(spot_path)
    out = {}
    for part, d, src, n in (('train', a, 'train', N_TRAIN), ('spot', b, 'test', N_SPOT)):
        for k in KEYS:
            out[f'{part}_{k}'] = np.asarray(d[f'{src}_{k}'][:n], dtype=np.float64)
        out[f'{part}_S'] = np.asarray(d[f'{src}_S'][:n], dtype=np.float64)
        out[f'{part}_hist'] = np.asarray(d[f'{src}_hist'][:n], dtype=np.int64)
    np.savez_compressed(OUT, **out)
    print(f'wrote {OUT} ({OUT.stat().st_size / 1024:.0f} KiB)')


if __name__ == '__main__':
    main(sys.argv[1], sys.argv[2])
