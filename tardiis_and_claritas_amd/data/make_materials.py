"""Regenerate materials.json from the reference's module-level tables.

The CLARITAS parameter set includes two measured particle-size distributions
and their exploratory complex-index profiles (CLARITAS.py:248-705).  They are
input data, kept identical so that `simulate('loess', ...)` means the same
population.  This script reads the reference source as text (ast literal
evaluation; nothing is imported or executed) and writes materials.json.

Usage: python make_materials.py /path/to/CLARITAS/CLARITAS.py
"""
import ast
import json
import sys
from pathlib import Path

NAMES = {
    'LOESS_DIAMETER_M': ('loess', 'diameter_m'),
    'LOESS_WEIGHTS': ('loess', 'weights'),
    'KAOLIN_DIAMETER_M': ('kaolin', 'diameter_m'),
    'KAOLIN_WEIGHTS': ('kaolin', 'weights'),
    'DEFAULT_LOESS_N_REAL': ('loess', 'n_real'),
    'DEFAULT_LOESS_K_IMAG': ('loess', 'k_imag'),
    'DEFAULT_KAOLIN_N_REAL': ('kaolin', 'n_real'),
    'DEFAULT_KAOLIN_K_IMAG': ('kaolin', 'k_imag'),
}


def _numbers(node):
    # np.array([...], dtype=...) -> first positional argument
    if isinstance(node, ast.Call) and node.args:
        node = node.args[0]
    return [float(v) for v in ast.literal_eval(node)]


def main(src):
    tree = ast.parse(Path(src).read_text())
    out = {'loess': {}, 'kaolin': {}}
    for stmt in tree.body:
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1:
            t = stmt.targets[0]
            if isinstance(t, ast.Name) and t.id in NAMES:
                mat, key = NAMES[t.id]
                out[mat][key] = _numbers(stmt.value)
    for mat, d in out.items():
        n = len(d['diameter_m'])
        assert all(len(v) == n for v in d.values()), (mat, {k: len(v) for k, v in d.items()})
    dst = Path(__file__).with_name('materials.json')
    dst.write_text(json.dumps(out, indent=1))
    print('wrote', dst, {m: len(d['diameter_m']) for m, d in out.items()})


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/CLARITAS/CLARITAS.py')
