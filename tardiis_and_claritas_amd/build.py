"""Compile the gfx950 HIP library libclaritas_hip.so in-tree.

    python -m tardiis_and_claritas_amd.build

Every translation unit under csrc/ is compiled by hipcc for gfx950 and linked
into one shared library next to this file; the ctypes layer loads it from
there.  Packed FP32 instructions are disabled (-packed-fp32-ops): hipcc would
otherwise form v_pk_* FP32 ops from scalar float code.  SimplifyCFG's
common-store sinking is disabled (-simplifycfg-sink-common=false): it merges
stores to different fields of the per-lane photon state into one store
through a selected pointer, which keeps that state in scratch memory instead
of VGPRs.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / 'csrc'
LIB = HERE / 'libclaritas_hip.so'
ARCH = os.environ.get('CLARITAS_OFFLOAD_ARCH', 'gfx950')
SOURCES = ['mie.hip', 'precompute.hip', 'meridional.hip', 'transport.hip', 'mlp.hip', 'mlp_static.hip', 'api.cpp',
           'causal_bin.cpp']
FLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-fPIC', '-std=c++17', '-munsafe-fp-atomics',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-mllvm', '-simplifycfg-sink-common=false',
    '-Wno-unused-result',
]


def hipcc() -> str:
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError('hipcc not found: the gfx950 library cannot be built')


def _compile(src: str, obj_dir: Path) -> Path:
    obj = obj_dir / (Path(src).stem + '.o')
    cmd = [hipcc(), *FLAGS, '-x', 'hip', '-c', str(CSRC / src), '-o', str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stderr[-6000:]}')
    return obj


def build(verbose: bool = False) -> Path:
    obj_dir = HERE / '_build'
    obj_dir.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = LIB.with_suffix('.so.tmp')
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr[-4000:]}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
