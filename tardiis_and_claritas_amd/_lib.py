"""Load libclaritas_hip.so and declare its C ABI for ctypes.

There is no CPU fallback: if the library is missing or no HIP device is
visible, `load()` / `Context()` raise.  Build the library with
`python -m tardiis_and_claritas_amd.build`.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import abi

LIB_PATH = Path(__file__).resolve().parent / 'libclaritas_hip.so'
_lib = None

P = C.POINTER
_SIGS = {
    'clr_abi_version': (C.c_int, []),
    'clr_sizeof': (C.c_size_t, [C.c_int]),
    'clr_create': (C.c_int, [P(C.c_void_p), C.c_int]),
    'clr_destroy': (None, [C.c_void_p]),
    'clr_last_error': (C.c_char_p, [C.c_void_p]),
    'clr_device_info': (C.c_int, [C.c_void_p, P(C.c_int64)]),
    'clr_mie_tables': (C.c_int, [C.c_void_p, P(abi.Particles), P(abi.MieOut)]),
    'clr_prepare': (C.c_int, [C.c_void_p, C.c_int, P(abi.Config), P(abi.Particles), P(abi.Sediment),
                              P(C.c_float), C.c_size_t]),
    'clr_n_slots': (C.c_int, [C.c_void_p]),
    'clr_get_tables': (C.c_int, [C.c_void_p, C.c_int, P(abi.TablesOut)]),
    'clr_run': (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, P(abi.Launch), P(abi.Tallies),
                          P(abi.Record), C.c_uint64]),
    'clr_run_batch': (C.c_int, [C.c_void_p, P(abi.BatchItem), C.c_int, P(abi.Launch), P(abi.Tallies)]),
    'clr_time_run': (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, P(abi.Launch), C.c_int, P(C.c_double)]),
    'clr_last_kernel_ms': (C.c_double, [C.c_void_p]),
    'clr_run_light': (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, P(abi.Launch), P(abi.LightConfig),
                                P(abi.Tallies), P(abi.LightOut)]),
    'clr_run_causal': (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, P(abi.Launch), P(abi.LightConfig),
                                 P(abi.Tallies), P(abi.LightOut), P(abi.CausalOut)]),
    'clr_causal_layout': (C.c_int, [C.c_int, C.c_int, C.c_char_p, C.c_int, P(C.c_int)]),
    'clr_causal_bin': (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                                 P(C.c_double)]),
    'clr_light_probe': (C.c_int, [C.c_void_p, C.c_int, C.c_int, P(C.c_float), P(abi.LightConfig), C.c_double,
                                  C.c_double, P(abi.LightOut)]),
    'clr_meridional_solve': (C.c_int, [C.c_void_p, P(abi.Config), P(abi.Particles), P(abi.Sediment),
                                       P(abi.Meridional), P(abi.MeridionalOut)]),
    'clr_probe_shape': (C.c_int, [C.c_int, P(C.c_int), P(C.c_int)]),
    'clr_probe': (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, P(C.c_float), P(C.c_double)]),
    'clr_mlp_init': (C.c_int, [C.c_void_p, P(abi.MlpShape), C.c_uint64]),
    'clr_mlp_param_count': (C.c_size_t, [C.c_void_p]),
    'clr_mlp_get_params': (C.c_int, [C.c_void_p, P(C.c_float), C.c_size_t]),
    'clr_mlp_set_params': (C.c_int, [C.c_void_p, P(C.c_float), C.c_size_t]),
    'clr_mlp_train': (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int64, C.c_int32,
                                C.c_int32, C.c_float, C.c_uint64, P(C.c_float)]),
    'clr_mlp_grad': (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), C.c_int32, P(C.c_float),
                               P(C.c_float)]),
    'clr_mlp_infer': (C.c_int, [C.c_void_p, P(C.c_float), C.c_int64, P(C.c_float), P(C.c_float)]),
    'clr_mlp_time_infer': (C.c_int, [C.c_void_p, C.c_int64, C.c_int, P(C.c_double)]),
    'clr_mlp_infer_static': (C.c_int, [C.c_void_p, P(C.c_float), C.c_int64, P(C.c_float), P(C.c_float)]),
}
EXPORTED = tuple(_SIGS)


def load(path: str | os.PathLike | None = None):
    """Return the loaded library (symbols typed); raise if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f'{p} is missing: build the gfx950 library with '
            '`python -m tardiis_and_claritas_amd.build` (there is no CPU fallback)')
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.clr_abi_version() != abi.ABI_VERSION:
        raise RuntimeError('libclaritas_hip.so ABI version mismatch; rebuild it')
    abi.check_sizes(lib)
    if path is None:
        _lib = lib
    return lib


class ClaritasError(RuntimeError):
    """Error reported by the HIP library (code and message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f'CLARITAS HIP error {code}: {msg}')
        self.code = code


class Context:
    """One device context of the HIP library (device buffers, stream)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.clr_create(C.byref(h), int(device))
        if rc != 0:
            raise ClaritasError(rc, f'clr_create(device={device}) failed: no usable HIP device')
        self.h = h
        self.device = int(device)

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.clr_last_error(self.h)
            raise ClaritasError(rc, f'{what}: {msg.decode() if msg else ""}')

    def device_info(self):
        info = (C.c_int64 * 8)()
        self.check(self.lib.clr_device_info(self.h, info), 'clr_device_info')
        keys = ['compute_units', 'wavefront', 'lds_per_block', 'clock_khz', 'gfx', 'hbm_mib',
                'plain_blocks_per_cu', 'generic_blocks_per_cu']
        return dict(zip(keys, list(info)))

    def probe(self, mode: int, items, slot: int = 0, seed: int = 0):
        """Run clr_probe `mode` (abi.PROBE_MODES name or code) on items
        [n][IN] of slot `slot`; returns [n][OUT] doubles."""
        import numpy as np
        code = abi.PROBE_MODES[mode] if isinstance(mode, str) else int(mode)
        si, so = C.c_int(), C.c_int()
        self.check(self.lib.clr_probe_shape(code, C.byref(si), C.byref(so)), 'clr_probe_shape')
        a = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, si.value)
        out = np.zeros((a.shape[0], so.value), dtype=np.float64)
        self.check(self.lib.clr_probe(self.h, int(slot), code, a.shape[0], int(seed) & 0xFFFFFFFFFFFFFFFF,
                                      a.ctypes.data_as(P(C.c_float)), out.ctypes.data_as(P(C.c_double))),
                   'clr_probe')
        return out

    def close(self):
        if getattr(self, 'h', None):
            self.lib.clr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
