"""The C ABI: the library loads without a GPU, exports every entry point
declared in include/claritas.h, its struct sizes match the ctypes mirror,
and it fails loudly (an error code, no fallback) when no device is present."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / 'include' / 'claritas.h').read_text()
    return sorted(set(re.findall(r'\b(clr_[a-z_0-9]+)\s*\(', text)))


def test_header_matches_ctypes_table():
    from tardiis_and_claritas_amd import _lib
    assert sorted(_lib.EXPORTED) == _declared()


def test_library_exports_and_sizes():
    from tardiis_and_claritas_amd import _lib, abi
    lib = _lib.load()  # raises on ABI version or struct-size mismatch
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.clr_abi_version() == abi.ABI_VERSION
    assert lib.clr_sizeof(99) == 0


def test_no_device_fails_loudly():
    from conftest import _hip_devices
    if _hip_devices() > 0:
        pytest.skip('a HIP device is present')
    from tardiis_and_claritas_amd import _lib
    with pytest.raises(_lib.ClaritasError):
        _lib.Context(0)


def test_model_validation_is_fail_closed():
    from tardiis_and_claritas_amd import ClaritasForwardModel, workloads
    kw = workloads.base_kwargs()
    # rough surfaces and the morphology channel are modelled; out-of-range values are not
    ClaritasForwardModel(**dict(kw, particle_surface_rms_slope_deg=25.0, wave_surface_rms_height_m=100e-9))
    for bad in (dict(particle_surface_rms_slope_deg=80.0), dict(wave_surface_rms_height_m=-1e-9),
                dict(wave_spheroid_aspect_ratio=5.5), dict(wave_spheroid_orientation_model='random'),
                dict(wave_spheroid_orientation_kappa=101.0)):
        with pytest.raises(ValueError):
            ClaritasForwardModel(**dict(kw, **bad))
    with pytest.raises(ValueError):
        ClaritasForwardModel(**dict(kw, particle_event_model='hybrid_mie_exact_rate_moment_constrained'))
    with pytest.raises(ValueError):
        ClaritasForwardModel(**dict(kw, sediment_transport_model='rz_field'))
