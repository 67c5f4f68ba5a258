"""MI355X-native photon-transport path of the CLARITAS sample-cell model.

`ClaritasForwardModel(...).simulate(material, concentration_g_per_L, ...)`
returns a `ForwardResult` computed by the gfx950 HIP library
libclaritas_hip.so (built with `python -m tardiis_and_claritas_amd.build`).
"""
from .physics import Material, lognormal, monodisperse, reference_material, cell_geometry  # noqa: F401
from .model import ClaritasForwardModel, ForwardResult  # noqa: F401

__all__ = ['ClaritasForwardModel', 'ForwardResult', 'Material', 'lognormal', 'monodisperse',
           'reference_material', 'cell_geometry']
