"""MI355X-native distillation elemental generation.

Drop-in for the ``ElementalGenerator`` path of EasyDistillation
(lattice/generator/elemental.py) computing on gfx950 through the HIP library
``lib/libelemental_hip.so`` (C ABI in ``include/elemental_hip.h``).
"""

from .backend import get_backend, set_backend
from .constant import Nc, Nd, Ns
from .elemental import ElementalGenerator, blending_weights
from .displacement import DisplacementElementalGenerator
from .insertion import MomentumPhase, derivative, derivative_index, num_derivative, output
from .preset import (
    ArrayFileData,
    EigenvectorHostmem,
    EigenvectorNpy,
    EigenvectorTimeSlice,
    ElementalBinary,
    ElementalNpy,
    GaugeFieldBinary,
    GaugeFieldHostmem,
    GaugeFieldIldg,
    GaugeFieldNpy,
    GaugeFieldTimeSlice,
    save_elemental,
)

__all__ = [
    "get_backend",
    "set_backend",
    "Nc",
    "Nd",
    "Ns",
    "ElementalGenerator",
    "DisplacementElementalGenerator",
    "blending_weights",
    "MomentumPhase",
    "derivative",
    "derivative_index",
    "num_derivative",
    "output",
    "ArrayFileData",
    "EigenvectorHostmem",
    "EigenvectorNpy",
    "EigenvectorTimeSlice",
    "ElementalBinary",
    "ElementalNpy",
    "GaugeFieldBinary",
    "GaugeFieldIldg",
    "GaugeFieldHostmem",
    "GaugeFieldNpy",
    "GaugeFieldTimeSlice",
    "save_elemental",
]
