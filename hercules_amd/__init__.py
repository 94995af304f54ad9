"""hercules_amd — MI355X-native HERCULES time-step hot path.

The product is the HIP library ``libhercules_amd.so`` (C ABI in
``include/hercules_amd.h``); this package is its Python host mirror.
"""
from .params import Params, read_namelist
from .channel import Channel, Ensemble, HerculesError, load_library, FIELDS, EQUATIONS
from . import grid, cases

__all__ = ["Params", "read_namelist", "Channel", "Ensemble", "HerculesError", "load_library", "FIELDS",
           "EQUATIONS", "grid", "cases"]
