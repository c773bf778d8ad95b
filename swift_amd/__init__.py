"""swift_amd -- MI355X-native SPHENIX neighbour pass (density, ghost,
gradient, extra ghost, force) behind SWIFT's task interface.

The compute path is libswsph.so (hand-written HIP for gfx950) behind the C ABI
of include/swsph.h; this package is the Python host mirror of the reference's
task entry points plus synthetic initial conditions.
"""
from ._lib import SwsphError, lib, LIB_PATH
from .engine import Engine, FIELDS
from .server import Server
from .ics import HydroProps
from . import ics

__all__ = ["Engine", "FIELDS", "HydroProps", "Server", "SwsphError", "ics", "lib",
           "LIB_PATH"]
