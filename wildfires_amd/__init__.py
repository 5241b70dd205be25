"""wildfires_amd -- MI355X-native (gfx950 HIP) time-step library for the
dsanmartin/wildfires coupled atmosphere-fire solver.

The product is libwildfires_amd.so (include/wildfires.h); this package holds
its sources (csrc/), the in-tree build (build.py), a ctypes mirror of the C-ABI
(api.py) and the workload definitions (configs.py).
"""
from .api import (Params, Solver, WildfireError, initial_conditions, lib, pack_state, read_parameters,
                  synchronize)

__all__ = ["Params", "Solver", "WildfireError", "initial_conditions", "lib", "pack_state", "read_parameters",
           "synchronize"]
