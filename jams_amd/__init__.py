"""jams_amd: an MI355X (gfx950) implementation of JAMS's atomistic Heun-LLG
time step behind the reference's Solver/Hamiltonian/Thermostat plugin API.

Native parts (built in-tree by ``jams_amd.build``):
  lib/libjamship.so  -- HIP engine, C ABI in include/jams_hip.h
  lib/libjamshost.so -- config/lattice/neighbour-list/plugin host layer,
                        C ABI in include/jams_host.h
  lib/jams           -- command-line driver
"""
from .engine import Engine, EngineError, load_library  # noqa: F401

__all__ = ["Engine", "EngineError", "load_library"]
