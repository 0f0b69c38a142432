"""bonzo_amd -- MI355X-native Van-Leer MHD step with the host API of
sergekomarov/bonzo's Simulation/BnzIntegr (see simulation.py).

The numerical path runs in libbnzgpu.so (HIP, gfx950) through the C ABI of
include/bnzgpu.h; importing works without a GPU, running does not.
"""
from . import problems
from ._lib import BnzError, Context
from .simulation import BnzGrid, BnzIntegr, Reconstr, RSolver, Simulation, TIntegr, read_param

__all__ = ["problems", "Context", "BnzError", "Simulation", "BnzGrid", "BnzIntegr",
           "RSolver", "Reconstr", "TIntegr", "read_param"]
