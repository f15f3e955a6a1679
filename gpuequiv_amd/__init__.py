"""gpuequiv on MI355X: spectroscopy energy games solved by hand-written HIP kernels.

The public names mirror the Rust crate Gobbel2000/gpuequiv v1.0.0; the solver
behind ``EnergyGame.run`` is the C ABI of ``libgpuequiv_hip.so``
(include/gpuequiv.h).
"""
from ._lib import Error
from .energy import (Energy, EnergyArray, EnergyConf, Upd, Update, UpdateArray, std_equivalences, update)
from .energygame import EnergyGame, GameGraph

__all__ = [
    "Energy", "EnergyArray", "EnergyConf", "Upd", "Update", "UpdateArray", "std_equivalences", "update",
    "EnergyGame", "GameGraph", "Error",
]
