"""gpuequiv on MI355X: spectroscopy energy games solved by hand-written HIP kernels.

The public names mirror the Rust crate Gobbel2000/gpuequiv v1.0.0; the device
work behind ``EnergyGame.run``, ``GameBuild`` and ``TransitionSystem.compare``
/ ``equivalences`` is the C ABI of ``libgpuequiv_hip.so`` (include/gpuequiv.h).
"""
from ._lib import Error
from .energy import (Energy, EnergyArray, EnergyConf, Upd, Update, UpdateArray, std_equivalences, update)
from .energygame import EnergyGame, GameGraph
from .equivalence import Equivalence, EquivalenceRelation, StartInfo
from .gamebuild import AttackPosition, DefendPosition, GameBuild, Position, SingletonPosition
from .lts import Transition, TransitionSystem

__all__ = [
    "Energy", "EnergyArray", "EnergyConf", "Upd", "Update", "UpdateArray", "std_equivalences", "update",
    "EnergyGame", "GameGraph", "Error", "Equivalence", "EquivalenceRelation", "StartInfo",
    "AttackPosition", "DefendPosition", "GameBuild", "Position", "SingletonPosition", "Transition",
    "TransitionSystem",
]
