"""lbmengine_hip — LBMengine.jl's time step on MI355X (gfx950).

The reference's model-construction and step API (modelInit, addSquirmer!,
addBead!, LBMpropagate!, hydroVariablesUpdate!) over liblbm_hip.so, whose
C ABI is include/lbm_hip.h.
"""
from ._capi import LBMError, LIB_PATH
from .model import (DEFAULT, LBMbatch, LBMmodel, ResolvedModel, addBead, addSquirmer, guard,
                    hydroVariablesUpdate, julia_range, LBMpropagate, modelInit, propagate_ticks, resolve_model,
                    resolve_particle, sphere_count)
from .interactions import (addBistableBond, addDipoles, addLennardJones, addLinearBond, addPolarBond,
                           addWCA)
from .rheology import viscousStressTensor
from .savedata import writeTensor

__all__ = [
    "DEFAULT", "LBMError", "LBMbatch", "LBMmodel", "LIB_PATH", "ResolvedModel", "addBead", "addSquirmer",
    "addBistableBond", "addDipoles", "addLennardJones", "addLinearBond", "addPolarBond", "addWCA",
    "guard", "hydroVariablesUpdate", "julia_range", "LBMpropagate", "modelInit", "propagate_ticks",
    "resolve_model", "resolve_particle", "sphere_count", "viscousStressTensor", "writeTensor",
]
