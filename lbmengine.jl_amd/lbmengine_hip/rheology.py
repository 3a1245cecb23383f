"""viscousStressTensor (cpu/fluids.jl:507-520) on the device."""
from __future__ import annotations

import numpy as np

from . import _capi


def viscousStressTensor(model):
    """The viscous stress of the current populations as the reference
    returns it: a dims x dims nested list whose [a][b] entry is the field
    sigma_ab (shape of the lattice), sigma_ab = -(1 - dt/(2 tau))
    sum_i c_ia c_ib (f_i - feq_i)."""
    rm = model.resolved
    d = rm.dims
    out = np.zeros(tuple(rm.n[:d]) + (d, d), order="F")
    _capi.check(model._fn["lbm_viscous_stress"](model._h, _capi.as_dptr(out)))
    return [[np.asfortranarray(out[..., a, b]) for b in range(d)] for a in range(d)]
