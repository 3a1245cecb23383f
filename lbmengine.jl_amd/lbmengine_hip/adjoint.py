"""Gradients of a swimming-speed loss through the discrete adjoint of the
step (liblbm_hip.so's lbm_adjoint_run), and its autograd wrapper.

    L = (mean of |V_swimmer| over the last `window` ticks - target_speed)^2
        + force_penalty * sum F^2

with respect to every particle's B1, B2 and the Guo force field, from the
model's current state over `ticks` ticks (checkpointed every
`checkpoint_every` ticks).  The reference has no differentiable step
(gpu/fluids.jl:132-226 updates Float64 arrays in place); this is new.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _capi


class LbmAdjointConfig(C.Structure):
    _fields_ = [
        ("ticks", C.c_int64),
        ("checkpoint_every", C.c_int64),
        ("window", C.c_int64),
        ("particle", C.c_int32),
        ("reserved", C.c_int32),
        ("target_speed", C.c_double),
        ("force_penalty", C.c_double),
    ]


@dataclass
class SwimmingGradient:
    loss: float
    dB1: np.ndarray                 # one per particle
    dB2: np.ndarray
    dF: Optional[np.ndarray]        # shape (n..., 3), or None without a force field
    forward_ms: float
    reverse_ms: float


def swimming_loss_gradient(model, *, ticks: int, checkpoint_every: int = 100, window: int | None = None,
                           particle: int = 0, target_speed: float = 1e-4, force_penalty: float = 0.0,
                           force_gradient: bool = True) -> SwimmingGradient:
    """Loss and gradients by the adjoint; the model ends `ticks` ahead."""
    fn = _capi.load()
    rm = model.resolved
    cfg = LbmAdjointConfig(int(ticks), int(checkpoint_every), int(ticks if window is None else window),
                           int(particle), 0, float(target_speed), float(force_penalty))
    npart = len(model.particles)
    loss = C.c_double()
    dB1 = np.zeros(max(npart, 1))
    dB2 = np.zeros(max(npart, 1))
    dF = None
    if force_gradient and rm.force_mode == _capi.FORCE_FIELD:
        dF = np.zeros(tuple(rm.n[: rm.dims]) + (rm.dims,), order="F")
    _capi.check(fn["lbm_adjoint_run"](model._h, C.byref(cfg), C.byref(loss), _capi.as_dptr(dB1),
                                      _capi.as_dptr(dB2), None if dF is None else _capi.as_dptr(dF)))
    fms, rms = C.c_double(), C.c_double()
    _capi.check(fn["lbm_adjoint_last_times"](model._h, C.byref(fms), C.byref(rms)))
    return SwimmingGradient(loss.value, dB1[:npart], dB2[:npart], dF, fms.value, rms.value)


def swimming_loss(model, *, ticks: int, window: int | None = None, particle: int = 0,
                  target_speed: float = 1e-4, force_penalty: float = 0.0, force_field=None,
                  trajectory: list | None = None) -> float:
    """The same loss by a plain forward run (for finite differences).  With
    `trajectory`, every particle's position before each tick is appended to
    it (the geometry the tick used).  `force_field` is the field the penalty
    is taken on: the model's own, which is zero at wall nodes."""
    w = ticks if window is None else window
    speeds = []
    if trajectory is None:
        model.run(ticks - w)
        start = ticks - w
    else:
        start = 0
    for t in range(start, ticks):
        if trajectory is not None:
            trajectory.append([np.array(p.position) for p in model.particles])
        model.run(1)
        if t >= ticks - w:
            speeds.append(float(np.linalg.norm(model.particles[particle].velocity)))
    model.sync()
    L = (np.mean(speeds) - target_speed) ** 2
    if force_field is not None:
        L += force_penalty * float(np.sum(np.asarray(force_field) ** 2))
    return float(L)


def link_switches(model, traj_a: list, traj_b: list) -> int:
    """Ticks at which two runs' link sets differ: a node solid in one run's
    sphere (getSphere's test on the tick's position) and not in the other's.
    Finite differences across such a tick see the jump of the piecewise-
    constant geometry, which the gradient does not contain."""
    rm = model.resolved
    coords = [np.asarray(c, dtype=np.float64) for c in rm.coords[: rm.dims]]
    radii = [p.particleParams["radius"] for p in model.particles]
    count = 0
    for pa, pb in zip(traj_a, traj_b):
        differ = False
        for Ra, Rb, r in zip(pa, pb, radii):
            r2 = r * r
            # nodes within r + 2 of the first run's centre
            sel = [np.nonzero(np.abs(c - Ra[k]) <= r + 2.0)[0] for k, c in enumerate(coords)]
            grids = np.meshgrid(*[coords[k][sel[k]] for k in range(rm.dims)], indexing="ij")
            da = sum((g + (-Ra[k])) ** 2 for k, g in enumerate(grids)) if rm.dims == 2 else \
                ((grids[0] + (-Ra[0])) ** 2 + (grids[1] + (-Ra[1])) ** 2) + (grids[2] + (-Ra[2])) ** 2
            db = sum((g + (-Rb[k])) ** 2 for k, g in enumerate(grids)) if rm.dims == 2 else \
                ((grids[0] + (-Rb[0])) ** 2 + (grids[1] + (-Rb[1])) ** 2) + (grids[2] + (-Rb[2])) ** 2
            if np.any((da < r2) != (db < r2)):
                differ = True
        count += int(differ)
    return count


def _autograd_function():
    import torch

    class SwimmingSpeedLoss(torch.autograd.Function):
        """torch.autograd counterpart of a ChainRulesCore rrule: forward runs
        the model built by `make_model(B1, B2, F)` through the adjoint and
        keeps the gradients; backward scales them by the incoming gradient."""

        @staticmethod
        def forward(ctx, B1, B2, F, make_model: Callable, options: dict):
            model = make_model(B1.detach().cpu().numpy(), B2.detach().cpu().numpy(), F.detach().cpu().numpy())
            g = swimming_loss_gradient(model, **options)
            ctx.save_for_backward(torch.as_tensor(g.dB1, dtype=B1.dtype), torch.as_tensor(g.dB2, dtype=B2.dtype),
                                  torch.as_tensor(g.dF if g.dF is not None else np.zeros(F.shape), dtype=F.dtype))
            return torch.tensor(g.loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            dB1, dB2, dF = ctx.saved_tensors
            return grad * dB1, grad * dB2, grad * dF, None, None

    return SwimmingSpeedLoss


def swimming_speed_loss(B1, B2, F, make_model: Callable, **options):
    """Autograd entry: loss tensor whose backward gives dL/dB1, dL/dB2, dL/dF."""
    return _autograd_function().apply(B1, B2, F, make_model, options)
