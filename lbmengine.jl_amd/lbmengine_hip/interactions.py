"""Particle-particle interactions: the reference's addLinearBond!,
addPolarBond!, addBistableBond!, addDipoles!, addLennardJones! and addWCA!
(gpu/initMethods.jl:229-414) over model.particleInteractions.

Parameters are resolved here, on the host, as the reference resolves them at
construction; the forces themselves (applyInteraction!,
gpu/particles.jl:110-197) are evaluated by the step kernels, which apply the
expanded terms in order before every particle move of a tick.  Particle ids
are 1-based, as in the reference.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from . import _capi
from .model import DEFAULT


# --- the reference's interaction structs (gpu/structs.jl:55-93) -----------
@dataclass
class LinearInteraction:
    id1: int
    id2: int
    equilibriumDisplacement: float
    stiffness: float


@dataclass
class BistableInteraction:
    id1: int
    id2: int
    trapRadius: float
    trapWidth: float
    hillHeight: float
    a: float
    b: float


@dataclass
class PolarInteraction:
    id1: int
    id2: int
    id3: int
    equilibriumAngle: float
    stiffness: float


@dataclass
class DipoleDipoleInteraction:
    pairs: list
    dipoleConstant: float
    magneticField: object  # a constant vector, a number (a 2D field along z), or a function of time giving either


@dataclass
class LennardJonesInteraction:
    pairs: list
    epsilon: float
    sigma: float
    cutoff: float


# --- Julia arithmetic the resolution relies on ------------------------------
def _fma(a: float, b: float, c: float) -> float:
    """a*b + c rounded once (exact rational arithmetic, then one rounding)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def julia_pow(x: float, n: int) -> float:
    """x^n for an integer n >= 1 as Julia's ^(::Float64, ::Integer) computes
    it (Base.Math.pow_body: binary powering whose products carry their
    rounding errors, folded back at the end); n == 3 is the literal x*x*x."""
    x = float(x)
    if n == 3:
        return x * x * x
    y, xnlo, ynlo = 1.0, 0.0, 0.0
    while n > 1:
        if n & 1:
            err = _fma(y, xnlo, x * ynlo)
            hi = x * y
            ynlo = _fma(x, y, -hi) + err
            y = hi
        err = x * 2.0 * xnlo
        hi = x * x
        xnlo = _fma(x, x, -hi) + err
        x = hi
        n >>= 1
    err = _fma(y, xnlo, x * ynlo)
    return _fma(x, y, err) if math.isfinite(x) and math.isfinite(err) else x * y


def _norm(v) -> float:
    """`v |> Array |> norm` for a short vector: squares summed in order."""
    s = 0.0
    for i, x in enumerate(v):
        s = x * x if i == 0 else s + x * x
    return math.sqrt(s)


def _position(model, i: int) -> np.ndarray:
    return np.asarray(model.particles[i - 1].position, dtype=np.float64)


def _check_id(model, *ids) -> None:
    for i in ids:
        assert 1 <= i <= len(model.particles), f"there is no particle {i}"


def _dx_dt(model):
    st = model.spaceTime
    return float(st["latticeParameter"]), float(st["timeStep"])


# --- constructors -------------------------------------------------------------
def addLinearBond(model, id1: int, id2: int, *, equilibriumDisplacement=DEFAULT, stiffness=1) -> None:
    """addLinearBond! (gpu/initMethods.jl:229-245)."""
    _check_id(model, id1, id2)
    if equilibriumDisplacement is DEFAULT:
        equilibriumDisplacement = _norm(_position(model, id2) - _position(model, id1))
    assert isinstance(equilibriumDisplacement, (int, float)), "equilibriumDisplacement must be a number!"
    assert isinstance(stiffness, (int, float)), "stiffness must be a number!"
    model.particleInteractions.append(LinearInteraction(id1, id2, float(equilibriumDisplacement), float(stiffness)))
    sync_interactions(model)


def addPolarBond(model, id1: int, id2: int, id3: int, *, equilibriumAngle=DEFAULT, stiffness=1) -> None:
    """addPolarBond! (gpu/initMethods.jl:247-273)."""
    _check_id(model, id1, id2, id3)
    vecA = _position(model, id1) - _position(model, id2)
    vecB = _position(model, id3) - _position(model, id2)
    normA, normB = _norm(vecA), _norm(vecB)
    s = 0.0
    for k in range(len(vecA)):
        s = vecA[k] * vecB[k] if k == 0 else s + vecA[k] * vecB[k]
    cosAlpha = s / (normA * normB)
    if equilibriumAngle is DEFAULT:
        equilibriumAngle = math.acos(cosAlpha)
    assert isinstance(equilibriumAngle, (int, float)), "equilibriumAngle must be a number!"
    assert isinstance(stiffness, (int, float)), "stiffness must be a number!"
    model.particleInteractions.append(PolarInteraction(id1, id2, id3, float(equilibriumAngle), float(stiffness)))
    sync_interactions(model)


def addBistableBond(model, id1: int, id2: int, *, lowDisp=DEFAULT, highDisp=DEFAULT, hillHeight=1) -> None:
    """addBistableBond! (gpu/initMethods.jl:282-308)."""
    _check_id(model, id1, id2)
    dx, _ = _dx_dt(model)
    if lowDisp is DEFAULT:
        lowDisp = _norm(_position(model, id2) - _position(model, id1))
    if highDisp is DEFAULT:
        highDisp = lowDisp + dx
    assert isinstance(lowDisp, (int, float)), "lowDisp must be a number!"
    assert isinstance(highDisp, (int, float)), "highDisp must be a number!"
    assert isinstance(hillHeight, (int, float)), "hillHeight must be a number!"
    trapRadius = (highDisp + lowDisp) / 2
    trapWidth = float(highDisp - lowDisp)
    h = float(hillHeight)
    a = 64.0 * h / julia_pow(trapWidth, 4)        # 4 * 16 * hillHeight / trapWidth^4
    b = 16.0 * h / (trapWidth * trapWidth)        # 2 * 8 * hillHeight / trapWidth^2
    model.particleInteractions.append(BistableInteraction(id1, id2, float(trapRadius), trapWidth, h, a, b))
    sync_interactions(model)


def addDipoles(model, *ids, magneticField=(1, 0), susceptibility=0.5, permeability=400 * math.pi,
               rewrite: bool = False) -> None:
    """addDipoles! (gpu/initMethods.jl:316-359): one dipole-dipole interaction
    over every pair of the given particles (all of them when none is given);
    a later call adds its ids to the existing pairs unless rewrite.  The field
    is a vector, in 2D a number (a field along z), or a function of the model
    time returning either; the function is evaluated at every tick's time
    (gpu/particles.jl:161), its vectors rounded to Float32 as the
    reference's `magneticField(t) |> cu` does (gpu/initMethods.jl:339)."""
    ids = list(ids) if ids else list(range(1, len(model.particles) + 1))
    _check_id(model, *ids)
    at = next((i for i, t in enumerate(model.particleInteractions) if isinstance(t, DipoleDipoleInteraction)), None)
    write_new = at is None or rewrite
    if not write_new:
        old = [i for pair in model.particleInteractions[at].pairs for i in pair]
        ids = sorted(set(ids + old))
    assert len(ids) > 1, "There must be at least two dipoles!"
    pairs = [(ids[i], ids[j]) for i in range(len(ids)) for j in range(len(ids)) if i < j]
    if write_new:
        if rewrite and at is not None:
            del model.particleInteractions[at]
        if callable(magneticField):
            B = magneticField
            field_value(B, model.time, model.dims)  # shape check at the current time
        elif np.ndim(magneticField) == 0:
            if model.dims != 2:
                raise NotImplementedError("a number as magneticField is a 2D field")
            B = float(magneticField)
        else:
            B = np.array(magneticField, dtype=np.float64)
            assert B.shape == (model.dims,), f"magneticField must have {model.dims} components"
        r = float(model.particles[ids[0] - 1].particleParams["radius"])
        chi, mu = float(susceptibility), float(permeability)
        dc = 4.0 * math.pi * julia_pow(r, 6) * (chi * chi) / (3.0 * mu)
        model.particleInteractions.append(DipoleDipoleInteraction(pairs, dc, B))
    else:
        model.particleInteractions[at].pairs = pairs
    sync_interactions(model)


def _default_epsilon_sigma(model, epsilon, sigma):
    dx, dt = _dx_dt(model)
    if epsilon is DEFAULT:
        rho = np.asarray(model.massDensity, dtype=np.float64).ravel(order="F")
        rho = rho[rho != 0.0]
        epsilon = float(np.sum(rho)) / rho.size * julia_pow(dx, 5) / (dt * dt)
    if sigma is DEFAULT:
        r = float(model.particles[0].particleParams["radius"])
        sigma = 2 * r + dx
    return float(epsilon), float(sigma)


def addLennardJones(model, *, epsilon=DEFAULT, sigma=DEFAULT, cutoff=DEFAULT) -> None:
    """addLennardJones! (gpu/initMethods.jl:361-396): every pair of particles;
    replaces an earlier Lennard-Jones interaction.  The default epsilon
    averages the nonzero densities (a sum of unspecified order in the
    reference)."""
    ids = list(range(1, len(model.particles) + 1))
    at = next((i for i, t in enumerate(model.particleInteractions) if isinstance(t, LennardJonesInteraction)), None)
    if at is not None:
        del model.particleInteractions[at]
    pairs = [(ids[i], ids[j]) for i in range(len(ids)) for j in range(len(ids)) if i < j]
    epsilon, sigma = _default_epsilon_sigma(model, epsilon, sigma)
    cutoff = math.inf if cutoff is DEFAULT else float(cutoff)
    model.particleInteractions.append(LennardJonesInteraction(pairs, epsilon, sigma, cutoff))
    sync_interactions(model)


def addWCA(model, *, epsilon=DEFAULT, sigma=DEFAULT) -> None:
    """addWCA! (gpu/initMethods.jl:398-414): Lennard-Jones cut at its minimum."""
    epsilon, sigma = _default_epsilon_sigma(model, epsilon, sigma)
    addLennardJones(model, epsilon=epsilon, sigma=sigma, cutoff=2.0 ** (1 / 6) * sigma)


def field_value(fn, t: float, dims: int):
    """(B, scalar) of a time-dependent field at time t: a number (2D only)
    stays a Float64 number; a vector is rounded to Float32, as `cu` rounds a
    Float64 array (gpu/initMethods.jl:339)."""
    v = fn(t)
    if np.ndim(v) == 0:
        if dims != 2:
            raise NotImplementedError("a number as magneticField is a 2D field")
        return np.array([float(v), 0.0, 0.0]), True
    a = np.asarray(v)
    assert a.shape == (dims,), f"magneticField must have {dims} components"
    if np.issubdtype(a.dtype, np.floating):
        a = a.astype(np.float32)
    out = np.zeros(3)
    out[:dims] = a.astype(np.float64)
    return out, False


def field_schedule(fn, time0: float, dt: float, count: int, dims: int) -> np.ndarray:
    """The field of the next `count` ticks, (count, 3): tick k ends at the
    time tick! reaches by adding dt k times to time0 (gpu/fluids.jl:212)."""
    out = np.zeros((count, 3))
    t = time0
    for k in range(count):
        t = t + dt
        out[k] = field_value(fn, t, dims)[0]
    return out


def scheduled_field(interactions):
    """The time-dependent field function of the interactions, or None."""
    for it in interactions:
        if isinstance(it, DipoleDipoleInteraction) and callable(it.magneticField):
            return it.magneticField
    return None


# --- expansion into the library's terms ---------------------------------------
def expand(interactions, dims: int):
    """The terms molecularDynamicsTick! applies, in its order
    (gpu/particles.jl:100-108): one per bond, one per pair of a pair list.
    A dipole interaction whose field is zero adds nothing (`norm(B) ≈ 0`)."""
    terms = []

    def term(kind, ids, params=(), fieldv=()):
        t = _capi.LbmInteraction()
        t.kind = _capi.INTERACTION[kind]
        for k, i in enumerate(ids):
            t.ids[k] = i - 1
        for k, v in enumerate(params):
            t.params[k] = v
        for k, v in enumerate(fieldv):
            t.field[k] = v
        terms.append(t)

    for it in interactions:
        if isinstance(it, LinearInteraction):
            term("linear", (it.id1, it.id2), (it.equilibriumDisplacement, it.stiffness))
        elif isinstance(it, BistableInteraction):
            term("bistable", (it.id1, it.id2), (it.trapRadius, it.a, it.b))
        elif isinstance(it, PolarInteraction):
            term("polar", (it.id1, it.id2, it.id3), (it.equilibriumAngle, it.stiffness))
        elif isinstance(it, DipoleDipoleInteraction):
            B = it.magneticField
            if callable(B):
                # the field comes from the handle's schedule, tick by tick
                scalar = field_value(B, 0.0, dims)[1]
                for a, b in it.pairs:
                    term("dipole", (a, b), (it.dipoleConstant, 1.0 if scalar else 0.0, 1.0))
                continue
            scalar = np.ndim(B) == 0
            Bv = [float(B)] if scalar else [float(x) for x in B]
            if all(x == 0.0 for x in Bv):
                continue
            for a, b in it.pairs:
                term("dipole", (a, b), (it.dipoleConstant, 1.0 if scalar else 0.0), Bv)
        elif isinstance(it, LennardJonesInteraction):
            for a, b in it.pairs:
                term("lennardJones", (a, b), (it.epsilon, it.sigma, it.cutoff))
        else:
            raise TypeError(f"unknown interaction {it!r}")
    if len(terms) > _capi.MAX_INTERACTIONS:
        raise ValueError(f"more than {_capi.MAX_INTERACTIONS} interaction terms")
    arr = (_capi.LbmInteraction * max(1, len(terms)))(*terms)
    return arr, len(terms)


def sync_interactions(model) -> None:
    """Hand model.particleInteractions, expanded, to the model's backend."""
    arr, n = expand(model.particleInteractions, model.dims)
    model._set_interactions(arr, n)
    model._field_fn = scheduled_field(model.particleInteractions)
