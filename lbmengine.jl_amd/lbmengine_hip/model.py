"""Host front end of the hip backend: the reference's model-construction and
step API over liblbm_hip.so.

Names and keyword arguments follow LBMengine.jl's GPU backend
(src/gpu/initMethods.jl, src/gpu/fluids.jl); Julia's `!` suffix is dropped
(`addSquirmer!` -> `addSquirmer`).  Parameter resolution (units, relaxation
times, B1/B2, masses) happens here on the host exactly as in the
reference; the time step runs in the HIP library.

Backend-specific keywords of `modelInit`: `storage` ("f64", "f32", "f16"
or "f16c", a 16-bit float of 1 sign, 4 exponent and 11 mantissa bits: how
populations are stored on the device), `semantics` ("gpu" to
follow src/gpu, "cpu" to follow src/cpu) and `rho0` (shift of the
deviation storage).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Any, Callable, Optional

import numpy as np

from . import _capi
from .lattice import velocity_set

DEFAULT = None  # stands for the reference's `:default`


def _sym(s):
    return s.lstrip(":") if isinstance(s, str) else s


# ------------------------------------------------------------- coordinates
def _rat(x) -> Fraction:
    if isinstance(x, (int, np.integer)):
        return Fraction(int(x))
    fr = Fraction(float(x))
    small = fr.limit_denominator(1 << 20)
    return small if float(small) == float(x) else fr


def julia_range(start, stop, step) -> np.ndarray:
    """Values of Julia's `range(start, stop=stop, step=step)` in Float64.

    The reference builds coordinates this way (gpu/initMethods.jl:501-505).
    Julia lifts float endpoints to nearby rationals when it can; this
    evaluates start + i*step exactly in rationals and rounds once, which is
    Julia's result for rational endpoints (integer limits with dx = 1 in
    particular)."""
    a, s, b = _rat(start), _rat(step), _rat(stop)
    if s == 0:
        raise ValueError("range step cannot be zero")
    n = math.floor((b - a) / s) + 1
    if n < 1:
        raise ValueError("empty coordinate range")
    return np.array([float(a + i * s) for i in range(n)], dtype=np.float64)


# ------------------------------------------------------------ resolution
@dataclass
class ResolvedModel:
    """modelInit after unit, scheme and initial-condition resolution."""
    dims: int
    velocities: str
    cs: np.ndarray
    ws: np.ndarray
    opp: np.ndarray
    n: tuple
    coords: list
    collision: str
    storage: str
    semantics: str
    walled: int
    force_mode: int
    force_uniform: np.ndarray
    force_field: Optional[np.ndarray]
    dt: float
    dx: float
    tau: float
    tau_minus: float
    viscosity: float
    kinetic: float
    rho0: float
    latticeSpeed: float
    schemes: list
    init_rho: Any          # number or (N...) field
    init_u: Any            # (d,) vector or (N..., d) field
    fluidParams: dict = field(default_factory=dict)
    incompressible: bool = False
    wall_region: Optional[np.ndarray] = None    # the `walls` function's region (N...)
    wall_velocity: Optional[np.ndarray] = None  # the `movingWalls` function's field (N..., d)

    @property
    def N(self) -> int:
        return int(np.prod(self.n))

    @property
    def Q(self) -> int:
        return len(self.cs)

    def config(self):
        """(lbm_config, keep-alive list) for this model."""
        cfg = _capi.LbmConfig()
        cfg.dims = self.dims
        cfg.q = self.Q
        for k in range(3):
            cfg.n[k] = int(self.n[k]) if k < self.dims else 1
        cfg.collision = _capi.COLLISION[self.collision]
        cfg.storage = _capi.STORAGE[self.storage]
        cfg.semantics = _capi.SEMANTICS[self.semantics]
        cfg.walled = self.walled
        cfg.force = self.force_mode
        cfg.flags = _capi.FLAG_INCOMPRESSIBLE if self.incompressible else 0
        for k in range(3):
            cfg.force_uniform[k] = float(self.force_uniform[k]) if k < len(self.force_uniform) else 0.0
        cfg.dt, cfg.dx, cfg.tau, cfg.tau_minus, cfg.rho0 = self.dt, self.dx, self.tau, self.tau_minus, self.rho0
        keep = []
        for k in range(self.dims):
            arr = np.ascontiguousarray(self.coords[k], dtype=np.float64)
            keep.append(arr)
            cfg.coords[k] = arr.ctypes.data_as(C.POINTER(C.c_double))
        return cfg, keep

    def wall_mask(self) -> np.ndarray:
        """wallRegion: wallNodes (gpu/misc.jl:7-31, the first and last layer of
        each walled dim), the `walls` region and the moving walls' nodes
        (gpu/initMethods.jl:580-620)."""
        m = np.zeros(self.n[: self.dims], dtype=bool, order="F")
        for k in range(self.dims):
            if (self.walled >> k) & 1:
                idx = [slice(None)] * self.dims
                idx[k] = 0
                m[tuple(idx)] = True
                idx[k] = self.n[k] - 1
                m[tuple(idx)] = True
        if self.wall_region is not None:
            m |= self.wall_region
        if self.wall_velocity is not None:
            m |= np.sqrt((self.wall_velocity ** 2).sum(axis=-1)) > 0
        return m

    def initial_populations(self) -> np.ndarray:
        """findInitialConditions (gpu/initMethods.jl:417-438) for the
        resolved initial rho, u and force; (N..., Q) column-major FP64."""
        d, shp = self.dims, tuple(self.n[: self.dims])
        wall = self.wall_mask()
        rho = np.broadcast_to(np.asarray(self.init_rho, dtype=np.float64), shp).copy(order="F")
        u = np.asarray(self.init_u, dtype=np.float64)
        u = np.broadcast_to(u, shp + (d,)).copy(order="F") if u.ndim == 1 else u.copy(order="F")
        rho[wall] = 0.0
        u[wall] = 0.0
        if self.force_mode != _capi.FORCE_NONE:
            F = self.force_array()
            ct = np.zeros_like(u)
            nz = rho != 0.0
            for k in range(d):
                ct[..., k][nz] = F[..., k][nz] / rho[nz]
            u = u - (0.5 * self.dt) * ct
        f = np.zeros(shp + (self.Q,), order="F")
        for q in range(self.Q):
            ci = self.cs[q].astype(np.float64) * self.latticeSpeed
            udc = np.zeros(shp)
            for k in range(d):
                udc = udc + u[..., k] * ci[k]
            uu = np.zeros(shp)
            for k in range(d):
                uu = uu + u[..., k] * u[..., k]
            inv2, inv4 = self.fluidParams["invC2_s"], self.fluidParams["invC4_s"]
            first = udc * inv2 + (udc * udc) * (0.5 * inv4)
            second = first - uu * (0.5 * inv2) + 1.0
            f[..., q] = second * (self.ws[q] * rho)
        return f

    def force_array(self) -> np.ndarray:
        shp = tuple(self.n[: self.dims])
        if self.force_mode == _capi.FORCE_FIELD:
            F = np.array(self.force_field, dtype=np.float64, order="F")
        else:
            F = np.broadcast_to(self.force_uniform[: self.dims], shp + (self.dims,)).copy(order="F")
        F[self.wall_mask()] = 0.0
        return F

    def uniform_init(self) -> bool:
        return np.ndim(self.init_rho) == 0 and np.ndim(self.init_u) == 1


def resolve_model(*, massDensity=DEFAULT, fluidVelocity=DEFAULT, velocities=DEFAULT,
                  collisionModel=DEFAULT, relaxationTimeRatio=DEFAULT, viscosity=DEFAULT,
                  timeStep=DEFAULT, latticeParameter=DEFAULT, kineticParameter=0.25,
                  xlims=DEFAULT, ylims=DEFAULT, zlims=DEFAULT, dims=2, walledDimensions=DEFAULT,
                  dampenEcho=False, walls=DEFAULT, movingWalls=DEFAULT, isFluidCompressible=True,
                  forceDensity=DEFAULT, forcingScheme=DEFAULT, saveData=False,
                  storage="f64", semantics="gpu", rho0=1.0) -> ResolvedModel:
    """Unit and scheme resolution of modelInit (gpu/initMethods.jl:440-702)."""
    given = [x is not DEFAULT for x in (timeStep, relaxationTimeRatio, viscosity)]
    assert not all(given), "timeStep, relaxationTimeRatio, and viscosity cannot be all simultaneously defined!"
    assert timeStep is DEFAULT or latticeParameter is DEFAULT, \
        "timeStep, and latticeParameter cannot be both simultaneously defined!"
    if latticeParameter is not DEFAULT:
        timeStep = latticeParameter
    if timeStep is DEFAULT:
        relaxationTimeRatio = 15 if relaxationTimeRatio is DEFAULT else relaxationTimeRatio
        viscosity = 1 if viscosity is DEFAULT else viscosity
        timeStep = 3 * viscosity / (relaxationTimeRatio - 0.5)
    elif relaxationTimeRatio is DEFAULT:
        viscosity = 1 if viscosity is DEFAULT else viscosity
        relaxationTimeRatio = 3 * viscosity / timeStep + 0.5
    else:
        viscosity = timeStep / 3 * (relaxationTimeRatio - 0.5)
    latticeParameter = timeStep
    dt, dx = float(timeStep), float(latticeParameter)

    xlims = (-1, 1) if xlims is DEFAULT else xlims
    ylims = xlims if ylims is DEFAULT else ylims
    zlims = xlims if zlims is DEFAULT else zlims
    assert all(t[1] != t[0] for t in (xlims, ylims, zlims)), "Coordinate lower and upper limits cannot be equal!"
    coords = [julia_range(l[0], l[1], latticeParameter) for l in (xlims, ylims, zlims)][:dims]
    n = tuple(len(c) for c in coords) + (1,) * (3 - dims)
    shp = n[:dims]

    if dampenEcho:
        raise NotImplementedError(":cbc is not implemented in the gpu modules")
    forcingScheme = "guo" if forcingScheme is DEFAULT else _sym(forcingScheme)
    if forcingScheme != "guo":
        raise NotImplementedError(f"forcing scheme {forcingScheme} is not implemented in the hip backend")

    if massDensity is DEFAULT:
        init_rho = 1.0
    elif np.ndim(massDensity) == 0:
        init_rho = float(massDensity)
    else:
        assert np.shape(massDensity) == shp, f"The mass density given has the wrong size! {shp} expected."
        init_rho = np.array(massDensity, dtype=np.float64, order="F")
    if fluidVelocity is DEFAULT:
        init_u = np.zeros(dims)
    elif np.ndim(fluidVelocity) == 1:
        init_u = np.array(fluidVelocity, dtype=np.float64)
    else:
        assert np.shape(fluidVelocity)[:dims] == shp, f"The fluid velocity given has the wrong size! {shp} expected."
        init_u = np.array(fluidVelocity, dtype=np.float64, order="F")

    if velocities is DEFAULT:
        assert dims > 1, "1 dimensional LBM is not implemented in the hip backend!"
        assert dims <= 3
        velocities = ["D1Q3", "D2Q9", "D3Q15"][dims - 1]
    velocities = _sym(velocities)
    cs, ws, opp = velocity_set(velocities)
    assert cs.shape[1] == dims, f"The chosen velocity must be {dims}-dimensional!"

    latticeSpeed = float(latticeParameter / timeStep)
    c2_s, c4_s = latticeSpeed ** 2 / 3, latticeSpeed ** 4 / 9
    invC2_s, invC4_s = 3 / latticeSpeed ** 2, 9 / latticeSpeed ** 4
    collisionModel = "trt" if collisionModel is DEFAULT else _sym(collisionModel)
    assert collisionModel in ("bgk", "trt"), f"Collision model {collisionModel} is not implemented!"
    relaxationTime = relaxationTimeRatio * timeStep
    schemes = [collisionModel]
    tau_minus = float("nan")
    kin = float("nan")
    if collisionModel == "trt":
        kin = kineticParameter
        if _sym(kineticParameter) == "debug":
            kin = viscosity ** 2 / (c4_s * timeStep ** 2)
        tau_minus = kin * c2_s * timeStep ** 2 / viscosity + timeStep / 2
    fluidParams = dict(c_s=latticeSpeed / math.sqrt(3), c2_s=c2_s, c4_s=c4_s,
                       invC_s=math.sqrt(3) / latticeSpeed, invC2_s=invC2_s, invC4_s=invC4_s,
                       relaxationTime=float(relaxationTime), viscosity=float(viscosity),
                       isFluidCompressible=True)
    if collisionModel == "trt":
        fluidParams.update(relaxationTimePlus=float(relaxationTime), relaxationTimeMinus=float(tau_minus))

    walled = 0
    if walledDimensions is not DEFAULT and len(walledDimensions) != 0:
        for d_ in walledDimensions:
            assert 1 <= d_ <= dims, "walledDimensions are 1-based dimensions of the model"
            walled |= 1 << (d_ - 1)
        schemes.append("bounceBack")

    # walls and moving walls given as functions of the node coordinates
    # (gpu/initMethods.jl:591-626), evaluated over the grid x fastest
    wall_region = wall_velocity = None
    if walls is not DEFAULT:
        assert callable(walls), "walls must be a function!"
        vals = [walls(*pt) for pt in _grid_points(coords, shp)]
        assert all(isinstance(v, (bool, np.bool_)) for v in vals), "walls must return boolean values!"
        wall_region = np.array(vals, dtype=bool).reshape(shp, order="F")
        schemes.append("bounceBack")
    if movingWalls is not DEFAULT:
        assert callable(movingWalls), "movingWalls must be a function!"
        vals = [np.asarray(movingWalls(*pt), dtype=np.float64).ravel() for pt in _grid_points(coords, shp)]
        assert len(vals[0]) == dims, f"movingWalls must return a vector of dimension {dims}!"
        wall_velocity = np.stack(vals).reshape(shp + (dims,), order="F")
        schemes += ["movingWalls", "bounceBack"]

    incompressible = not isFluidCompressible
    fluidParams["isFluidCompressible"] = bool(isFluidCompressible)
    if incompressible:
        # equilibriumDistribution's w rho + w rho0 (...) (gpu/fluids.jl:51)
        # with rho0 the initial density, uniform here (the deviation shift)
        if np.ndim(init_rho) != 0:
            raise NotImplementedError("the incompressible equilibrium needs a uniform initial density in the hip backend")
        rho0 = float(init_rho)

    force_mode, force_uniform, force_field = _capi.FORCE_NONE, np.zeros(3), None
    if forceDensity is not DEFAULT:
        fd = np.asarray(forceDensity, dtype=np.float64)
        if fd.ndim == 1 and fd.shape[0] == dims:
            force_mode = _capi.FORCE_UNIFORM
            force_uniform[:dims] = fd
        elif fd.shape == shp + (dims,):
            force_mode = _capi.FORCE_FIELD
            force_field = np.array(fd, order="F")
        else:
            raise ValueError("force density does not have consistent dimensions!")
        schemes.append(forcingScheme)

    if saveData:
        schemes.append("saveData")

    storage, semantics = _sym(storage), _sym(semantics)
    assert storage in _capi.STORAGE, f"storage must be one of {list(_capi.STORAGE)}"
    assert semantics in _capi.SEMANTICS, f"semantics must be one of {list(_capi.SEMANTICS)}"
    return ResolvedModel(
        dims=dims, velocities=velocities, cs=cs, ws=ws, opp=opp, n=n, coords=coords,
        collision=collisionModel, storage=storage, semantics=semantics, walled=walled,
        force_mode=force_mode, force_uniform=force_uniform, force_field=force_field,
        dt=dt, dx=dx, tau=float(relaxationTime), tau_minus=float(tau_minus),
        viscosity=float(viscosity), kinetic=float(kin) if not isinstance(kin, str) else float("nan"),
        rho0=float(rho0), latticeSpeed=latticeSpeed, schemes=schemes, init_rho=init_rho, init_u=init_u,
        fluidParams=fluidParams, incompressible=incompressible, wall_region=wall_region,
        wall_velocity=wall_velocity)


def _grid_points(coords, shp):
    """The node coordinates of the grid in Julia's order (x fastest)."""
    grids = np.meshgrid(*coords, indexing="ij")
    flat = [g.ravel(order="F") for g in grids]
    return [tuple(float(c[i]) for c in flat) for i in range(int(np.prod(shp)))]


def sphere_count(rm: ResolvedModel, radius: float) -> int:
    """sum(getSphere(radius, X)): nodes with X_1^2 + X_2^2 [+ X_3^2] < r^2
    around the coordinate origin (gpu/initMethods.jl:108, kernels.jl:335-353)."""
    r2 = radius * radius if isinstance(radius, float) else radius ** 2
    x = rm.coords[0]
    s = (x * x)[:, None] + (rm.coords[1] * rm.coords[1])[None, :]
    if rm.dims == 2:
        return int(np.count_nonzero(s < r2))
    z = rm.coords[2]
    return int(sum(np.count_nonzero(s + zk * zk < r2) for zk in z))


def resolve_particle(rm: ResolvedModel, *, kind: str, massDensity=1.0, radius=0.1, position=DEFAULT,
                     velocity=DEFAULT, angularVelocity=DEFAULT, slipSpeed=DEFAULT, swimmingSpeed=DEFAULT,
                     swimmingDirection=DEFAULT, B1=DEFAULT, B2=DEFAULT, beta=DEFAULT,
                     coupleTorques=False, coupleForces=False, scheme=DEFAULT) -> _capi.LbmParticle:
    """addSquirmer!/addBead! parameter resolution (gpu/initMethods.jl:7-222)."""
    d = rm.dims
    mass = massDensity * sphere_count(rm, radius) * rm.dx ** d
    position = np.zeros(d) if position is DEFAULT else np.asarray(position, dtype=np.float64)
    velocity = np.zeros(d) if velocity is DEFAULT else np.asarray(velocity, dtype=np.float64)
    if len(position) != len(velocity) or len(position) != d:
        raise ValueError(f"The position and velocity dimensions must match the dimensionality of the model! dims = {d}")
    if d == 2:
        moi = 0.5 * mass * radius ** 2
        if angularVelocity is DEFAULT:
            angularVelocity = 0.0
        assert np.ndim(angularVelocity) == 0, "In two dimensions, angularVelocity must be a number!"
        omega = np.array([float(angularVelocity), 0.0, 0.0])
    else:
        moi = 0.4 * mass * radius ** 2
        omega = np.zeros(3) if angularVelocity is DEFAULT else np.asarray(angularVelocity, dtype=np.float64)
        assert omega.shape == (3,), "In three dimensions, angularVelocity must be an array!"
    scheme = "ladd" if scheme is DEFAULT else _sym(scheme)
    assert scheme in ("psm", "ladd"), f"{scheme} cannot be used as a particle-fluid collision scheme!"
    assert scheme == "ladd", f"{scheme} is not implemented in the gpu modules!"

    p = _capi.LbmParticle()
    p.kind = _capi.SQUIRMER if kind == "squirmer" else _capi.BEAD
    p.couple_forces = int(bool(coupleForces))
    p.couple_torques = int(bool(coupleTorques))
    p.radius = float(radius)
    for k in range(d):
        p.position[k] = position[k]
        p.velocity[k] = velocity[k]
    for k in range(3):
        p.angular_velocity[k] = omega[k]
    p.inverse_mass = 1 / mass
    p.inverse_moment_of_inertia = 1 / moi
    if kind != "squirmer":
        return p

    defaults = [x is DEFAULT for x in (B1, B2, beta)]
    assert any(defaults), "B1, B2, and beta cannot be all simultaneously defined!"
    if beta is DEFAULT:
        B1 = 1 if B1 is DEFAULT else B1
        B2 = 0 if B2 is DEFAULT else B2
        beta = B2 / B1
    elif B2 is DEFAULT:
        B1 = 1 if B1 is DEFAULT else B1
        B2 = B1 * beta
    else:
        assert beta != 0, "if beta = 0, then B2 can only be zero! In this case, do not declare both."
        B1 = B2 / beta
    if swimmingDirection is DEFAULT:
        swimmingDirection = np.zeros(d)
        swimmingDirection[0] = 1
    e = np.asarray(swimmingDirection, dtype=np.float64)
    assert len(e) == d, f"the swimming direction must be {d}-dimensional!"
    e = e / np.sqrt(np.sum(e * e))
    assert slipSpeed is DEFAULT or swimmingSpeed is DEFAULT, \
        "slipSpeed and swimmingSpeed cannot be all simultaneously defined!"
    if slipSpeed is DEFAULT:
        swimmingSpeed = 1e-3 if swimmingSpeed is DEFAULT else swimmingSpeed
        e = e * np.sign(swimmingSpeed)
        B1 = abs(3 / 2 * swimmingSpeed)
        B2 = B1 * beta
    else:
        thetas = np.linspace(-math.pi, math.pi, 100)
        norm = max(abs((B1 + B2 * math.cos(t)) * math.sin(t)) for t in thetas)
        B1 = B1 * (slipSpeed / norm)
        B2 = B2 * (slipSpeed / norm)
    p.B1, p.B2 = float(B1), float(B2)
    for k in range(d):
        p.swimming_direction[k] = e[k]
    return p


# ------------------------------------------------------------------ model
class _Particle:
    """Read-through view of one particle (position, velocity, ...)."""

    def __init__(self, model: "LBMmodel", pid: int):
        self._m, self.id = model, pid

    def _state(self) -> _capi.LbmParticle:
        p = _capi.LbmParticle()
        _capi.check(self._m._fn["lbm_get_particle"](self._m._h, self.id, C.byref(p)))
        return p

    @property
    def position(self):
        return np.array(self._state().position[: self._m.dims])

    @property
    def velocity(self):
        return np.array(self._state().velocity[: self._m.dims])

    @property
    def angularVelocity(self):
        w = self._state().angular_velocity
        return float(w[0]) if self._m.dims == 2 else np.array(w[:3])

    @property
    def swimmingDirection(self):
        return np.array(self._state().swimming_direction[: self._m.dims])

    @property
    def particleParams(self) -> dict:
        p = self._state()
        return dict(radius=p.radius, B1=p.B1, B2=p.B2, inverseMass=p.inverse_mass,
                    inverseMomentOfInertia=p.inverse_moment_of_inertia,
                    swimmingDirection=np.array(p.swimming_direction[: self._m.dims]),
                    coupleForces=bool(p.couple_forces), coupleTorques=bool(p.couple_torques),
                    properties=["spherical", "squirmer" if p.kind == _capi.SQUIRMER else "bead"])

    def load(self):
        """(forceInput, torqueInput) consumed by the last particle move."""
        F = np.zeros(3)
        T = np.zeros(3)
        _capi.check(self._m._fn["lbm_get_particle_load"](self._m._h, self.id, _capi.as_dptr(F), _capi.as_dptr(T)))
        return F[: self._m.dims], (T[0] if self._m.dims == 2 else T)

    def links(self):
        """Boundary links (node linear index, direction j) in findall order."""
        fn, h = self._m._fn, self._m._h
        cnt = C.c_int64(0)
        _capi.check(fn["lbm_get_links"](h, self.id, None, None, 0, C.byref(cnt)))
        node = np.zeros(cnt.value, dtype=np.int64)
        dirs = np.zeros(cnt.value, dtype=np.int32)
        _capi.check(fn["lbm_get_links"](h, self.id, node.ctypes.data_as(C.POINTER(C.c_int64)),
                                        dirs.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)))
        return node, dirs


def set_walls(fn: dict, h, rm: ResolvedModel) -> None:
    """The `walls` region and the `movingWalls` velocities of a resolved
    model, through lbm_set_walls / lbm_set_wall_velocity (before the
    populations)."""
    if rm.wall_region is not None:
        w = np.asfortranarray(rm.wall_region, dtype=np.uint8)
        _capi.check(fn["lbm_set_walls"](h, w.ctypes.data_as(C.POINTER(C.c_uint8))), fn)
    if rm.wall_velocity is not None:
        u = np.asfortranarray(rm.wall_velocity, dtype=np.float64)
        _capi.check(fn["lbm_set_wall_velocity"](h, _capi.as_dptr(u)), fn)


# ticks per uploaded chunk of a time-dependent magnetic field (24 B a tick)
SCHEDULE_CHUNK = 1 << 16


class LBMmodel:
    """A lattice on one GPU; mirrors the fields of the reference's LBMmodel
    (gpu/structs.jl:207-231) as read-through properties."""

    def __init__(self, rm: ResolvedModel):
        self._fn = _capi.load()
        self.resolved = rm
        cfg, keep = rm.config()
        h = C.c_void_p()
        _capi.check(self._fn["lbm_create"](C.byref(cfg), C.byref(h)))
        self._h = h
        self.particles: list[_Particle] = []
        self.particleInteractions: list = []
        self.schemes = list(rm.schemes)
        set_walls(self._fn, h, rm)
        if rm.force_mode == _capi.FORCE_FIELD:
            F = np.asfortranarray(rm.force_field, dtype=np.float64)
            _capi.check(self._fn["lbm_set_force_field"](h, _capi.as_dptr(F)))
        if rm.uniform_init():
            u = np.zeros(3)
            u[: rm.dims] = rm.init_u
            _capi.check(self._fn["lbm_init_uniform"](h, float(rm.init_rho), _capi.as_dptr(u)))
        else:
            f0 = np.asfortranarray(rm.initial_populations())
            _capi.check(self._fn["lbm_set_populations"](h, _capi.as_dptr(f0)))
        if "saveData" in self.schemes:
            from . import savedata
            savedata.start(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._fn["lbm_destroy"](h)
            except Exception:
                pass
            self._h = None

    # --- reference fields ----------------------------------------------
    @property
    def dims(self) -> int:
        return self.resolved.dims

    @property
    def spaceTime(self) -> dict:
        rm = self.resolved
        return dict(coordinates=rm.coords, latticeParameter=rm.dx, timeStep=rm.dt,
                    latticeSpeed=rm.latticeSpeed, dims=rm.dims)

    @property
    def fluidParams(self) -> dict:
        return dict(self.resolved.fluidParams)

    @property
    def tick(self) -> int:
        t = C.c_int64()
        _capi.check(self._fn["lbm_get_tick"](self._h, C.byref(t), None))
        return t.value

    @property
    def time(self) -> float:
        t = C.c_double()
        _capi.check(self._fn["lbm_get_tick"](self._h, None, C.byref(t)))
        return t.value

    @property
    def distributions(self) -> np.ndarray:
        rm = self.resolved
        f = np.zeros(tuple(rm.n[: rm.dims]) + (rm.Q,), order="F")
        _capi.check(self._fn["lbm_get_populations"](self._h, _capi.as_dptr(f)))
        return f

    @distributions.setter
    def distributions(self, f) -> None:
        f = _capi.fortran_f64(f)
        rm = self.resolved
        assert f.shape == tuple(rm.n[: rm.dims]) + (rm.Q,)
        _capi.check(self._fn["lbm_set_populations"](self._h, _capi.as_dptr(f)))

    def _hydro(self):
        rm = self.resolved
        shp = tuple(rm.n[: rm.dims])
        rho = np.zeros(shp, order="F")
        u = np.zeros(shp + (rm.dims,), order="F")
        _capi.check(self._fn["lbm_get_hydro"](self._h, _capi.as_dptr(rho), _capi.as_dptr(u)))
        return rho, u

    @property
    def massDensity(self) -> np.ndarray:
        return self._hydro()[0]

    @property
    def fluidVelocity(self) -> np.ndarray:
        return self._hydro()[1]

    # --- state snapshots (no counterpart in the reference) -----------------
    def save_state(self, path=None) -> np.ndarray:
        """The state a run continues from: the populations in the storage's
        own encoding, every particle's state record, the tick and the time,
        as a byte array (also written to `path` when given).  `load_state`
        on a model built the same way resumes from it bit for bit."""
        n = C.c_int64()
        _capi.check(self._fn["lbm_state_size"](self._h, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _capi.check(self._fn["lbm_save_state"](self._h, buf.ctypes.data_as(C.c_void_p), n.value))
        if path is not None:
            buf.tofile(path)
        return buf

    def load_state(self, src) -> None:
        """Resume from a `save_state` snapshot: a byte array or a file path."""
        if isinstance(src, (str, os.PathLike)):
            buf = np.fromfile(src, np.uint8)
        else:
            buf = np.ascontiguousarray(np.frombuffer(src, np.uint8))
        _capi.check(self._fn["lbm_load_state"](self._h, buf.ctypes.data_as(C.c_void_p), buf.size))

    # --- backend controls ------------------------------------------------
    def set_block(self, bx: int, by: int, bz: int = 1) -> None:
        _capi.check(self._fn["lbm_set_block"](self._h, bx, by, bz))

    def get_block(self):
        b = [C.c_int32() for _ in range(3)]
        _capi.check(self._fn["lbm_get_block"](self._h, *[C.byref(x) for x in b]))
        return tuple(x.value for x in b)

    def set_tile_order(self, xcd_remap: int) -> None:
        _capi.check(self._fn["lbm_set_tile_order"](self._h, int(xcd_remap)))

    def get_tile_order(self) -> int:
        v = C.c_int32()
        _capi.check(self._fn["lbm_get_tile_order"](self._h, C.byref(v)))
        return v.value

    def set_occupancy(self, low: int) -> None:
        """1: the gathering kernels built for half the waves per SIMD and
        twice the registers (faster box path); 0: the default."""
        _capi.check(self._fn["lbm_set_occupancy"](self._h, int(low)))

    def get_occupancy(self) -> int:
        v = C.c_int32()
        _capi.check(self._fn["lbm_get_occupancy"](self._h, C.byref(v)))
        return v.value

    def autotune(self, reps: int = 20):
        """Pick the fastest block shape of the stream-collide kernel."""
        best = (C.c_int32 * 3)()
        ms = C.c_double()
        _capi.check(self._fn["lbm_autotune"](self._h, int(reps), best, C.byref(ms)))
        return tuple(best), ms.value

    def set_graph(self, ticks_per_graph: int) -> None:
        """Ticks per captured hipGraph (even), or 0 for direct launches."""
        _capi.check(self._fn["lbm_set_graph"](self._h, int(ticks_per_graph)))

    def sync(self) -> None:
        _capi.check(self._fn["lbm_sync"](self._h))

    def check(self):
        """(non-finite populations, populations at the storage's clamp) of
        the fluid nodes."""
        bad, sat = C.c_int64(), C.c_int64()
        _capi.check(self._fn["lbm_check_populations"](self._h, C.byref(bad), C.byref(sat)))
        return bad.value, sat.value

    def run(self, ticks: int) -> None:
        ticks = int(ticks)
        if getattr(self, "_field_fn", None) is None:
            _capi.check(self._fn["lbm_run"](self._h, ticks))
            return
        # a time-dependent magnetic field: its values for the ticks of each
        # chunk go to the device before the chunk runs
        done = 0
        while done < ticks:
            n = min(ticks - done, SCHEDULE_CHUNK)
            self._upload_schedule(n)
            _capi.check(self._fn["lbm_run"](self._h, n))
            done += n

    def _upload_schedule(self, n: int) -> None:
        from .interactions import field_schedule
        t = C.c_int64()
        tm = C.c_double()
        _capi.check(self._fn["lbm_get_tick"](self._h, C.byref(t), C.byref(tm)))
        B = np.ascontiguousarray(field_schedule(self._field_fn, tm.value, self.resolved.dt, n, self.dims))
        _capi.check(self._fn["lbm_set_field_schedule"](self._h, _capi.as_dptr(B), t.value + 1, n))

    def _set_interactions(self, terms, n: int) -> None:
        _capi.check(self._fn["lbm_set_interactions"](self._h, terms, int(n)))

    def set_timing(self, on: bool) -> None:
        _capi.check(self._fn["lbm_set_timing"](self._h, int(bool(on))))

    def last_run_times(self):
        k, s = C.c_double(), C.c_double()
        _capi.check(self._fn["lbm_last_run_times"](self._h, C.byref(k), C.byref(s)))
        return k.value, s.value


class LBMbatch:
    """Independent models (one lattice, storage, collision and force kind;
    any sizes, parameters and particles) advanced together: every kernel of a
    run covers all of them in one launch.  Each model ends where
    LBMpropagate(model, ticks=n) would leave it.  No counterpart in the
    reference, where a parameter sweep loops LBMpropagate! over models
    (gpu/fluids.jl:228-264)."""

    def __init__(self, models):
        self._fn = _capi.load()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._h.value for m in self.models])
        b = C.c_void_p()
        _capi.check(self._fn["lbm_batch_create"](arr, len(self.models), C.byref(b)))
        self._b = b

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            try:
                self._fn["lbm_batch_destroy"](b)
            except Exception:
                pass
            self._b = None

    def run(self, ticks: int) -> None:
        for m in self.models:
            if getattr(m, "_field_fn", None) is not None:
                m._upload_schedule(int(ticks))
        _capi.check(self._fn["lbm_batch_run"](self._b, int(ticks)))

    def sync(self) -> None:
        _capi.check(self._fn["lbm_batch_sync"](self._b))

    def set_block(self, bx: int, by: int, bz: int = 1) -> None:
        _capi.check(self._fn["lbm_batch_set_block"](self._b, bx, by, bz))

    def get_block(self):
        b = [C.c_int32() for _ in range(3)]
        _capi.check(self._fn["lbm_batch_get_block"](self._b, *[C.byref(x) for x in b]))
        return tuple(x.value for x in b)

    def set_row_tiles(self, threads: int) -> None:
        """Blocks of `threads` threads; every box a tile shape of that many
        threads spanning whole rows of the box."""
        _capi.check(self._fn["lbm_batch_set_row_tiles"](self._b, int(threads)))

    def get_row_tiles(self) -> int:
        """Threads of the row tiles, 0 while the boxes share one block shape."""
        v = C.c_int32()
        _capi.check(self._fn["lbm_batch_get_row_tiles"](self._b, C.byref(v)))
        return v.value

    def set_tile_order(self, xcd_remap: int) -> None:
        _capi.check(self._fn["lbm_batch_set_tile_order"](self._b, int(xcd_remap)))

    def get_tile_order(self) -> int:
        v = C.c_int32()
        _capi.check(self._fn["lbm_batch_get_tile_order"](self._b, C.byref(v)))
        return v.value

    def set_occupancy(self, low: int) -> None:
        _capi.check(self._fn["lbm_batch_set_occupancy"](self._b, int(low)))

    def get_occupancy(self) -> int:
        v = C.c_int32()
        _capi.check(self._fn["lbm_batch_get_occupancy"](self._b, C.byref(v)))
        return v.value

    def autotune(self, reps: int = 10):
        """Pick the fastest block shape of the batched stream-collide kernel:
        returns (shape, ms) with shape (threads, 0, 0) for row tiles."""
        best = (C.c_int32 * 3)()
        ms = C.c_double()
        _capi.check(self._fn["lbm_batch_autotune"](self._b, int(reps), best, C.byref(ms)))
        return tuple(best), ms.value

    def set_graph(self, ticks_per_graph: int) -> None:
        _capi.check(self._fn["lbm_batch_set_graph"](self._b, int(ticks_per_graph)))

    def set_timing(self, on: bool) -> None:
        _capi.check(self._fn["lbm_batch_set_timing"](self._b, int(bool(on))))

    def last_run_times(self):
        k, s = C.c_double(), C.c_double()
        _capi.check(self._fn["lbm_batch_last_run_times"](self._b, C.byref(k), C.byref(s)))
        return k.value, s.value


def modelInit(**kw) -> LBMmodel:
    """modelInit (gpu/initMethods.jl:440-702) on the hip backend."""
    return LBMmodel(resolve_model(**kw))


def _add(model: LBMmodel, kind: str, kw) -> None:
    p = resolve_particle(model.resolved, kind=kind, **kw)
    pid = C.c_int32()
    _capi.check(model._fn["lbm_add_particle"](model._h, C.byref(p), C.byref(pid)))
    model.particles.append(_Particle(model, pid.value))
    for s in ("ladd", "bounceBack"):
        if s not in model.schemes:
            model.schemes.append(s)
    if "saveData" in model.schemes:
        from . import savedata
        savedata.write_particle(model, pid.value)


def addSquirmer(model: LBMmodel, **kw) -> None:
    """addSquirmer! (gpu/initMethods.jl:91-222)."""
    _add(model, "squirmer", kw)


def addBead(model: LBMmodel, **kw) -> None:
    """addBead! (gpu/initMethods.jl:7-89)."""
    _add(model, "bead", kw)


def hydroVariablesUpdate(model: LBMmodel) -> None:
    """hydroVariablesUpdate! (gpu/fluids.jl:32-37).  The hip backend derives
    rho and u from the populations whenever they are read, so there is no
    cached state to refresh; kept for API parity."""
    return None


def propagate_ticks(dt: float, simulationTime=DEFAULT, ticks=DEFAULT) -> int:
    """Number of ticks LBMpropagate! runs (gpu/fluids.jl:231-241)."""
    assert simulationTime is DEFAULT or ticks is DEFAULT, \
        "simulationTime and ticks cannot be simultaneously chosen, as the time step is defined already in the model!"
    if simulationTime is DEFAULT and ticks is DEFAULT:
        return 100
    if ticks is DEFAULT:
        return max(0, math.floor(_rat(simulationTime) / _rat(dt) - 1) + 1)
    return int(ticks)


def LBMpropagate(model, *, simulationTime=DEFAULT, ticks=DEFAULT, verbose=False,
                 ticksSaved=DEFAULT) -> None:
    """LBMpropagate! (gpu/fluids.jl:228-264): `ticks` calls of tick!.  Given
    an LBMbatch, every model of it advances (simulationTime is then read in
    the time step of the batch's first model, which all must share)."""
    if isinstance(model, LBMbatch):
        if any("saveData" in m.schemes for m in model.models):
            raise NotImplementedError("saveData runs one model at a time: call LBMpropagate on the model")
        dts = {m.resolved.dt for m in model.models}
        if simulationTime is not DEFAULT and len(dts) != 1:
            raise ValueError("simulationTime needs one time step across the batch; give ticks instead")
        n = propagate_ticks(model.models[0].resolved.dt, simulationTime, ticks)
        model.run(n)
        model.sync()
        for m in model.models:
            guard(m)
        if verbose:
            print(f"\r t = {round(model.models[0].time, 2)}")
        return
    n = propagate_ticks(model.resolved.dt, simulationTime, ticks)
    if "saveData" in model.schemes:
        from . import savedata
        savedata.propagate(model, n, None if ticksSaved is DEFAULT else ticksSaved)
    else:
        model.run(n)
        model.sync()
    guard(model)
    if verbose:
        print(f"\r t = {round(model.time, 2)}")


def guard(model) -> None:
    """After a run: a NaN or infinite population raises FloatingPointError
    (the run went unstable, e.g. tau <= 1/2 or flow beyond the lattice's
    range); populations held at the f16c storage's clamp warn.  The
    reference checks nothing while it runs."""
    bad, sat = model.check()
    if bad:
        raise FloatingPointError(f"{bad} non-finite populations after tick {model.tick}: the run went unstable")
    if sat:
        import warnings
        warnings.warn(f"{sat} populations at the f16c storage's clamp after tick {model.tick}: deviations beyond "
                      "w/8 are not representable; use f32 storage for this flow", RuntimeWarning)
