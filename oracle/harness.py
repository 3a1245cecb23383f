"""ctypes driver of oracle/liblbm_oracle.so (and of the dual-number
restatement oracle/liblbm_oracle_dual.so) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker.  It takes the same resolved model as the hip backend
(lbmengine_hip.resolve_model / resolve_particle, the host-side restatement
of modelInit and addSquirmer!) and exposes the same read accessors as
lbmengine_hip.LBMmodel, so a parity test builds both from one set of
keyword arguments and compares field by field.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liblbm_oracle.so"
sys.path.insert(0, str(ROOT / "lbmengine.jl_amd"))

from lbmengine_hip import _capi  # noqa: E402
from lbmengine_hip.model import (DEFAULT, ResolvedModel, propagate_ticks, resolve_model,  # noqa: E402
                                 resolve_particle, set_walls)

_LIB = None
_FN: dict = {}
_FIELD_CB = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.c_void_p)


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_LIB


def load():
    global _LIB, _FN
    if _LIB is None:
        if not ORACLE_LIB.exists():
            build_oracle()
        _LIB = C.CDLL(str(ORACLE_LIB))
        _FN = _capi.bind(_LIB, "lbm_", "lbmo_")
        fn = _LIB.lbmo_get_solid
        fn.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint8)]
        fn.restype = C.c_int
        _FN["lbm_get_solid"] = fn
        _LIB.lbmo_set_field_fn.argtypes = [C.c_void_p, _FIELD_CB, C.c_void_p]
        _LIB.lbmo_set_field_fn.restype = C.c_int
    return _FN


def _check(rc):
    if rc != 0:
        raise _capi.LBMError(rc, _FN["lbm_last_error"]().decode())


class _OParticle:
    def __init__(self, m: "OracleModel", pid: int):
        self._m, self.id = m, pid

    def _state(self):
        p = _capi.LbmParticle()
        _check(self._m._fn["lbm_get_particle"](self._m._h, self.id, C.byref(p)))
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
        return dict(radius=self._state().radius)

    def load(self):
        F, T = np.zeros(3), np.zeros(3)
        _check(self._m._fn["lbm_get_particle_load"](self._m._h, self.id, _capi.as_dptr(F), _capi.as_dptr(T)))
        return F[: self._m.dims], (T[0] if self._m.dims == 2 else T)

    def links(self):
        fn, h = self._m._fn, self._m._h
        cnt = C.c_int64(0)
        _check(fn["lbm_get_links"](h, self.id, None, None, 0, C.byref(cnt)))
        node = np.zeros(cnt.value, dtype=np.int64)
        dirs = np.zeros(cnt.value, dtype=np.int32)
        _check(fn["lbm_get_links"](h, self.id, node.ctypes.data_as(C.POINTER(C.c_int64)),
                                   dirs.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)))
        return node, dirs

    def solid(self):
        rm = self._m.resolved
        out = np.zeros(rm.N, dtype=np.uint8)
        _check(self._m._fn["lbm_get_solid"](self._m._h, self.id, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out.reshape(tuple(rm.n[: rm.dims]), order="F").astype(bool)


class OracleModel:
    """The oracle run of a resolved model (FP64, CPU)."""

    def __init__(self, rm: ResolvedModel):
        self._fn = load()
        self.resolved = rm
        cfg, keep = rm.config()
        h = C.c_void_p()
        _check(self._fn["lbm_create"](C.byref(cfg), C.byref(h)))
        self._h = h
        self.particles = []
        self.particleInteractions = []
        set_walls(self._fn, h, rm)
        if rm.force_mode == _capi.FORCE_FIELD:
            F = np.asfortranarray(rm.force_field, dtype=np.float64)
            _check(self._fn["lbm_set_force_field"](h, _capi.as_dptr(F)))
        if rm.uniform_init():
            u = np.zeros(3)
            u[: rm.dims] = rm.init_u
            _check(self._fn["lbm_init_uniform"](h, float(rm.init_rho), _capi.as_dptr(u)))
        else:
            f0 = np.asfortranarray(rm.initial_populations())
            _check(self._fn["lbm_set_populations"](h, _capi.as_dptr(f0)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._fn["lbm_destroy"](h)
            self._h = None

    @property
    def dims(self):
        return self.resolved.dims

    @property
    def spaceTime(self) -> dict:
        rm = self.resolved
        return dict(latticeParameter=rm.dx, timeStep=rm.dt, dims=rm.dims)

    def _set_interactions(self, terms, n: int) -> None:
        _check(self._fn["lbm_set_interactions"](self._h, terms, int(n)))

    @property
    def _field_fn(self):
        return getattr(self, "_field_py", None)

    @_field_fn.setter
    def _field_fn(self, fn) -> None:
        """A time-dependent magnetic field: the oracle calls back with
        model.time at every tick's molecularDynamicsTick! and receives
        magneticField(model.time) |> cu (gpu/initMethods.jl:339,
        gpu/particles.jl:161): a number as it is, a Float64 vector rounded to
        Float32."""
        self._field_py = fn
        if fn is None:
            self._field_cb = None
            _check(_LIB.lbmo_set_field_fn(self._h, C.cast(None, _FIELD_CB), None))
            return

        def cb(t, out, _ctx):
            v = fn(t)
            if np.ndim(v) == 0:
                out[0], out[1], out[2] = float(v), 0.0, 0.0
                return
            for k, x in enumerate(list(v)):
                out[k] = float(np.float32(x)) if isinstance(x, (float, np.floating)) else float(x)

        self._field_cb = _FIELD_CB(cb)
        _check(_LIB.lbmo_set_field_fn(self._h, self._field_cb, None))

    def add(self, kind: str, **kw):
        p = resolve_particle(self.resolved, kind=kind, **kw)
        pid = C.c_int32()
        _check(self._fn["lbm_add_particle"](self._h, C.byref(p), C.byref(pid)))
        self.particles.append(_OParticle(self, pid.value))

    def run(self, ticks: int):
        _check(self._fn["lbm_run"](self._h, int(ticks)))

    @property
    def tick(self):
        t, tm = C.c_int64(), C.c_double()
        _check(self._fn["lbm_get_tick"](self._h, C.byref(t), C.byref(tm)))
        return t.value

    @property
    def time(self):
        t, tm = C.c_int64(), C.c_double()
        _check(self._fn["lbm_get_tick"](self._h, C.byref(t), C.byref(tm)))
        return tm.value

    @property
    def distributions(self):
        rm = self.resolved
        f = np.zeros(tuple(rm.n[: rm.dims]) + (rm.Q,), order="F")
        _check(self._fn["lbm_get_populations"](self._h, _capi.as_dptr(f)))
        return f

    @distributions.setter
    def distributions(self, f):
        f = _capi.fortran_f64(f)
        _check(self._fn["lbm_set_populations"](self._h, _capi.as_dptr(f)))

    def _hydro(self):
        rm = self.resolved
        shp = tuple(rm.n[: rm.dims])
        rho = np.zeros(shp, order="F")
        u = np.zeros(shp + (rm.dims,), order="F")
        _check(self._fn["lbm_get_hydro"](self._h, _capi.as_dptr(rho), _capi.as_dptr(u)))
        return rho, u

    @property
    def massDensity(self):
        return self._hydro()[0]

    @property
    def fluidVelocity(self):
        return self._hydro()[1]

    def viscous_stress(self):
        """sigma as (N..., d, d), column major."""
        rm = self.resolved
        out = np.zeros(tuple(rm.n[: rm.dims]) + (rm.dims, rm.dims), order="F")
        _check(self._fn["lbm_viscous_stress"](self._h, _capi.as_dptr(out)))
        return out


def oracle_model(**kw) -> OracleModel:
    return OracleModel(resolve_model(**kw))


ORACLE_DUAL_LIB = ORACLE_DIR / "liblbm_oracle_dual.so"
_DUAL = None


def load_dual():
    global _DUAL
    if _DUAL is None:
        if not ORACLE_DUAL_LIB.exists():
            build_oracle()
        _DUAL = C.CDLL(str(ORACLE_DUAL_LIB))
        fn = _DUAL.lbmd_loss_jvp
        dp = C.POINTER(C.c_double)
        fn.argtypes = [C.POINTER(_capi.LbmConfig), dp, dp, C.c_int32, C.POINTER(_capi.LbmParticle), dp, dp, dp,
                       C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_double, dp, dp, dp]
        fn.restype = C.c_int
    return _DUAL


def dual_loss_jvp(rm: ResolvedModel, particles, f0, *, ticks: int, window: int, F=None, dB1=None, dB2=None,
                  dF=None, swimmer: int = 0, target: float = 1e-4, penalty: float = 0.0):
    """Loss (mean swimmer speed over the last `window` ticks - target)^2 +
    penalty sum F^2 from populations f0 and its derivative along (dB1, dB2,
    dF), by the dual-number restatement of the oracle's tick.  Returns
    (loss, dloss, speeds)."""
    lib = load_dual()
    cfg, keep = rm.config()
    np_ = len(particles)
    parr = (_capi.LbmParticle * np_)(*particles)
    f0 = np.asfortranarray(f0, dtype=np.float64)

    def opt(a, n):
        if a is None:
            return None
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        assert a.size == n
        keep.append(a)
        return _capi.as_dptr(a)

    d = rm.dims
    loss, dl = C.c_double(), C.c_double()
    speeds = np.zeros(window)
    rc = lib.lbmd_loss_jvp(C.byref(cfg), _capi.as_dptr(f0), opt(F, rm.N * d), np_, parr, opt(dB1, np_),
                           opt(dB2, np_), opt(dF, rm.N * d), int(ticks), int(window), int(swimmer), float(target),
                           float(penalty), C.byref(loss), C.byref(dl), _capi.as_dptr(speeds))
    if rc != 0:
        raise _capi.LBMError(rc, "lbmd_loss_jvp rejected its arguments")
    return loss.value, dl.value, speeds


__all__ = ["OracleModel", "oracle_model", "build_oracle", "dual_loss_jvp", "DEFAULT", "propagate_ticks"]
