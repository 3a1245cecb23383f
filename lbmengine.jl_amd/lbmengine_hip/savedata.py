"""saveData: the reference's output files (gpu/aux.jl:271-362) written while
the step runs on the device.

modelInit(saveData=true) empties output.lbm/ in the working directory and
writes fluidTrj_<tick>.csv of the initial state (gpu/initMethods.jl:671-698);
each particle added writes its row of particlesTrj.csv
(gpu/initMethods.jl:86,219); LBMpropagate! writes fluidTrj_<tick>.csv at
`ticksSaved` checkpoint ticks and a particlesTrj.csv row per particle per
tick (gpu/fluids.jl:243-259).

The run is split at the checkpoint ticks only: between them the step kernels
run back to back (graph replays included) and every particle move also
stores the moved state in a device log, read back after each stretch.  The
fluid files are written by the library's native writer.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import shutil
from fractions import Fraction

import numpy as np

from . import _capi

OUTPUT_DIR = "output.lbm"
DEFAULT_TICKS_SAVED = 100
LOG_TICKS = 4096  # device log ring: ticks between two reads at most

_AX = ("x", "y", "z")


def julia_float(x: float) -> str:
    """A Float64 as Julia prints it: the shortest decimal that reads back to
    the same value, plain for 1e-4 <= |x| < 1e6 (whole numbers end in
    ".0"), d.ddde<exp> outside."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-Inf" if x < 0 else "Inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    x = abs(x)
    if x == 0.0:
        return sign + "0.0"
    # shortest round-trip digits and their decimal exponent
    mant, _, exp = np.format_float_scientific(x, unique=True, trim="-").partition("e")
    digits, e10 = mant.replace(".", ""), int(exp)
    nd, pt = len(digits), e10 + 1
    if -4 < pt <= 6:
        if pt <= 0:
            body = "0." + "0" * (-pt) + digits
        elif pt < nd:
            body = digits[:pt] + "." + digits[pt:]
        else:
            body = digits + "0" * (pt - nd) + ".0"
    else:
        body = digits[0] + "." + (digits[1:] or "0") + f"e{e10}"
    return sign + body


def _out(name: str) -> str:
    return os.path.join(OUTPUT_DIR, name)


def start(model) -> None:
    """modelInit's saveData branch (gpu/initMethods.jl:671-674, 698)."""
    shutil.rmtree(OUTPUT_DIR, ignore_errors=True)
    os.mkdir(OUTPUT_DIR)
    write_fluid(model)


def write_fluid(model) -> None:
    """writeTrajectories (gpu/aux.jl:271-317) of the model's current state."""
    rm = model.resolved
    d = rm.dims
    rho, u = model._hydro()
    f = model.distributions
    n = (C.c_int64 * 3)(*[int(rm.n[k]) if k < d else 1 for k in range(3)])
    coords = [np.ascontiguousarray(rm.coords[k], dtype=np.float64) for k in range(d)]
    cptr = (_capi._dp * 3)(*[c.ctypes.data_as(_capi._dp) for c in coords], *([None] * (3 - d)))
    path = _out(f"fluidTrj_{model.tick}.csv").encode()
    rho = np.asfortranarray(rho)
    u = np.asfortranarray(u)
    _capi.check(model._fn["lbm_write_fluid_csv"](path, d, n, model.tick, model.time, cptr, _capi.as_dptr(rho),
                                                 _capi.as_dptr(u), rm.Q, _capi.as_dptr(f)))


def writeTensor(model, T, name: str) -> None:
    """writeTensor (gpu/aux.jl:364-400): output.lbm/<name>.csv of a dims x
    dims tensor field T (T[a][b] an array over the lattice), with the node
    indices and coordinates."""
    rm = model.resolved
    d = rm.dims
    os.makedirs(OUTPUT_DIR, exist_ok=True)
    shp = tuple(rm.n[:d])
    t = np.zeros(shp + (d, d), order="F")
    for a in range(d):
        for b in range(d):
            assert np.shape(T[a][b]) == shp, f"tensor components must have the lattice's shape {shp}"
            t[..., a, b] = T[a][b]
    n = (C.c_int64 * 3)(*[int(rm.n[k]) if k < d else 1 for k in range(3)])
    coords = [np.ascontiguousarray(rm.coords[k], dtype=np.float64) for k in range(d)]
    cptr = (_capi._dp * 3)(*[c.ctypes.data_as(_capi._dp) for c in coords], *([None] * (3 - d)))
    _capi.check(model._fn["lbm_write_tensor_csv"](_out(f"{name}.csv").encode(), d, n, model.tick, model.time, cptr,
                                                  _capi.as_dptr(t)))


def _particle_header(d: int) -> str:
    cols = ["tick", "time", "particleId"]
    cols += [f"position_{a}" for a in _AX[:d]] + [f"velocity_{a}" for a in _AX[:d]]
    cols += ["angularVelocity"] if d == 2 else [f"angularVelocity_{a}" for a in _AX]
    return ",".join(cols) + "\n"


def _particle_row(d: int, tick: int, time: float, pid: int, R, V, W) -> str:
    vals = [str(int(tick)), julia_float(time), str(int(pid))]
    vals += [julia_float(R[k]) for k in range(d)] + [julia_float(V[k]) for k in range(d)]
    vals += [julia_float(W[0])] if d == 2 else [julia_float(W[k]) for k in range(3)]
    return ",".join(vals) + "\n"


def append_particle_rows(model, rows) -> None:
    """writeParticleTrajectory's CSV.write(...; append) (gpu/aux.jl:349-355):
    the header only when the file is new."""
    path = _out("particlesTrj.csv")
    new = not os.path.isfile(path)
    with open(path, "a") as fh:
        if new:
            fh.write(_particle_header(model.dims))
        fh.writelines(rows)


def write_particle(model, pid: int) -> None:
    """writeParticleTrajectory (gpu/aux.jl:319-356) of particle pid (0-based)."""
    p = model.particles[pid]
    W = p.angularVelocity
    W = [W] if model.dims == 2 else W
    append_particle_rows(model, [_particle_row(model.dims, model.tick, model.time, pid + 1, p.position,
                                               p.velocity, W)])


def checkpoints(t0: int, ticks: int, dt: float, ticks_saved) -> set:
    """The ticks LBMpropagate! writes fluidTrj files at (gpu/fluids.jl:243-247):
    round.(range(tick, stop = tick + floor(time[end] / dt), length = ticksSaved)),
    ties to even; time[end] is ticks * dt with dt read as the decimal it
    was written as (Julia's float ranges), rounded once."""
    from .model import _rat
    ts = DEFAULT_TICKS_SAVED if ticks_saved is None else int(ticks_saved)
    t_end = float(_rat(dt) * ticks)
    total = math.floor(t_end / dt)
    if ts < 1:
        raise ValueError("ticksSaved must be at least 1")
    if ts == 1:
        if total != 0:
            raise ValueError(f"range({float(t0)}, stop={float(t0 + total)}, length=1): endpoints differ")
        return {t0}
    step = Fraction(total, ts - 1)
    return {t0 + round(i * step) for i in range(ts)}


def propagate(model, ticks: int, ticks_saved=None) -> None:
    """LBMpropagate! with :saveData (gpu/fluids.jl:243-259)."""
    t0 = model.tick
    ckpt = checkpoints(t0, ticks, model.resolved.dt, ticks_saved)
    np_ = len(model.particles)
    d = model.dims
    if np_:
        _capi.check(model._fn["lbm_set_trajectory_log"](model._h, min(max(ticks, 1), LOG_TICKS)))
    try:
        done = 0
        while done < ticks:
            a = t0 + done
            b = min(t0 + ticks, a + LOG_TICKS) if np_ else t0 + ticks
            nxt = [c for c in ckpt if a < c <= b]
            if nxt:
                b = min(nxt)
            time_a = model.time
            model.run(b - a)
            if np_:
                logs = []
                for p in range(np_):
                    buf = np.zeros((b - a, 9))
                    _capi.check(model._fn["lbm_get_trajectory"](model._h, p, a + 1, b - a, _capi.as_dptr(buf)))
                    logs.append(buf)
                rows, t = [], time_a
                for i in range(b - a):
                    t = t + model.resolved.dt
                    for p in range(np_):
                        e = logs[p][i]
                        rows.append(_particle_row(d, a + 1 + i, t, p + 1, e[0:3], e[3:6], e[6:9]))
                append_particle_rows(model, rows)
            else:
                model.sync()
            if b in ckpt:
                write_fluid(model)
            done = b - t0
    finally:
        if np_:
            _capi.check(model._fn["lbm_set_trajectory_log"](model._h, 0))
