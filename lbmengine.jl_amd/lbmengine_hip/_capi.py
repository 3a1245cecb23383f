"""ctypes binding of liblbm_hip.so (include/lbm_hip.h).

The structs mirror `lbm_config` and `lbm_particle`; every call is checked
and a failing one raises LBMError with the library's message.  Loading
fails loudly when the library has not been built: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LBM_OK = 0
LBM_EINVAL, LBM_ENOMEM, LBM_EDEVICE, LBM_ESTATE, LBM_ERANGE = -1, -2, -3, -4, -5

COLLISION = {"bgk": 0, "trt": 1}
STORAGE = {"f64": 0, "f32": 1, "f16": 2, "f16c": 3}
SEMANTICS = {"gpu": 0, "cpu": 1}
FORCE_NONE, FORCE_UNIFORM, FORCE_FIELD = 0, 1, 2
FLAG_INCOMPRESSIBLE = 1
BEAD, SQUIRMER = 0, 1
MAX_PARTICLES = 64
LOCAL_PARTICLES = 8  # particles within reach of one box block; also the adjoint's, a batch box's and lbm_run_noalloc's limit
INTERACTION = {"linear": 0, "bistable": 1, "polar": 2, "dipole": 3, "lennardJones": 4}
MAX_INTERACTIONS = 256

LIB_PATH = Path(os.environ.get("LBM_HIP_LIB", Path(__file__).resolve().parent.parent / "liblbm_hip.so"))


class LBMError(RuntimeError):
    """Raised when a liblbm_hip call returns a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class LbmConfig(C.Structure):
    _fields_ = [
        ("dims", C.c_int32), ("q", C.c_int32), ("n", C.c_int64 * 3),
        ("collision", C.c_int32), ("storage", C.c_int32), ("semantics", C.c_int32),
        ("walled", C.c_int32), ("force", C.c_int32), ("flags", C.c_int32),
        ("force_uniform", C.c_double * 3),
        ("dt", C.c_double), ("dx", C.c_double), ("tau", C.c_double), ("tau_minus", C.c_double),
        ("rho0", C.c_double),
        ("coords", C.POINTER(C.c_double) * 3),
    ]


class LbmParticle(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("couple_forces", C.c_int32), ("couple_torques", C.c_int32),
        ("reserved0", C.c_int32),
        ("radius", C.c_double),
        ("position", C.c_double * 3), ("velocity", C.c_double * 3),
        ("angular_velocity", C.c_double * 3), ("swimming_direction", C.c_double * 3),
        ("B1", C.c_double), ("B2", C.c_double),
        ("inverse_mass", C.c_double), ("inverse_moment_of_inertia", C.c_double),
    ]


class LbmInteraction(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("ids", C.c_int32 * 3),
        ("params", C.c_double * 4), ("field", C.c_double * 3),
    ]


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "lbm_last_error": [],
    "lbm_version": [],
    "lbm_device_count": [C.POINTER(C.c_int32)],
    "lbm_create": [C.POINTER(LbmConfig), C.POINTER(_vp)],
    "lbm_destroy": [_vp],
    "lbm_init_uniform": [_vp, C.c_double, _dp],
    "lbm_set_force_field": [_vp, _dp],
    "lbm_set_walls": [_vp, C.POINTER(C.c_uint8)],
    "lbm_set_wall_velocity": [_vp, _dp],
    "lbm_set_populations": [_vp, _dp],
    "lbm_get_populations": [_vp, _dp],
    "lbm_get_hydro": [_vp, _dp, _dp],
    "lbm_add_particle": [_vp, C.POINTER(LbmParticle), C.POINTER(C.c_int32)],
    "lbm_get_particle": [_vp, C.c_int32, C.POINTER(LbmParticle)],
    "lbm_get_particle_load": [_vp, C.c_int32, _dp, _dp],
    "lbm_get_links": [_vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int64,
                      C.POINTER(C.c_int64)],
    "lbm_set_interactions": [_vp, C.POINTER(LbmInteraction), C.c_int32],
    "lbm_set_field_schedule": [_vp, _dp, C.c_int64, C.c_int64],
    "lbm_set_occupancy": [_vp, C.c_int32],
    "lbm_get_occupancy": [_vp, C.POINTER(C.c_int32)],
    "lbm_batch_set_occupancy": [_vp, C.c_int32],
    "lbm_batch_get_occupancy": [_vp, C.POINTER(C.c_int32)],
    "lbm_set_trajectory_log": [_vp, C.c_int64],
    "lbm_get_trajectory": [_vp, C.c_int32, C.c_int64, C.c_int64, _dp],
    "lbm_write_fluid_csv": [C.c_char_p, C.c_int32, C.POINTER(C.c_int64), C.c_int64, C.c_double,
                            C.POINTER(_dp), _dp, _dp, C.c_int32, _dp],
    "lbm_write_tensor_csv": [C.c_char_p, C.c_int32, C.POINTER(C.c_int64), C.c_int64, C.c_double,
                             C.POINTER(_dp), _dp],
    "lbm_viscous_stress": [_vp, _dp],
    "lbm_check_populations": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "lbm_run": [_vp, C.c_int64],
    "lbm_run_noalloc": [_vp, C.c_int64],
    "lbm_sync": [_vp],
    "lbm_get_tick": [_vp, C.POINTER(C.c_int64), _dp],
    "lbm_state_size": [_vp, C.POINTER(C.c_int64)],
    "lbm_save_state": [_vp, _vp, C.c_int64],
    "lbm_load_state": [_vp, _vp, C.c_int64],
    "lbm_set_block": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "lbm_get_block": [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "lbm_set_tile_order": [_vp, C.c_int32],
    "lbm_get_tile_order": [_vp, C.POINTER(C.c_int32)],
    "lbm_batch_set_tile_order": [_vp, C.c_int32],
    "lbm_batch_get_tile_order": [_vp, C.POINTER(C.c_int32)],
    "lbm_set_timing": [_vp, C.c_int],
    "lbm_last_run_ms": [_vp, _dp],
    "lbm_last_run_times": [_vp, _dp, _dp],
    "lbm_measure_copy": [C.c_int64, C.c_int32, _dp],
    "lbm_measure_pull": [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp],
    "lbm_autotune": [_vp, C.c_int32, C.POINTER(C.c_int32), _dp],
    "lbm_set_graph": [_vp, C.c_int32],
    "lbm_batch_create": [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)],
    "lbm_batch_destroy": [_vp],
    "lbm_batch_run": [_vp, C.c_int64],
    "lbm_batch_sync": [_vp],
    "lbm_batch_set_block": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "lbm_batch_get_block": [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "lbm_batch_set_row_tiles": [_vp, C.c_int32],
    "lbm_batch_get_row_tiles": [_vp, C.POINTER(C.c_int32)],
    "lbm_batch_autotune": [_vp, C.c_int32, C.POINTER(C.c_int32), _dp],
    "lbm_batch_set_graph": [_vp, C.c_int32],
    "lbm_batch_set_timing": [_vp, C.c_int],
    "lbm_batch_last_run_times": [_vp, _dp, _dp],
    "lbm_adjoint_run": [_vp, _vp, _dp, _dp, _dp, _dp],
    "lbm_adjoint_last_times": [_vp, _dp, _dp],
}
_RESTYPES = {"lbm_last_error": C.c_char_p, "lbm_version": C.c_char_p}


def bind(lib: C.CDLL, prefix_from: str = "lbm_", prefix_to: str = "lbm_") -> dict:
    """Attach prototypes to lib for every entry point; returns name -> function."""
    out = {}
    for name, args in SIGNATURES.items():
        sym = prefix_to + name[len(prefix_from):]
        if not hasattr(lib, sym):
            continue
        fn = getattr(lib, sym)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
        out[name] = fn
    return out


_LIB = None
_FN: dict = {}


def load() -> dict:
    """Load liblbm_hip.so once; raises if it is missing."""
    global _LIB, _FN
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python __graft_entry__.py build` "
                "(or lbmengine_hip.build.build()) first")
        _LIB = C.CDLL(str(LIB_PATH))
        _FN = bind(_LIB)
    return _FN


def check(rc: int, fn: dict | None = None) -> None:
    if rc != LBM_OK:
        fns = fn or load()
        msg = fns["lbm_last_error"]().decode(errors="replace")
        raise LBMError(rc, msg)


def as_dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def fortran_f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def measure_copy_ms(nbytes: int, reps: int = 10) -> float:
    """Average ms of a plain device copy of nbytes (read nbytes, write nbytes)."""
    ms = C.c_double(0.0)
    check(load()["lbm_measure_copy"](int(nbytes) // 16 * 16, int(reps), C.byref(ms)))
    return ms.value


def measure_pull_ms(storage: int, nx: int, ny: int, nz: int, reps: int = 10) -> float:
    """Average ms of the D3Q19 step's gathers and stores alone (no collision)
    on a periodic nx x ny x nz lattice of the given storage."""
    ms = C.c_double(0.0)
    check(load()["lbm_measure_pull"](int(storage), int(nx), int(ny), int(nz), int(reps), C.byref(ms)))
    return ms.value


def device_count() -> int:
    """Visible HIP devices, asked through liblbm_hip.so (0 when none)."""
    n = C.c_int32(0)
    check(load()["lbm_device_count"](C.byref(n)))
    return n.value
