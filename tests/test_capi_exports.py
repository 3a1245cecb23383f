"""liblbm_hip.so loads on the CPU and exports every entry point that
include/lbm_hip.h declares (no compute calls: there is no GPU here)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "lbm_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_step_api():
    names = declared()
    for must in ("lbm_create", "lbm_destroy", "lbm_run", "lbm_sync", "lbm_add_particle",
                 "lbm_get_links", "lbm_get_hydro", "lbm_get_populations", "lbm_set_populations"):
        assert must in names


def test_library_exports_every_declared_symbol(hiplib):
    lib = ctypes.CDLL(str(hiplib))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_text(hiplib):
    from lbmengine_hip import _capi
    fn = _capi.load()
    assert b"gfx950" in fn["lbm_version"]()
    assert isinstance(fn["lbm_last_error"](), bytes)
    # a null configuration is rejected before any device work
    h = ctypes.c_void_p()
    rc = fn["lbm_create"](None, ctypes.byref(h))
    assert rc == _capi.LBM_EINVAL and b"null" in fn["lbm_last_error"]()


def test_bad_configuration_is_rejected_without_device(hiplib):
    import numpy as np
    from lbmengine_hip import _capi, resolve_model
    fn = _capi.load()
    rm = resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, xlims=(0, 7))
    cfg, keep = rm.config()
    cfg.q = 13
    h = ctypes.c_void_p()
    assert fn["lbm_create"](ctypes.byref(cfg), ctypes.byref(h)) == _capi.LBM_EINVAL
    assert b"D3Q19" in fn["lbm_last_error"]()
    cfg.q = 19
    cfg.flags = 4  # no such flag
    assert fn["lbm_create"](ctypes.byref(cfg), ctypes.byref(h)) == _capi.LBM_EINVAL
    assert b"flags" in fn["lbm_last_error"]()


def test_measurements_reject_bad_arguments(hiplib):
    """The copy and pull measurements check their arguments before any
    device call."""
    from lbmengine_hip import _capi
    fn = _capi.load()
    ms = ctypes.c_double()
    assert fn["lbm_measure_copy"](17, 1, ctypes.byref(ms)) == _capi.LBM_EINVAL
    assert fn["lbm_measure_pull"](7, 8, 8, 8, 1, ctypes.byref(ms)) == _capi.LBM_EINVAL
    assert fn["lbm_measure_pull"](3, 0, 8, 8, 1, ctypes.byref(ms)) == _capi.LBM_EINVAL
    assert fn["lbm_measure_pull"](3, 8, 8, 8, 0, ctypes.byref(ms)) == _capi.LBM_EINVAL
