"""saveData host pieces, without a GPU: the library's native fluidTrj writer
(lbm_write_fluid_csv) against the column layout of writeTrajectories
(gpu/aux.jl:271-317), its numbers against a round trip and the Python
formatter of the particle rows, and the checkpoint ticks of LBMpropagate!
(gpu/fluids.jl:243-247).

The text form of a Float64 (Julia's print through CSV.jl, a dependency absent
here) is parity unpinned: the files hold the shortest decimal that reads
back to the same double, which is what Julia prints, in Julia's notation as
restated in savedata.julia_float."""
import ctypes as C
import math

import numpy as np
import pytest


def write(tmp_path, dims, n, tick, time, coords, rho, u, f):
    from lbmengine_hip import _capi
    fn = _capi.load()
    nn = (C.c_int64 * 3)(*[n[k] if k < dims else 1 for k in range(3)])
    cs = [np.ascontiguousarray(c, dtype=np.float64) for c in coords]
    cp = (_capi._dp * 3)(*[c.ctypes.data_as(_capi._dp) for c in cs], *([None] * (3 - dims)))
    rho, u, f = (np.asfortranarray(a, dtype=np.float64) for a in (rho, u, f))
    path = tmp_path / f"fluidTrj_{tick}.csv"
    _capi.check(fn["lbm_write_fluid_csv"](str(path).encode(), dims, nn, tick, time, cp, _capi.as_dptr(rho),
                                          _capi.as_dptr(u), f.shape[-1], _capi.as_dptr(f)))
    return path.read_text()


def values(rng, shape):
    """Values over many binades, with the awkward ones mixed in."""
    v = rng.standard_normal(shape) * 10.0 ** rng.integers(-9, 9, shape)
    flat = v.reshape(-1)
    special = [0.0, -0.0, 1.0, 1e-4, 9.999e-5, 1e6, 999999.9, 123456.0, 0.1, 1 / 3, 5e-324, -2.5e-8]
    flat[: len(special)] = special
    return v


@pytest.mark.parametrize("dims", [2, 3])
def test_fluid_csv_layout_and_values(hiplib, tmp_path, dims):
    from lbmengine_hip.savedata import julia_float
    rng = np.random.default_rng(dims)
    n = (5, 4, 3)[:dims]
    Q = 9 if dims == 2 else 19
    coords = [np.arange(n[k]) - 2.0 for k in range(dims)]
    rho = values(rng, n)
    u = values(rng, n + (dims,))
    f = values(rng, n + (Q,))
    text = write(tmp_path, dims, n, 37, 3.7, coords, rho, u, f)
    lines = text.splitlines()
    ax = "xyz"[:dims]
    want = (["tick", "time"] + [f"id_{a}" for a in ax] + [f"coordinate_{a}" for a in ax] + ["massDensity"]
            + [f"fluidVelocity_{a}" for a in ax] + [f"f{i}" for i in range(1, Q + 1)])
    assert lines[0].split(",") == want
    N = int(np.prod(n))
    assert len(lines) == N + 1
    R = rho.reshape(-1, order="F")
    U = u.reshape(N, dims, order="F")
    F = f.reshape(N, Q, order="F")
    for x in range(N):
        idx = np.unravel_index(x, n, order="F")
        row = [str(37), julia_float(3.7)] + [str(i + 1) for i in idx]
        row += [julia_float(coords[k][idx[k]]) for k in range(dims)] + [julia_float(R[x])]
        row += [julia_float(v) for v in U[x]] + [julia_float(v) for v in F[x]]
        assert lines[x + 1] == ",".join(row)
        cells = lines[x + 1].split(",")
        back = np.array([float(c) for c in cells[2 + 2 * dims:]])
        assert np.array_equal(back, np.concatenate([[R[x]], U[x], F[x]]))


def test_julia_float_is_shortest_and_round_trips():
    from lbmengine_hip.savedata import julia_float
    rng = np.random.default_rng(5)
    for x in values(rng, 4000).reshape(-1):
        s = julia_float(x)
        assert float(s) == x and math.copysign(1, float(s)) == math.copysign(1, x)
        digits = s.lstrip("-").split("e")[0].replace(".", "").strip("0") or "0"
        assert len(digits) <= len(repr(abs(x)).split("e")[0].replace(".", "").strip("0") or "0")
    assert [julia_float(v) for v in (1e-5, 1e-4, 1e6, 123456.7, 2.0 ** 20, -96.0, 0.0, -0.0)] == \
        ["1.0e-5", "0.0001", "1.0e6", "123456.7", "1.048576e6", "-96.0", "0.0", "-0.0"]


def test_checkpoint_ticks():
    from lbmengine_hip.savedata import checkpoints
    # range(0, stop = 100, length = 100): 100 distinct ticks of 0..100, 50 skipped
    c = checkpoints(0, 100, 1.0, None)
    assert min(c) == 0 and max(c) == 100 and len(c) == 100 and 50 not in c
    # half-integer points round to even: 0, .5, 1, ..., 5
    assert checkpoints(0, 5, 1.0, 11) == {0, 1, 2, 3, 4, 5}
    assert checkpoints(10, 6, 1.0, 5) == {10, 12, 13, 14, 16}  # 10, 11.5, 13, 14.5, 16
    # a time step of 0.1: 3 ticks end at time 0.3 and floor(0.3 / 0.1) = 2,
    # so the reference's last checkpoint is tick 2 of 3
    assert checkpoints(0, 3, 0.1, 2) == {0, 2}
    assert max(checkpoints(0, 30, 0.1, 4)) == 30
    with pytest.raises(ValueError):
        checkpoints(0, 5, 1.0, 1)
