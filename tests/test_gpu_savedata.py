"""saveData on the device: the files LBMpropagate! writes (gpu/fluids.jl:
243-259, gpu/aux.jl:271-362) against the oracle run tick by tick (FP64 to
<= 1e-12 relative), and the device's per-tick particle log against a twin
model stepped one tick per run (bit-exact), across log-ring wraps.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12
KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16,
           xlims=(-10, 9))
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, xlims=(-16, 15))
SQ3 = dict(radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3, beta=-3.0, swimmingDirection=[1, 1, 0],
           coupleForces=True, coupleTorques=True)


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def read_csv(path):
    lines = open(path).read().splitlines()
    head = lines[0].split(",")
    rows = np.array([[float(c) for c in ln.split(",")] for ln in lines[1:]])
    return head, rows


def test_savedata_files_match_the_oracle(gpu, oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    os.mkdir("output.lbm")
    open("output.lbm/stale.csv", "w").close()
    g = gpu.modelInit(**KW3, saveData=True)
    assert not os.path.exists("output.lbm/stale.csv")  # modelInit empties the folder
    gpu.addSquirmer(g, **SQ3)
    o = oracle.oracle_model(**KW3)
    o.add("squirmer", **SQ3)
    gpu.LBMpropagate(g, ticks=25, ticksSaved=6)
    assert sorted(os.listdir("output.lbm")) == sorted(
        ["particlesTrj.csv"] + [f"fluidTrj_{t}.csv" for t in (0, 5, 10, 15, 20, 25)])
    head, prow = read_csv("output.lbm/particlesTrj.csv")
    assert head == ["tick", "time", "particleId", "position_x", "position_y", "position_z", "velocity_x",
                    "velocity_y", "velocity_z", "angularVelocity_x", "angularVelocity_y", "angularVelocity_z"]
    assert prow.shape == (26, 12) and np.array_equal(prow[:, 0], np.arange(26))
    N = 20 ** 3
    for t in range(26):
        if t:
            o.run(1)
        p = o.particles[0]
        assert prow[t, 1] == o.time and prow[t, 2] == 1
        for cols, want in ((slice(3, 6), p.position), (slice(6, 9), p.velocity), (slice(9, 12), p.angularVelocity)):
            assert rel(prow[t, cols], want) <= RTOL
        if t % 5 == 0:
            fh, fr = read_csv(f"output.lbm/fluidTrj_{t}.csv")
            assert fr.shape == (N, 2 + 3 + 3 + 1 + 3 + 19)
            assert np.all(fr[:, 0] == t) and np.all(fr[:, 1] == o.time)
            idx = np.stack(np.unravel_index(np.arange(N), (20, 20, 20), order="F"), axis=1)
            assert np.array_equal(fr[:, 2:5], idx + 1) and np.array_equal(fr[:, 5:8], idx - 10.0)
            assert rel(fr[:, 8], o.massDensity.reshape(-1, order="F")) <= RTOL
            assert rel(fr[:, 9:12], o.fluidVelocity.reshape(N, 3, order="F")) <= RTOL
            assert rel(fr[:, 12:], o.distributions.reshape(N, 19, order="F")) <= RTOL
    # a second propagate appends to the particle file and adds its checkpoints
    gpu.LBMpropagate(g, ticks=4, ticksSaved=2)
    assert os.path.exists("output.lbm/fluidTrj_29.csv")
    _, prow2 = read_csv("output.lbm/particlesTrj.csv")
    assert prow2.shape[0] == 30 and np.array_equal(prow2[:26], prow)


def test_trajectory_log_across_ring_wraps(gpu, tmp_path, monkeypatch):
    """Two particles, a 7-tick ring: the saved rows equal a twin model's
    states read after every single-tick run."""
    from lbmengine_hip import savedata
    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr(savedata, "LOG_TICKS", 7)
    parts = [dict(radius=4.0, position=[-6.3, 0.2], swimmingSpeed=2e-3, beta=2.0, swimmingDirection=[1, 0.3],
                  coupleForces=True, coupleTorques=True),
             dict(radius=3.5, position=[6.1, -0.4], swimmingSpeed=1e-3, beta=-1.0, swimmingDirection=[-0.2, 1],
                  coupleForces=True, coupleTorques=True)]
    g = gpu.modelInit(**KW2, saveData=True)
    twin = gpu.modelInit(**KW2)
    for pk in parts:
        gpu.addSquirmer(g, **pk)
        gpu.addSquirmer(twin, **pk)
    gpu.LBMpropagate(g, ticks=40, ticksSaved=3)
    head, rows = read_csv("output.lbm/particlesTrj.csv")
    assert head[-1] == "angularVelocity" and rows.shape == (2 + 80, 8)
    for t in range(1, 41):
        twin.run(1)
        for p in range(2):
            r = rows[2 + 2 * (t - 1) + p]
            q = twin.particles[p]
            assert r[0] == t and r[1] == twin.time and r[2] == p + 1
            assert np.array_equal(r[3:5], q.position) and np.array_equal(r[5:7], q.velocity)
            assert r[7] == q.angularVelocity
    assert np.array_equal(g.distributions, twin.distributions)
    assert sorted(os.listdir("output.lbm")) == sorted(
        ["particlesTrj.csv", "fluidTrj_0.csv", "fluidTrj_20.csv", "fluidTrj_40.csv"])


def test_trajectory_log_bounds(gpu):
    from lbmengine_hip import _capi
    m = gpu.modelInit(**KW3)
    gpu.addSquirmer(m, **SQ3)
    fn = m._fn
    buf = np.zeros((4, 9))
    assert fn["lbm_get_trajectory"](m._h, 0, 1, 1, _capi.as_dptr(buf)) == _capi.LBM_ESTATE
    m.run(3)
    _capi.check(fn["lbm_set_trajectory_log"](m._h, 4))
    m.run(6)
    m.sync()
    # ticks 4..9 were logged; the ring keeps the last 4
    assert fn["lbm_get_trajectory"](m._h, 0, 5, 2, _capi.as_dptr(buf)) == _capi.LBM_EINVAL
    assert fn["lbm_get_trajectory"](m._h, 0, 9, 2, _capi.as_dptr(buf)) == _capi.LBM_EINVAL
    _capi.check(fn["lbm_get_trajectory"](m._h, 0, 6, 4, _capi.as_dptr(buf)))
    assert np.array_equal(buf[3, 0:3], m.particles[0].position)
    _capi.check(fn["lbm_set_trajectory_log"](m._h, 0))
