"""viscousStressTensor on the device (cpu/fluids.jl:507-520) against the
oracle in FP64 (<= 1e-12 relative to the field's scale), the FP32 storage
against FP64, and writeTensor's file (gpu/aux.jl:364-400) read back."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16,
           xlims=(-10, 9))
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.9, collisionModel="bgk",
           xlims=(0, 31), ylims=(0, 23), walledDimensions=[2], forceDensity=[1e-5, 0.0])
SQ3 = dict(radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3, beta=-3.0, swimmingDirection=[1, 1, 0],
           coupleForces=True, coupleTorques=True)


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def stack(T):
    d = len(T)
    return np.stack([np.stack([T[a][b] for b in range(d)], axis=-1) for a in range(d)], axis=-2)


def test_stress_3d_squirmer_force_field(gpu, oracle):
    rng = np.random.default_rng(2)
    F = rng.normal(0.0, 1e-6, (20, 20, 20, 3))
    g = gpu.modelInit(**KW3, forceDensity=F)
    o = oracle.oracle_model(**KW3, forceDensity=F)
    gpu.addSquirmer(g, **SQ3)
    o.add("squirmer", **SQ3)
    gpu.LBMpropagate(g, ticks=30)
    o.run(30)
    assert rel(stack(gpu.viscousStressTensor(g)), o.viscous_stress()) <= 1e-12


def test_stress_2d_walls_uniform_force(gpu, oracle):
    g = gpu.modelInit(**KW2)
    o = oracle.oracle_model(**KW2)
    gpu.LBMpropagate(g, ticks=60)
    o.run(60)
    assert rel(stack(gpu.viscousStressTensor(g)), o.viscous_stress()) <= 1e-12


def test_fp32_stress_against_fp64(gpu):
    out = []
    for storage in ("f64", "f32"):
        m = gpu.modelInit(**KW3, storage=storage)
        gpu.addSquirmer(m, **SQ3)
        gpu.LBMpropagate(m, ticks=200)
        out.append(stack(gpu.viscousStressTensor(m)))
    assert rel(out[1], out[0]) <= 1e-3


def test_write_tensor(gpu, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    m = gpu.modelInit(**KW2)
    gpu.LBMpropagate(m, ticks=7)
    T = gpu.viscousStressTensor(m)
    gpu.writeTensor(m, T, "stress")
    lines = open(os.path.join("output.lbm", "stress.csv")).read().splitlines()
    assert lines[0] == ("tick,time,id_x,id_y,coordinate_x,coordinate_y,"
                        "component_xx,component_yx,component_xy,component_yy")
    rows = np.array([[float(c) for c in ln.split(",")] for ln in lines[1:]])
    assert rows.shape == (32 * 24, 10) and np.all(rows[:, 0] == 7) and np.all(rows[:, 1] == m.time)
    for j, (a, b) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1)]):
        assert np.array_equal(rows[:, 6 + j], T[a][b].reshape(-1, order="F"))
