"""Workloads at their full sizes, checked through properties that do not
need the oracle (which is too slow there): link counts, the transient
Poiseuille series, agreement of the reduced-precision storages with FP64."""
import math

import numpy as np
import pytest

import lbmengine_hip as lh

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, kineticParameter=3 / 16)


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def poiseuille_transient(Nz, g, nu, t):
    """u_x(z, t) of a channel started from rest, walls at z = 1/2 and
    z = Nz - 3/2 (halfway bounce-back), driven by a uniform force g."""
    H = Nz - 2.0
    zeta = np.arange(Nz) - 0.5
    u = g / (2 * nu) * zeta * (H - zeta)
    for n in range(1, 400, 2):
        u -= 4 * g * H * H / (nu * math.pi ** 3 * n ** 3) * np.sin(n * math.pi * zeta / H) * \
             math.exp(-nu * (n * math.pi / H) ** 2 * t)
    u[0] = u[-1] = 0.0
    return u


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_config2_poiseuille_channel_20000_steps(gpu, storage):
    """D3Q19 TRT (Lambda 3/16) channel 256 x 64 x 64, walls at z = 0 and 63,
    periodic in x and y, uniform Guo force 1e-6 along x, tau 0.9, 20000 steps."""
    g, tau, steps = 1e-6, 0.9, 20000
    m = lh.modelInit(**KW3, relaxationTimeRatio=tau, xlims=(0, 255), ylims=(0, 63), zlims=(0, 63),
                     walledDimensions=[3], forceDensity=[g, 0, 0], storage=storage)
    lh.LBMpropagate(m, ticks=steps)
    u = m.fluidVelocity
    nu = (tau - 0.5) / 3
    exact = poiseuille_transient(64, g, nu, steps)
    prof = u[:, :, :, 0].mean(axis=(0, 1))
    assert np.max(np.abs(prof - exact)) < 5e-4 * np.max(exact)
    # translation invariance in x and y, no cross flow
    assert np.max(np.abs(u[:, :, :, 0] - prof[None, None, :])) < 1e-6 * np.max(exact)
    assert np.max(np.abs(u[:, :, :, 1:])) < 1e-6 * np.max(exact)
    if storage == "f32":
        ref = lh.modelInit(**KW3, relaxationTimeRatio=tau, xlims=(0, 255), ylims=(0, 63), zlims=(0, 63),
                           walledDimensions=[3], forceDensity=[g, 0, 0], storage="f64")
        lh.LBMpropagate(ref, ticks=steps)
        assert rel_l2(u, ref.fluidVelocity) < 1e-3


def test_headline_geometry_192_cubed(gpu):
    """192^3 periodic, squirmer r = 12 at the centre: 10146 boundary links,
    7123 solid nodes in the origin-centred mass, FP16 deviation storage."""
    m = lh.modelInit(**KW3, relaxationTimeRatio=0.8, xlims=(-96, 95), storage="f16")
    lh.addSquirmer(m, radius=12.0, position=[0.0, 0.0, 0.0], swimmingSpeed=2e-3 / 3, beta=0.0,
                   swimmingDirection=[0, 0, 1], coupleForces=True, coupleTorques=True)
    n, d = m.particles[0].links()
    assert len(n) == 10146
    assert math.isclose(1 / m.particles[0].particleParams["inverseMass"], 7123.0)
    lh.LBMpropagate(m, ticks=30)
    v = m.particles[0].velocity
    assert v[2] > 0 and abs(v[0]) < 1e-3 * v[2] and abs(v[1]) < 1e-3 * v[2]
    rho = m.massDensity
    assert np.all(np.isfinite(rho)) and abs(rho.mean() - 1) < 1e-3


def run_squirmer(storage, steps, n=48, r=6.0, beta=-5.0, seed=3):
    rng = np.random.default_rng(seed)
    m = lh.modelInit(**KW3, relaxationTimeRatio=0.8, xlims=(-n // 2, n // 2 - 1), storage=storage)
    e = rng.standard_normal(3)
    lh.addSquirmer(m, radius=r, position=rng.uniform(-1, 1, 3), swimmingSpeed=2e-3 / 3, beta=beta,
                   swimmingDirection=e, coupleForces=True, coupleTorques=True)
    speeds = []
    for _ in range(steps // 50):
        lh.LBMpropagate(m, ticks=50)
        speeds.append(np.linalg.norm(m.particles[0].velocity))
    return m, np.array(speeds)


@pytest.mark.parametrize("storage", ["f32", "f16", "f16c"])
@pytest.mark.parametrize("beta", [-5.0, 5.0])
def test_reduced_precision_against_fp64(gpu, storage, beta):
    ref, s_ref = run_squirmer("f64", 1000, beta=beta)
    low, s_low = run_squirmer(storage, 1000, beta=beta)
    eu = rel_l2(low.fluidVelocity, ref.fluidVelocity)
    es = float(np.max(np.abs(s_low - s_ref)) / np.max(np.abs(s_ref)))
    print(f"{storage} beta={beta}: u-field rel L2 err {eu:.3e}, swimming-speed rel err {es:.3e}")
    assert eu < 1e-3 and es < 1e-3
