"""Known-answer tests that pin the oracle (oracle/lbm_oracle.c) before it is
trusted as the checker of the HIP step.

The reference ships no tests, fixtures or golden vectors (test/runtests.jl:1),
so each part of the step is pinned by what the mathematics fixes: closed-form
moments, exact conservation laws, the steady Poiseuille profile that TRT with
Lambda = 3/16 and halfway bounce-back reproduces exactly, TRT == BGK at
Lambda = (tau - 1/2)^2, a brute-force restatement of the boundary-link
definition, link/solid counts of centred spheres, and closed-form particle
updates (kick, drift, Rodrigues).
"""
import math

import numpy as np
import pytest

from lbmengine_hip import sphere_count

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0)
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0)


def bare_momentum(rm, f):
    return (f.reshape(-1, f.shape[-1]) @ rm.cs.astype(float)).sum(axis=0)


# --------------------------------------------------------------- moments
@pytest.mark.parametrize("kw", [KW2, KW3])
def test_equilibrium_moments_closed_form(oracle, kw):
    u0 = [0.02, -0.013, 0.007][: kw["dims"]]
    m = oracle.oracle_model(**kw, relaxationTimeRatio=0.8, xlims=(0, 3), fluidVelocity=u0, massDensity=1.3)
    f = m.distributions
    rho = f.sum(axis=-1)
    j = bare_momentum(m.resolved, f) / rho.size
    # sum_i feq_i = rho and sum_i feq_i c_i = rho u (compressible feq)
    assert np.allclose(rho, 1.3, rtol=1e-14, atol=0)
    assert np.allclose(j, 1.3 * np.array(u0), rtol=1e-13, atol=0)
    assert np.allclose(m.fluidVelocity, np.broadcast_to(u0, m.fluidVelocity.shape), rtol=1e-13, atol=0)


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("kw", [KW2, KW3])
def test_periodic_conservation(oracle, kw, coll):
    rng = np.random.default_rng(1)
    m = oracle.oracle_model(**kw, relaxationTimeRatio=0.7, collisionModel=coll, xlims=(0, 5))
    f = m.distributions
    f = f * (1 + 0.05 * rng.standard_normal(f.shape))
    m.distributions = f
    M0, J0 = f.sum(), bare_momentum(m.resolved, f)
    m.run(25)
    f1 = m.distributions
    assert abs(f1.sum() - M0) < 1e-12 * M0
    assert np.allclose(bare_momentum(m.resolved, f1), J0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_uniform_force_momentum_growth(oracle, coll):
    # Guo forcing: the bare momentum sum_x sum_i f_i c_i grows by N F dt per tick
    F = np.array([1e-5, -2e-6, 3e-6])
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.9, collisionModel=coll, kineticParameter=3 / 16,
                            xlims=(0, 5), forceDensity=F)
    J0 = bare_momentum(m.resolved, m.distributions)
    m.run(10)
    J1 = bare_momentum(m.resolved, m.distributions)
    N = m.resolved.N
    assert np.allclose(J1 - J0, 10 * N * F, rtol=1e-10, atol=0)


def test_trt_equals_bgk_at_debug_lambda(oracle):
    rng = np.random.default_rng(7)
    kw = dict(**KW3, relaxationTimeRatio=0.83, xlims=(0, 5), forceDensity=[2e-6, 1e-6, 0.0])
    a = oracle.oracle_model(**kw, collisionModel="bgk")
    b = oracle.oracle_model(**kw, collisionModel="trt", kineticParameter="debug")
    f = a.distributions * (1 + 0.02 * rng.standard_normal(a.distributions.shape))
    a.distributions = f
    b.distributions = f
    a.run(15)
    b.run(15)
    assert np.allclose(a.distributions, b.distributions, rtol=1e-13, atol=1e-16)


def test_shear_wave_decay_rate(oracle):
    # u_x(y) = U sin(k y) decays as exp(-nu k^2 t), nu = c_s^2 (tau - 1/2) dt
    L, tau, steps = 32, 0.8, 200
    y = np.arange(L)
    u = np.zeros((4, L, 2))
    u[:, :, 0] = 1e-4 * np.sin(2 * np.pi * y / L)[None, :]
    m = oracle.oracle_model(**KW2, relaxationTimeRatio=tau, collisionModel="bgk", xlims=(0, 3), ylims=(0, L - 1),
                            fluidVelocity=u)
    a0 = np.abs(np.fft.rfft(m.fluidVelocity[0, :, 0])[1])
    m.run(steps)
    a1 = np.abs(np.fft.rfft(m.fluidVelocity[0, :, 0])[1])
    nu, k = (tau - 0.5) / 3, 2 * np.pi / L
    rate = -np.log(a1 / a0) / steps
    assert abs(rate / (nu * k * k) - 1) < 0.01


def test_poiseuille_steady_profile_exact_with_magic_trt(oracle):
    # walls at z = 0 and z = Nz-1 (0-based); halfway bounce-back puts the no-slip
    # planes at z = 1/2 and z = Nz - 3/2, H = Nz - 2; TRT with Lambda = 3/16
    # reproduces u(z) = F/(2 nu) (z - 1/2)(Nz - 3/2 - z) exactly.
    Nz, g, tau = 12, 1e-6, 0.9
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=tau, kineticParameter=3 / 16, xlims=(0, 3), ylims=(0, 3),
                            zlims=(0, Nz - 1), walledDimensions=[3], forceDensity=[g, 0, 0])
    m.run(3000)
    ux = m.fluidVelocity[1, 1, :, 0]
    nu = (tau - 0.5) / 3
    z = np.arange(Nz)
    exact = g / (2 * nu) * (z - 0.5) * (Nz - 1.5 - z)
    exact[0] = exact[-1] = 0.0
    assert np.max(np.abs(ux - exact)) < 1e-11 * np.max(exact)
    # wall layers hold zero populations
    f = m.distributions
    assert np.all(f[:, :, 0, :] == 0) and np.all(f[:, :, -1, :] == 0)


# ----------------------------------------------------------------- links
def brute_links(rm, R, r):
    """Boundary links straight from the reference's definitions:
    solid = X_1^2 + X_2^2 [+ X_3^2] < r^2 around R (gpu/kernels.jl:335-353),
    SIR_j = (circshift(solid, -c_j) .|| solid) .⊻ solid (gpu/aux.jl:190),
    listed by findall over (N..., Q)."""
    d = rm.dims
    grids = np.meshgrid(*[rm.coords[k] for k in range(d)], indexing="ij")
    s = None
    for k in range(d):
        a = grids[k] + (-R[k])
        s = a * a if s is None else s + a * a
    solid = s < r * r
    nodes, dirs = [], []
    for j, c in enumerate(rm.cs):
        shifted = np.roll(solid, shift=tuple(-int(x) for x in c), axis=tuple(range(d)))
        sir = (shifted | solid) ^ solid
        idx = np.flatnonzero(sir.ravel(order="F"))
        nodes.append(idx)
        dirs.append(np.full(idx.size, j))
    return solid, np.concatenate(nodes), np.concatenate(dirs)


@pytest.mark.parametrize("kw,lims,r,R", [
    (KW3, (-12, 11), 4.3, [0.37, -1.21, 0.5]),
    (KW3, (-10, 9), 3.0, [0.0, 0.0, 0.0]),
    (KW2, (-16, 15), 5.5, [1.25, -0.75]),
])
def test_links_match_brute_force(oracle, kw, lims, r, R):
    m = oracle.oracle_model(**kw, relaxationTimeRatio=0.8, xlims=lims)
    m.add("squirmer", radius=r, position=R)
    solid, bn, bd = brute_links(m.resolved, R, r)
    n, d = m.particles[0].links()
    assert np.array_equal(n, bn) and np.array_equal(d, bd)
    assert np.array_equal(m.particles[0].solid(), solid)


@pytest.mark.parametrize("kw,lims,r,nsolid,nlinks", [
    (KW3, (-24, 23), 12.0, 7123, 10146),   # headline squirmer, r = 12
    (KW3, (-20, 19), 10.0, 4139, 7146),    # adjoint squirmer, r = 10
    (KW3, (-12, 11), 6.0, 895, 2514),      # smallest batched squirmer, r = 6
    (KW2, (-64, 63), 8.0, 193, 152),       # D2Q9 disk, r = 8
])
def test_centred_sphere_link_counts(oracle, kw, lims, r, nsolid, nlinks):
    m = oracle.oracle_model(**kw, relaxationTimeRatio=0.8, xlims=lims)
    m.add("squirmer", radius=r)
    n, d = m.particles[0].links()
    assert int(m.particles[0].solid().sum()) == nsolid
    assert len(n) == nlinks


# ------------------------------------------------------------- particles
def test_kick_drift_closed_form(oracle):
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    m.add("squirmer", radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3 / 3,
          swimmingDirection=[1, 0.5, 0], coupleForces=True, coupleTorques=True)
    p = m.particles[0]
    m.run(3)
    v0, x0 = p.velocity, p.position
    m.run(1)
    F, T = p.load()
    inv_m = 1 / sphere_count(m.resolved, 4.0)
    # v += dt F / m ; x += dt v (gpu/particles.jl:17-28), dt = 1
    assert np.allclose(p.velocity, v0 + F * inv_m, rtol=1e-14, atol=1e-20)
    assert np.allclose(p.position, x0 + p.velocity, rtol=1e-15, atol=0)


def test_rodrigues_rotation_of_swimming_direction(oracle):
    from scipy.spatial.transform import Rotation
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    m.add("squirmer", radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3 / 3, beta=5,
          swimmingDirection=[1, 0.5, 0.2], coupleForces=True, coupleTorques=True)
    p = m.particles[0]
    m.run(5)
    e0 = p.swimmingDirection
    m.run(1)
    w, e1 = p.angularVelocity, p.swimmingDirection
    expect = Rotation.from_rotvec(w * 1.0).apply(e0)
    assert np.linalg.norm(w) > 0
    assert np.allclose(e1, expect, rtol=0, atol=1e-15)
    assert abs(np.linalg.norm(e1) - 1) < 1e-15


def test_cpu_semantics_does_not_rotate(oracle):
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11),
                            semantics="cpu")
    m.add("squirmer", radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3 / 3, beta=5,
          swimmingDirection=[1, 0.5, 0.2], coupleForces=True, coupleTorques=True)
    e0 = m.particles[0].swimmingDirection
    m.run(6)
    assert np.linalg.norm(m.particles[0].angularVelocity) > 0
    assert np.array_equal(m.particles[0].swimmingDirection, e0)


def test_squirmer_swims_along_its_axis(oracle):
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    e = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    m.add("squirmer", radius=4.0, position=[0.0, 0.0, 0.0], swimmingSpeed=5e-3, swimmingDirection=e,
          coupleForces=True, coupleTorques=True)
    m.run(150)
    v = m.particles[0].velocity
    assert v @ e > 0 and np.linalg.norm(v - (v @ e) * e) < 0.05 * (v @ e)


def test_drag_on_held_bead_points_downstream(oracle):
    m = oracle.oracle_model(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-10, 9),
                            forceDensity=[1e-5, 0, 0])
    m.add("bead", radius=3.0, massDensity=1e12, coupleForces=True)
    m.run(60)
    F, _ = m.particles[0].load()
    assert F[0] > 0 and abs(F[1]) < 1e-3 * F[0] and abs(F[2]) < 1e-3 * F[0]
