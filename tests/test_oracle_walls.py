"""Known answers that pin the oracle's other schemes: the D3Q15 and D3Q27
velocity sets, walls and moving walls given as functions
(gpu/initMethods.jl:580-626, gpu/fluids.jl:149-174) and the incompressible
equilibrium (gpu/fluids.jl:51).

* Plane Couette flow: a wall layer moving at U over a resting one gives the
  linear profile U (z - 1/2) / (Nz - 2) between the halfway planes, which
  halfway bounce-back with the moving-wall term reproduces exactly at steady
  state, for every lattice and both collisions.
* One tick next to a moving wall from rest: each bounced population is
  w_i + 6 w_i u_w . c_i (the moving-wall term, with rho = 1).
* A `walls` function that draws the layers of a walled dimension gives the
  same populations, bit for bit, as walledDimensions.
* The steady Poiseuille profile, exact with TRT Lambda = 3/16, on D3Q15 and
  D3Q27; mass and momentum conservation and the Guo momentum growth.
* Incompressible equilibrium: sum feq = rho and sum feq c = rho0 u, so one
  BGK tick of a uniform state at density rho != rho0 relaxes the momentum
  j to j - (dt/tau) (j - rho0 j / rho) (TRT: tau-), which the compressible
  form leaves unchanged.
"""
import numpy as np
import pytest

LAT3 = ["D3Q15", "D3Q19", "D3Q27"]


def kw3(vel, **kw):
    return dict(dims=3, velocities=vel, latticeParameter=1.0, kineticParameter=3 / 16, **kw)


def bare_momentum(rm, f):
    return (f.reshape(-1, f.shape[-1]) @ rm.cs.astype(float)).sum(axis=0)


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("vel", LAT3)
def test_couette_profile_exact(oracle, vel, coll):
    Nz, U = 10, 1e-3
    m = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=0.8, collisionModel=coll, xlims=(0, 3), ylims=(0, 2),
                            zlims=(0, Nz - 1), walledDimensions=[3],
                            movingWalls=lambda x, y, z: [U, 0.0, 0.0] if z == Nz - 1 else [0.0, 0.0, 0.0])
    m.run(4000)
    u = m.fluidVelocity
    z = np.arange(Nz)
    exact = U * (z - 0.5) / (Nz - 2)
    exact[0] = exact[-1] = 0.0
    assert np.max(np.abs(u[..., 0] - exact[None, None, :])) < 2e-12 * U
    assert np.max(np.abs(u[..., 1:])) < 1e-15


def test_couette_d2q9_moving_side(oracle):
    Ny, U = 12, -2e-3
    m = oracle.oracle_model(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.9,
                            collisionModel="bgk", xlims=(0, 4), ylims=(0, Ny - 1), walledDimensions=[2],
                            movingWalls=lambda x, y: [U, 0.0] if y == 0 else [0.0, 0.0])
    m.run(5000)
    y = np.arange(Ny)
    exact = U * (Ny - 1.5 - y) / (Ny - 2)
    exact[0] = exact[-1] = 0.0
    assert np.max(np.abs(m.fluidVelocity[..., 0] - exact[None, :])) < 2e-12 * abs(U)


@pytest.mark.parametrize("vel", ["D3Q15", "D3Q27"])
def test_moving_wall_term_one_tick(oracle, vel):
    # fluid at rest next to a moving wall layer: after one tick every
    # population bounced off the wall is w_i + 6 w_i u_w . c_i
    uw = np.array([1.5e-3, -0.7e-3, 0.0])
    m = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=0.8, xlims=(0, 3), ylims=(0, 3), zlims=(0, 5),
                            walledDimensions=[3], movingWalls=lambda x, y, z: list(uw) if z == 0 else [0.0, 0.0, 0.0])
    rm = m.resolved
    m.run(1)
    f = m.distributions
    for i, c in enumerate(rm.cs):
        expect_lo = rm.ws[i] + (6 * rm.ws[i] * float(uw @ c) if c[2] > 0 else 0.0)  # node z = 1
        expect_hi = rm.ws[i]  # node z = 4 next to the resting wall: plain bounce-back
        assert np.allclose(f[:, :, 1, i], expect_lo, rtol=0, atol=2e-16)
        assert np.allclose(f[:, :, 4, i], expect_hi, rtol=0, atol=2e-16)
    assert np.all(f[:, :, 0, :] == 0) and np.all(f[:, :, -1, :] == 0)


@pytest.mark.parametrize("vel", ["D3Q19", "D3Q27"])
def test_walls_function_equals_walled_dimensions(oracle, vel):
    rng = np.random.default_rng(5)
    common = dict(relaxationTimeRatio=0.7, xlims=(0, 4), ylims=(0, 5), zlims=(0, 6), forceDensity=[1e-5, 2e-6, 0.0])
    a = oracle.oracle_model(**kw3(vel), **common, walledDimensions=[2])
    b = oracle.oracle_model(**kw3(vel), **common, walls=lambda x, y, z: y == 0 or y == 5)
    f = a.distributions
    f = f * (1 + 0.02 * rng.standard_normal(f.shape))
    f[a.distributions == 0] = 0.0
    a.distributions = f
    b.distributions = f
    a.run(20)
    b.run(20)
    assert np.array_equal(a.distributions, b.distributions)


@pytest.mark.parametrize("vel", ["D3Q15", "D3Q27"])
def test_poiseuille_exact_other_lattices(oracle, vel):
    Nz, g, tau = 12, 1e-6, 0.9
    m = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=tau, xlims=(0, 3), ylims=(0, 3), zlims=(0, Nz - 1),
                            walledDimensions=[3], forceDensity=[g, 0, 0])
    m.run(3000)
    ux = m.fluidVelocity[1, 1, :, 0]
    nu = (tau - 0.5) / 3
    z = np.arange(Nz)
    exact = g / (2 * nu) * (z - 0.5) * (Nz - 1.5 - z)
    exact[0] = exact[-1] = 0.0
    assert np.max(np.abs(ux - exact)) < 1e-11 * np.max(exact)


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("vel", ["D3Q15", "D3Q27"])
def test_conservation_and_guo_growth_other_lattices(oracle, vel, coll):
    rng = np.random.default_rng(2)
    m = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=0.7, collisionModel=coll, xlims=(0, 4))
    f = m.distributions * (1 + 0.05 * rng.standard_normal(m.distributions.shape))
    m.distributions = f
    M0, J0 = f.sum(), bare_momentum(m.resolved, f)
    m.run(20)
    f1 = m.distributions
    assert abs(f1.sum() - M0) < 1e-12 * M0
    assert np.allclose(bare_momentum(m.resolved, f1), J0, rtol=0, atol=1e-12)
    F = np.array([1e-5, -2e-6, 3e-6])
    g = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=0.9, collisionModel=coll, xlims=(0, 4), forceDensity=F)
    J0 = bare_momentum(g.resolved, g.distributions)
    g.run(10)
    assert np.allclose(bare_momentum(g.resolved, g.distributions) - J0, 10 * g.resolved.N * F, rtol=1e-10, atol=0)


@pytest.mark.parametrize("kw", [dict(dims=2, velocities="D2Q9"), dict(dims=3, velocities="D3Q19"),
                                dict(dims=3, velocities="D3Q27")])
@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_incompressible_equilibrium_moments(oracle, kw, coll):
    rho0, rho, tau = 1.3, 1.1, 0.8
    d = kw["dims"]
    u = np.array([0.02, -0.013, 0.007][:d])
    m = oracle.oracle_model(**kw, latticeParameter=1.0, relaxationTimeRatio=tau, collisionModel=coll,
                            kineticParameter=3 / 16, xlims=(0, 3), massDensity=rho0, isFluidCompressible=False)
    rm = m.resolved
    # a uniform state at rho with velocity u (compressible equilibrium)
    f = np.zeros(tuple(rm.n[:d]) + (rm.Q,), order="F")
    for i, c in enumerate(rm.cs):
        cu = float(c @ u)
        f[..., i] = rm.ws[i] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * float(u @ u))
    m.distributions = f
    j = rho * u
    m.run(1)
    f1 = m.distributions.reshape(-1, rm.Q)
    assert np.allclose(f1.sum(axis=1), rho, rtol=1e-14, atol=0)
    # feq carries rho0 u: the momentum relaxes towards it (BGK 1/tau; TRT
    # through the antisymmetric part, 1/tau-)
    rate = 1 / tau if coll == "bgk" else 1 / rm.tau_minus
    expect = j - rate * (j - rho0 * j / rho)
    assert np.allclose(f1 @ rm.cs.astype(float), expect[None, :], rtol=1e-12, atol=0)
    # the compressible model keeps it
    c = oracle.oracle_model(**kw, latticeParameter=1.0, relaxationTimeRatio=tau, collisionModel=coll,
                            kineticParameter=3 / 16, xlims=(0, 3), massDensity=rho0)
    c.distributions = f
    c.run(1)
    assert np.allclose(c.distributions.reshape(-1, rm.Q) @ rm.cs.astype(float), j[None, :], rtol=1e-13, atol=0)


@pytest.mark.parametrize("vel,r,R", [("D3Q15", 3.3, [0.4, -0.2, 0.3]), ("D3Q27", 3.0, [0.0, 0.0, 0.0])])
def test_links_brute_force_other_lattices(oracle, vel, r, R):
    from test_oracle_kat import brute_links
    m = oracle.oracle_model(**kw3(vel), relaxationTimeRatio=0.8, xlims=(-10, 9))
    m.add("squirmer", radius=r, position=R)
    solid, bn, bd = brute_links(m.resolved, R, r)
    n, d = m.particles[0].links()
    assert np.array_equal(n, bn) and np.array_equal(d, bd)
