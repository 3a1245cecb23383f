"""Host-side restatement of modelInit / addSquirmer! parameter resolution
(gpu/initMethods.jl) and of LBMpropagate!'s tick count."""
import math

import numpy as np
import pytest

from lbmengine_hip import julia_range, propagate_ticks, resolve_model, resolve_particle, sphere_count
from lbmengine_hip import _capi


def test_julia_range_integer_limits():
    x = julia_range(-96, 95, 1.0)
    assert len(x) == 192 and x[0] == -96.0 and x[-1] == 95.0
    assert np.all(np.diff(x) == 1.0)


def test_julia_range_decimal_step_is_rational():
    # Julia: collect(range(0, stop=1, step=0.1))[4] == 0.3 (not 0.30000000000000004)
    x = julia_range(0, 1, 0.1)
    assert len(x) == 11
    assert x[3] == 0.3 and x[7] == 0.7 and x[-1] == 1.0


def test_units_with_lattice_parameter():
    rm = resolve_model(dims=3, velocities=":D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                       kineticParameter=3 / 16, xlims=(-8, 7))
    assert rm.dt == 1.0 and rm.dx == 1.0 and rm.latticeSpeed == 1.0
    assert rm.tau == 0.8
    assert math.isclose(rm.viscosity, 0.1, rel_tol=1e-15)
    # tau- = Lambda c_s^2 dt^2 / nu + dt/2 = (3/16)(1/3)/0.1 + 1/2
    assert math.isclose(rm.tau_minus, 1.125, rel_tol=1e-15)
    assert rm.n == (16, 16, 16) and rm.collision == "trt"


def test_units_defaults():
    # no units given: relaxationTimeRatio 15, viscosity 1, dt = 3 nu/(tau - 1/2)
    rm = resolve_model(dims=2, velocities="D2Q9", xlims=(-1, 1))
    assert math.isclose(rm.dt, 3 / 14.5, rel_tol=1e-15)
    assert rm.dx == rm.dt and rm.tau == 15 * rm.dt


def test_debug_kinetic_parameter_makes_trt_bgk():
    rm = resolve_model(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.9,
                       kineticParameter="debug", xlims=(0, 7))
    assert math.isclose(rm.tau_minus, rm.tau, rel_tol=1e-14)


def test_default_velocity_set_is_d3q15():
    # gpu/initMethods.jl:534: D3Q15 unless velocities is given
    rm = resolve_model(dims=3, latticeParameter=1.0, xlims=(0, 7))
    assert rm.velocities == "D3Q15" and rm.Q == 15


def test_walls_functions_resolution():
    # walls(x, y, z) and movingWalls(x, y, z) over the grid, x fastest
    # (gpu/initMethods.jl:591-626); moving nodes join the wall region
    rm = resolve_model(dims=3, velocities="D3Q27", latticeParameter=1.0, xlims=(0, 5), ylims=(0, 4), zlims=(0, 3),
                       walls=lambda x, y, z: x == 0 and y == 2,
                       movingWalls=lambda x, y, z: [0.1, 0.0, 0.0] if z == 3 else [0.0, 0.0, 0.0])
    assert rm.wall_region.shape == (6, 5, 4) and rm.wall_region[0, 2, :].all() and rm.wall_region.sum() == 4
    assert rm.wall_velocity.shape == (6, 5, 4, 3) and np.all(rm.wall_velocity[:, :, 3, 0] == 0.1)
    m = rm.wall_mask()
    assert m[:, :, 3].all() and m[0, 2, :].all() and m.sum() == 6 * 5 + 3
    assert "movingWalls" in rm.schemes and "bounceBack" in rm.schemes
    with pytest.raises(AssertionError):
        resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, xlims=(0, 3), walls=lambda x, y, z: 1)
    with pytest.raises(AssertionError):
        resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, xlims=(0, 3),
                      movingWalls=lambda x, y, z: [0.0, 0.0])


def test_incompressible_resolution():
    rm = resolve_model(dims=2, velocities="D2Q9", latticeParameter=1.0, xlims=(0, 7), massDensity=1.3,
                       isFluidCompressible=False)
    cfg, _ = rm.config()
    assert rm.incompressible and cfg.flags == _capi.FLAG_INCOMPRESSIBLE and cfg.rho0 == 1.3
    with pytest.raises(NotImplementedError):
        resolve_model(dims=2, velocities="D2Q9", latticeParameter=1.0, xlims=(0, 3),
                      massDensity=np.ones((4, 4)), isFluidCompressible=False)


def test_walls_and_force_resolution():
    rm = resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.9,
                       kineticParameter=3 / 16, xlims=(0, 15), ylims=(0, 7), zlims=(0, 9),
                       walledDimensions=[3], forceDensity=[1e-6, 0, 0])
    assert rm.walled == 4 and "bounceBack" in rm.schemes and "guo" in rm.schemes
    assert rm.force_mode == _capi.FORCE_UNIFORM
    F = rm.force_array()
    assert np.all(F[:, :, 0, :] == 0) and np.all(F[:, :, -1, :] == 0) and F[3, 3, 3, 0] == 1e-6


def test_sphere_mass_counts():
    # nodes strictly inside the origin-centred sphere (gpu/initMethods.jl:108)
    rm3 = resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                        xlims=(-96, 95))
    assert sphere_count(rm3, 12.0) == 7123
    assert sphere_count(rm3, 10.0) == 4139
    assert sphere_count(rm3, 6.0) == 895
    rm2 = resolve_model(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8,
                        xlims=(-64, 63), collisionModel="bgk")
    assert sphere_count(rm2, 8.0) == 193


def test_squirmer_b1_override_and_inertia():
    rm = resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                       xlims=(-16, 15))
    p = resolve_particle(rm, kind="squirmer", radius=4.0, swimmingSpeed=2e-3 / 3, beta=-5,
                         swimmingDirection=[0, 2, 0], coupleForces=True, coupleTorques=True)
    # B1 := 3/2 |swimmingSpeed| (gpu/initMethods.jl:161-167), B2 = beta B1
    assert math.isclose(p.B1, 1e-3, rel_tol=1e-15) and math.isclose(p.B2, -5e-3, rel_tol=1e-15)
    assert list(p.swimming_direction) == [0.0, 1.0, 0.0]
    m = sphere_count(rm, 4.0)
    assert math.isclose(1 / p.inverse_mass, m, rel_tol=1e-15)
    assert math.isclose(1 / p.inverse_moment_of_inertia, 0.4 * m * 16, rel_tol=1e-15)
    # negative swimming speed flips the direction
    p2 = resolve_particle(rm, kind="squirmer", radius=4.0, swimmingSpeed=-1e-3, swimmingDirection=[1, 0, 0])
    assert p2.swimming_direction[0] == -1.0 and p2.B1 == 1.5e-3


def test_squirmer_slip_speed_normalisation():
    rm = resolve_model(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8,
                       xlims=(-16, 15))
    p = resolve_particle(rm, kind="squirmer", radius=4.0, slipSpeed=1e-3, B1=1, B2=0)
    # max |(B1 + B2 cos t) sin t| over 100 points of [-pi, pi] rescaled to slipSpeed
    th = np.linspace(-math.pi, math.pi, 100)
    assert math.isclose(p.B1, 1e-3 / np.max(np.abs(np.sin(th))), rel_tol=1e-14)


def test_particle_argument_errors():
    rm = resolve_model(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                       xlims=(-16, 15))
    with pytest.raises(AssertionError):
        resolve_particle(rm, kind="squirmer", radius=4.0, B1=1, B2=1, beta=1)
    with pytest.raises(AssertionError):
        resolve_particle(rm, kind="squirmer", radius=4.0, slipSpeed=1e-3, swimmingSpeed=1e-3)
    with pytest.raises(ValueError):
        resolve_particle(rm, kind="squirmer", radius=4.0, position=[0, 0])


def test_propagate_ticks():
    assert propagate_ticks(1.0) == 100
    assert propagate_ticks(1.0, ticks=7) == 7
    assert propagate_ticks(0.5, simulationTime=2.0) == 4
    with pytest.raises(AssertionError):
        propagate_ticks(1.0, simulationTime=2.0, ticks=3)
