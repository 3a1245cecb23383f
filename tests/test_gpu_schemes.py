"""The other schemes on the device against the oracle, FP64, on identical
seeded inputs: the D3Q15 and D3Q27 velocity sets (with particles), walls
and moving walls given as functions (gpu/initMethods.jl:580-626,
gpu/fluids.jl:149-174), the incompressible equilibrium (gpu/fluids.jl:51),
and walled lattices large enough for the tiles away from the walls to take
the kernel's periodic fast path.  Fields <= 1e-12 relative, particles <=
1e-12, links bit-exact; the reduced-precision storages against FP64."""
import numpy as np
import pytest

from test_gpu_parity import add_both, compare_fields, compare_particles, pair, rel

pytestmark = pytest.mark.gpu


def kw3(vel, **kw):
    return dict(dims=3, velocities=vel, latticeParameter=1.0, **kw)


@pytest.mark.parametrize("coll", ["trt", "bgk"])
@pytest.mark.parametrize("vel", ["D3Q15", "D3Q27"])
def test_other_lattices_random_state_with_force_and_walls(gpu, oracle, vel, coll):
    g, o = pair(gpu, oracle, seed=13, **kw3(vel), relaxationTimeRatio=0.83, collisionModel=coll,
                kineticParameter=3 / 16, xlims=(0, 13), ylims=(0, 10), zlims=(0, 9), walledDimensions=[2],
                forceDensity=[2e-6, 0.0, -1e-6])
    g.run(25)
    o.run(25)
    compare_fields(g, o)


@pytest.mark.parametrize("vel,beta", [("D3Q15", -5.0), ("D3Q27", 5.0), ("D3Q27", 0.0)])
def test_other_lattices_squirmer(gpu, oracle, vel, beta):
    rng = np.random.default_rng(40 + int(beta))
    g, o = pair(gpu, oracle, **kw3(vel), relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    add_both(g, o, "squirmer", radius=5.0, position=rng.uniform(-2, 2, 3), swimmingSpeed=2e-3 / 3, beta=beta,
             swimmingDirection=rng.standard_normal(3), coupleForces=True, coupleTorques=True)
    compare_particles(g, o)
    g.run(40)
    o.run(40)
    compare_fields(g, o)
    compare_particles(g, o)


@pytest.mark.parametrize("kw", [dict(dims=2, velocities="D2Q9", xlims=(-16, 15)),
                                dict(dims=3, velocities="D3Q19", xlims=(-10, 9))])
def test_incompressible_with_particle_and_force(gpu, oracle, kw):
    d = kw["dims"]
    g, o = pair(gpu, oracle, seed=6, **kw, latticeParameter=1.0, relaxationTimeRatio=0.75, kineticParameter=3 / 16,
                massDensity=1.2, isFluidCompressible=False, forceDensity=[2e-6, -1e-6, 5e-7][:d])
    add_both(g, o, "squirmer", radius=4.0, position=[0.3, -0.4, 0.2][:d], swimmingSpeed=2e-3 / 3, beta=2.0,
             swimmingDirection=[1.0, 0.5, -0.3][:d], coupleForces=True, coupleTorques=True)
    g.run(30)
    o.run(30)
    compare_fields(g, o)
    compare_particles(g, o)
    s = gpu.viscousStressTensor(g)
    sg = np.stack([np.stack([s[a][b] for b in range(d)], axis=-1) for a in range(d)], axis=-2)
    assert rel(sg, o.viscous_stress()) <= 1e-12


def obstacle(x, y, z):
    return bool((x - 3.0) ** 2 + (y + 2.0) ** 2 < 9.0)  # a cylinder along z


@pytest.mark.parametrize("vel", ["D3Q19", "D3Q27"])
def test_walls_function_obstacle_and_particle(gpu, oracle, vel):
    g, o = pair(gpu, oracle, seed=9, **kw3(vel), relaxationTimeRatio=0.85, kineticParameter=3 / 16,
                xlims=(-12, 11), ylims=(-10, 9), zlims=(-8, 7), walledDimensions=[3], walls=obstacle,
                forceDensity=[3e-6, 0.0, 0.0])
    add_both(g, o, "squirmer", radius=3.5, position=[-6.2, 3.1, 0.4], swimmingSpeed=2e-3 / 3, beta=-3.0,
             swimmingDirection=[1.0, -0.2, 0.1], coupleForces=True, coupleTorques=True)
    g.run(35)
    o.run(35)
    compare_fields(g, o)
    compare_particles(g, o)
    # wall nodes hold zero populations, rho and u
    rm = g.resolved
    w = rm.wall_mask()
    assert np.all(g.distributions[w] == 0) and np.all(g.massDensity[w] == 0)


@pytest.mark.parametrize("vel", ["D3Q15", "D3Q19", "D3Q27"])
def test_lid_driven_cavity(gpu, oracle, vel):
    # walls on every side, the lid (z top layer) moving along x and y
    n = 12
    lid = lambda x, y, z: [1e-3, 4e-4, 0.0] if z == n - 1 and 0 < x < n - 1 and 0 < y < n - 1 else [0.0] * 3
    g, o = pair(gpu, oracle, **kw3(vel), relaxationTimeRatio=0.7, kineticParameter=3 / 16, xlims=(0, n - 1),
                walledDimensions=[1, 2, 3], movingWalls=lid)
    g.run(60)
    o.run(60)
    compare_fields(g, o)
    assert np.max(np.abs(g.fluidVelocity)) > 1e-4


def test_moving_walls_d2q9_with_bead(gpu, oracle):
    g, o = pair(gpu, oracle, dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.9,
                collisionModel="bgk", xlims=(-16, 15), ylims=(-10, 9), walledDimensions=[2],
                movingWalls=lambda x, y: [2e-3, 0.0] if y == 9 else [0.0, 0.0])
    add_both(g, o, "bead", radius=3.0, position=[0.4, 0.3], coupleForces=True, coupleTorques=True)
    g.run(50)
    o.run(50)
    compare_fields(g, o)
    compare_particles(g, o)


def test_walled_channel_interior_tiles(gpu, oracle):
    # 64 x 16 x 24 with walls at z = 0 and 23 and 64 x 1 x 1 blocks: the
    # tiles of planes 2 .. 21 take the periodic fast path, the others the
    # general one; graph replays of both
    g, o = pair(gpu, oracle, seed=2, **kw3("D3Q19"), relaxationTimeRatio=0.9, kineticParameter=3 / 16,
                xlims=(0, 63), ylims=(0, 15), zlims=(0, 23), walledDimensions=[3], forceDensity=[1e-6, 0.0, 0.0])
    g.set_block(64, 1, 1)
    g.run(70)
    o.run(70)
    compare_fields(g, o)


@pytest.mark.parametrize("storage,bound", [("f32", 1e-4), ("f16c", 2e-3)])
def test_reduced_precision_moving_walls(gpu, storage, bound):
    import lbmengine_hip as lh
    kw = dict(**kw3("D3Q19"), relaxationTimeRatio=0.7, kineticParameter=3 / 16, xlims=(0, 15),
              walledDimensions=[1, 2, 3], movingWalls=lambda x, y, z: [1e-3, 0.0, 0.0] if z == 15 else [0.0] * 3)
    ref = lh.modelInit(**kw, storage="f64")
    low = lh.modelInit(**kw, storage=storage)
    ref.run(400)
    low.run(400)
    e = float(np.linalg.norm(low.fluidVelocity - ref.fluidVelocity) / np.linalg.norm(ref.fluidVelocity))
    print(f"{storage}: lid-driven cavity u rel L2 {e:.3e}")
    assert e < bound
