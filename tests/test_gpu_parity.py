"""HIP FP64 step against the oracle on identical seeded inputs.

Fields (populations, rho, u) must agree to <= 1e-12 relative to the field's
scale; particle trajectories to <= 1e-12 relative (the momentum-exchange sums
run in a different order on the device, which the tolerance absorbs); the
boundary-link lists bit-exactly.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0)
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0)
RTOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def perturb(model, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    f = model.distributions
    f = f * (1 + amp * rng.standard_normal(f.shape))
    return np.asfortranarray(f)


def pair(gpu, oracle, seed=None, **kw):
    g = gpu.modelInit(**kw)
    o = oracle.oracle_model(**kw)
    if seed is not None:
        f = perturb(o, seed)
        # walls keep zero populations
        f[o.distributions == 0] = 0.0
        g.distributions = f
        o.distributions = f
    return g, o


def compare_fields(g, o):
    assert rel(g.distributions, o.distributions) <= RTOL
    assert rel(g.massDensity, o.massDensity) <= RTOL
    assert rel(g.fluidVelocity, o.fluidVelocity) <= RTOL


def compare_particles(g, o, rtol=RTOL):
    for pg, po in zip(g.particles, o.particles):
        assert rel(pg.position, po.position) <= rtol
        assert rel(pg.velocity, po.velocity) <= rtol
        assert rel(pg.angularVelocity, po.angularVelocity) <= rtol
        assert rel(pg.swimmingDirection, po.swimmingDirection) <= rtol
        ng, dg = pg.links()
        no, do = po.links()
        assert np.array_equal(ng, no) and np.array_equal(dg, do)


@pytest.mark.parametrize("coll", ["trt", "bgk"])
def test_periodic_d3q19_random_state(gpu, oracle, coll):
    g, o = pair(gpu, oracle, seed=11, **KW3, relaxationTimeRatio=0.77, collisionModel=coll,
                kineticParameter=3 / 16, xlims=(0, 15), ylims=(0, 11), zlims=(0, 9))
    g.run(20)
    g.sync()
    o.run(20)
    compare_fields(g, o)
    assert g.tick == o.tick == 20


@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_periodic_d2q9_random_state(gpu, oracle, coll):
    g, o = pair(gpu, oracle, seed=5, **KW2, relaxationTimeRatio=0.66, collisionModel=coll,
                xlims=(0, 19), ylims=(0, 15))
    g.run(25)
    o.run(25)
    compare_fields(g, o)


def test_walls_and_uniform_guo_force(gpu, oracle):
    g, o = pair(gpu, oracle, seed=3, **KW3, relaxationTimeRatio=0.9, kineticParameter=3 / 16,
                xlims=(0, 11), ylims=(0, 9), zlims=(0, 13), walledDimensions=[3], forceDensity=[1e-5, 2e-6, 0])
    g.run(30)
    o.run(30)
    compare_fields(g, o)


def test_walls_two_dims_and_force_field(gpu, oracle):
    rng = np.random.default_rng(21)
    F = 1e-5 * rng.standard_normal((10, 12, 9, 3))
    g, o = pair(gpu, oracle, seed=4, **KW3, relaxationTimeRatio=0.85, kineticParameter=0.25,
                xlims=(0, 9), ylims=(0, 11), zlims=(0, 8), walledDimensions=[1, 2], forceDensity=F)
    g.run(30)
    o.run(30)
    compare_fields(g, o)


def test_d2q9_bgk_walls_force(gpu, oracle):
    g, o = pair(gpu, oracle, seed=8, **KW2, relaxationTimeRatio=0.8, collisionModel="bgk",
                xlims=(0, 17), ylims=(0, 12), walledDimensions=[2], forceDensity=[3e-6, 0.0])
    g.run(40)
    o.run(40)
    compare_fields(g, o)


def add_both(g, o, kind, **kw):
    import lbmengine_hip
    if kind == "squirmer":
        lbmengine_hip.addSquirmer(g, **kw)
    else:
        lbmengine_hip.addBead(g, **kw)
    o.add(kind, **kw)


@pytest.mark.parametrize("beta", [-5.0, 0.0, 5.0])
def test_squirmer_d3q19_gpu_semantics(gpu, oracle, beta):
    rng = np.random.default_rng(int(10 + beta))
    e = rng.standard_normal(3)
    g, o = pair(gpu, oracle, **KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    kw = dict(radius=5.0, position=rng.uniform(-2, 2, 3), swimmingSpeed=2e-3 / 3, beta=beta,
              swimmingDirection=e, coupleForces=True, coupleTorques=True)
    add_both(g, o, "squirmer", **kw)
    compare_particles(g, o)
    g.run(40)
    g.sync()
    o.run(40)
    compare_fields(g, o)
    compare_particles(g, o)


def test_d2q9_cpu_semantics_squirmer(gpu, oracle):
    # the reference CPU path's plain method: D2Q9 BGK, a squirmer disk, Ladd
    rng = np.random.default_rng(1)
    th = rng.uniform(0, 2 * math.pi)
    g, o = pair(gpu, oracle, **KW2, relaxationTimeRatio=0.8, collisionModel="bgk", xlims=(-24, 23),
                semantics="cpu")
    add_both(g, o, "squirmer", radius=8.0, position=[0.0, 0.0], swimmingSpeed=2e-3 / 3, beta=0.0,
             swimmingDirection=[math.cos(th), math.sin(th)], coupleForces=True, coupleTorques=True)
    g.run(60)
    o.run(60)
    compare_fields(g, o)
    compare_particles(g, o)


def test_bead_in_driven_flow_and_two_particles(gpu, oracle):
    g, o = pair(gpu, oracle, **KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-14, 13),
                forceDensity=[2e-6, 0, 0])
    add_both(g, o, "bead", radius=3.5, position=[-5.3, 0.2, 0.4], coupleForces=True, coupleTorques=True)
    add_both(g, o, "squirmer", radius=4.0, position=[5.1, -0.3, 1.2], swimmingSpeed=1e-3, beta=2.0,
             swimmingDirection=[0, 1, 1], coupleForces=True, coupleTorques=True)
    g.run(30)
    o.run(30)
    compare_fields(g, o)
    compare_particles(g, o)


def test_uncoupled_pumper_is_static(gpu, oracle):
    g, o = pair(gpu, oracle, **KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-10, 9))
    add_both(g, o, "squirmer", radius=4.0, position=[0.5, 0.25, 0.0], swimmingSpeed=1e-3, beta=-3.0)
    g.run(25)
    o.run(25)
    compare_fields(g, o)
    compare_particles(g, o)
    assert np.array_equal(g.particles[0].position, [0.5, 0.25, 0.0])


def test_config1_d2q9_squirmer_1000_steps(gpu, oracle):
    """128^2 periodic D2Q9 BGK, tau 0.8, squirmer disk r = 8 at the centre,
    random orientation from the seed, B1 = 1e-3, beta = 0, Ladd with momentum
    exchange, 1000 steps, on the reference CPU path's semantics."""
    rng = np.random.default_rng(2024)
    th = rng.uniform(0, 2 * math.pi)
    kw = dict(**KW2, relaxationTimeRatio=0.8, collisionModel="bgk", xlims=(-64, 63), semantics="cpu")
    g, o = pair(gpu, oracle, **kw)
    add_both(g, o, "squirmer", radius=8.0, position=[0.0, 0.0], swimmingSpeed=2e-3 / 3, beta=0.0,
             swimmingDirection=[math.cos(th), math.sin(th)], coupleForces=True, coupleTorques=True)
    g.run(1000)
    g.sync()
    o.run(1000)
    compare_fields(g, o)
    compare_particles(g, o)


def test_links_after_motion_bit_exact(gpu, oracle):
    g, o = pair(gpu, oracle, **KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-12, 11))
    add_both(g, o, "squirmer", radius=4.6, position=[0.49, 0.51, -0.5], swimmingSpeed=4e-2, beta=1.0,
             swimmingDirection=[1, 1, 1], coupleForces=True, coupleTorques=True)
    for _ in range(4):
        g.run(10)
        o.run(10)
        ng, dg = g.particles[0].links()
        no, do = o.particles[0].links()
        assert np.array_equal(ng, no) and np.array_equal(dg, do)


def test_particle_near_edge_is_reported(gpu):
    import lbmengine_hip
    m = gpu.modelInit(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-8, 7))
    with pytest.raises(lbmengine_hip.LBMError):
        lbmengine_hip.addSquirmer(m, radius=3.0, position=[5.5, 0, 0])


@pytest.mark.parametrize("ticks_per_graph", [2, 6])
def test_graph_replays_match_direct_launches(gpu, oracle, ticks_per_graph):
    """Captured hipGraphs replayed several times, with eager work and syncs
    between the replays, give bit-identical states to direct launches."""
    import lbmengine_hip
    kw = dict(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-10, 9), forceDensity=[1e-6, 0, 0])
    sq = dict(radius=4.0, position=[0.3, -0.4, 0.2], swimmingSpeed=1e-3, beta=-2.0, swimmingDirection=[1, 2, 0],
              coupleForces=True, coupleTorques=True)
    a = gpu.modelInit(**kw)
    b = gpu.modelInit(**kw)
    lbmengine_hip.addSquirmer(a, **sq)
    lbmengine_hip.addSquirmer(b, **sq)
    a.set_graph(ticks_per_graph)
    b.set_graph(0)
    for steps in (7, ticks_per_graph * 3, 5, ticks_per_graph * 2 + 1):
        a.run(steps)
        b.run(steps)
        fa, fb = a.distributions, b.distributions  # eager work between replays
        assert np.array_equal(fa, fb)
        assert np.array_equal(a.particles[0].position, b.particles[0].position)
        assert np.array_equal(a.particles[0].velocity, b.particles[0].velocity)
    assert a.tick == b.tick


def test_autotune_keeps_state_and_results(gpu, oracle):
    kw = dict(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(0, 47), ylims=(0, 31), zlims=(0, 23))
    g, o = pair(gpu, oracle, seed=9, **kw)
    f0 = g.distributions
    shape, ms = g.autotune(3)
    assert ms > 0 and np.prod(shape) >= 64 and tuple(g.get_block()) == tuple(shape)
    assert np.array_equal(g.distributions, f0)
    g.run(12)
    o.run(12)
    compare_fields(g, o)


@pytest.mark.parametrize("block", [(64, 1, 1), (16, 4, 1), (8, 4, 2), (32, 2, 4)])
def test_particle_tiles_any_block_shape(gpu, oracle, block):
    """Two squirmers whose link boxes overlap (a one-node gap, so some link
    slots belong to both): every block shape splits their boxes into a
    different set of tiles, and the results match the oracle."""
    kw = dict(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-13, 12), ylims=(-9, 10),
              zlims=(-8, 9))
    g, o = pair(gpu, oracle, seed=21, **kw)
    g.set_block(*block)
    add_both(g, o, "squirmer", radius=4.0, position=[-4.6, 0.3, 0.1], swimmingSpeed=2e-3, beta=1.0,
             swimmingDirection=[1, 0.2, 0], coupleForces=True, coupleTorques=True)
    add_both(g, o, "bead", radius=4.0, position=[4.6, -0.2, 0.3], coupleForces=True, coupleTorques=True)
    g.run(16)
    o.run(16)
    compare_fields(g, o)
    compare_particles(g, o)


def test_d2q9_particle_tiles_block_shape(gpu, oracle):
    kw = dict(**KW2, relaxationTimeRatio=0.8, collisionModel="trt", xlims=(-30, 29), ylims=(-20, 21))
    g, o = pair(gpu, oracle, seed=4, **kw)
    g.set_block(16, 8, 1)
    add_both(g, o, "squirmer", radius=6.5, position=[0.4, 0.7], swimmingSpeed=3e-3, beta=-1.0,
             swimmingDirection=[0.3, 1.0], coupleForces=True, coupleTorques=True)
    g.run(30)
    o.run(30)
    compare_fields(g, o)
    compare_particles(g, o)


def test_autotune_with_particle_keeps_state(gpu, oracle):
    kw = dict(**KW3, relaxationTimeRatio=0.8, kineticParameter=3 / 16, xlims=(-16, 15), ylims=(-12, 11),
              zlims=(-10, 9))
    g, o = pair(gpu, oracle, **kw)
    sq = dict(radius=5.0, position=[0.3, 0.2, -0.1], swimmingSpeed=2e-3, beta=-1.0, swimmingDirection=[0, 0, 1],
              coupleForces=True, coupleTorques=True)
    add_both(g, o, "squirmer", **sq)
    g.run(3)
    o.run(3)
    f0 = g.distributions
    p0 = g.particles[0].position
    g.autotune(2)
    assert np.array_equal(g.distributions, f0)
    assert np.array_equal(g.particles[0].position, p0)
    g.run(10)
    o.run(10)
    compare_fields(g, o)
    compare_particles(g, o)
