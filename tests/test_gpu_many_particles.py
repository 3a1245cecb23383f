"""Lattices with more than 8 particles (the reference's particle vector has no
fixed length, gpu/structs.jl:22, gpu/fluids.jl:158-207): a move kernel
follows every step kernel, and each box block works with the particles
whose owned regions meet its own.  FP64 populations and particle
trajectories against the oracle to <= 1e-12 relative, boundary links
bit-exactly, over runs with graph replays; clusters of up to 8 particles
within reach of one another (links shared between particles, taken in
particle order); the limits refused with an error."""
import numpy as np
import pytest

from test_gpu_parity import add_both, compare_fields, compare_particles, pair

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)


def grid_positions(rng, spacing, shape, jitter=0.4):
    out = []
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                c = (np.array([i, j, k]) - (np.array(shape) - 1) / 2) * spacing
                out.append(c + rng.uniform(-jitter, jitter, 3))
    return out


def add_mix(g, o, rng, positions, r):
    for i, R in enumerate(positions):
        if i % 3 == 0:
            add_both(g, o, "bead", radius=r, position=R, coupleForces=True, coupleTorques=True)
        else:
            add_both(g, o, "squirmer", radius=r, position=R, swimmingSpeed=2e-3 / 3, beta=rng.uniform(-4, 4),
                     swimmingDirection=rng.standard_normal(3), coupleForces=True, coupleTorques=True)


def run_both(g, o, plan=(5, 40)):
    for ticks in plan:  # the second run replays captured graphs
        g.run(ticks)
        g.sync()
        o.run(ticks)
        compare_fields(g, o)
        compare_particles(g, o)


def test_twelve_separate_particles(gpu, oracle):
    rng = np.random.default_rng(21)
    g, o = pair(gpu, oracle, **KW3, xlims=(-32, 31), ylims=(-24, 23), zlims=(-24, 23),
                forceDensity=[1e-6, -5e-7, 0.0])
    add_mix(g, o, rng, grid_positions(rng, 20.0, (3, 2, 2)), 3.5)
    assert len(g.particles) == 12
    run_both(g, o)


def test_clusters_share_links(gpu, oracle):
    # two clusters of 8 (2 x 2 x 2, gaps of about one node between spheres),
    # far enough apart that no block reaches both
    rng = np.random.default_rng(22)
    g, o = pair(gpu, oracle, **KW3, xlims=(-28, 27), ylims=(-14, 13), zlims=(-14, 13))
    pos = [c + np.array([-13.0, 0.0, 0.0]) for c in grid_positions(rng, 7.6, (2, 2, 2), 0.2)]
    pos += [c + np.array([13.0, 0.0, 0.0]) for c in grid_positions(rng, 7.6, (2, 2, 2), 0.2)]
    add_mix(g, o, rng, pos, 3.0)
    links = sum(len(p.links()[0]) for p in g.particles)
    assert links > 0
    run_both(g, o, plan=(4, 36))


def test_interactions_with_many_particles(gpu, oracle):
    from lbmengine_hip import interactions as it
    rng = np.random.default_rng(23)
    g, o = pair(gpu, oracle, **KW3, xlims=(-30, 29), ylims=(-12, 11), zlims=(-12, 11))
    pos = [np.array([x, 0.5 * (-1) ** i, 0.0]) + rng.uniform(-0.3, 0.3, 3)
           for i, x in enumerate(np.linspace(-24, 24, 10))]
    add_mix(g, o, rng, pos, 2.5)
    for m in (g, o):
        for i in range(1, 10):
            it.addLinearBond(m, i, i + 1, equilibriumDisplacement=5.0, stiffness=0.02)
        it.addPolarBond(m, 4, 5, 6, equilibriumAngle=2.8, stiffness=0.05)
        it.addDipoles(m, 1, 10, magneticField=lambda t: [0.2, 0.3 * np.cos(0.05 * t), 0.1], permeability=1.0)
    run_both(g, o, plan=(6, 40))


def test_d2q9_many_disks(gpu, oracle):
    rng = np.random.default_rng(24)
    g, o = pair(gpu, oracle, dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.9,
                collisionModel="bgk", xlims=(-40, 39), ylims=(-30, 29))
    pos = [np.array([x, y]) + rng.uniform(-0.3, 0.3, 2) for x in (-27, -9, 9, 27) for y in (-16, 0, 16)]
    for i, R in enumerate(pos):
        if i % 2:
            add_both(g, o, "bead", radius=3.0, position=R, coupleForces=True, coupleTorques=True)
        else:
            th = rng.uniform(0, 2 * np.pi)
            add_both(g, o, "squirmer", radius=3.0, position=R, swimmingSpeed=2e-3 / 3, beta=1.0,
                     swimmingDirection=[np.cos(th), np.sin(th)], coupleForces=True, coupleTorques=True)
    run_both(g, o)


def test_reduced_precision_many_particles(gpu):
    import lbmengine_hip as lh
    rng = np.random.default_rng(25)
    pos = grid_positions(rng, 18.0, (3, 3, 1))
    runs = {}
    for st in ("f64", "f32"):
        m = lh.modelInit(**KW3, xlims=(-30, 29), ylims=(-30, 29), zlims=(-12, 11), storage=st)
        r2 = np.random.default_rng(26)
        for R in pos:
            lh.addSquirmer(m, radius=4.0, position=R, swimmingSpeed=2e-3 / 3, beta=r2.uniform(-3, 3),
                           swimmingDirection=r2.standard_normal(3), coupleForces=True, coupleTorques=True)
        m.autotune(3)
        lh.LBMpropagate(m, ticks=200)
        runs[st] = (m.fluidVelocity, [p.velocity for p in m.particles])
    e = np.linalg.norm(runs["f32"][0] - runs["f64"][0]) / np.linalg.norm(runs["f64"][0])
    assert e < 1e-4
    for a, b in zip(runs["f32"][1], runs["f64"][1]):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4


def test_limits_are_refused(gpu):
    import lbmengine_hip as lh
    from lbmengine_hip import adjoint
    # nine beads within reach of one another
    m = lh.modelInit(**KW3, xlims=(-24, 23))
    for R in grid_positions(np.random.default_rng(27), 6.0, (3, 3, 1), 0.1):
        lh.addBead(m, radius=2.5, position=R, coupleForces=True, coupleTorques=True)
    m.run(2)
    with pytest.raises(lh.LBMError, match="within reach"):
        m.sync()
    # the adjoint, a box of a batch and the allocation-free run take at most 8
    m = lh.modelInit(**KW3, xlims=(-32, 31))
    for R in grid_positions(np.random.default_rng(28), 20.0, (3, 3, 1)):
        lh.addSquirmer(m, radius=3.0, position=R, swimmingSpeed=2e-3 / 3, coupleForces=True, coupleTorques=True)
    with pytest.raises(lh.LBMError):
        adjoint.swimming_loss_gradient(m, ticks=4, checkpoint_every=2)
    with pytest.raises(lh.LBMError):
        b = lh.LBMbatch([m])
        b.run(2)
    assert m._fn["lbm_run_noalloc"](m._h, 2) == -1
    m.run(3)  # the lattice itself runs
    m.sync()
