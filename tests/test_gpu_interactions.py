"""Particle-particle interactions on the device (the step kernels apply every
term before the tick's moves) against the oracle's molecularDynamicsTick!
(gpu/particles.jl:100-197): FP64 populations and particle trajectories to
<= 1e-12 relative, boundary links bit-exactly, over runs long enough for the
bonds to move the particles across nodes and for graph replays to run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12
KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16,
           xlims=(-16, 15))
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, xlims=(-24, 23))


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def build(gpu, oracle, kw, parts, add):
    from lbmengine_hip import interactions as it
    g = gpu.modelInit(**kw)
    o = oracle.oracle_model(**kw)
    for kind, pk in parts:
        (gpu.addSquirmer if kind == "squirmer" else gpu.addBead)(g, **pk)
        o.add(kind, **pk)
    add(it, g)
    add(it, o)
    return g, o


def compare(g, o):
    assert rel(g.distributions, o.distributions) <= RTOL
    for pg, po in zip(g.particles, o.particles):
        for attr in ("position", "velocity", "angularVelocity", "swimmingDirection"):
            assert rel(getattr(pg, attr), getattr(po, attr)) <= RTOL, attr
        ng, dg = pg.links()
        no, do = po.links()
        assert np.array_equal(ng, no) and np.array_equal(dg, do)


def bead(pos, r=3.0):
    return ("bead", dict(radius=r, position=pos, coupleForces=True, coupleTorques=True))


def squirmer(pos, e, r=3.0, beta=1.0):
    return ("squirmer", dict(radius=r, position=pos, swimmingSpeed=2e-3, beta=beta, swimmingDirection=e,
                             coupleForces=True, coupleTorques=True))


CASES = {
    "linear_3d": (KW3, [bead([-4.2, 0.3, 0.1]), squirmer([5.1, -1.2, 0.7], [1, 0, 0])],
                  lambda it, m: it.addLinearBond(m, 1, 2, equilibriumDisplacement=7.0, stiffness=0.05)),
    "bistable_2d": (KW2, [bead([-4.0, 0.2]), bead([4.5, 0.9])],
                    lambda it, m: it.addBistableBond(m, 1, 2, lowDisp=7.0, highDisp=9.5, hillHeight=0.02)),
    "polar_3d": (KW3, [bead([-7.0, 3.1, 0.8]), bead([0.2, -0.4, 0.3]), bead([7.0, 2.0, -1.1])],
                 lambda it, m: it.addPolarBond(m, 1, 2, 3, equilibriumAngle=1.3, stiffness=0.2)),
    "polar_2d": (KW2, [bead([-7.0, 3.1]), bead([0.2, -0.4]), bead([7.0, 2.0])],
                 lambda it, m: it.addPolarBond(m, 1, 2, 3, equilibriumAngle=2.0, stiffness=0.1)),
    "dipoles_3d": (KW3, [bead([-5.0, 1.0, 0.5]), bead([4.0, -1.5, 2.0]), bead([0.5, 7.0, -4.0])],
                   lambda it, m: it.addDipoles(m, magneticField=[0.3, -0.5, 0.8], susceptibility=0.5,
                                               permeability=1.0)),
    "dipoles_2d_scalar": (KW2, [bead([-5.0, 1.0]), bead([4.0, -1.5])],
                          lambda it, m: it.addDipoles(m, 1, 2, magneticField=0.9, permeability=1.0)),
    "lennard_jones_3d": (KW3, [bead([-3.9, 0.5, 0.0]), bead([4.0, 0.0, 0.6]), bead([0.0, 8.5, -0.5])],
                         lambda it, m: it.addLennardJones(m, epsilon=0.02, sigma=7.8, cutoff=9.0)),
    "wca_2d": (KW2, [bead([-3.6, 0.0]), bead([3.6, 0.3])], lambda it, m: it.addWCA(m, epsilon=0.05)),
    # fields that change with time: read tick by tick from the device's
    # schedule (the oracle calls the function at every tick)
    "dipoles_3d_rotating": (KW3, [bead([-5.0, 1.0, 0.5]), bead([4.0, -1.5, 2.0]), bead([0.5, 7.0, -4.0])],
                            lambda it, m: it.addDipoles(m, magneticField=lambda t: [0.6 * np.cos(0.07 * t),
                                                                                     0.6 * np.sin(0.07 * t), 0.3],
                                                        susceptibility=0.5, permeability=1.0)),
    "dipoles_2d_pulsed_scalar": (KW2, [bead([-5.0, 1.0]), bead([4.0, -1.5])],
                                 lambda it, m: it.addDipoles(m, 1, 2, magneticField=lambda t: 0.9 if t % 8 < 4 else 0.0,
                                                             permeability=1.0)),
    "all_3d": (KW3, [squirmer([-5.0, 0.5, 0.2], [0, 1, 0]), bead([3.5, -1.0, 0.4]), bead([0.5, 7.0, -3.0])],
               lambda it, m: (it.addLinearBond(m, 1, 2, stiffness=0.03, equilibriumDisplacement=8.0),
                              it.addBistableBond(m, 2, 3, hillHeight=0.2),
                              it.addPolarBond(m, 1, 2, 3, equilibriumAngle=1.0, stiffness=0.1),
                              it.addDipoles(m, magneticField=[0, 0.2, 0.4], permeability=1.0),
                              it.addLennardJones(m, epsilon=1e-3, sigma=8.0))),
}


@pytest.mark.parametrize("case", list(CASES))
def test_interactions_against_oracle(gpu, oracle, case):
    kw, parts, add = CASES[case]
    g, o = build(gpu, oracle, kw, parts, add)
    x0 = [p.position.copy() for p in g.particles]
    for ticks in (5, 40):  # the second run replays captured graphs
        gpu.LBMpropagate(g, ticks=ticks)
        o.run(ticks)
        compare(g, o)
    moved = max(np.max(np.abs(p.position - x)) for p, x in zip(g.particles, x0))
    assert moved > 1e-3, moved


def test_interaction_terms_are_replaced(gpu, oracle):
    """Re-adding Lennard-Jones replaces the old term list on the device too."""
    kw, parts, _ = CASES["lennard_jones_3d"]

    def add(it, m):
        it.addLennardJones(m, epsilon=0.5, sigma=3.0)
        it.addLennardJones(m, epsilon=0.02, sigma=7.8, cutoff=9.0)
    g, o = build(gpu, oracle, kw, parts, add)
    gpu.LBMpropagate(g, ticks=20)
    o.run(20)
    compare(g, o)


def test_interactions_need_gpu_semantics(gpu):
    from lbmengine_hip import interactions as it
    m = gpu.modelInit(**KW2, semantics="cpu")
    gpu.addBead(m, radius=3.0, position=[-4.0, 0.0])
    gpu.addBead(m, radius=3.0, position=[4.0, 0.0])
    with pytest.raises(gpu.LBMError):
        it.addLinearBond(m, 1, 2)


def test_time_dependent_field_in_chunks_and_refused_outside(gpu, oracle, monkeypatch):
    """Runs longer than one uploaded chunk of the field re-upload it between
    chunks; a bare lbm_run past the uploaded ticks is refused."""
    from lbmengine_hip import model as gm
    kw, parts, add = CASES["dipoles_3d_rotating"]
    g, o = build(gpu, oracle, kw, parts, add)
    monkeypatch.setattr(gm, "SCHEDULE_CHUNK", 16)
    gpu.LBMpropagate(g, ticks=45)
    o.run(45)
    compare(g, o)
    # the last chunk uploaded ticks 33 .. 45: tick 46 is outside
    assert g._fn["lbm_run"](g._h, 1) == -4


def test_time_dependent_field_in_a_batch(gpu, oracle):
    kw, parts, add = CASES["dipoles_2d_pulsed_scalar"]
    g1, o1 = build(gpu, oracle, kw, parts, add)
    g2, o2 = build(gpu, oracle, kw, parts, lambda it, m: it.addDipoles(m, magneticField=lambda t: 0.5 + 0.01 * t,
                                                                      permeability=1.0))
    b = gpu.LBMbatch([g1, g2])
    gpu.LBMpropagate(b, ticks=37)
    o1.run(37)
    o2.run(37)
    compare(g1, o1)
    compare(g2, o2)
