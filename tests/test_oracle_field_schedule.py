"""A magnetic field that is a function of time (addDipoles! with a Function,
gpu/initMethods.jl:337-346; evaluated as magneticField(model.time) at
gpu/particles.jl:161, after tick! advanced the time, gpu/fluids.jl:212-218).

The oracle calls the function back at every tick; these runs pin that path
against the oracle's constant-field path, which the potential-gradient
checks of test_oracle_interactions.py pin:
  * a function that is constant in time gives the constant field's run, bit
    for bit;
  * a field switched on between two tick times gives the run without dipoles
    up to the switch and the constant-field run after it (an off-by-one in
    the time the field is read at fails);
  * a Float64 vector is used rounded to Float32 (`|> cu`), a number as it is;
  * the times the function sees are dt added once per tick.
The host's schedule (the fields the device reads, tick by tick) is checked
against the same times."""
import numpy as np
import pytest

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16,
           xlims=(-12, 11))
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=0.5, relaxationTimeRatio=0.8, xlims=(-10, 9.5))
POS3 = [[-4.0, 1.0, 0.5], [4.0, -1.5, 2.0]]
POS2 = [[-4.0, 1.0], [4.0, -1.5]]


def beads(oracle, kw, pos):
    m = oracle.oracle_model(**kw)
    for R in pos:
        m.add("bead", radius=2.5, position=R, coupleForces=True, coupleTorques=True)
    return m


def state(m):
    return [np.concatenate([p.position, p.velocity, np.atleast_1d(p.angularVelocity)]) for p in m.particles]


def same(a, b):
    for x, y in zip(state(a), state(b)):
        assert np.array_equal(x, y), (x, y)


@pytest.mark.parametrize("kw,pos,B", [(KW3, POS3, [0.5, -0.25, 0.75]), (KW2, POS2, [0.5, -0.25])],
                         ids=["3d", "2d"])
def test_constant_function_is_the_constant_field(oracle, kw, pos, B):
    from lbmengine_hip import interactions as it
    a, b = beads(oracle, kw, pos), beads(oracle, kw, pos)
    it.addDipoles(a, magneticField=lambda t: list(B))
    it.addDipoles(b, magneticField=B)
    a.run(4)
    b.run(4)
    same(a, b)
    d = len(pos[0])
    assert np.any(state(a)[0][:d] != np.array(pos[0]))  # the beads moved


def test_switch_on_between_ticks(oracle):
    from lbmengine_hip import interactions as it
    B = [0.0, 0.5, -0.5]
    a = beads(oracle, KW3, POS3)
    it.addDipoles(a, magneticField=lambda t: B if t > 3.5 else [0.0, 0.0, 0.0])
    a.run(6)
    b = beads(oracle, KW3, POS3)
    b.run(3)
    it.addDipoles(b, magneticField=B)
    b.run(3)
    same(a, b)
    c = beads(oracle, KW3, POS3)
    c.run(4)  # one tick late: differs
    it.addDipoles(c, magneticField=B)
    c.run(2)
    assert not np.array_equal(state(a)[0], state(c)[0])


def test_vectors_are_rounded_to_float32_numbers_are_not(oracle):
    from lbmengine_hip import interactions as it
    B = [0.1, 0.2, 0.3]
    a, b, c = (beads(oracle, KW3, POS3) for _ in range(3))
    it.addDipoles(a, magneticField=lambda t: B)
    it.addDipoles(b, magneticField=[float(np.float32(x)) for x in B])
    it.addDipoles(c, magneticField=B)
    for m in (a, b, c):
        m.run(2)
    same(a, b)
    assert not np.array_equal(state(a)[0], state(c)[0])
    # a number (2D, a field along z) stays Float64
    d, e = beads(oracle, KW2, POS2), beads(oracle, KW2, POS2)
    it.addDipoles(d, magneticField=lambda t: 0.1)
    it.addDipoles(e, magneticField=0.1)
    d.run(2)
    e.run(2)
    same(d, e)


def test_times_seen_and_schedule(oracle):
    from lbmengine_hip import interactions as it
    seen = []

    def field(t):
        seen.append(t)
        return [0.0, 0.0] if t < 1.2 else [0.3, 0.1]
    m = beads(oracle, KW2, POS2)
    it.addDipoles(m, magneticField=field)
    seen.clear()
    m.run(5)
    # one call per tick (one dipole pair), at dt = 0.5 added once per tick
    t, want = 0.0, []
    for _ in range(5):
        t = t + 0.5
        want.append(t)
    assert seen == want
    assert m.time == want[-1]
    # the schedule the host uploads holds the same fields for the same ticks
    S = it.field_schedule(field, 0.0, 0.5, 5, 2)
    assert S.shape == (5, 3)
    assert np.array_equal(S[:2], np.zeros((2, 3)))
    assert np.array_equal(S[2:, :2], np.tile([np.float32(0.3), np.float32(0.1)], (3, 1)).astype(np.float64))


def test_scheduled_terms_carry_the_flag():
    from lbmengine_hip import interactions as it
    terms, n = it.expand([it.DipoleDipoleInteraction([(1, 2), (1, 3)], 2.0, lambda t: [1.0, 0.0, 0.0])], 3)
    assert n == 2
    assert all(terms[i].params[2] == 1.0 and terms[i].params[1] == 0.0 for i in range(n))
    terms, n = it.expand([it.DipoleDipoleInteraction([(1, 2)], 2.0, lambda t: 0.5)], 2)
    assert terms[0].params[1] == 1.0 and terms[0].params[2] == 1.0
    terms, n = it.expand([it.DipoleDipoleInteraction([(1, 2)], 2.0, [0.0, 1.0, 0.0])], 3)
    assert terms[0].params[2] == 0.0
