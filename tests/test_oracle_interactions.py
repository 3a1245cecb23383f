"""The oracle's particle-particle interactions (applyInteraction!,
gpu/particles.jl:110-197) against the physics each one encodes: every
force is minus the gradient of its potential, taken here by central
differences of the potential at the particles' positions, so a dropped term,
a flipped sign or a swapped particle fails.  The force an interaction adds is
read as the difference of forceInput between two oracle runs of one tick from
the same state, with and without the interaction.

Potentials (r = |R2 - R1|, r^ = (R2 - R1)/r):
  linear      k/2 (r - r0)^2
  bistable    a x^4/4 - b x^2/2,  x = r - (low + high)/2, minima at low, high
  polar       k/2 (theta - theta0)^2, theta the angle at particle 2
  dipole      (C/3) (B^2 - 3 (B.r^)^2) / r^3   (induced dipoles along B)
  LJ          4 eps ((sigma/r)^12 - (sigma/r)^6), zero beyond the cutoff
"""
import math

import numpy as np
import pytest

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16,
           xlims=(-16, 15))
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, xlims=(-24, 23))


def forces(oracle, kw, positions, add):
    """Force each particle receives from the interactions `add` installs."""
    from lbmengine_hip import interactions as it
    runs = []
    for with_it in (True, False):
        m = oracle.oracle_model(**kw)
        for R in positions:
            m.add("bead", radius=3.0, position=R, coupleForces=True, coupleTorques=False)
        if with_it:
            add(it, m)
        m.run(1)
        runs.append([p.load()[0] for p in m.particles])
    return [a - b for a, b in zip(*runs)]


def grad(U, positions, h=1e-5):
    """-dU/dR_i by central differences."""
    P = [np.array(p, dtype=float) for p in positions]
    out = []
    for i in range(len(P)):
        g = np.zeros(len(P[i]))
        for k in range(len(P[i])):
            Pp = [q.copy() for q in P]
            Pm = [q.copy() for q in P]
            Pp[i][k] += h
            Pm[i][k] -= h
            g[k] = -(U(Pp) - U(Pm)) / (2 * h)
        out.append(g)
    return out


def close(got, want, rel=1e-6):
    scale = max(np.max(np.abs(w)) for w in want)
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= rel * scale, (g, w)


def dist(P, i, j):
    return float(np.linalg.norm(P[j] - P[i]))


@pytest.mark.parametrize("kw", [KW3, KW2], ids=["3d", "2d"])
def test_linear_bond(oracle, kw):
    d = kw["dims"]
    pos = [np.array([-4.2, 0.3, 0.1][:d]), np.array([5.1, -1.2, 0.7][:d])]
    k, r0 = 2e-3, 7.0
    got = forces(oracle, kw, pos, lambda it, m: it.addLinearBond(m, 1, 2, equilibriumDisplacement=r0, stiffness=k))
    close(got, grad(lambda P: k / 2 * (dist(P, 0, 1) - r0) ** 2, pos))
    # the default rest length is the current distance: no force
    got = forces(oracle, kw, pos, lambda it, m: it.addLinearBond(m, 1, 2, stiffness=k))
    assert all(np.max(np.abs(g)) < 1e-16 for g in got)


@pytest.mark.parametrize("kw", [KW3, KW2], ids=["3d", "2d"])
def test_bistable_bond(oracle, kw):
    d = kw["dims"]
    pos = [np.array([-4.0, 0.2, -0.3][:d]), np.array([4.5, 0.9, 0.4][:d])]
    low, high, hill = 7.0, 9.5, 3e-3
    got = forces(oracle, kw, pos,
                 lambda it, m: it.addBistableBond(m, 1, 2, lowDisp=low, highDisp=high, hillHeight=hill))
    w, r0 = high - low, (high + low) / 2
    a, b = 64 * hill / w ** 4, 16 * hill / w ** 2
    assert math.isclose(math.sqrt(b / a), w / 2)  # the two wells sit at low and high

    def U(P):
        x = dist(P, 0, 1) - r0
        return a * x ** 4 / 4 - b * x ** 2 / 2
    close(got, grad(U, pos))


@pytest.mark.parametrize("kw", [KW3, KW2], ids=["3d", "2d"])
def test_polar_bond(oracle, kw):
    d = kw["dims"]
    pos = [np.array([-6.0, 3.1, 0.8][:d]), np.array([0.2, -0.4, 0.3][:d]), np.array([5.5, 2.0, -1.1][:d])]
    k, th0 = 5e-3, 1.2

    def angle(P):
        a, c = P[0] - P[1], P[2] - P[1]
        return math.acos(float(np.dot(a, c) / (np.linalg.norm(a) * np.linalg.norm(c))))
    got = forces(oracle, kw, pos, lambda it, m: it.addPolarBond(m, 1, 2, 3, equilibriumAngle=th0, stiffness=k))
    close(got, grad(lambda P: k / 2 * (angle(P) - th0) ** 2, pos))
    assert np.max(np.abs(sum(got))) < 1e-15  # momentum is conserved


def test_dipoles_3d(oracle):
    pos = [np.array([-5.0, 1.0, 0.5]), np.array([4.0, -1.5, 2.0]), np.array([0.5, 6.0, -4.0])]
    B = np.array([0.3, -0.5, 0.8])
    chi, mu = 0.4, 200.0
    got = forces(oracle, KW3, pos,
                 lambda it, m: it.addDipoles(m, magneticField=B, susceptibility=chi, permeability=mu))
    C = 4 * math.pi * 3.0 ** 6 * chi ** 2 / (3 * mu)

    def U(P):
        tot = 0.0
        for i in range(3):
            for j in range(i + 1, 3):
                r = P[j] - P[i]
                n = np.linalg.norm(r)
                tot += C / 3 * (B @ B - 3 * (B @ r / n) ** 2) / n ** 3
        return tot
    close(got, grad(U, pos))


def test_dipoles_2d_field_along_z(oracle):
    pos = [np.array([-5.0, 1.0]), np.array([4.0, -1.5])]
    Bz = 0.7
    got = forces(oracle, KW2, pos, lambda it, m: it.addDipoles(m, 1, 2, magneticField=Bz))
    C = 4 * math.pi * 3.0 ** 6 * 0.5 ** 2 / (3 * 400 * math.pi)
    close(got, grad(lambda P: C / 3 * Bz ** 2 / dist(P, 0, 1) ** 3, pos))


def test_dipoles_zero_field_adds_nothing(oracle):
    pos = [np.array([-5.0, 1.0, 0.5]), np.array([4.0, -1.5, 2.0])]
    got = forces(oracle, KW3, pos, lambda it, m: it.addDipoles(m, magneticField=[0.0, 0.0, 0.0]))
    assert all(np.all(g == 0) for g in got)


@pytest.mark.parametrize("kw", [KW3, KW2], ids=["3d", "2d"])
def test_lennard_jones_and_cutoff(oracle, kw):
    d = kw["dims"]
    pos = [np.array([-4.0, 0.5, 0.0][:d]), np.array([4.2, 0.0, 0.6][:d]), np.array([0.0, 9.0, -0.5][:d])]
    eps, sig = 1e-3, 7.5
    got = forces(oracle, kw, pos, lambda it, m: it.addLennardJones(m, epsilon=eps, sigma=sig))

    def U(P, cut=math.inf):
        tot = 0.0
        for i in range(3):
            for j in range(i + 1, 3):
                r = dist(P, i, j)
                if r < cut:
                    tot += 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
        return tot
    close(got, grad(U, pos))
    # a cutoff between the pair distances keeps only the closer pairs
    r = sorted(dist(pos, i, j) for i in range(3) for j in range(i + 1, 3))
    cut = (r[1] + r[2]) / 2
    got = forces(oracle, kw, pos, lambda it, m: it.addLennardJones(m, epsilon=eps, sigma=sig, cutoff=cut))
    close(got, grad(lambda P: U(P, cut), pos))


def test_wca_cutoff_is_the_minimum(oracle):
    from lbmengine_hip import interactions as it
    m = oracle.oracle_model(**KW3)
    m.add("bead", radius=3.0, position=[-4.0, 0.0, 0.0])
    m.add("bead", radius=3.0, position=[4.0, 0.0, 0.0])
    it.addWCA(m)
    lj = m.particleInteractions[-1]
    assert lj.sigma == 7.0 and math.isclose(lj.epsilon, 1.0, rel_tol=1e-14)  # 2r + dx; mean density 1, dx = dt = 1
    assert math.isclose(lj.cutoff, 2 ** (1 / 6) * 7.0)


def test_interactions_accumulate_in_order(oracle):
    """Several interactions on the same particles add up."""
    pos = [np.array([-4.2, 0.3, 0.1]), np.array([5.1, -1.2, 0.7]), np.array([0.3, 6.5, -2.0])]

    def add(it, m):
        it.addLinearBond(m, 1, 2, equilibriumDisplacement=8.0, stiffness=1e-3)
        it.addBistableBond(m, 2, 3, lowDisp=7.0, highDisp=9.0, hillHeight=1e-3)
        it.addPolarBond(m, 1, 2, 3, equilibriumAngle=1.0, stiffness=2e-3)
    got = forces(oracle, KW3, pos, add)
    parts = [forces(oracle, KW3, pos, lambda it, m: it.addLinearBond(m, 1, 2, equilibriumDisplacement=8.0,
                                                                      stiffness=1e-3)),
             forces(oracle, KW3, pos, lambda it, m: it.addBistableBond(m, 2, 3, lowDisp=7.0, highDisp=9.0,
                                                                        hillHeight=1e-3)),
             forces(oracle, KW3, pos, lambda it, m: it.addPolarBond(m, 1, 2, 3, equilibriumAngle=1.0,
                                                                     stiffness=2e-3))]
    close(got, [sum(p[i] for p in parts) for i in range(3)], rel=1e-12)


def test_dipole_pairs_follow_the_reference_order():
    """addDipoles! pairs: given order for a new interaction, sorted unique ids
    when added to an existing one (gpu/initMethods.jl:316-354)."""
    from lbmengine_hip import interactions as it

    class M:
        dims = 3
        particleInteractions = []
        particles = [type("P", (), {"particleParams": {"radius": 2.0}, "position": np.zeros(3)})()] * 4

        def _set_interactions(self, arr, n):
            self.n = n
    m = M()
    it.addDipoles(m, 3, 1, 2, magneticField=[0, 0, 1])
    assert m.particleInteractions[0].pairs == [(3, 1), (3, 2), (1, 2)]
    it.addDipoles(m, 4, magneticField=[0, 0, 1])
    assert m.particleInteractions[0].pairs == [(1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4)]
    assert m.n == 6


def test_julia_pow_is_within_an_ulp():
    from fractions import Fraction

    from lbmengine_hip.interactions import julia_pow
    rng = np.random.default_rng(1)
    for x in rng.uniform(0.1, 3.0, 200):
        for n in (2, 4, 5, 6, 12):
            exact = float(Fraction(float(x)) ** n)
            assert abs(julia_pow(float(x), n) - exact) <= math.ulp(exact)
    assert julia_pow(1.1, 3) == 1.1 * 1.1 * 1.1
