"""Closed-form checks that pin the oracle's particle boundary: the Ladd
moving-wall term, the momentum exchange and the squirmer surface slip
(gpu/fluids.jl:177-207, gpu/aux.jl:229-263; CPU backend cpu/fluids.jl:419-442,
cpu/particles.jl:20-43).

Each expected value comes from the physics, not from the oracle's formulas:

* A wall moving with the fluid leaves the equilibrium untouched: with the
  populations at feq(rho, V) and a bead whose surface moves at V, halfway
  bounce-back plus the Ladd term must give back feq(rho(x), V) at every link,
  with rho taken at the link's fluid node.  (A factor-2 error in the Ladd
  coefficient, a wrong sign or rho0 in place of rho(x) breaks this.)
* The first momentum exchange on a held bead in uniform flow u is, for a
  sphere centred on a node (point-symmetric link set),
  F = 6 rho sum_links w_j (c_j . u) c_j, with zero torque for u in a lattice
  plane (mirror symmetry).
* One tick of an uncoupled squirmer in fluid at rest writes
  w_i + 6 w_i u_s . c_i into each link population, where u_s is Blake's
  squirmer slip (B1 + B2 cos th) sin th th_hat evaluated in the spherical
  basis of the swimming axis (scipy rotation, arccos / atan2), not by the
  reference's vector form.
* The CPU backend's variants: its node velocity lives on the solid nodes
  within 2 dx of the surface only (cpu/particles.jl:92-96) and its Ladd term
  adds nodeVelocity(x) + nodeVelocity(x - c) (cpu/fluids.jl:429); at a link
  the fluid end contributes zero, so the same closed forms hold.
* Sanity (not parity): a free squirmer swims at about 2 B1 / 3.
"""
import numpy as np
import pytest

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, collisionModel="bgk")


def feq(rm, rho, u):
    """Textbook second-order equilibrium w_i rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u.u)
    for lattice units (c_s^2 = 1/3), rho (N...) and u (N..., d) or (d,)."""
    d = rm.dims
    shp = tuple(rm.n[:d])
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), shp)
    u = np.broadcast_to(np.asarray(u, dtype=np.float64), shp + (d,))
    uu = (u * u).sum(axis=-1)
    f = np.zeros(shp + (rm.Q,), order="F")
    for q in range(rm.Q):
        cu = (u * rm.cs[q].astype(float)).sum(axis=-1)
        f[..., q] = rm.ws[q] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu)
    return f


def solid_mask(rm, R, r):
    """getSphere's nodes: sum_k (X_k - R_k)^2 < r^2, no periodic image."""
    grids = np.meshgrid(*[rm.coords[k] for k in range(rm.dims)], indexing="ij")
    s = sum((grids[k] - R[k]) ** 2 for k in range(rm.dims))
    return s < r * r


def link_slots(rm, solid):
    """(mask (N..., Q)) of the populations i whose upstream node x - c_i is
    solid while x is fluid: the slots halfway bounce-back rewrites."""
    d = rm.dims
    m = np.zeros(solid.shape + (rm.Q,), dtype=bool)
    for i, c in enumerate(rm.cs):
        up = np.roll(solid, shift=tuple(int(x) for x in c), axis=tuple(range(d)))  # up[x] = solid[x - c]
        m[..., i] = up & ~solid
    return m


def node_coords(rm):
    grids = np.meshgrid(*[rm.coords[k] for k in range(rm.dims)], indexing="ij")
    return np.stack(grids, axis=-1)


def blake_slip(r_vec, e, B1, B2):
    """Squirmer slip u_s = (B1 + B2 cos th) sin th th_hat at points r_vec
    (..., 3) relative to the centre, for swimming axis e: polar angle and
    its unit vector from the spherical basis of a frame whose z axis is e."""
    from scipy.spatial.transform import Rotation
    e = np.asarray(e, dtype=float)
    rot, _ = Rotation.align_vectors([e], [[0.0, 0.0, 1.0]])  # frame z -> e
    local = rot.inv().apply(r_vec.reshape(-1, 3))
    th = np.arccos(np.clip(local[:, 2] / np.linalg.norm(local, axis=1), -1, 1))
    ph = np.arctan2(local[:, 1], local[:, 0])
    th_hat = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], axis=1)
    us = ((B1 + B2 * np.cos(th)) * np.sin(th))[:, None] * th_hat
    return rot.apply(us).reshape(r_vec.shape)


# ------------------------------------------------------------- Ladd term
@pytest.mark.parametrize("semantics", ["gpu", "cpu"])
def test_comoving_bead_keeps_equilibrium(oracle, semantics):
    V = np.array([1.5e-3, -0.8e-3, 0.6e-3])
    m = oracle.oracle_model(**KW3, xlims=(-10, 9), fluidVelocity=V, semantics=semantics)
    m.add("bead", radius=4.3, position=[0.3, -0.4, 0.2], velocity=V)
    m.run(5)
    f = m.distributions
    ref = feq(m.resolved, 1.0, V)
    assert np.max(np.abs(f - ref)) < 1e-14
    # control: a held bead disturbs the same populations
    h = oracle.oracle_model(**KW3, xlims=(-10, 9), fluidVelocity=V, semantics=semantics)
    h.add("bead", radius=4.3, position=[0.3, -0.4, 0.2])
    h.run(5)
    assert np.max(np.abs(h.distributions - ref)) > 1e-5


@pytest.mark.parametrize("semantics", ["gpu", "cpu"])
def test_ladd_term_uses_the_fluid_nodes_density(oracle, semantics):
    # populations at feq(rho(x), V) with a random rho field (rho0 = 1): after
    # one tick each link population is feq_i(rho(x), V) with x its fluid node
    rng = np.random.default_rng(3)
    V = np.array([1.0e-3, 0.5e-3, -1.2e-3])
    R, r = [0.4, 0.1, -0.3], 3.6
    m = oracle.oracle_model(**KW3, xlims=(-8, 7), semantics=semantics)
    rm = m.resolved
    rho = 1.0 + 0.2 * rng.random(tuple(rm.n[:3]))
    m.distributions = feq(rm, rho, V)
    m.add("bead", radius=r, position=R, velocity=V)
    m.run(1)
    slots = link_slots(rm, solid_mask(rm, R, r))
    f, ref = m.distributions, feq(rm, rho, V)
    assert slots.sum() > 300
    assert np.max(np.abs(f[slots] - ref[slots])) < 1e-15
    # rho0 in place of rho(x) would be off by 6 w (rho(x) - 1) V.c at a link
    off = np.abs(feq(rm, 1.0, V) - feq(rm, 1.0, -V))[slots] * np.abs(rho[..., None] - 1).repeat(rm.Q, -1)[slots]
    assert np.max(off) > 1e-5


# ----------------------------------------------------- momentum exchange
@pytest.mark.parametrize("semantics", ["gpu", "cpu"])
@pytest.mark.parametrize("kw,lims,r,u", [
    (KW3, (-10, 9), 4.0, [2.0e-3, -1.3e-3, 0.0]),
    (KW3, (-10, 9), 3.2, [0.0, 1.1e-3, 1.7e-3]),
    (KW2, (-16, 15), 5.5, [1.4e-3, 0.9e-3]),
])
def test_held_bead_first_tick_load(oracle, semantics, kw, lims, r, u):
    u = np.array(u)
    m = oracle.oracle_model(**kw, xlims=lims, fluidVelocity=u, semantics=semantics)
    rm = m.resolved
    m.add("bead", radius=r, massDensity=1e12, coupleForces=True, coupleTorques=True)
    m.run(1)
    F, T = m.particles[0].load()
    slots = link_slots(rm, solid_mask(rm, np.zeros(rm.dims), r))
    # population i at a link bounces c_i off the wall: F = -sum_slots 2 feq_ibar c_i
    # = 6 rho sum_slots w_i (c_i . u) c_i (the even part cancels in +-c pairs)
    expect = np.zeros(rm.dims)
    for i, c in enumerate(rm.cs):
        n = int(slots[..., i].sum())
        expect += 6.0 * rm.ws[i] * n * float(c @ u) * c
    assert np.allclose(F, expect, rtol=1e-12, atol=1e-18)
    assert np.linalg.norm(expect) > 1e-3
    assert np.max(np.abs(T)) < 1e-12 * np.linalg.norm(expect) * r
    # the same bead moving with the fluid feels no force: s = f*_ibar + f_i is
    # even in c_i once the Ladd term is in, and the links pair up
    c = oracle.oracle_model(**kw, xlims=lims, fluidVelocity=u, semantics=semantics)
    c.add("bead", radius=r, velocity=u, massDensity=1e12, coupleForces=True, coupleTorques=True)
    c.run(1)
    Fc, Tc = c.particles[0].load()
    assert np.max(np.abs(Fc)) < 1e-12 * np.linalg.norm(expect)
    assert np.max(np.abs(Tc)) < 1e-12 * np.linalg.norm(expect) * r


# ----------------------------------------------------------- squirmer slip
@pytest.mark.parametrize("semantics", ["gpu", "cpu"])
@pytest.mark.parametrize("beta", [-5.0, 0.0, 5.0])
def test_squirmer_slip_after_one_tick(oracle, beta, semantics):
    R, r, B1 = np.array([0.31, -0.47, 0.18]), 4.2, 1e-3
    e = np.array([0.3, -0.8, 0.5]) if semantics == "gpu" else np.array([-0.6, 0.2, 0.7])
    e = e / np.linalg.norm(e)
    m = oracle.oracle_model(**KW3, xlims=(-10, 9), semantics=semantics)
    rm = m.resolved
    m.add("squirmer", radius=r, position=R, swimmingDirection=e, swimmingSpeed=2 * B1 / 3, beta=beta)
    B1 = abs(3 / 2 * (2 * B1 / 3))  # addSquirmer!'s B1 from swimmingSpeed
    m.run(1)
    f = m.distributions
    slots = link_slots(rm, solid_mask(rm, R, r))
    X = node_coords(rm)
    expect = np.broadcast_to(rm.ws, f.shape).copy(order="F")
    for i, c in enumerate(rm.cs):
        xb = X - c.astype(float)               # the solid end x - c_i
        us = blake_slip(xb - R, e, B1, beta * B1)
        add = 6.0 * rm.ws[i] * (us @ c.astype(float))
        expect[..., i] = np.where(slots[..., i], rm.ws[i] + add, rm.ws[i])
    assert np.max(np.abs(f - expect)) < 1e-15
    assert np.max(np.abs(f - rm.ws)) > 1e-4


# ------------------------------------------------------- physics sanity
def test_free_squirmer_speed_near_two_thirds_B1(oracle):
    # Blake's squirmer swims at U = 2 B1 / 3; in a periodic box of 40 nodes
    # with radius 5 the images and the Ladd boundary's effective radius
    # shift it by a few per cent
    B1 = 1.5e-3
    m = oracle.oracle_model(**KW3, xlims=(-20, 19))
    m.add("squirmer", radius=5.0, swimmingSpeed=2 * B1 / 3, swimmingDirection=[1, 0, 0], coupleForces=True,
          coupleTorques=True)
    m.run(600)
    v = m.particles[0].velocity
    assert abs(v[0] / (2 * B1 / 3) - 1) < 0.1
    assert np.linalg.norm(v[1:]) < 1e-3 * v[0]
