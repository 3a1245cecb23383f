"""The oracle's viscousStressTensor (cpu/fluids.jl:507-520) against what the
mathematics fixes: zero on equilibrium populations (so a dropped or wrong
equilibrium term shows), symmetry, and the Chapman-Enskog closed form
sigma = rho nu (grad u + grad u^T) on a decaying shear wave (so the
coefficient, its sign and the pairing of the c_a c_b products show)."""
import numpy as np
import pytest

KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, collisionModel="bgk")
KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)


def textbook_feq(rho, u, cs, ws):
    """w_i rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u.u), lattice units."""
    cu = np.tensordot(u, cs.T.astype(float), axes=([-1], [0]))  # (..., Q)
    uu = np.sum(u * u, axis=-1)[..., None]
    return ws * rho[..., None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * uu)


@pytest.mark.parametrize("kw", [KW2, KW3], ids=["d2q9", "d3q19"])
def test_equilibrium_has_no_stress(oracle, kw):
    n = 12
    m = oracle.oracle_model(**kw, xlims=(0, n - 1))
    rm = m.resolved
    d = rm.dims
    rng = np.random.default_rng(3)
    shp = (n,) * d
    rho = 1 + 0.01 * rng.standard_normal(shp)
    u = 0.02 * rng.standard_normal(shp + (d,))
    m.distributions = np.asfortranarray(textbook_feq(rho, u, rm.cs, rm.ws))
    s = m.viscous_stress()
    assert np.max(np.abs(s)) < 1e-15
    # the same populations off equilibrium carry stress, symmetric in a, b
    f = m.distributions * (1 + 0.01 * rng.standard_normal(shp + (rm.Q,)))
    m.distributions = np.asfortranarray(f)
    s = m.viscous_stress()
    assert np.max(np.abs(s)) > 1e-5
    for a in range(d):
        for b in range(d):
            assert np.array_equal(s[..., a, b], s[..., b, a])


@pytest.mark.parametrize("kw", [KW2, KW3], ids=["d2q9", "d3q19"])
def test_shear_wave_stress_is_rho_nu_du_dy(oracle, kw):
    d = kw["dims"]
    L = 32
    lims = dict(xlims=(0, 7), ylims=(0, L - 1), zlims=(0, 3))
    y = np.arange(L)
    U = 1e-3
    shp = (8, L, 4)[:d]
    u = np.zeros(shp + (d,))
    prof = U * np.sin(2 * np.pi * y / L)
    u[..., 0] = prof.reshape((1, L) + (1,) * (d - 2))
    m = oracle.oracle_model(**kw, **lims, fluidVelocity=u)
    m.run(150)
    nu = (0.8 - 0.5) / 3
    ux = m.fluidVelocity[..., 0]
    rho = m.massDensity
    k = (2 * np.pi * np.fft.fftfreq(L)).reshape((1, L) + (1,) * (d - 2))
    dux = np.real(np.fft.ifft(1j * k * np.fft.fft(ux, axis=1), axis=1))
    s = m.viscous_stress()
    want = rho * nu * dux
    scale = np.max(np.abs(want))
    assert scale > 1e-7
    assert np.max(np.abs(s[..., 0, 1] - want)) <= 5e-3 * scale
    assert np.max(np.abs(s[..., 0, 0])) <= 5e-3 * scale
