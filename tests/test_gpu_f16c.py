"""The custom 16-bit storage (1 sign, 4 exponent, 11 mantissa bits over
h = (f - w rho0) / w, |h| < 1/8) on the device.

* Round trip through lbm_set_populations / lbm_get_populations against a
  quantizer written from the format's definition with frexp (12 significant
  bits for 2^-18 <= |h| < 2^-3, an absolute step of 2^-29 below, clamped to
  the largest value above), not from the kernel's bit manipulation.
* The headline workload (192^3, squirmer r = 12, beta in {-5, 0, 5}) against
  FP64 after 600 ticks: the velocity field outside the sphere within 1e-3
  relative L2 (the bound the headline is quoted at), FP32 storage within 1e-4.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import lbmengine_hip as lh

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)
HMAX = 2.0 ** -3 - 2.0 ** -15  # largest magnitude: (2 - 2^-11) 2^-4


def ulp_of(h):
    """Spacing of the format around |h| (12 significant bits, 2^-29 below 2^-18)."""
    a = np.abs(h)
    _, e = np.frexp(a)  # a = m 2^e, 0.5 <= m < 1
    return np.where(a >= 2.0 ** -18, np.ldexp(1.0, e - 12), 2.0 ** -29)


def quantize(h):
    a = np.minimum(np.abs(h), HMAX)
    u = ulp_of(a)
    q = np.floor(a / u + 0.5) * u  # nearest, ties away from zero
    return np.sign(h) * np.minimum(q, HMAX)


def test_round_trip_matches_the_format(gpu):
    rng = np.random.default_rng(11)
    m = lh.modelInit(**KW3, xlims=(0, 15), storage="f16c")
    rm = m.resolved
    shape = tuple(rm.n[:3]) + (rm.Q,)
    n = int(np.prod(shape))
    # on-grid values of every binade and the gradual-underflow range, random
    # values over the whole range (log-uniform), and values beyond the clamp
    grid = []
    for e in range(-17, -2):  # binades [2^(e-1), 2^e)
        k = rng.integers(2 ** 11, 2 ** 12, 64)
        grid.append(np.ldexp(k.astype(float), e - 12))
    grid.append(rng.integers(0, 2 ** 11, 64) * 2.0 ** -29)
    grid = np.concatenate(grid) * rng.choice([-1.0, 1.0], sum(len(g) for g in grid))
    rand = np.exp(rng.uniform(np.log(1e-12), np.log(0.12), 4000)) * rng.choice([-1.0, 1.0], 4000)
    big = np.array([0.125, -0.2, 0.5, -1.0, 0.13])
    h = np.zeros(n)
    vals = np.concatenate([grid, rand, big])
    h[: len(vals)] = vals
    h[len(vals):] = rng.uniform(-0.01, 0.01, n - len(vals))
    h = h.reshape(shape, order="F")
    w = rm.ws[None, None, None, :]
    m.distributions = w * (1.0 + h)
    back = (m.distributions / w - 1.0).reshape(-1, order="F")[: len(vals)]
    # (get_populations scales by the FP32 weight and subtracts in FP64: equal
    # to within 2^-22 relative plus a few FP64 ulps of 1)
    tol = 2.0 ** -22 * np.abs(vals) + 4e-16
    ng, nr, nb = len(grid), len(rand), len(big)
    # on-grid values come back unchanged
    assert np.all(np.abs(back[:ng] - vals[:ng]) <= tol[:ng])
    # others: the nearest value of the format (the FP32 conversions on the
    # way in can only tip exact half-way cases)
    r = slice(ng, ng + nr)
    ref = quantize(vals[r])
    assert np.all(np.abs(back[r] - vals[r]) <= 0.5 * ulp_of(vals[r]) * (1 + 2.0 ** -10) + tol[r])
    assert np.mean(np.abs(back[r] - ref) <= tol[r]) > 0.99
    # beyond the range: clamped to the largest magnitude, sign kept
    assert np.allclose(np.abs(back[-nb:]), HMAX, rtol=2.0 ** -22)
    assert np.array_equal(np.sign(back[-nb:]), np.sign(big))


sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


@pytest.mark.parametrize("beta", [-5.0, 0.0, 5.0])
def test_headline_workload_against_fp64(gpu, beta):
    import bench

    class A:
        n, radius, seed = 192, 12.0, 0
    A.beta = beta
    print(f"192^3 squirmer beta={beta}: f64, f32, f16c for 600 ticks", flush=True)
    ref = bench.build_model(lh, "f64", A)
    ref.run(600)
    ref.sync()
    u64 = ref.fluidVelocity
    R = ref.particles[0].position
    c = [np.asarray(x) for x in ref.resolved.coords]
    d2 = (c[0][:, None, None] - R[0]) ** 2 + (c[1][None, :, None] - R[1]) ** 2 + (c[2][None, None, :] - R[2]) ** 2
    ext = d2 >= A.radius ** 2
    v64 = ref.particles[0].velocity
    del ref
    for storage, bound in (("f32", 1e-4), ("f16c", 1e-3)):
        m = bench.build_model(lh, storage, A)
        m.run(600)
        m.sync()
        u = m.fluidVelocity
        e = float(np.linalg.norm((u - u64)[ext]) / np.linalg.norm(u64[ext]))
        es = float(abs(np.linalg.norm(m.particles[0].velocity) - np.linalg.norm(v64)) / np.linalg.norm(v64))
        print(f"  {storage}: exterior u rel L2 {e:.3e}, speed rel {es:.3e}", flush=True)
        assert e < bound and es < bound
        del m


def test_population_guard(gpu):
    """lbm_check_populations and LBMpropagate's guard: a NaN planted from the
    host is reported and raises; an f16c population beyond the storage's
    range is clamped and reported; wall nodes are not counted."""
    import warnings
    for storage in ("f64", "f32", "f16", "f16c"):
        m = lh.modelInit(**KW3, xlims=(0, 7), storage=storage, walledDimensions=[3])
        assert m.check() == (0, 0)
        f = m.distributions
        f[3, 4, 2, 5] = np.nan
        m.distributions = f
        bad, sat = m.check()
        # (f16c stores no NaN: the encoder's clamp turns it into the clamp value)
        assert (bad, sat) == ((0, 1) if storage == "f16c" else (1, 0))
        if storage != "f16c":
            with pytest.raises(FloatingPointError):
                lh.LBMpropagate(m, ticks=2)
    m = lh.modelInit(**KW3, xlims=(0, 7), storage="f16c")
    f = m.distributions
    f[1, 2, 3, 7] = f[1, 2, 3, 7] * 1.5  # h = 0.5, beyond 1/8
    m.distributions = f
    assert m.check() == (0, 1)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        lh.guard(m)
    assert any("clamp" in str(x.message) for x in w)
