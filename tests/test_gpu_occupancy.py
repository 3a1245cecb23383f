"""The low-occupancy build of the gathering kernels (lbm_set_occupancy: twice
the registers, half the waves per SIMD) runs the same arithmetic as the
default build: populations and particle states agree bit for bit after runs
with particles (box path, moves, graph replays), for one lattice and for a
batch; the autotuner's choice is one of the two."""
import numpy as np
import pytest

import lbmengine_hip as lh

pytestmark = pytest.mark.gpu


def squirmer_model(storage, n=40, r=6.0, seed=3, **kw):
    rng = np.random.default_rng(seed)
    m = lh.modelInit(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                     kineticParameter=3 / 16, xlims=(-n // 2, n // 2 - 1), storage=storage, **kw)
    lh.addSquirmer(m, radius=r, position=rng.uniform(-3, -1, 3), swimmingSpeed=2e-3 / 3, beta=-2.0,
                   swimmingDirection=rng.standard_normal(3), coupleForces=True, coupleTorques=True)
    lh.addBead(m, radius=4.0, position=[0.25 * n + 0.3, 0.25 * n - 0.2, 0.25 * n + 0.1], coupleForces=True,
               coupleTorques=True)
    return m


def state(m):
    parts = [np.concatenate([p.position, p.velocity, p.angularVelocity, p.swimmingDirection]) for p in m.particles]
    return m.distributions, np.stack(parts)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16c"])
def test_low_occupancy_is_bit_identical(gpu, storage):
    a = squirmer_model(storage, forceDensity=[1e-6, 0.0, -5e-7])
    b = squirmer_model(storage, forceDensity=[1e-6, 0.0, -5e-7])
    a.set_block(64, 2, 1)
    b.set_block(64, 2, 1)
    b.set_occupancy(1)
    assert (a.get_occupancy(), b.get_occupancy()) == (0, 1)
    for ticks in (3, 70):  # the second run replays captured graphs
        a.run(ticks)
        b.run(ticks)
        fa, pa = state(a)
        fb, pb = state(b)
        assert np.array_equal(fa, fb)
        assert np.array_equal(pa, pb)
    # set_block keeps the occupancy variant
    b.set_block(128, 1, 1)
    assert b.get_occupancy() == 1


def test_low_occupancy_batch_is_bit_identical(gpu):
    def boxes():
        return [squirmer_model("f32", n=n, r=r, seed=s) for n, r, s in ((32, 5.0, 1), (40, 6.0, 2), (48, 7.0, 3))]
    xa, xb = boxes(), boxes()
    ba, bb = lh.LBMbatch(xa), lh.LBMbatch(xb)
    ba.set_block(64, 1, 1)
    bb.set_block(64, 1, 1)
    bb.set_occupancy(1)
    ba.run(50)
    bb.run(50)
    ba.sync()
    bb.sync()
    for ma, mb in zip(xa, xb):
        fa, pa = state(ma)
        fb, pb = state(mb)
        assert np.array_equal(fa, fb) and np.array_equal(pa, pb)


def test_autotune_picks_a_variant(gpu):
    m = squirmer_model("f32")
    m.autotune(3)
    assert m.get_occupancy() in (0, 1)
    m.run(20)
    m.sync()
    assert m.check() == (0, 0)


# Odd tile counts, tiles that stick out of the lattice and tiles next to walls
# and particles, in both builds.
@pytest.mark.parametrize("storage,n,block,kw", [
    ("f32", 35, (64, 1, 1), {}),                              # odd tile count, x partial
    ("f16c", 33, (32, 1, 2), {}),                              # x and z partial
    ("f16c", 40, (32, 2, 1), {"walledDimensions": [3]}),      # walls
    ("f16", 40, (64, 2, 1), {"walledDimensions": [1, 2]}),
])
def test_low_occupancy_tile_pairs(gpu, storage, n, block, kw):
    a = squirmer_model(storage, n=n, forceDensity=[1e-6, 0.0, -5e-7], **kw)
    b = squirmer_model(storage, n=n, forceDensity=[1e-6, 0.0, -5e-7], **kw)
    a.set_block(*block)
    b.set_block(*block)
    b.set_occupancy(1)
    for ticks in (3, 41):
        a.run(ticks)
        b.run(ticks)
        fa, pa = state(a)
        fb, pb = state(b)
        assert np.array_equal(fa, fb)
        assert np.array_equal(pa, pb)


def test_low_occupancy_adjoint_is_bit_identical(gpu):
    """The adjoint's recomputation runs the keeping kernels of the handle's
    build: gradients agree bit for bit."""
    from lbmengine_hip import adjoint
    n = 24
    rng = np.random.default_rng(12)
    F = rng.normal(0.0, 1e-7, (n, n, n, 3))
    e = rng.standard_normal(3)
    res = []
    for occ in (0, 1):
        m = lh.modelInit(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                         kineticParameter=3 / 16, xlims=(-n // 2, n // 2 - 1), storage="f32",
                         forceDensity=np.asfortranarray(F))
        lh.addSquirmer(m, radius=5.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3 / 3, beta=1.0,
                       swimmingDirection=e / np.linalg.norm(e), coupleForces=True, coupleTorques=True)
        m.set_block(64, 1, 1)
        m.set_occupancy(occ)
        res.append(adjoint.swimming_loss_gradient(m, ticks=40, checkpoint_every=10, window=10, force_penalty=1.0))
    a, b = res
    assert a.loss == b.loss
    assert np.array_equal(a.dB1, b.dB1) and np.array_equal(a.dB2, b.dB2)
    assert np.array_equal(a.dF, b.dF)


def test_autotune_leaves_the_state_as_it_was(gpu):
    """The autotuner's trials step copies of the populations and move the
    particles through both record parities: afterwards a run from the tuned
    handle equals a run of a twin given the same shape, bit for bit, and the
    per-tick log holds only the run's ticks."""
    from lbmengine_hip import _capi
    a = squirmer_model("f32")
    b = squirmer_model("f32")
    for m in (a, b):
        _capi.check(m._fn["lbm_set_trajectory_log"](m._h, 64))
        m.run(7)
    a.autotune(3)
    b.set_block(*a.get_block())
    b.set_tile_order(a.get_tile_order())
    b.set_occupancy(a.get_occupancy())
    a.run(30)
    b.run(30)
    fa, pa = state(a)
    fb, pb = state(b)
    assert np.array_equal(fa, fb) and np.array_equal(pa, pb)
    la, lb = np.zeros((37, 9)), np.zeros((37, 9))
    _capi.check(a._fn["lbm_get_trajectory"](a._h, 0, 1, 37, _capi.as_dptr(la)))
    _capi.check(b._fn["lbm_get_trajectory"](b._h, 0, 1, 37, _capi.as_dptr(lb)))
    assert np.array_equal(la, lb)


def test_batch_autotune_leaves_the_state_as_it_was(gpu):
    def boxes():
        return [squirmer_model("f32", n=n, r=r, seed=s) for n, r, s in ((32, 5.0, 1), (40, 6.0, 2))]
    xa, xb = boxes(), boxes()
    ba, bb = lh.LBMbatch(xa), lh.LBMbatch(xb)
    ba.run(5)
    bb.run(5)
    ba.autotune(2)
    rows = ba.get_row_tiles()
    if rows:
        bb.set_row_tiles(rows)
    else:
        bb.set_block(*ba.get_block())
    bb.set_tile_order(ba.get_tile_order())
    bb.set_occupancy(ba.get_occupancy())
    ba.run(20)
    bb.run(20)
    ba.sync()
    bb.sync()
    for ma, mb in zip(xa, xb):
        fa, pa = state(ma)
        fb, pb = state(mb)
        assert np.array_equal(fa, fb) and np.array_equal(pa, pb)
