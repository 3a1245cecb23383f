"""State snapshots (lbm_save_state / lbm_load_state): a model built the same
way and given a snapshot runs on bit for bit as the saved one does, in
memory or through a file, with walls, with more than 8 particles and with
bonds; snapshots of another lattice, storage or particle set are refused;
the per-tick log refuses the ticks before the restored one.  The reference
has no resume, so the uninterrupted run is the check."""
import numpy as np
import pytest

import lbmengine_hip as lh
from lbmengine_hip import _capi

pytestmark = pytest.mark.gpu


def model(storage, n=36, walls=False, particles=2, seed=4):
    rng = np.random.default_rng(seed)
    kw = dict(walledDimensions=[3]) if walls else {}
    m = lh.modelInit(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                     kineticParameter=3 / 16, xlims=(-n // 2, n // 2 - 1), storage=storage,
                     forceDensity=[2e-6, -1e-6, 0.0], **kw)
    if particles >= 1:
        lh.addSquirmer(m, radius=5.0, position=rng.uniform(-3, -1, 3), swimmingSpeed=2e-3 / 3, beta=3.0,
                       swimmingDirection=rng.standard_normal(3), coupleForces=True, coupleTorques=True)
    if particles >= 2:
        lh.addBead(m, radius=3.5, position=[0.25 * n, 0.2 * n, 0.22 * n], coupleForces=True, coupleTorques=True)
    return m


def state(m):
    parts = [np.concatenate([p.position, p.velocity, p.angularVelocity, p.swimmingDirection]) for p in m.particles]
    return m.distributions, (np.stack(parts) if parts else np.zeros(0)), m.tick, m.time


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(state(a), state(b)))


@pytest.mark.parametrize("storage,walls", [("f16c", False), ("f32", False), ("f64", True)])
def test_resume_is_bit_identical(gpu, storage, walls, tmp_path):
    a = model(storage, walls=walls)
    a.run(23)
    snap = a.save_state(tmp_path / "state.bin")
    a.run(31)
    b = model(storage, walls=walls)
    b.load_state(snap)
    b.run(31)
    assert same(a, b)
    c = model(storage, walls=walls)
    c.run(5)  # (its own state is overwritten)
    c.load_state(tmp_path / "state.bin")
    assert c.tick == 23
    c.run(31)
    assert same(a, c)


def test_snapshots_of_other_models_are_refused(gpu):
    snap = model("f32").save_state()
    for other in (model("f32", n=40), model("f16c"), model("f32", particles=1)):
        with pytest.raises(_capi.LBMError):
            other.load_state(snap)
    with pytest.raises(_capi.LBMError):
        model("f32").load_state(snap[:100])
    bad = snap.copy()
    bad[0] ^= 1
    with pytest.raises(_capi.LBMError):
        model("f32").load_state(bad)


def test_log_after_resume(gpu):
    a = model("f32")
    a.run(12)
    snap = a.save_state()
    b = model("f32")
    _capi.check(b._fn["lbm_set_trajectory_log"](b._h, 64))
    b.run(4)
    b.load_state(snap)
    out = np.zeros((3, 9))
    assert b._fn["lbm_get_trajectory"](b._h, 0, 2, 3, _capi.as_dptr(out)) == _capi.LBM_EINVAL
    b.run(5)
    b.sync()
    _capi.check(b._fn["lbm_get_trajectory"](b._h, 0, 15, 3, _capi.as_dptr(out)))
    assert np.array_equal(out[-1, :3], b.particles[0].position)


def test_resume_with_many_particles_and_bonds(gpu):
    """Twelve particles (the move kernel's path) tied by linear bonds: the
    resumed run equals the saved one's continuation bit for bit."""
    from lbmengine_hip import interactions as it

    def build():
        rng = np.random.default_rng(31)
        m = lh.modelInit(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8,
                         kineticParameter=3 / 16, xlims=(-32, 31), ylims=(-24, 23), zlims=(-24, 23),
                         storage="f32", forceDensity=[1e-6, -5e-7, 0.0])
        for i in range(3):
            for j in range(2):
                for k in range(2):
                    R = (np.array([i - 1.0, j - 0.5, k - 0.5]) * 20.0) + rng.uniform(-0.4, 0.4, 3)
                    if (i + j + k) % 3 == 0:
                        lh.addBead(m, radius=3.5, position=R, coupleForces=True, coupleTorques=True)
                    else:
                        lh.addSquirmer(m, radius=3.5, position=R, swimmingSpeed=2e-3 / 3, beta=rng.uniform(-4, 4),
                                       swimmingDirection=rng.standard_normal(3), coupleForces=True,
                                       coupleTorques=True)
        it.addLinearBond(m, 1, 2, equilibriumDisplacement=20.0, stiffness=0.01)
        it.addLinearBond(m, 5, 6, equilibriumDisplacement=20.0, stiffness=0.02)
        return m

    a = build()
    assert len(a.particles) == 12
    a.run(9)
    snap = a.save_state()
    a.run(40)
    b = build()
    b.load_state(snap)
    b.run(40)
    assert same(a, b)
