"""lbm_run_noalloc (the allocation-free run, static device scratch) against
lbm_run through the C ABI: two calls in a row, the second on a smaller
lattice with fewer particles, each bit-identical to lbm_run on a twin handle;
a configuration whose scratch exceeds the fixed array is refused before any
launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)


def twins(gpu, storage, lims, particles, **kw):
    out = []
    for _ in range(2):
        m = gpu.modelInit(**KW3, xlims=lims, storage=storage, **kw)
        for p in particles:
            gpu.addSquirmer(m, **p)
        out.append(m)
    return out


def run_noalloc(m, ticks):
    from lbmengine_hip import _capi
    _capi.check(m._fn["lbm_run_noalloc"](m._h, int(ticks)))


def same_state(a, b):
    assert np.array_equal(a.distributions, b.distributions)
    for pa, pb in zip(a.particles, b.particles):
        for name in ("position", "velocity", "angularVelocity", "swimmingDirection"):
            assert np.array_equal(getattr(pa, name), getattr(pb, name)), name
        fa, ta = pa.load()
        fb, tb = pb.load()
        assert np.array_equal(fa, fb) and np.array_equal(ta, tb)


SQ = dict(swimmingSpeed=2e-3 / 3, coupleForces=True, coupleTorques=True)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_noalloc_run_matches_run_twice(gpu, storage):
    print(f"no-alloc: storage {storage}, 40^3 with two squirmers, then 24^3 with one", flush=True)
    big = [dict(radius=6.0, position=[-6.3, 2.1, 0.4], beta=-5.0, swimmingDirection=[1, 0.3, -0.2], **SQ),
           dict(radius=4.5, position=[8.2, -5.6, 3.3], beta=5.0, swimmingDirection=[-0.2, 1, 0.5], **SQ)]
    a, b = twins(gpu, storage, (-20, 19), big)
    a.run(37)
    run_noalloc(b, 37)
    b.sync()
    same_state(a, b)
    # the second call: a smaller lattice and one particle (the array keeps
    # the first call's contents)
    small = [dict(radius=3.5, position=[0.3, -0.2, 0.4], beta=2.0, swimmingDirection=[0.5, -1, 0.25], **SQ)]
    c, d = twins(gpu, storage, (-12, 11), small)
    c.run(11)
    run_noalloc(d, 11)
    d.sync()
    same_state(c, d)
    # runs continue from either entry point alike
    a.run(5)
    run_noalloc(b, 5)
    b.sync()
    same_state(a, b)


def test_noalloc_refuses_scratch_beyond_its_array(gpu):
    from lbmengine_hip import _capi
    print("no-alloc: oversized particle scratch is refused", flush=True)
    m = gpu.modelInit(**KW3, xlims=(-64, 63), storage="f32")
    gpu.addSquirmer(m, radius=52.0, **SQ)
    f0 = m.distributions
    t0 = m.tick
    rc = m._fn["lbm_run_noalloc"](m._h, 3)
    assert rc == _capi.LBM_EINVAL
    assert b"exceeds" in m._fn["lbm_last_error"]()
    assert m.tick == t0 and np.array_equal(m.distributions, f0)
